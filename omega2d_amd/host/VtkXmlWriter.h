// VtkXmlWriter.h - the on-disk format of the reference's particle and panel output: XML VTK
// unstructured-grid files (.vtu) with inline base64 data arrays (src/VtkXmlWriter.h,
// src/Points.h:556-740, src/Surfaces.h:1238-1379). Files written here are byte-identical to
// the reference's, so ParaView comparisons against reference runs see the same bytes:
//  - layout: an XML declaration, then one element per line, indented 4 spaces per level; an
//    element with no content closes as "/>"; a data array's text stays on its element's line;
//  - attributes of an element are written in ascending key order;
//  - every DataArray has format="binary" and text " " + base64(uint32 header) + base64(bytes)
//    + " ", where the header holds the length of the encoded byte string, as the reference's
//    files do (not the raw byte count).
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace omega2d {

// RFC 4648 base64 with '=' padding
inline std::string base64_encode(const void* data, size_t n) {
  static const char tbl[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  const unsigned char* p = static_cast<const unsigned char*>(data);
  std::string out;
  out.reserve(4 * ((n + 2) / 3));
  size_t i = 0;
  for (; i + 3 <= n; i += 3) {
    const uint32_t v = (uint32_t)p[i] << 16 | (uint32_t)p[i + 1] << 8 | p[i + 2];
    out += tbl[v >> 18];
    out += tbl[(v >> 12) & 63];
    out += tbl[(v >> 6) & 63];
    out += tbl[v & 63];
  }
  if (n - i == 1) {
    const uint32_t v = (uint32_t)p[i] << 16;
    out += tbl[v >> 18];
    out += tbl[(v >> 12) & 63];
    out += "==";
  } else if (n - i == 2) {
    const uint32_t v = (uint32_t)p[i] << 16 | (uint32_t)p[i + 1] << 8;
    out += tbl[v >> 18];
    out += tbl[(v >> 12) & 63];
    out += tbl[(v >> 6) & 63];
    out += '=';
  }
  return out;
}

using VtkAttrs = std::vector<std::pair<std::string, std::string>>;

// A .vtu file under construction: nested elements and inline base64 data arrays.
class VtuFile {
 public:
  explicit VtuFile(const std::string& path) : fp_(std::fopen(path.c_str(), "wb")) {
    if (!fp_) throw std::runtime_error("cannot open " + path + " for writing");
    put("<?xml version=\"1.0\"?>");
    open("VTKFile", {{"type", "UnstructuredGrid"}, {"version", "0.1"},
                     {"byte_order", "LittleEndian"}, {"header_type", "UInt32"}},
         false);
    open("UnstructuredGrid");
  }
  VtuFile(const VtuFile&) = delete;
  VtuFile& operator=(const VtuFile&) = delete;
  ~VtuFile() {
    if (fp_) std::fclose(fp_);
  }

  // start an element; attributes go out in ascending key order unless `sorted` is false
  void open(const std::string& name, VtkAttrs attrs = {}, bool sorted = true) {
    seal();
    if (sorted) std::sort(attrs.begin(), attrs.end());
    if (text_level_ < 0) put("\n");
    indent(level_);
    put("<" + name);
    for (const auto& a : attrs) put(" " + a.first + "=\"" + escape(a.second, true) + "\"");
    stack_.push_back(name);
    pending_ = true;
    ++level_;
  }

  // one DataArray element holding `bytes` raw bytes of `data`
  void data_array(VtkAttrs attrs, const void* data, size_t bytes) {
    std::sort(attrs.begin(), attrs.end());
    attrs.emplace_back("format", "binary");
    open("DataArray", std::move(attrs), false);
    const std::string enc = base64_encode(data, bytes);
    const uint32_t len = (uint32_t)enc.size();
    text(" " + base64_encode(&len, sizeof(len)) + enc + " ");
    close();
  }

  template <class T>
  void data_array(VtkAttrs attrs, const std::vector<T>& v) {
    data_array(std::move(attrs), v.data(), sizeof(T) * v.size());
  }

  void close() {
    --level_;
    const std::string name = stack_.back();
    stack_.pop_back();
    if (pending_) {
      put("/>");
    } else {
      if (text_level_ < 0) {
        put("\n");
        indent(level_);
      }
      put("</" + name + ">");
    }
    if (text_level_ == level_) text_level_ = -1;
    if (level_ == 0) put("\n");
    pending_ = false;
  }

  // close every open element and the file
  void finish() {
    while (!stack_.empty()) close();
    if (std::fclose(fp_) != 0) {
      fp_ = nullptr;
      throw std::runtime_error("write failed");
    }
    fp_ = nullptr;
  }

 private:
  void put(const std::string& s) { std::fwrite(s.data(), 1, s.size(), fp_); }
  void indent(int level) {
    for (int i = 0; i < level; ++i) put("    ");
  }
  void seal() {
    if (pending_) put(">");
    pending_ = false;
  }
  void text(const std::string& t) {
    text_level_ = level_ - 1;
    seal();
    put(escape(t, false));
  }
  static std::string escape(const std::string& s, bool attribute) {
    std::string o;
    for (char c : s) {
      if (c == '&') o += "&amp;";
      else if (c == '<') o += "&lt;";
      else if (c == '>') o += "&gt;";
      else if (attribute && c == '"') o += "&quot;";
      else if (attribute && c == '\'') o += "&apos;";
      else o += c;
    }
    return o;
  }

  std::FILE* fp_;
  std::vector<std::string> stack_;
  int level_ = 0;
  int text_level_ = -1;
  bool pending_ = false;  // an element's start tag is still open (no '>' yet)
};

// file name of collection `index` at output frame `frameno`: <prefix>NN_FFFFF.vtu
inline std::string vtu_name(const char* prefix, size_t index, size_t frameno) {
  char buf[96];
  std::snprintf(buf, sizeof(buf), "%s%02zu_%05zu.vtu", prefix, index, frameno);
  return buf;
}

// x, y (and z = 0) interleaved as Float32 triples
template <class S>
std::vector<float> xyz_float(const std::vector<S>& x, const std::vector<S>& y) {
  std::vector<float> v(3 * x.size());
  for (size_t i = 0; i < x.size(); ++i) {
    v[3 * i] = float(x[i]);
    v[3 * i + 1] = float(y[i]);
    v[3 * i + 2] = 0.0f;
  }
  return v;
}

template <class S>
std::vector<float> as_float(const std::vector<S>& v) {
  return std::vector<float>(v.begin(), v.end());
}

// the leading FieldData block with the simulation time
inline void vtu_time(VtuFile& f, double time) {
  f.open("FieldData");
  f.data_array({{"type", "Float64"}, {"Name", "TimeValue"}, {"NumberOfTuples", "1"}}, &time,
               sizeof(time));
  f.close();
}

}  // namespace omega2d
