// morton.hip - the Morton order of a set of points: 16-bit keys (8 bits per axis on a 256 x 256
// grid over the points' box), radix-sorted with their indices. Kernels that take targets in
// groups (the proximity scans, panels on points) use it so that a group of 64 targets is
// compact; no result depends on the order.
#include <hipcub/hipcub.hpp>

#include <cfloat>

#include "o2d_runtime.h"

namespace o2d {

namespace {

constexpr int kBlock = 256;
constexpr int kBoxBlocks = 256;

// part[4 b ...] = (x min, x max, y min, y max) of targets b, b + grid, ... (non-finite values
// only spoil the order, never a result)
__global__ __launch_bounds__(kBlock) void target_box_kernel(int64_t n, const double* __restrict__ x,
                                                            const double* __restrict__ y,
                                                            double* __restrict__ part) {
  __shared__ double sm[4][kBlock];
  double a = DBL_MAX, b = -DBL_MAX, c = DBL_MAX, e = -DBL_MAX;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    a = fmin(a, x[i]);
    b = fmax(b, x[i]);
    c = fmin(c, y[i]);
    e = fmax(e, y[i]);
  }
  const int t = threadIdx.x;
  sm[0][t] = a;
  sm[1][t] = b;
  sm[2][t] = c;
  sm[3][t] = e;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      sm[0][t] = fmin(sm[0][t], sm[0][t + w]);
      sm[1][t] = fmax(sm[1][t], sm[1][t + w]);
      sm[2][t] = fmin(sm[2][t], sm[2][t + w]);
      sm[3][t] = fmax(sm[3][t], sm[3][t + w]);
    }
    __syncthreads();
  }
  if (t < 4) part[4 * blockIdx.x + t] = sm[t][0];
}

__device__ __forceinline__ uint32_t spread16(uint32_t v) {  // bits of v to the even positions
  v &= 0xFFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// the box from the partials (once per block), then each target's Morton key on a 256 x 256 grid
// over it (16 bits: groups of 64 targets stay compact and the sort takes two digit passes)
constexpr int kKeyBits = 16;
static_assert(kBoxBlocks <= kBlock, "one partial per thread of a block");
__global__ __launch_bounds__(kBlock) void target_key_kernel(int64_t n, const double* __restrict__ x,
                                                            const double* __restrict__ y,
                                                            const double* __restrict__ part,
                                                            int nb, uint32_t* __restrict__ key,
                                                            uint32_t* __restrict__ val) {
  __shared__ double sm[4][kBlock];
  const int t = threadIdx.x;
  sm[0][t] = t < nb ? part[4 * t] : DBL_MAX;
  sm[1][t] = t < nb ? part[4 * t + 1] : -DBL_MAX;
  sm[2][t] = t < nb ? part[4 * t + 2] : DBL_MAX;
  sm[3][t] = t < nb ? part[4 * t + 3] : -DBL_MAX;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      sm[0][t] = fmin(sm[0][t], sm[0][t + w]);
      sm[1][t] = fmax(sm[1][t], sm[1][t + w]);
      sm[2][t] = fmin(sm[2][t], sm[2][t + w]);
      sm[3][t] = fmax(sm[3][t], sm[3][t + w]);
    }
    __syncthreads();
  }
  const int64_t i = (int64_t)blockIdx.x * kBlock + t;
  if (i >= n) return;
  const double a = sm[0][0], b = sm[1][0], c = sm[2][0], e = sm[3][0];
  double qx = (x[i] - a) / (b - a) * 255.0, qy = (y[i] - c) / (e - c) * 255.0;
  qx = qx > 0.0 ? (qx < 255.0 ? qx : 255.0) : 0.0;  // NaN to 0
  qy = qy > 0.0 ? (qy < 255.0 ? qy : 255.0) : 0.0;
  key[i] = spread16((uint32_t)qx) | (spread16((uint32_t)qy) << 1);
  val[i] = (uint32_t)i;
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

int morton_order16(Device* d, hipStream_t s, int64_t n, const double* x, const double* y,
                   const uint32_t** order) {
  void *kA, *vA, *kB, *vB, *bp;
  O2D_TRY(slot(d, kKeysA, n * sizeof(uint32_t), s, &kA));
  O2D_TRY(slot(d, kOrderA, n * sizeof(uint32_t), s, &vA));
  O2D_TRY(slot(d, kKeysB, n * sizeof(uint32_t), s, &kB));
  O2D_TRY(slot(d, kOrderB, n * sizeof(uint32_t), s, &vB));
  O2D_TRY(slot(d, kBBox, sizeof(double) * 4 * kBoxBlocks, s, &bp));
  target_box_kernel<<<kBoxBlocks, kBlock, 0, s>>>(n, x, y, static_cast<double*>(bp));
  O2D_LAUNCH_CHECK();
  target_key_kernel<<<blocks(n), kBlock, 0, s>>>(n, x, y, static_cast<const double*>(bp),
                                                 kBoxBlocks, static_cast<uint32_t*>(kA),
                                                 static_cast<uint32_t*>(vA));
  O2D_LAUNCH_CHECK();
  size_t tbytes = 0;
  O2D_HIP(hipcub::DeviceRadixSort::SortPairs(
      nullptr, tbytes, static_cast<uint32_t*>(kA), static_cast<uint32_t*>(kB),
      static_cast<uint32_t*>(vA), static_cast<uint32_t*>(vB), (int)n, 0, kKeyBits, s));
  void* tmp;
  O2D_TRY(slot(d, kSortTemp, tbytes, s, &tmp));
  O2D_HIP(hipcub::DeviceRadixSort::SortPairs(
      tmp, tbytes, static_cast<uint32_t*>(kA), static_cast<uint32_t*>(kB),
      static_cast<uint32_t*>(vA), static_cast<uint32_t*>(vB), (int)n, 0, kKeyBits, s));
  *order = static_cast<const uint32_t*>(vB);
  return 0;
}

}  // namespace o2d
