"""The .vtu files of the C++ host mirror (omega2d_amd/host/VtkXmlWriter.h, Points/Surfaces
write_vtk) are byte-identical to the files the reference's own writers produced for the same
data (tests/golden/vtk/, src/Points.h:556-740, src/Surfaces.h:1238-1379). No GPU work: the host
driver's vtk mode only writes files."""
import os
import subprocess

import numpy as np
import pytest

from helpers import circle_panels
from omega2d_amd import build as B

HERE = os.path.dirname(__file__)
GOLD = np.load(os.path.join(HERE, "golden", "reference_vectors.npz"))
VDIR = os.path.join(HERE, "golden", "vtk")


@pytest.mark.skipif(not os.path.exists(B.HOST_TEST_BIN), reason="host driver not built")
def test_vtu_files_match_reference_bytes(tmp_path):
    g = GOLD
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)  # noqa: E731  float storage
    n = len(g["vk_x"])
    nx, ny, idx = circle_panels(12)
    header = np.array([n, len(nx), len(idx) // 2, 0, 12, 0, 3], np.int64)
    payload = [np.array([0.125, 2.5])] + [f32(g[k]) for k in ("vk_x", "vk_y", "vk_s", "vk_r",
                                                              "vk_u", "vk_v", "vk_w")]
    payload += [nx, ny, idx, f32(g["vk_pvs"]), f32(g["vk_pss"])]
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(header.tobytes())
        for a in payload:
            f.write(np.ascontiguousarray(a).tobytes())
    out = tmp_path / "vtu"
    out.mkdir()
    r = subprocess.run([B.HOST_TEST_BIN, "vtk", str(inp), str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    names = sorted(os.listdir(VDIR))
    assert names == ["fldpt_01_00012.vtu", "panel_00_00003.vtu", "part_00_00012.vtu"]
    assert sorted(os.listdir(out)) == names
    for nm in names:
        with open(os.path.join(VDIR, nm), "rb") as a, open(out / nm, "rb") as b:
            assert a.read() == b.read(), nm


def test_golden_vtu_decodes_to_its_inputs():
    """The golden particle file holds the inputs it was written from (base64 decode)."""
    import base64
    import re
    text = open(os.path.join(VDIR, "part_00_00012.vtu")).read()
    arrays = dict(re.findall(r'Name="([^"]+)"[^>]*format="binary"> (\S+) </DataArray>', text))
    hdr_len = 8  # base64 of the 4-byte header
    circ = np.frombuffer(base64.b64decode(arrays["circulation"][hdr_len:]), np.float32)
    assert np.array_equal(circ, GOLD["vk_s"].astype(np.float32))
    vel = np.frombuffer(base64.b64decode(arrays["velocity"][hdr_len:]), np.float32).reshape(-1, 3)
    assert np.array_equal(vel[:, 0], GOLD["vk_u"].astype(np.float32))
    assert np.all(vel[:, 2] == 0.0)
    t = np.frombuffer(base64.b64decode(arrays["TimeValue"][hdr_len:]), np.float64)
    assert t[0] == 0.125
