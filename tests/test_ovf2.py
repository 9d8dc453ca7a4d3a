"""OVF 2.0 file I/O through the C-ABI (host code; no device needed).

Pinned by the reference's own data file Boris/vortex_hh_clk_up.ovf (copied verbatim into
tests/golden/): the reader must return its header and values, and the writer must reproduce
the file byte for byte from them (the reference's OVF2::Write_OVF2_VEC layout,
Boris/OVF2_Handlers.cpp:701-797).  Round trips cover binary 4, binary 8 and text, vector and
scalar data; hand-written files cover meshunit nm, CRLF line ends and the error paths of
OVF2::Read_OVF2_VEC (Boris/OVF2_Handlers.cpp:364-690)."""
from pathlib import Path

import numpy as np
import pytest

from boris2_amd import _lib, api

GOLD = Path(__file__).resolve().parent / "golden"
VORTEX = GOLD / "vortex_hh_clk_up.ovf"


def test_reads_reference_vortex_file():
    d = api.read_ovf2(VORTEX)
    assert d.n == (512, 128, 1)
    np.testing.assert_allclose(d.h, (5e-9, 5e-9, 4e-8), rtol=1e-15)
    np.testing.assert_allclose(d.rect, (0, 0, 0, 2.56e-6, 6.4e-7, 4e-8), rtol=1e-15)
    assert d.data_bytes == 8 and d.values.shape == (512 * 128, 3)
    ref = np.load(GOLD / "vortex_hh_clk_up.npz")["M"]
    assert np.array_equal(d.values.astype(np.float32), ref)
    # a vortex state of a permalloy-like film: |M| uniform
    mag = np.linalg.norm(d.values, axis=1)
    assert np.ptp(mag) <= 1e-6 * mag.max()


def test_writer_reproduces_reference_file_bytes(tmp_path):
    d = api.read_ovf2(VORTEX)
    out = tmp_path / VORTEX.name          # the title line is the file name
    api.write_ovf2(out, d.values, d.n, d.h, d.rect, "bin8")
    assert out.read_bytes() == VORTEX.read_bytes()


@pytest.mark.parametrize("dtype", ["bin4", "bin8", "text"])
@pytest.mark.parametrize("vd", [1, 3])
def test_round_trip(tmp_path, dtype, vd):
    rng = np.random.default_rng(vd)
    n, h = (7, 5, 3), (2e-9, 3e-9, 4.5e-9)
    v = rng.normal(size=(7 * 5 * 3, vd)) * 8e5
    f = tmp_path / f"rt_{dtype}_{vd}.ovf"
    api.write_ovf2(f, v if vd == 3 else v[:, 0], n, h, data_type=dtype, norm=2.0)
    d = api.read_ovf2(f)
    assert d.n == n and d.values.shape == (105, vd)
    np.testing.assert_allclose(d.h, h, rtol=1e-6)
    want = v / 2.0
    if dtype == "bin8":
        assert np.array_equal(d.values, want)
    elif dtype == "bin4":
        assert np.array_equal(d.values, want.astype(np.float32).astype(np.float64))
    else:
        np.testing.assert_allclose(d.values, want, rtol=5e-6)      # 6 significant digits


def test_hand_written_text_file_nm_unit_crlf(tmp_path):
    lines = ["# OOMMF OVF 2.0", "# Segment count: 1", "# Begin: Segment", "# Begin: Header", "# Title: hand",
             "# meshunit: nm", "# meshtype: rectangular", "# xmin: 0", "# ymin: 0", "# zmin: 0",
             "# xmax: 8", "# ymax: 4", "# zmax: 3", "# xnodes: 2", "# ynodes: 1", "# znodes: 1",
             "# xstepsize: 4", "# ystepsize: 4", "# zstepsize: 3", "# valuedim: 3", "# End: Header",
             "# Begin: Data Text", "1 2 3", "\t4\t5  6", "# End: Data Text", "# End: Segment"]
    f = tmp_path / "hand.ovf"
    f.write_bytes(("\r\n".join(lines) + "\r\n").encode())
    d = api.read_ovf2(f)
    assert d.n == (2, 1, 1) and d.data_bytes == 1
    np.testing.assert_allclose(d.h, (4e-9, 4e-9, 3e-9), rtol=1e-15)
    np.testing.assert_allclose(d.rect[3:], (8e-9, 4e-9, 3e-9), rtol=1e-15)
    assert np.array_equal(d.values, [[1, 2, 3], [4, 5, 6]])


@pytest.mark.parametrize("edit", ["version", "valuedim", "meshtype", "unit", "nodes", "check"])
def test_rejects_malformed_files(tmp_path, edit):
    good = tmp_path / "good.ovf"
    api.write_ovf2(good, np.ones((4, 3)), (2, 2, 1), (1e-9, 1e-9, 1e-9), data_type="bin4")
    raw = good.read_bytes()
    sub = {"version": (b"# OOMMF OVF 2.0", b"# OOMMF OVF 1.0"), "valuedim": (b"# valuedim: 3", b"# valuedim: 2"),
           "meshtype": (b"rectangular", b"irregular"), "unit": (b"# meshunit: m", b"# meshunit: mm"),
           "nodes": (b"# xnodes: 2", b"# xnodes: 3"), "check": (np.float32(1234567.0).tobytes(), np.float32(1.0).tobytes())}[edit]
    bad = tmp_path / "bad.ovf"
    bad.write_bytes(raw.replace(*sub))
    with pytest.raises(RuntimeError, match="OVF2"):
        api.read_ovf2(bad)


def test_read_rejects_wrong_count():
    L = _lib.load()
    buf = np.zeros(10)
    assert L.bm_ovf2_read(str(VORTEX).encode(), buf.ctypes.data, 10) == 1
