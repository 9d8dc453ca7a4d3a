"""Writes the committed fixtures under tests/golden/ (run in the build container, where the
reference checkout and oracle/_ref exist; the GPU test runs only read the results).

* vortex_hh_clk_up.npz — the reference-held OVF2 file Boris/vortex_hh_clk_up.ovf
  (512 x 128 x 1 cells of 5 x 5 x 40 nm, binary-8 data after the 123456789012345.0 check
  value, x fastest) converted to float32, used as a realistic demag input.
* vortex_hh_clk_up.ovf — the same reference data file copied verbatim (1.5 MB): the OVF2
  reader's input and the byte-level target of the OVF2 writer.
* trilayer_switch_mconv.txt — the reference-held data file "MLConv Paper/3D/trilayer_switch_mconv.txt"
  (a `savedata` table: header and <M> of three meshes vs time) copied verbatim: the target of
  the data-table writer's format, and the trajectory the GPU trilayer test must reproduce.
* tensor_golden.npz — FP64 Newell demag tensors written by oracle/_ref/ref_dump, i.e. by the
  reference's own Boris/DemagTFunc*.cpp compiled here: first-octant values (the rest are
  mirror images) for small meshes, and a seeded sample of entries for meshes large enough to
  use the far-field series (distance >= 40 cells); plus SelfDemag and single-value
  Ldia/Lodia.
* zshift_golden.npz — FP64 2D z-shifted source->destination tensors (multilayered demag)
  written by oracle/_ref/ref_dump from the reference's Boris/DemagTFunc_Shifted.cpp and
  DemagTFunc_Irregular.cpp: regular (equal cells) and irregular (thicknesses differ), with
  near-field and far-field (>= 40 cells) entries; all entries for small meshes, a seeded
  sample for large ones.
* shift_golden.npz — FP64 2D tensors shifted in-plane as well as along z (the general branch
  of the reference's Boris/DemagTFunc_Shifted.cpp and DemagTFunc_Irregular.cpp, used by its
  multilayered demag for meshes whose rectangles are offset in-plane), written by
  oracle/_ref/ref_dump: regular and irregular, fractional offsets, near and far field.
  (`python tests/golden/make_fixtures.py shift` writes this file alone.)
* zshift_pbc_golden.npz — FP64 2D z-shifted tensors with periodic images along x and y
  (multilayered demag with PBC) written by oracle/_ref/ref_dump from the reference's
  Boris/DemagTFunc_Shifted_PBC.cpp and DemagTFunc_Irregular_PBC.cpp: regular and irregular,
  images along x, y or both, near-field only and with far-field images.
  (`python tests/golden/make_fixtures.py zshift_pbc` writes this file alone.)
* pbc_golden.npz — FP64 periodic demag tensors (N = n along periodic axes, images summed)
  written by oracle/_ref/ref_dump from the reference's Boris/DemagTFunc_PBC.cpp: 2D and 3D,
  images along x, y and z, near-field only and with far-field images.
"""
from __future__ import annotations

import shutil
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
REF = Path("/root/reference")
REF_DUMP = ROOT / "oracle" / "_ref" / "ref_dump"


def read_ovf2(path: Path):
    raw = path.read_bytes()
    head_end = raw.index(b"# Begin: data binary 8\n") + len(b"# Begin: data binary 8\n")
    hdr = raw[:head_end].decode()
    kv = {}
    for line in hdr.splitlines():
        if ":" in line:
            k, v = line[1:].split(":", 1)
            kv[k.strip()] = v.strip()
    n = (int(kv["xnodes"]), int(kv["ynodes"]), int(kv["znodes"]))
    h = (float(kv["xstepsize"]), float(kv["ystepsize"]), float(kv["zstepsize"]))
    check = np.frombuffer(raw, dtype="<f8", count=1, offset=head_end)[0]
    assert check == 123456789012345.0, check
    data = np.frombuffer(raw, dtype="<f8", count=3 * n[0] * n[1] * n[2], offset=head_end + 8)
    return n, h, data.reshape(-1, 3)


def dump(*args) -> np.ndarray:
    out = Path("/tmp/_golden_dump.bin")
    subprocess.run([str(REF_DUMP), *[a if isinstance(a, str) else repr(a) if isinstance(a, float) else str(a) for a in args], str(out)], check=True)
    v = np.fromfile(out)
    out.unlink()
    return v


def main() -> None:
    n, h, M = read_ovf2(REF / "Boris" / "vortex_hh_clk_up.ovf")
    np.savez_compressed(HERE / "vortex_hh_clk_up.npz", M=M.astype(np.float32), n=np.array(n), h=np.array(h))
    shutil.copyfile(REF / "Boris" / "vortex_hh_clk_up.ovf", HERE / "vortex_hh_clk_up.ovf")
    shutil.copyfile(REF / "MLConv Paper" / "3D" / "trilayer_switch_mconv.txt", HERE / "trilayer_switch_mconv.txt")

    if not REF_DUMP.exists():
        sys.exit("oracle/_ref/ref_dump is not built: make -f oracle/Makefile.ref")
    rng = np.random.default_rng(20261202)
    out = {}
    cases3 = [("c3a", (6, 5, 3), (12, 10, 8), (4e-9, 4e-9, 4e-9)),
              ("c3b", (7, 4, 5), (14, 8, 10), (3.9e-9, 3.9e-9, 3e-9)),
              ("c3c", (50, 44, 3), (100, 88, 6), (4e-9, 4e-9, 4e-9)),
              ("c3d", (45, 3, 42), (90, 6, 128), (2e-9, 5e-9, 3e-9))]
    for name, nn, NN, hh in cases3:
        v = dump("tensor3d", *nn, *NN, *hh)
        m = NN[0] * NN[1] * NN[2]
        Dd = v[:3 * m].reshape(NN[2], NN[1], NN[0], 3)
        Do = v[3 * m:].reshape(NN[2], NN[1], NN[0], 3)
        if nn[0] * nn[1] * nn[2] <= 200:
            idx = np.array([(k, j, i) for k in range(NN[2]) for j in range(NN[1]) for i in range(NN[0])])
        else:
            idx = np.stack([rng.integers(0, NN[2], 1500), rng.integers(0, NN[1], 1500), rng.integers(0, NN[0], 1500)], axis=1)
        out[name + "_n"], out[name + "_N"], out[name + "_h"] = np.array(nn), np.array(NN), np.array(hh)
        out[name + "_idx"] = idx
        out[name + "_diag"] = Dd[idx[:, 0], idx[:, 1], idx[:, 2]]
        out[name + "_off"] = Do[idx[:, 0], idx[:, 1], idx[:, 2]]
    cases2 = [("c2a", (128, 32), (256, 64), (3.90625e-9, 3.90625e-9, 3e-9)),
              ("c2b", (41, 45), (82, 90), (4e-9, 4e-9, 8e-9))]
    for name, nn, NN, hh in cases2:
        v = dump("tensor2d", *nn, *NN, *hh)
        m = NN[0] * NN[1]
        Dd = v[:3 * m].reshape(NN[1], NN[0], 3)
        Do = v[3 * m:].reshape(NN[1], NN[0])
        idx = np.stack([rng.integers(0, NN[1], 1500), rng.integers(0, NN[0], 1500)], axis=1)
        out[name + "_n"], out[name + "_N"], out[name + "_h"] = np.array(nn), np.array(NN), np.array(hh)
        out[name + "_idx"] = idx
        out[name + "_diag"] = Dd[idx[:, 0], idx[:, 1]]
        out[name + "_off"] = Do[idx[:, 0], idx[:, 1]]
    selfs = [(1.0, 1.0, 1.0), (1.0, 1.0, 0.768), (0.5, 1.0, 0.25), (5e-9, 5e-9, 4e-8)]
    out["self_h"] = np.array(selfs)
    out["self_val"] = np.array([dump("self", *s) for s in selfs])
    singles = [((3.0, 2.0, 1.0), (1.0, 1.0, 1.0)), ((10.5, -4.0, 2.0), (1.0, 0.8, 0.6)), ((0.0, 7.0, -3.0), (2.0, 1.0, 1.0))]
    out["single_in"] = np.array([list(d) + list(hh) for d, hh in singles])
    out["single_val"] = np.array([dump("single", *d, *hh) for d, hh in singles])
    np.savez_compressed(HERE / "tensor_golden.npz", **out)
    zshift_fixtures()
    zshift_pbc_fixtures()
    shift_fixtures()
    pbc_fixtures()
    print("wrote", sorted(p.name for p in HERE.glob("*.npz")))


# (name, N, cell size, images per axis): periodic tensors, N = n along a periodic axis
PBC_CASES = [
    ("p2_x", (8, 12, 1), (4e-9, 4e-9, 4e-9), (3, 0, 0)),
    ("p2_yz", (16, 6, 1), (4e-9, 3e-9, 5e-9), (0, 4, 2)),
    ("p2_xy_far", (64, 48, 1), (4e-9, 4e-9, 8e-9), (10, 10, 0)),
    ("p3_x", (8, 10, 4), (4e-9, 4e-9, 4e-9), (2, 0, 0)),
    ("p3_yz", (12, 6, 8), (3e-9, 4e-9, 2e-9), (0, 3, 4)),
    ("p3_xyz_far", (16, 16, 16), (4e-9, 4e-9, 4e-9), (4, 4, 4)),
]


def pbc_fixtures() -> None:
    """pbc_golden.npz: the reference's periodic tensors (Boris/DemagTFunc_PBC.cpp compiled
    into oracle/_ref/ref_dump), whole for small N and at 1500 seeded points otherwise."""
    rng = np.random.default_rng(20261204)
    out = {}
    for name, NN, hh, img in PBC_CASES:
        m = NN[0] * NN[1] * NN[2]
        if NN[2] == 1:
            v = dump("pbc2d", NN[0], NN[1], *map(float, hh), *img)
            Dd, Do = v[:3 * m].reshape(1, NN[1], NN[0], 3), v[3 * m:].reshape(1, NN[1], NN[0], 1)
        else:
            v = dump("pbc3d", *NN, *map(float, hh), *img)
            Dd, Do = v[:3 * m].reshape(NN[2], NN[1], NN[0], 3), v[3 * m:].reshape(NN[2], NN[1], NN[0], 3)
        if m <= 1000:
            idx = np.array([(k, j, i) for k in range(NN[2]) for j in range(NN[1]) for i in range(NN[0])])
        else:
            idx = np.stack([rng.integers(0, NN[2], 1500), rng.integers(0, NN[1], 1500), rng.integers(0, NN[0], 1500)], axis=1)
        out[name + "_N"], out[name + "_h"], out[name + "_img"] = np.array(NN), np.array(hh), np.array(img)
        out[name + "_idx"] = idx
        out[name + "_diag"] = Dd[idx[:, 0], idx[:, 1], idx[:, 2]]
        out[name + "_off"] = Do[idx[:, 0], idx[:, 1], idx[:, 2]]
    # SelfDemag_PBC: cell size in metres (as the evaluation speedup calls it), mesh cells, images
    selfs = [((4e-9, 4e-9, 4e-9), (32, 24, 1), (3, 0, 0)), ((5e-9, 4e-9, 3e-9), (16, 16, 8), (2, 2, 1)),
             ((4e-9, 4e-9, 8e-9), (64, 48, 1), (10, 10, 0)), ((3e-9, 3e-9, 3e-9), (8, 8, 4), (0, 0, 5))]
    out["selfpbc_in"] = np.array([list(hh) + list(nn) + list(img) for hh, nn, img in selfs], dtype=np.float64)
    out["selfpbc_val"] = np.array([dump("selfpbc", *map(float, hh), *nn, *img) for hh, nn, img in selfs])
    np.savez_compressed(HERE / "pbc_golden.npz", **out)


# (name, n, N, source cell, destination cell, shift), cells and shift normalised
ZSHIFT_CASES = [
    ("z_irr_overlap", (8, 6), (16, 12), (0.5, 0.5, 1.0), (0.5, 0.5, 0.5), 0.75),
    ("z_reg_touch", (8, 6), (16, 12), (0.5, 0.5, 1.0), (0.5, 0.5, 1.0), 1.0),
    ("z_irr_below", (9, 7), (18, 14), (0.5, 0.5, 0.25), (0.5, 0.5, 1.0), -0.625),
    ("z_irr_far", (50, 45), (100, 90), (0.25, 0.25, 1.0), (0.25, 0.25, 0.5), 0.75),
    ("z_reg_far", (50, 45), (100, 90), (0.25, 0.25, 0.5), (0.25, 0.25, 0.5), -0.5),
    ("z_irr_distant", (12, 10), (24, 20), (0.25, 0.25, 0.5), (0.25, 0.25, 1.0), 45.0),
    ("z_cfg5_a", (64, 64), (128, 128), (0.25, 0.25, 0.5), (0.25, 0.25, 0.25), 0.5),
    ("z_cfg5_b", (64, 64), (128, 128), (0.25, 0.25, 1.0), (0.25, 0.25, 0.5), -1.0),
]


def zshift_fixtures() -> None:
    rng = np.random.default_rng(20261203)
    out = {}
    for name, nn, NN, s, d, sh in ZSHIFT_CASES:
        v = dump("zshift2d", *nn, *NN, *map(float, s), *map(float, d), float(sh))
        m = NN[0] * NN[1]
        Dd = v[:3 * m].reshape(NN[1], NN[0], 3)
        Do = v[3 * m:].reshape(NN[1], NN[0], 3)
        if m <= 400:
            idx = np.array([(j, i) for j in range(NN[1]) for i in range(NN[0])])
        else:
            idx = np.stack([rng.integers(0, NN[1], 1500), rng.integers(0, NN[0], 1500)], axis=1)
        out[name + "_n"], out[name + "_N"] = np.array(nn), np.array(NN)
        out[name + "_s"], out[name + "_d"], out[name + "_shift"] = np.array(s), np.array(d), np.array(sh)
        out[name + "_idx"] = idx
        out[name + "_diag"] = Dd[idx[:, 0], idx[:, 1]]
        out[name + "_off"] = Do[idx[:, 0], idx[:, 1]]
    out["names"] = np.array([c[0] for c in ZSHIFT_CASES])
    np.savez_compressed(HERE / "zshift_golden.npz", **out)


# (name, N, source cell, destination cell, shift, x images, y images), normalised
ZSHIFT_PBC_CASES = [
    ("zp_reg_x", (8, 12), (1.0, 1.0, 0.5), (1.0, 1.0, 0.5), 0.5, 2, 0),
    ("zp_irr_y", (16, 8), (1.0, 0.8, 0.5), (1.0, 0.8, 0.25), 0.9, 0, 3),
    ("zp_irr_xy", (10, 10), (0.5, 0.5, 1.0), (0.5, 0.5, 0.5), -1.25, 3, 3),
    ("zp_reg_xy_far", (64, 64), (0.5, 0.5, 1.0), (0.5, 0.5, 1.0), 1.0, 2, 2),
    ("zp_irr_x_far", (96, 12), (1.0, 1.0, 0.25), (1.0, 1.0, 0.5), 2.5, 1, 0),
]


def zshift_pbc_fixtures() -> None:
    rng = np.random.default_rng(20261205)
    out = {}
    for name, NN, s, d, sh, px, py in ZSHIFT_PBC_CASES:
        v = dump("zshift2dpbc", *NN, *map(float, s), *map(float, d), float(sh), px, py)
        m = NN[0] * NN[1]
        Dd = v[:3 * m].reshape(NN[1], NN[0], 3)
        Do = v[3 * m:].reshape(NN[1], NN[0], 3)
        if m <= 400:
            idx = np.array([(j, i) for j in range(NN[1]) for i in range(NN[0])])
        else:
            idx = np.stack([rng.integers(0, NN[1], 1500), rng.integers(0, NN[0], 1500)], axis=1)
        out[name + "_N"], out[name + "_img"] = np.array(NN), np.array([px, py])
        out[name + "_s"], out[name + "_d"], out[name + "_shift"] = np.array(s), np.array(d), np.array(sh)
        out[name + "_idx"] = idx
        out[name + "_diag"] = Dd[idx[:, 0], idx[:, 1]]
        out[name + "_off"] = Do[idx[:, 0], idx[:, 1]]
    out["names"] = np.array([c[0] for c in ZSHIFT_PBC_CASES])
    np.savez_compressed(HERE / "zshift_pbc_golden.npz", **out)


# (name, n, N, source cell, destination cell, shift), normalised
SHIFT_CASES = [
    ("s_reg_x", (8, 6), (16, 12), (1.0, 1.0, 0.5), (1.0, 1.0, 0.5), (0.3, 0.0, 0.75)),
    ("s_irr_xy", (8, 6), (16, 12), (1.0, 0.8, 1.0), (1.0, 0.8, 0.5), (0.4, -0.25, 0.75)),
    ("s_reg_far", (50, 45), (100, 90), (0.25, 0.25, 0.5), (0.25, 0.25, 0.5), (0.1, 0.05, -0.5)),
    ("s_irr_far", (50, 45), (100, 90), (0.25, 0.25, 1.0), (0.25, 0.25, 0.5), (-0.125, 0.2, 0.75)),
    ("s_reg_plane", (12, 10), (24, 20), (0.5, 0.5, 0.5), (0.5, 0.5, 0.5), (0.2, 0.3, 0.0)),
    ("s_irr_big", (20, 16), (40, 32), (0.5, 0.5, 0.25), (0.5, 0.5, 1.0), (3.6, -1.3, -1.5)),
]


def shift_fixtures() -> None:
    rng = np.random.default_rng(20261206)
    out = {}
    for name, nn, NN, s, d, sh in SHIFT_CASES:
        v = dump("shift2d", *nn, *NN, *map(float, s), *map(float, d), *map(float, sh))
        m = NN[0] * NN[1]
        Dd = v[:3 * m].reshape(NN[1], NN[0], 3)
        Do = v[3 * m:].reshape(NN[1], NN[0], 3)
        if m <= 800:
            idx = np.array([(j, i) for j in range(NN[1]) for i in range(NN[0])])
        else:
            idx = np.stack([rng.integers(0, NN[1], 1500), rng.integers(0, NN[0], 1500)], axis=1)
        out[name + "_n"], out[name + "_N"] = np.array(nn), np.array(NN)
        out[name + "_s"], out[name + "_d"], out[name + "_shift"] = np.array(s), np.array(d), np.array(sh)
        out[name + "_idx"] = idx
        out[name + "_diag"] = Dd[idx[:, 0], idx[:, 1]]
        out[name + "_off"] = Do[idx[:, 0], idx[:, 1]]
    out["names"] = np.array([c[0] for c in SHIFT_CASES])
    np.savez_compressed(HERE / "shift_golden.npz", **out)


if __name__ == "__main__":
    if sys.argv[1:] == ["shift"]:
        shift_fixtures()
    elif sys.argv[1:] == ["zshift_pbc"]:
        zshift_pbc_fixtures()
    else:
        main()
