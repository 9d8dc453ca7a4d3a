"""GPU tests of the radix schedules of the batched FFT passes: every alternative plan of every
y, z and 2D y pass length the batch below holds gives the same trajectories as the default
plans to FP32 rounding (a wrong stage order, twiddle or LDS index would differ at order one),
the radix-tuning autotuner's choices are reported and can be re-applied to a new batch, which
then reproduces the tuned batch bit for bit, and bad schedules are refused."""
import numpy as np
import pytest
import torch

from boris2_amd import api

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def separate_yzy_passes(monkeypatch):
    """These tests exercise the alternative radix schedules of the separate y and z passes:
    the batches' demag handles are created with the fused thin-mesh pass off."""
    monkeypatch.setenv("BM_YZY", "0")

H = (4e-9, 4e-9, 4e-9)
# y lengths 2 ny and z lengths 2 nz covering every length with an alternative plan:
# 3D y 16..2048 and 384 / 96 / 768, z 4..128; 2D (y2d) 16 ... 2048
SIZES = [(16, 8, 2), (32, 16, 2), (32, 32, 4), (16, 48, 2), (64, 64, 8), (16, 96, 2), (8, 128, 16),
         (8, 192, 2), (8, 256, 32), (4, 384, 2), (4, 512, 2), (8, 8, 64), (4, 1024, 2),
         (16, 8, 1), (32, 16, 1), (32, 32, 1), (16, 48, 1), (64, 64, 1), (16, 96, 1), (8, 128, 1), (8, 192, 1),
         (8, 256, 1), (4, 384, 1), (4, 512, 1), (4, 1024, 1)]


def meshes_and_M0():
    rng = np.random.default_rng(21)
    meshes = []
    for n in SIZES:
        p = api.FMParams(Ms=float(rng.uniform(5e5, 1.1e6)), A=1.3e-11, alpha=float(rng.uniform(0.02, 0.3)),
                         modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE)
        meshes.append((n, H, p))
    M0 = [api.random_magnetisation(n[0] * n[1] * n[2], p.Ms, 40 + b).astype(np.float32) for b, (n, _, p) in enumerate(meshes)]
    return meshes, M0


def groups():
    """(pass, length) of every FFT pass-size group of the batch."""
    g = set()
    for n in SIZES:
        if n[2] == 1:
            g.add(("y2d", 2 * n[1]))
        else:
            g.add(("y_fwd", 2 * n[1])); g.add(("y_inv", 2 * n[1])); g.add(("z", 2 * n[2]))
    return sorted(g)


def variants(n):
    return {4: 2, 8: 3, 16: 3, 32: 3, 64: 3, 96: 2, 128: 3, 192: 2, 256: 3, 384: 2, 512: 3, 768: 2, 1024: 3, 2048: 2}.get(n, 1)


def run(bs, steps=5):
    bs.set_dt(40e-15)
    bs.run(steps)
    torch.cuda.synchronize()
    return bs.M.cpu().numpy().astype(np.float64)


def per_mesh_err(bs, a, b):
    e = []
    for k in range(len(SIZES)):
        s = slice(int(bs.offsets[k]), int(bs.offsets[k + 1]))
        e.append(float(np.linalg.norm(a[s] - b[s]) / np.linalg.norm(b[s])))
    return e


@pytest.mark.parametrize("v", [1, 2])
def test_alternative_radix_schedules_match_default(gpu, v):
    meshes, M0 = meshes_and_M0()
    ref_bs = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0)
    ref = run(ref_bs)
    bs = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0)
    pinned = []
    for pass_, n in groups():
        if variants(n) > v:
            bs.set_schedule(pass_, n, radix=v)
            pinned.append((pass_, n))
    assert len(pinned) >= (20 if v == 1 else 12), pinned
    rep = bs.autotune_report()
    assert sum(f" radix={v} " in l for l in rep) == len(pinned), rep
    out = run(bs)
    e = per_mesh_err(bs, out, ref)
    print(f"radix {v}: {len(pinned)} groups, worst mesh {SIZES[int(np.argmax(e))]} rel L2 {max(e):.2e}")
    assert max(e) <= 1e-6, list(zip(SIZES, e))
    # and it differs from the default plans (a schedule that silently fell back would not)
    assert max(e) > 0
    ref_bs.close(); bs.close()


def test_radix_autotune_report_reapplies(gpu):
    meshes, M0 = meshes_and_M0()
    ref_bs = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0)
    ref = run(ref_bs)
    a = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0, autotune="radix")
    out = run(a)
    rep = a.autotune_report()
    kinds = {l.split()[0] for l in rep}
    assert {"x_fwd", "x_inv", "y_fwd", "y_inv", "z", "y2d"} <= kinds, rep
    for l in rep:
        w = l.split()
        assert w[4].startswith("radix=") and float(w[-2]) > 0, l
        if w[0][0] in "yz":
            r = [int(x) for x in w[5].split("x")]
            assert int(np.prod(r)) == int(w[1]), l     # the schedule's radices multiply to the length
    print("\n".join(rep))
    assert max(per_mesh_err(a, out, ref)) <= 1e-6
    b = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0)
    b.apply_report(rep)
    assert [l.rsplit(" ", 2)[0] for l in b.autotune_report()] == [l.rsplit(" ", 2)[0] for l in rep]
    out_b = run(b)
    assert np.array_equal(out_b, out)
    ref_bs.close(); a.close(); b.close()


def test_set_schedule_refuses_bad_requests(gpu):
    meshes, M0 = meshes_and_M0()
    bs = api.BatchSolver(api.BM_EVAL_RK4, meshes[:3], M0[:3])
    with pytest.raises(RuntimeError, match="error 1"):
        bs.set_schedule("y_fwd", 2048, radix=1)        # no such group in this batch
    with pytest.raises(RuntimeError, match="error 1"):
        bs.set_schedule("z", 4, radix=3)               # length 4 has two schedules
    with pytest.raises(RuntimeError, match="error 1"):
        bs.set_schedule("x_fwd", 16, f=3)
    with pytest.raises(RuntimeError, match="error 1"):
        bs.set_schedule("w", 16)
    bs.set_schedule("z", 4, per_cu=2, radix=1)
    bs.close()


# row lengths with a half-height x tile (nx 128..768), ragged ny against the 8-row spectra
# tiles and the 4-row half tiles of nx 512 and 768
HALF = [(128, 24, 2), (192, 40, 3), (256, 17, 2), (384, 12, 1), (512, 9, 2), (768, 8, 1)]


def test_half_height_x_tiles_match_default_bit_for_bit(gpu):
    rng = np.random.default_rng(5)
    meshes = [(n, H, api.FMParams(Ms=float(rng.uniform(5e5, 1.1e6)), A=1.3e-11, alpha=0.1,
                                  modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE)) for n in HALF]
    M0 = [api.random_magnetisation(n[0] * n[1] * n[2], p.Ms, 70 + b).astype(np.float32) for b, (n, _, p) in enumerate(meshes)]
    ref_bs = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0)
    for n in HALF:
        for pass_ in ("x_fwd", "x_inv"):
            ref_bs.set_schedule(pass_, n[0], f=1, per_cu=2)
    ref = run(ref_bs, 3)
    bs = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0)
    for n in HALF:
        for pass_ in ("x_fwd", "x_inv"):
            bs.set_schedule(pass_, n[0], f=-2, per_cu=4)
    rep = bs.autotune_report()
    assert sum(" f=-2 " in l for l in rep) == 2 * len(HALF), rep
    out = run(bs, 3)
    # the tile height only groups rows differently: the same arithmetic on every row
    assert np.array_equal(out, ref)
    assert np.abs(out - M0_flat(M0)).max() > 0
    with pytest.raises(RuntimeError, match="error 1"):
        bs.set_schedule("x_fwd", 128, f=-3)
    ref_bs.close(); bs.close()
    small = api.BatchSolver(api.BM_EVAL_RK4, meshes_and_M0()[0][:3], meshes_and_M0()[1][:3])
    with pytest.raises(RuntimeError, match="error 1"):
        small.set_schedule("x_fwd", 16, f=-2)          # nx 16: the tile is one row group
    small.close()


def M0_flat(M0):
    return np.concatenate([m.reshape(-1, 3) for m in M0]).astype(np.float64)
