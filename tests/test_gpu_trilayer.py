"""The reference's own multilayered-convolution example, against its own output file
(MLConv Paper/3D: trilayer_switch.bsm run by Boris, <M> per mesh every 10 ps in
trilayer_switch_mconv.txt, copied verbatim to tests/golden/): three permalloy meshes of
128 x 64 x {4, 2, 4} cells of 5 nm, 1 nm gaps, Ms 8e5 A/m, A 13 pJ/m, alpha 0.02; outer meshes
start along -x in 25.04 mT at 5 degrees, the middle one along +x in 6.26 mT; RK4 at 0.5 ps
(the file's evalMethod 2, dT 5e-13) for 3 ns.  All three meshes switch; every saved <M>
component must agree with the file to 1e-4 of Ms (measured: 1.3e-5).  The run writes its own
`savedata` table (boris2_amd.datatable) whose header equals the reference file's line for line
and whose rows carry the same times."""
from pathlib import Path

import numpy as np
import pytest

from boris2_amd import api

pytestmark = pytest.mark.gpu

DATA = Path(__file__).parent / "golden" / "trilayer_switch_mconv.txt"
# applied fields of the mumax3 twin of the problem (MLConv Paper/3D/mumax3/trilayer_switch.txt), in T
B_OUT = np.array([0.025037111148343075, 0.0021904566281595616, 0.0])
B_MID = np.array([0.006259271503900462, 0.0005476160419954826, 0.0])


def reference_rows():
    rows = []
    for line in DATA.read_text().splitlines():
        parts = line.strip().split("\t")
        if len(parts) == 10:
            try:
                rows.append([float(v) for v in parts])
            except ValueError:
                pass
    return np.array(rows)


def test_trilayer_switching_matches_the_reference_output(gpu, tmp_path):
    from boris2_amd import datatable as DT
    from test_datatable import meshes
    ref = reference_rows()
    ref_lines = DATA.read_text().split("\n")
    ms, sm = meshes()
    by = {m.name: m for m in ms}
    cols = [DT.time_column()] + [DT.avm_column(by[n]) for n in ("permalloy0", "permalloy", "permalloy2")]
    table = DT.DataTable(tmp_path / "trilayer.txt", ms, sm, cols, stamp=ref_lines[0])
    assert ref.shape == (300, 10)
    nx, ny = 128, 64
    hz, z0, grp = api.stack_of_meshes([(4, 5e-9, -21e-9), (2, 5e-9, 0.0), (4, 5e-9, 11e-9)])
    Ha = np.array([(B_OUT if g != 1 else B_MID) / api.MU0 for g in grp])
    p = api.FMParams(Ms=8e5, A=13e-12, alpha=0.02, modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN)
    nl = nx * ny
    M0 = np.zeros((len(grp) * nl, 3), np.float32)
    M0[:, 0] = np.repeat(np.where(grp == 1, 8e5, -8e5), nl)
    s = api.LayerStackSolver(api.BM_EVAL_RK4, nx, ny, 5e-9, 5e-9, hz, z0, p, M0=M0)
    s.set_layer_meshes(grp, Ha)
    s.set_dt(5e-13)
    worst, step = 0.0, 0
    for r in ref:
        target = int(round(r[0] / 5e-13))
        if target > step:
            s.run_to_iteration(target)
            step = target
        assert abs(s.state()["time"] - r[0]) <= 1e-6 * max(r[0], 1e-12)
        M = s.M.cpu().numpy().astype(np.float64)
        got = np.concatenate([M[np.repeat(grp == g, nl)].mean(0) for g in (0, 1, 2)])
        worst = max(worst, float(np.abs(got - r[1:]).max() / 8e5))
        table.add_row([s.state()["time"], got[0:3], got[3:6], got[6:9]])
    s.close()
    assert worst <= 1e-4, worst
    out = (tmp_path / "trilayer.txt").read_text().split("\n")
    assert out[:11] == ref_lines[:11]
    rows_out = [l for l in out[11:] if l.strip()]
    rows_ref = [l for l in ref_lines[11:] if l.strip()]
    assert len(rows_out) == len(rows_ref) == 300
    for a, b in zip(rows_out, rows_ref):
        assert a.split("\t")[0] == b.split("\t")[0]          # the same save times, as printed
    # the run is the switching event, not a quiet state: outer meshes end near +x, middle near -x
    assert ref[-1, 1] / 8e5 > 0.8 and ref[-1, 4] / 8e5 < -0.9
