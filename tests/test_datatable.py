"""The data-table writer against the reference-held `savedata` file
"MLConv Paper/3D/trilayer_switch_mconv.txt" (copied into tests/golden/): from the meshes'
rectangles, cell counts and cell sizes as the reference holds them (rectangles built from a
start and a size, cell sizes as size / cells) the header is reproduced line for line, and
every data row is reproduced from its values."""
from pathlib import Path

import numpy as np

from boris2_amd import datatable as DT

GOLD = Path(__file__).resolve().parent / "golden" / "trilayer_switch_mconv.txt"


def meshes():
    # lengths as parsed from "<value>nm" text (the literal doubles); permalloy0 starts at
    # -21 nm and is 20 nm thick, so its top face is -21 nm + 20 nm
    p = DT.MeshInfo("permalloy", (0, 0, 0, 640e-9, 320e-9, 10e-9), (128, 64, 2), (5e-9,) * 3)
    p2 = DT.MeshInfo("permalloy2", (0, 0, 11e-9, 640e-9, 320e-9, 11e-9 + 20e-9), (128, 64, 4), (5e-9,) * 3)
    p0 = DT.MeshInfo("permalloy0", (0, 0, -21e-9, 640e-9, 320e-9, -21e-9 + 20e-9), (128, 64, 4), (5e-9,) * 3)
    # the file prints the supermesh's 1 nm cell height as "1000pm", i.e. the value it held was
    # just below 1e-9 (how it was formed is not recorded); 1e-9 rounded to FP32 is such a value
    zs, ze = -21e-9, 11e-9 + 20e-9
    sm = DT.MeshInfo("supermesh", (0, 0, zs, 640e-9, 320e-9, ze), (128, 64, 52), (5e-9, 5e-9, float(np.float32(1e-9))))
    return [p, p2, p0], sm


def test_header_reproduces_reference_file(tmp_path):
    ref = GOLD.read_text().split("\n")
    ms, sm = meshes()
    by = {m.name: m for m in ms}
    cols = [DT.time_column()] + [DT.avm_column(by[n]) for n in ("permalloy0", "permalloy", "permalloy2")]
    out = tmp_path / "t.txt"
    DT.DataTable(out, ms, sm, cols, stamp=ref[0])
    got = out.read_text().split("\n")
    assert got[:11] == ref[:11]          # stamp, meshes, supermesh, labels, blank line
    assert got[11:] == [""]             # no rows yet


def test_rows_reproduce_reference_values():
    lines = GOLD.read_text().split("\n")[12:]
    ms, _ = meshes()
    by = {m.name: m for m in ms}
    cols = [DT.time_column()] + [DT.avm_column(by[n]) for n in ("permalloy0", "permalloy", "permalloy2")]
    count = 0
    for line in lines:
        if not line.strip():
            continue
        v = [float(x) for x in line.split("\t")]
        row = [v[0], v[1:4], v[4:7], v[7:10]]
        assert DT.DataTable.format_row(cols, row) == line
        count += 1
    assert count == 299


def test_engineering_units():
    assert DT.eng(-21e-9 + 20e-9, "m") == "-1000pm"    # the reference's own output for this value
    assert DT.eng(1e-9, "m") == "1nm" and DT.eng(float(np.float32(1e-9)), "m") == "1000pm"
    assert DT.eng(-1e-9, "m") == "-1nm"
    assert DT.eng(0.0, "m") == "0m"
    assert DT.eng(2.5e-3, "s") == "2.5ms"
    assert DT.eng(7.9577e5, "A/m") == "795.77kA/m"
