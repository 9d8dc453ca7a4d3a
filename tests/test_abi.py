"""The C-ABI library loads and exports every entry point declared in include/boris_mi355x.h
(no compute calls: this runs without a GPU); the product library does not link the oracle."""
import re
import subprocess
from pathlib import Path

from boris2_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared():
    txt = (ROOT / "include" / "boris_mi355x.h").read_text()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(bm_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_what_the_binding_lists():
    assert declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    for name in declared():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True, check=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if l.strip()}
    for name in declared():
        assert name in syms, name
    assert not any(s.startswith("or_") for s in syms), "product library must not contain oracle code"


def test_product_sources_do_not_reference_the_oracle():
    for p in (ROOT / "boris2_amd").rglob("*"):
        if p.suffix in (".py", ".hip", ".cpp", ".h"):
            txt = p.read_text()
            assert "oracle" not in txt.lower(), p


def test_unsupported_sizes_are_rejected_before_any_device_work():
    """Open axes take any nx <= 4096, ny <= 2048 and nz <= 128 (N padded up to a planned
    length); larger sizes return BM_ERR_SIZE and bad arguments BM_ERR_ARG, both checked
    before any allocation, so this runs without a device."""
    import ctypes as C
    L = _lib.load()
    h = C.c_void_p()
    cases = [((8, 8, 129), 2),      # N.z = 512 > 256
             ((4097, 8, 1), 2),     # nx above 4096
             ((8, 2049, 1), 2),     # N.y above 4096
             ((0, 8, 1), 1), ((8, 0, 1), 1), ((8, 8, 0), 1)]
    for n, code in cases:
        assert L.bm_demag_create(*n, 4e-9, 4e-9, 4e-9, 0, C.byref(h)) == code, n
        assert h.value is None
    assert L.bm_demag_create(8, 8, 1, -4e-9, 4e-9, 4e-9, 0, C.byref(h)) == 1


def test_pbc_sizes_are_checked_before_any_device_work():
    import ctypes as C
    L = _lib.load()
    h = C.c_void_p()
    cases = [((6, 8, 1, 1, 0, 0), 2),       # periodic nx = 6: N.x / 2 = 3 has no plan
             ((7, 8, 1, 1, 0, 0), 2),       # periodic nx odd
             ((8, 10, 1, 0, 1, 0), 2),      # periodic ny = 10 has no plan
             ((8, 8, 3, 0, 0, 1), 2),       # periodic nz = 3: N.z = 3 is not a power of two
             ((8, 8, 2, 0, 0, 1), 2),       # N.z = 2 below the smallest z plan
             ((8, 8, 512, 0, 0, 1), 2),     # periodic nz = 512 above the largest z plan
             ((8, 8192, 1, 0, 1, 0), 2),    # periodic ny = 8192 above the largest y plan
             ((8, 8, 1, -1, 0, 0), 1)]      # negative image count
    for (nx, ny, nz, px, py, pz), code in cases:
        assert L.bm_demag_create_pbc(nx, ny, nz, 4e-9, 4e-9, 4e-9, px, py, pz, 0, C.byref(h)) == code
        assert h.value is None
