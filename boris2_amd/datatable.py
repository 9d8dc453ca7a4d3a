"""Data tables in the reference's `savedata` text format, so ⟨M⟩ trajectories written here read
like (and diff against) the reference's own output files.

Follows Simulation::SaveData / SaveData_DiskBufferFlush (Boris/SimulationData.cpp:657-770):
a header with the creation time (ctime), one line per mesh and one for the supermesh
("<name> : Rectangle : s; e. Cells : nx, ny, nz. Cellsize : hx, hy, hz", lengths in
engineering notation with a metre unit), the saved-data labels ("dataname (unit) <mesh>
(box)" each followed by one tab per component), then one tab-separated row per save.
Numbers are printed at 6 significant digits (BorisLib ToString's default precision); lengths
with a unit use the engineering exponent of decexp_eng (BorisLib/Funcs_Math_base.h:52-80),
whose repeated scaling of the divisor is kept so values such as -21 nm + 20 nm print as the
reference prints them ("-1000pm").
"""
from __future__ import annotations

import time as _time
from dataclasses import dataclass

_PREFIX = {-18: "a", -15: "f", -12: "p", -9: "n", -6: "u", -3: "m", 0: "", 3: "k", 6: "M", 9: "G", 12: "T", 15: "P"}


def num(v: float) -> str:
    """A number at 6 significant digits, as an iostream with setprecision(6) prints it."""
    return "%.6g" % v


def decexp_eng(value: float, minimum: float = 1e-18, maximum: float = 1e18) -> tuple[float, int]:
    """value = mantissa * 10^exp with exp a multiple of 3 and 1 <= |mantissa| < 1000."""
    av = abs(value)
    if av < minimum or av >= maximum:
        return value, 0
    exp, div = 0, 1.0
    while True:
        if av / div >= 1e3:
            exp += 3
            div *= 1e3
            continue
        if av / div < 1:
            exp -= 3
            div *= 1e-3
            continue
        break
    return value / div, exp


def eng(value: float, unit: str) -> str:
    m, e = decexp_eng(value)
    return num(m) + _PREFIX.get(e, "") + unit


def vec(values, unit: str = "") -> str:
    return ", ".join(eng(v, unit) if unit else num(v) for v in values)


def rect(r, unit: str = "m") -> str:
    """Rectangle (sx, sy, sz, ex, ey, ez) as "s; e"."""
    return vec(r[:3], unit) + "; " + vec(r[3:], unit)


@dataclass
class MeshInfo:
    """A mesh of the header: name, rectangle (sx, sy, sz, ex, ey, ez) in m, cells, cell size in m."""
    name: str
    rect: tuple
    cells: tuple
    cellsize: tuple

    def line(self) -> str:
        return (f"<{self.name}> : Rectangle : {rect(self.rect)}. Cells : {vec(self.cells)}. "
                f"Cellsize : {vec(self.cellsize, 'm')}")


@dataclass
class Column:
    """A saved quantity: name (e.g. "time", "<M>"), unit, components, and for mesh quantities the
    mesh name and the box (relative rectangle) it is averaged over."""
    name: str
    unit: str = ""
    components: int = 1
    mesh: str | None = None
    box: tuple | None = None

    def label(self) -> str:
        s = self.name
        if self.unit:
            s += f" ({self.unit})"
        if self.mesh is not None:
            s += f" <{self.mesh}>"
        if self.box is not None:
            s += f" ({rect(self.box)})"
        return s + "\t" * self.components


def time_column() -> Column:
    return Column("time", "s", 1)


def avm_column(mesh: MeshInfo) -> Column:
    """<M> (A/m) of a whole mesh (box = the mesh's rectangle relative to its origin)."""
    r = mesh.rect
    return Column("<M>", "A/m", 3, mesh.name, (0.0, 0.0, 0.0, r[3] - r[0], r[4] - r[1], r[5] - r[2]))


class DataTable:
    """A `savedata` file: header on creation, add_row() appends one row of values (a number per
    single-component column, a 3-sequence per vector column)."""

    def __init__(self, path, meshes: list[MeshInfo], supermesh: MeshInfo, columns: list[Column], stamp: str | None = None):
        self.path = str(path)
        self.columns = list(columns)
        stamp = stamp if stamp is not None else _time.ctime()
        lines = [stamp, ""] + [m.line() for m in meshes] + [supermesh.line(), "",
                 "Saved data (dataname (unit) <meshname> (cells_rectangle)) : ", "",
                 "".join(c.label() for c in self.columns), ""]
        with open(self.path, "w") as f:
            f.write("\n".join(lines) + "\n")

    @staticmethod
    def format_row(columns: list[Column], values) -> str:
        out = []
        for c, v in zip(columns, values):
            out.extend(num(x) for x in (v if c.components > 1 else [v]))
        return "\t".join(out)

    def add_row(self, values) -> None:
        if len(values) != len(self.columns):
            raise ValueError("one value per column")
        with open(self.path, "a") as f:
            f.write(self.format_row(self.columns, values) + "\n")
