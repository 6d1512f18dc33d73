"""Wall-mask geometries and the legacy-VTK formats of the sim_tcpc path.

The reference reads its geometry from a legacy VTK RECTILINEAR_GRID file with a
CELL_DATA array ``wall`` (value 1 = wall, anything else = open), written by its
STL pipeline (geometry/tcpc_geometry.py:619-661) and parsed by ``readGeometry``
(solver/sim_tcpc/sim_tcpc.cu:134-185): cell counts are the point dimensions
minus one and the spacing is the first positive x-coordinate step.  This module
writes and reads that format with numpy only, and generates the synthetic
geometries used by the tests and the benchmark:

* ``channel_mask``: a straight cylindrical channel along x (the 64x32x32,
  radius-12 case), wall = cells with (y-cy)^2+(z-cz)^2 > R^2, cy = cz = the
  grid centre (15.5 for 32 cells).
* ``tcpc_mask``: the synthetic total cavopulmonary junction voxelised from
  analytic cylinders: an IVC tube from x-min, an SVC tube from x-max with their
  axes offset in y, and a pulmonary tube along y to the label-5/6 planes.

The solver output file (``results_<id>/vtk3D/rank000_data_<n>.vtk``, written
by solver/lbm3d/state.hpp:782-851 through vtk_writer.hpp:1-60) is parsed by
``read_output_vtk``.
"""
from __future__ import annotations

import dataclasses
import math
from pathlib import Path

import numpy as np

__all__ = [
    "Geometry",
    "channel_mask",
    "tcpc_mask",
    "write_wall_vtk",
    "read_wall_vtk",
    "read_output_vtk",
]


@dataclasses.dataclass
class Geometry:
    """A wall mask: ``labels[x, y, z]`` int32, 1 = wall, and the spacing in m."""

    labels: np.ndarray
    dl: float
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def dims(self) -> tuple[int, int, int]:
        return tuple(int(v) for v in self.labels.shape)  # type: ignore[return-value]

    def flat(self) -> np.ndarray:
        """x-fastest flat copy (the VTK cell order), int32 contiguous."""
        return np.ascontiguousarray(self.labels.transpose(2, 1, 0).reshape(-1), dtype=np.int32)


def channel_mask(nx: int = 64, ny: int = 32, nz: int = 32, radius: float = 12.0, dl: float = 1.0e-3) -> Geometry:
    """Straight channel along x: open where (y-cy)^2 + (z-cz)^2 <= R^2."""
    cy = (ny - 1) / 2.0
    cz = (nz - 1) / 2.0
    y = np.arange(ny, dtype=np.float64)[:, None]
    z = np.arange(nz, dtype=np.float64)[None, :]
    open_yz = (y - cy) ** 2 + (z - cz) ** 2 <= radius * radius
    lab = np.where(open_yz[None, :, :], 0, 1).astype(np.int32)
    lab = np.broadcast_to(lab, (nx, ny, nz)).copy()
    return Geometry(lab, dl, {"kind": "channel", "radius": radius, "center": (cy, cz)})


def tcpc_mask(
    scale: float = 1.0,
    r_ivc: float = 20.0,
    r_svc: float = 16.0,
    r_pa: float = 14.0,
    offset: float = 6.0,
    base: tuple[int, int, int] = (256, 192, 128),
    dl: float = 0.25e-3,
    dims: tuple[int, int, int] | None = None,
) -> Geometry:
    """Synthetic TCPC junction on a ``base * scale`` grid (256x192x128 at 1.0).

    Cell (x,y,z) is open when it lies within radius R of a tube axis, measured
    from integer cell indices:
      IVC: axis along x at (y = Y/2 - off/2, z = Z/2), x in [0, X/2]
      SVC: axis along x at (y = Y/2 + off/2, z = Z/2), x in [X/2, X)
      PA : axis along y at (x = X/2, z = Z/2), all y
    Everything else is wall (value 1), so no exterior label reaches the planes.
    Radii and the caval offset scale with ``scale``.
    """
    if dims is None:
        X, Y, Z = (int(round(b * scale)) for b in base)
    else:
        X, Y, Z = dims
    s = scale
    R_i, R_s, R_p, off = r_ivc * s, r_svc * s, r_pa * s, offset * s
    xc, zc = X // 2, Z // 2
    yi = Y // 2 - off / 2.0
    ys = Y // 2 + off / 2.0
    x = np.arange(X, dtype=np.float64)[:, None, None]
    y = np.arange(Y, dtype=np.float64)[None, :, None]
    z = np.arange(Z, dtype=np.float64)[None, None, :]
    ivc = ((y - yi) ** 2 + (z - zc) ** 2 <= R_i * R_i) & (x <= xc)
    svc = ((y - ys) ** 2 + (z - zc) ** 2 <= R_s * R_s) & (x >= xc)
    pa = (x - xc) ** 2 + (z - zc) ** 2 <= R_p * R_p
    open_ = ivc | svc | pa
    lab = np.where(open_, 0, 1).astype(np.int32)
    meta = {
        "kind": "tcpc",
        "scale": s,
        "ivc": {"center": (yi, float(zc)), "radius": R_i},
        "svc": {"center": (ys, float(zc)), "radius": R_s},
        "pa": {"center": (float(xc), float(zc)), "radius": R_p},
    }
    return Geometry(lab, dl * (1.0 / s if s != 0 else 1.0), meta)


# ----------------------------------------------------------------------------
# legacy VTK
def _coords(n: int, dl: float) -> np.ndarray:
    return (np.arange(n, dtype=np.float64) * dl).astype(">f4")


def write_wall_vtk(path: str | Path, geom: Geometry, binary: bool = True, coords: str = "float") -> None:
    """Write the wall mask as the reference's geometry.vtk: RECTILINEAR_GRID
    with point coordinates 0..N*dl (``coords`` "float" or "double") and
    CELL_DATA ``wall`` (double)."""
    if coords not in ("float", "double"):
        raise ValueError(f"coords must be 'float' or 'double', not {coords!r}")
    X, Y, Z = geom.dims
    path = Path(path)
    with path.open("wb") as f:
        f.write(b"# vtk DataFile Version 3.0\n")
        f.write(b"lbm355 wall mask\n")
        f.write(b"BINARY\n" if binary else b"ASCII\n")
        f.write(b"DATASET RECTILINEAR_GRID\n")
        f.write(f"DIMENSIONS {X + 1} {Y + 1} {Z + 1}\n".encode())
        for name, n in (("X", X), ("Y", Y), ("Z", Z)):
            c = (np.arange(n + 1, dtype=np.float64) * geom.dl).astype(">f8" if coords == "double" else ">f4")
            f.write(f"{name}_COORDINATES {n + 1} {coords}\n".encode())
            if binary:
                f.write(c.tobytes())
                f.write(b"\n")
            else:
                f.write((" ".join(f"{v:.9g}" for v in c.astype(np.float64)) + "\n").encode())
        f.write(f"CELL_DATA {X * Y * Z}\n".encode())
        f.write(b"SCALARS wall double 1\nLOOKUP_TABLE default\n")
        flat = geom.flat().astype(">f8")
        if binary:
            f.write(flat.tobytes())
            f.write(b"\n")
        else:
            f.write((" ".join(str(int(v)) for v in flat) + "\n").encode())


_VTK_TYPES = {
    "float": ">f4",
    "double": ">f8",
    "int": ">i4",
    "unsigned_int": ">u4",
    "short": ">i2",
    "unsigned_short": ">u2",
    "char": ">i1",
    "unsigned_char": ">u1",
    "long": ">i8",
}


class _Reader:
    def __init__(self, data: bytes):
        self.data = data
        self.pos = 0

    def line(self) -> str:
        end = self.data.index(b"\n", self.pos)
        s = self.data[self.pos:end].decode("latin-1").strip()
        self.pos = end + 1
        return s

    def next_nonempty(self) -> str:
        while self.pos < len(self.data):
            s = self.line()
            if s:
                return s
        raise EOFError

    def array(self, n: int, vtype: str, binary: bool) -> np.ndarray:
        dt = np.dtype(_VTK_TYPES[vtype])
        if binary:
            nb = n * dt.itemsize
            arr = np.frombuffer(self.data, dtype=dt, count=n, offset=self.pos).astype(dt.newbyteorder("="))
            self.pos += nb
            # a newline usually follows binary blocks
            if self.pos < len(self.data) and self.data[self.pos:self.pos + 1] == b"\n":
                self.pos += 1
            return arr
        vals: list[float] = []
        while len(vals) < n:
            vals.extend(float(t) for t in self.line().split())
        return np.asarray(vals[:n], dtype=np.float64)


def _parse(path: str | Path) -> dict:
    data = Path(path).read_bytes()
    r = _Reader(data)
    r.line()  # version
    r.line()  # title
    fmt = r.line().upper()
    binary = fmt.startswith("BINARY")
    ds = r.next_nonempty().split()
    if len(ds) < 2 or ds[1].upper() != "RECTILINEAR_GRID":
        raise ValueError(f"{path} is not a legacy VTK RectilinearGrid file.")
    out: dict = {"coords": {}, "point": {}, "cell": {}, "field": {}}
    section = None
    npts = ncells = 0
    while True:
        try:
            s = r.next_nonempty()
        except EOFError:
            break
        tok = s.split()
        key = tok[0].upper()
        if key == "DIMENSIONS":
            out["dims"] = tuple(int(v) for v in tok[1:4])
        elif key.endswith("_COORDINATES"):
            n = int(tok[1])
            out["coords"][key[0]] = r.array(n, tok[2], binary).astype(np.float64)
        elif key == "CELL_DATA":
            section, ncells = "cell", int(tok[1])
        elif key == "POINT_DATA":
            section, npts = "point", int(tok[1])
        elif key == "FIELD":
            nfields = int(tok[2])
            for _ in range(nfields):
                ft = r.next_nonempty().split()
                ncomp, ntup = int(ft[1]), int(ft[2])
                out["field"][ft[0]] = r.array(ncomp * ntup, ft[3], binary)
        elif key == "SCALARS":
            name, vtype = tok[1], tok[2]
            ncomp = int(tok[3]) if len(tok) > 3 else 1
            lt = r.next_nonempty()
            if not lt.upper().startswith("LOOKUP_TABLE"):
                raise ValueError("SCALARS without LOOKUP_TABLE")
            n = (ncells if section == "cell" else npts) * ncomp
            out[section][name] = r.array(n, vtype, binary)
        elif key == "VECTORS":
            name, vtype = tok[1], tok[2]
            n = (ncells if section == "cell" else npts) * 3
            out[section][name] = r.array(n, vtype, binary).reshape(-1, 3)
        else:
            raise ValueError(f"unsupported VTK keyword {tok[0]!r}")
    return out


def read_wall_vtk(path: str | Path) -> Geometry:
    """readGeometry (sim_tcpc.cu:134-185): dims = point dims - 1, dl = first
    positive x spacing, labels = int(wall) per cell."""
    p = _parse(path)
    X, Y, Z = (d - 1 for d in p["dims"])
    xc = p["coords"].get("X")
    if xc is None or len(xc) < 2:
        raise ValueError("Missing rectilinear x coordinates.")
    dl = 0.0
    for a, b in zip(xc[:-1], xc[1:]):
        if abs(b - a) > 0:
            dl = abs(b - a)
            break
    if dl <= 0:
        raise ValueError("Cannot infer positive rectilinear x spacing.")
    if "wall" not in p["cell"]:
        raise ValueError(f"{path} does not contain a cell-data array named 'wall'.")
    flat = p["cell"]["wall"].astype(np.float64).astype(np.int32)
    lab = flat.reshape(Z, Y, X).transpose(2, 1, 0).copy()
    return Geometry(lab, float(dl))


def read_output_vtk(path: str | Path) -> dict:
    """Parse a solver output file; arrays come back as [x, y, z(, 3)]."""
    p = _parse(path)
    X, Y, Z = p["dims"]
    res = {"dims": (X, Y, Z), "coords": p["coords"], "field": p["field"]}
    for name, arr in p["point"].items():
        if arr.ndim == 2:
            res[name] = arr.reshape(Z, Y, X, 3).transpose(2, 1, 0, 3).copy()
        else:
            res[name] = arr.reshape(Z, Y, X).transpose(2, 1, 0).copy()
    return res


def cylinder_area(radius: float) -> float:
    return math.pi * radius * radius
