"""STL surface -> wall-mask VTK of the TCPC25_03 case (the solver's geometry.vtk).

Stands in for geometry/tcpc_geometry.py:main (with tcpc_taper_ends.py and
meshgen/voxels.py), run by scripts/generate_geometry.py.  The reference does
this with trimesh (end detection, loft, voxelisation) and vtk (writer), neither
of which is present here; this module restates the steps they perform with
numpy and scipy.ndimage:

1. read the STL, weld identical vertices;
2. find the four planar end caps (two facing +-x, two facing +-y) and
   extend ends 0, 1, 2 along their outward normals (TAPER_PARAMS
   (100, 10.0) clamps to scale 1, i.e. a straight 10 mm extension,
   tcpc_geometry.py:167-172, tcpc_taper_ends.py:535-621);
3. voxelise the surface at pitch = longest extent / (520 - 1): subdivide
   every triangle until its edges are at most pitch / 2 and mark the voxel of
   each vertex (trimesh's "subdivide" voxeliser, the default of
   Trimesh.voxelized), fill the interior (scipy binary_fill_holes) and crop
   or pad to ceil(extent / pitch) + 1 voxels per axis (voxels.py:571-610);
4. label: 0 outside, 1 fluid, 2 the one-voxel solid shell, 3/4/5 the
   near-wall bands (voxels.py:814-1056), then the end planes 13/14/16/15 on
   the y-max / y-min / x-max / x-min planes of the fluid extent
   (tcpc_geometry.py:375-449) and one wall layer on z-min and z-max
   (tcpc_geometry.py:452-490);
5. the PE changes (tcpc_geometry.py:681-870): swap the first two axes, extend
   both pulmonary ends by 25 mm plus the padding that makes the new y extent a
   multiple of 32, narrow the extensions to a 2 mm radius with a tanh profile,
   drop the end tags inside the domain, turn near-wall cells with no
   second-band neighbour into wall, and map 3/4/5 to fluid and the exterior
   to wall;
6. write geometry.vtk and the pre-PE label file test.vtk with the value map
   of tcpc_geometry.py:663-679 (wall 1, fluid 0, end tags 3/4/5/6).

In test.vtk the exterior keeps the value 0, which the solver reads as fluid
(only 1 is wall, sim_tcpc.cu:518-519): exterior cells on the boundary planes
become inflow/outflow cells there.  geometry.vtk has no exterior fluid.

Parity with the reference's output files is unpinned: trimesh and vtk are not
available, so the checks are invariants (voxel volume against the mesh's
enclosed volume, cap areas against the labelled end planes, the y extent a
multiple of 32, the reader round trip).  Deviations from the reference's
steps: the end caps are chosen as the two largest planar facets per axis
(the reference takes four candidates per axis and drops two by centroid
order), and the cap of an extended end reuses the original cap's triangles
moved along the normal instead of a new ear-clipped triangulation of the
same rim polygon.
"""
from __future__ import annotations

import argparse
import math
import sys
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .geometry import Geometry, write_wall_vtk

# ---------------------------------------------------------------------------
# parameters of the TCPC25_03 geometry (tcpc_geometry.py:148-226 and
# configs/simulation.yaml)
LONGEST_AXIS_VOXELS = 520
EXTEND_ENDS = (0, 1, 2)
EXTENSION_LENGTH = 10.0          # STL units (mm)
EXTENSION_SEGMENTS = 24          # rings of the loft ("smooth" profile)
END_NORMAL_THRESHOLD = 0.95
PE_EXTENSION_LEFT_M = 0.025
PE_EXTENSION_RIGHT_M = 0.025
PE_OUTLET_RADIUS_M = 0.002
PE_BLOCK_SIZE = 32
LABEL_Y_MAX, LABEL_Y_MIN, LABEL_X_MIN, LABEL_X_MAX = 13, 14, 15, 16
STL_UNIT_M = 1.0e-3


# ---------------------------------------------------------------------------
# mesh
@dataclass
class Mesh:
    vertices: np.ndarray   # [n, 3] float64
    faces: np.ndarray      # [m, 3] int64, counter-clockwise seen from outside

    @property
    def bounds(self) -> np.ndarray:
        return np.stack([self.vertices.min(0), self.vertices.max(0)])

    def face_normals(self) -> tuple[np.ndarray, np.ndarray]:
        """(unit normals, areas) of the faces."""
        a, b, c = (self.vertices[self.faces[:, k]] for k in range(3))
        cr = np.cross(b - a, c - a)
        area2 = np.linalg.norm(cr, axis=1)
        return cr / np.maximum(area2, 1e-300)[:, None], 0.5 * area2

    def volume(self) -> float:
        """Enclosed volume (divergence theorem; positive for outward faces)."""
        a, b, c = (self.vertices[self.faces[:, k]] for k in range(3))
        return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)

    def area(self) -> float:
        return float(self.face_normals()[1].sum())

    def edge_use(self) -> np.ndarray:
        """How many faces use each undirected edge (2 everywhere = closed)."""
        e = np.sort(self.faces[:, [0, 1, 1, 2, 2, 0]].reshape(-1, 2), axis=1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        return cnt


def read_stl(path: str | Path) -> Mesh:
    """Binary or ASCII STL; identical vertex positions are merged."""
    raw = Path(path).read_bytes()
    tri = None
    if len(raw) >= 84:
        n = int(np.frombuffer(raw[80:84], "<u4")[0])
        if len(raw) == 84 + 50 * n:
            rec = np.frombuffer(raw[84:], np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]))
            tri = rec["v"].astype(np.float64)
    if tri is None:
        pts = [line.split()[1:4] for line in raw.decode("ascii", "replace").splitlines()
               if line.strip().startswith("vertex")]
        if not pts or len(pts) % 3:
            raise ValueError(f"{path}: not a binary or ASCII STL")
        tri = np.asarray(pts, dtype=np.float64).reshape(-1, 3, 3)
    uniq, inv = np.unique(tri.reshape(-1, 3), axis=0, return_inverse=True)
    return Mesh(uniq, inv.reshape(-1, 3).astype(np.int64))


def write_stl(path: str | Path, mesh: Mesh) -> None:
    """Binary STL (float32 coordinates)."""
    nrm, _ = mesh.face_normals()
    rec = np.zeros(len(mesh.faces), np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]))
    rec["n"] = nrm
    rec["v"] = mesh.vertices[mesh.faces]
    with Path(path).open("wb") as f:
        f.write(b"lbm355 binary STL".ljust(80, b"\0"))
        f.write(np.uint32(len(rec)).tobytes())
        f.write(rec.tobytes())


# ---------------------------------------------------------------------------
# end caps and their extension (tcpc_taper_ends.py)
@dataclass
class End:
    label: int
    faces: np.ndarray            # face indices of the planar cap
    normal: np.ndarray           # outward unit normal
    centroid: np.ndarray         # mean of the rim vertices
    area: float
    rim: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))  # ordered rim vertex loop


def planar_facets(mesh: Mesh, angle_tol: float = 1e-6, dist_tol: float = 1e-6) -> list[np.ndarray]:
    """Groups of two or more edge-connected faces that lie in one plane."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    nrm, _ = mesh.face_normals()
    m = len(mesh.faces)
    e = np.sort(mesh.faces[:, [0, 1, 1, 2, 2, 0]].reshape(-1, 2), axis=1)
    owner = np.repeat(np.arange(m), 3)
    order = np.lexsort((e[:, 1], e[:, 0]))
    e, owner = e[order], owner[order]
    same = np.all(e[1:] == e[:-1], axis=1)
    fa, fb = owner[:-1][same], owner[1:][same]
    scale = float(np.ptp(mesh.vertices, axis=0).max())
    cen = mesh.vertices[mesh.faces].mean(axis=1)
    ok = (np.einsum("ij,ij->i", nrm[fa], nrm[fb]) > 1.0 - angle_tol) & \
         (np.abs(np.einsum("ij,ij->i", nrm[fa], cen[fb] - cen[fa])) < dist_tol * scale)
    g = coo_matrix((np.ones(int(ok.sum())), (fa[ok], fb[ok])), shape=(m, m))
    _, comp = connected_components(g, directed=False)
    counts = np.bincount(comp)
    order = np.argsort(comp, kind="stable")
    groups = np.split(order, np.cumsum(counts)[:-1])
    return [grp for grp in groups if len(grp) >= 2]


def _rim(mesh: Mesh, faces: np.ndarray) -> np.ndarray:
    """Ordered boundary loop of a set of faces (edges used once within it),
    walked along the faces' orientation."""
    f = mesh.faces[faces]
    directed = f[:, [0, 1, 1, 2, 2, 0]].reshape(-1, 2)
    key = np.sort(directed, axis=1)
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    bnd = directed[cnt[inv.ravel()] == 1]
    nxt = {int(a): int(b) for a, b in bnd}
    if len(nxt) != len(bnd):
        raise ValueError("cap rim is not a simple loop")
    start = int(bnd[0, 0])
    loop = [start]
    while True:
        v = nxt[loop[-1]]
        if v == start:
            break
        loop.append(v)
        if len(loop) > len(bnd):
            raise ValueError("cap rim is not a single loop")
    if len(loop) != len(bnd):
        raise ValueError("cap has more than one boundary loop")
    return np.asarray(loop, np.int64)


def find_ends(mesh: Mesh) -> list[End]:
    """The four planar end caps, labelled as tcpc_geometry.py:919-951 does:
    the two x-facing caps first (sorted by centroid (y, x, z)), then the
    y-facing caps by increasing y."""
    nrm, area = mesh.face_normals()
    facets = planar_facets(mesh)
    chosen: list[End] = []
    for axis in (0, 1):
        cands = []
        for grp in facets:
            n = nrm[grp[0]]
            if abs(n[axis]) > END_NORMAL_THRESHOLD:
                cands.append((float(area[grp].sum()), grp))
        cands.sort(key=lambda t: -t[0])
        if len(cands) < 2:
            raise RuntimeError(f"found {len(cands)} planar end caps facing the {'xy'[axis]} axis, expected 2")
        picked = []
        for a, grp in cands[:2]:
            rim = _rim(mesh, grp)
            picked.append(End(-1, grp, nrm[grp[0]].copy(), mesh.vertices[rim].mean(0), a, rim))
        if axis == 0:
            picked.sort(key=lambda e: (e.centroid[1], e.centroid[0], e.centroid[2]))
        else:
            picked.sort(key=lambda e: e.centroid[1])
        chosen += picked
    for k, e in enumerate(chosen):
        e.label = k
    return chosen


def extend_ends(mesh: Mesh, ends: list[End], which=EXTEND_ENDS, length: float = EXTENSION_LENGTH,
                segments: int = EXTENSION_SEGMENTS) -> Mesh:
    """Extend the listed ends straight along their outward normals by
    `length`: the cap is removed, `segments` rings of the rim are lofted out
    (side quads as two triangles, tcpc_taper_ends.py:584-596) and the cap is
    put back at the far end."""
    V = [mesh.vertices]
    keep = np.ones(len(mesh.faces), bool)
    F = []
    nv = len(mesh.vertices)
    for e in ends:
        if e.label not in which or length <= 0:
            continue
        keep[e.faces] = False
        rim = e.rim
        n = len(rim)
        prev = rim
        for s in range(1, segments + 1):
            ring = mesh.vertices[rim] + e.normal * (length * s / segments)
            idx = np.arange(nv, nv + n)
            V.append(ring)
            nv += n
            j = np.roll(np.arange(n), -1)
            F.append(np.stack([prev, prev[j], idx[j]], axis=1))
            F.append(np.stack([prev, idx[j], idx], axis=1))
            prev = idx
        # the cap at the far end: the original cap's faces on the last ring
        remap = -np.ones(len(mesh.vertices), np.int64)
        remap[rim] = prev
        cap = mesh.faces[e.faces]
        inner = np.setdiff1d(np.unique(cap), rim)
        if len(inner):
            remap[inner] = np.arange(nv, nv + len(inner))
            V.append(mesh.vertices[inner] + e.normal * length)
            nv += len(inner)
        F.append(remap[cap])
    return Mesh(np.concatenate(V), np.concatenate([mesh.faces[keep]] + F))


# ---------------------------------------------------------------------------
# voxelisation (trimesh voxel.creation.voxelize_subdivide + voxels.py:571-610)
def surface_points(mesh: Mesh, max_edge: float, max_iter: int = 10) -> np.ndarray:
    """Vertices of the mesh after midpoint subdivision of every face until its
    longest edge is at most max_edge: a face split k times carries the
    barycentric lattice of 2^k divisions."""
    a, b, c = (mesh.vertices[mesh.faces[:, k]] for k in range(3))
    longest = np.max(np.stack([np.linalg.norm(b - a, axis=1), np.linalg.norm(c - b, axis=1),
                               np.linalg.norm(a - c, axis=1)]), axis=0)
    k = np.ceil(np.log2(np.maximum(longest / max_edge, 1.0))).astype(int)
    if k.max() > max_iter:
        raise ValueError(f"a face needs {k.max()} subdivisions (more than {max_iter})")
    out = []
    for lev in np.unique(k):
        sel = k == lev
        n = 1 << int(lev)
        i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
        m = (i + j) <= n
        u = (i[m] / n)[None, :, None]
        v = (j[m] / n)[None, :, None]
        A, B, C = a[sel][:, None, :], b[sel][:, None, :], c[sel][:, None, :]
        out.append((A + u * (B - A) + v * (C - A)).reshape(-1, 3))
    return np.concatenate(out)


def voxelize(mesh: Mesh, pitch: float) -> np.ndarray:
    """Boolean occupancy on the voxel lattice of the mesh bounds: surface
    voxels (each subdivided vertex rounded to the nearest lattice point at
    integer multiples of the pitch), interior filled, shaped
    ceil(extent / pitch) + 1 per axis with the lowest occupied voxel at 0."""
    from scipy import ndimage
    pts = surface_points(mesh, pitch / 2.0)
    hit = np.round(pts / pitch).astype(np.int64)
    hit -= hit.min(axis=0)
    shape = tuple(int(s) for s in hit.max(axis=0) + 1)
    occ = np.zeros(shape, bool)
    occ[hit[:, 0], hit[:, 1], hit[:, 2]] = True
    occ |= ndimage.binary_fill_holes(occ)
    ext = mesh.bounds[1] - mesh.bounds[0]
    want = tuple(int(v) for v in np.ceil(ext / pitch).astype(int) + 1)
    out = np.zeros(want, bool)
    s = tuple(slice(0, min(a, b)) for a, b in zip(occ.shape, want))
    out[s] = occ[s]
    return out


# ---------------------------------------------------------------------------
# labels
_N26 = np.ones((3, 3, 3), bool)


def _touches(mask: np.ndarray) -> np.ndarray:
    """Cells with a 26-neighbour in mask (outside the array counts as not)."""
    from scipy import ndimage
    return ndimage.binary_dilation(mask, structure=_N26)


def label_voxels(occ: np.ndarray) -> np.ndarray:
    """0 outside, 1 fluid, 2 outside cells next to fluid (the wall shell), then
    the near-wall bands: fluid next to 2 -> 3, fluid next to 3 -> 4, fluid
    next to 4 -> 5 (voxels.py:814-943, with no domain-face tags)."""
    lab = np.zeros(occ.shape, np.int32)
    lab[occ] = 1
    lab[~occ & _touches(occ)] = 2
    for band, nxt in ((2, 3), (3, 4), (4, 5)):
        lab[(lab == 1) & _touches(lab == band)] = nxt
    return lab


def label_end_planes(lab: np.ndarray) -> np.ndarray:
    """Tag the largest 8-connected fluid component on the y-max, y-min,
    x-max and x-min planes of the fluid extent with 13, 14, 16, 15, in that
    order (tcpc_geometry.py:375-449)."""
    from scipy import ndimage
    fluid = np.isin(lab, (1, 3, 4, 5))
    if not fluid.any():
        raise RuntimeError("no fluid voxels")
    ys = np.flatnonzero(fluid.any(axis=(0, 2)))
    xs = np.flatnonzero(fluid.any(axis=(1, 2)))
    planes = ((1, int(ys[-1]), LABEL_Y_MAX), (1, int(ys[0]), LABEL_Y_MIN),
              (0, int(xs[-1]), LABEL_X_MAX), (0, int(xs[0]), LABEL_X_MIN))
    out = lab.copy()
    for axis, idx, tag in planes:
        sl = [slice(None)] * 3
        sl[axis] = idx
        comp, n = ndimage.label(fluid[tuple(sl)], structure=np.ones((3, 3), bool))
        if n == 0:
            raise RuntimeError(f"no fluid on the end plane {'xy'[axis]}={idx}")
        sizes = np.bincount(comp.ravel())[1:]
        view = out[tuple(sl)]
        view[comp == 1 + int(np.argmax(sizes))] = tag
    return out


def pad_z(lab: np.ndarray, value: int = 2) -> np.ndarray:
    """One layer of `value` below z-min and above z-max (the faces that carry
    no end, tcpc_geometry.py:452-490)."""
    return np.pad(lab, ((0, 0), (0, 0), (1, 1)), constant_values=value)


def pe_changes(lab: np.ndarray, dx_m: float, left_m: float = PE_EXTENSION_LEFT_M, right_m: float = PE_EXTENSION_RIGHT_M,
               outlet_radius_m: float = PE_OUTLET_RADIUS_M, block: int = PE_BLOCK_SIZE) -> np.ndarray:
    """makePEChanges (tcpc_geometry.py:681-870) on the padded labels.

    The first two axes swap (the new x is the STL's y).  Along the new y the
    labels are extended by edge replication: n_left = floor(left / dx) cells
    before, n_right = floor(right / dx) plus the padding to the next multiple
    of `block` (a whole block when it already is one) after.  In the
    extensions the pulmonary end tags (15 before, 16 after) outside a radius
    that falls from the end's half-width to `outlet_radius` over the first
    2/3 n_left cells, r = r_out + (r_end - r_out) (1 - tanh(2 s - 2)) / 2 with
    s = 2 d / (2/3 n_left) at distance d into the extension, become wall.  On
    the cells off the x and y boundary planes the end tags become fluid and
    a near-wall cell (3) without a second-band (4) 26-neighbour becomes wall;
    finally 3/4/5 are fluid and the exterior (0) is wall."""
    from scipy import ndimage
    n_left = int(math.floor(left_m / dx_m))
    n_right = int(math.floor(right_m / dx_m))
    ny0 = lab.shape[0]
    n_right += block - (ny0 + n_left + n_right) % block
    pe = np.pad(lab, ((n_left, n_right), (0, 0), (0, 0)), mode="edge").transpose(1, 0, 2).copy()
    Nx, Ny, Nz = pe.shape
    n3 = int(math.floor(2 * n_left / 3.0))
    x = np.arange(Nx)[:, None]
    z = np.arange(Nz)[None, :]
    ends = pe[:, [0, Ny - 1], :]

    def extent(tag):
        xs, _, zs = np.nonzero(ends == tag)
        if xs.size == 0:
            raise RuntimeError(f"end tag {tag} missing on the y end planes")
        hx = (int(xs.max()) - int(xs.min())) // 2
        hz = (int(zs.max()) - int(zs.min())) // 2
        return hx * dx_m, hx + int(xs.min()), hz + int(zs.min())

    def radius2(d, r_end):
        if d < n3:
            rf = (1.0 - math.tanh(2.0 * (2.0 * d / n3) - 2.0)) / 2.0
            return (outlet_radius_m + rf * (r_end - outlet_radius_m)) ** 2
        return outlet_radius_m ** 2

    for tag, ys in ((LABEL_X_MAX, [(y, y - (Ny - n_right)) for y in range(Ny - n_right, Ny)]),
                    (LABEL_X_MIN, [(y, n_left - 1 - y) for y in range(n_left)])):
        r_end, xc, zc = extent(tag)
        d2 = ((x - xc) ** 2 + (z - zc) ** 2) * dx_m * dx_m
        for y, d in ys:
            sl = pe[:, y, :]
            sl[(sl == tag) & (d2 > radius2(d, r_end))] = 2
    inner = np.zeros(pe.shape, bool)
    inner[1:-1, 1:-1, :] = True
    pe[inner & np.isin(pe, (13, 14, 15, 16))] = 1
    near4 = ndimage.binary_dilation(pe == 4, structure=_N26)
    pe[inner & (pe == 3) & ~near4] = 2
    pe[np.isin(pe, (3, 4, 5))] = 1
    pe[pe == 0] = 2
    return pe


VTK_VALUE = {2: 1, 1: 0, 13: 3, 14: 4, 15: 5, 16: 6}


def vtk_values(lab: np.ndarray) -> np.ndarray:
    """The value map of the written files (newIndex2, tcpc_geometry.py:663-679):
    wall 2 -> 1, fluid 1 -> 0, end tags 13/14/15/16 -> 3/4/5/6, others kept."""
    out = lab.copy()
    for k, v in VTK_VALUE.items():
        out[lab == k] = v
    return out


# ---------------------------------------------------------------------------
@dataclass
class Result:
    geometry: Geometry       # geometry.vtk content (after the PE changes)
    test: Geometry           # test.vtk content (pre-PE labels)
    mesh: Mesh               # the extended surface
    ends: list
    pitch: float             # STL units
    dx_m: float
    occupancy: np.ndarray


def build(stl: str | Path, longest_axis_voxels: int = LONGEST_AXIS_VOXELS, extend=EXTEND_ENDS,
          length: float = EXTENSION_LENGTH, left_m: float = PE_EXTENSION_LEFT_M, right_m: float = PE_EXTENSION_RIGHT_M,
          outlet_radius_m: float = PE_OUTLET_RADIUS_M, block: int = PE_BLOCK_SIZE, unit_m: float = STL_UNIT_M,
          log=print) -> Result:
    """The whole pipeline of tcpc_geometry.py:main on one STL."""
    if int(longest_axis_voxels) < 2:
        raise ValueError("longest_axis_voxels must be an integer >= 2")
    mesh = read_stl(stl)
    log(f"mesh: {len(mesh.vertices)} vertices, {len(mesh.faces)} faces, closed={bool((mesh.edge_use() == 2).all())}")
    ends = find_ends(mesh)
    for e in ends:
        log(f"  end {e.label}: centroid {np.round(e.centroid, 3).tolist()}, normal {np.round(e.normal, 3).tolist()}, "
            f"area {e.area:.3f}, rim {len(e.rim)}")
    out = extend_ends(mesh, ends, extend, length)
    ext = out.bounds[1] - out.bounds[0]
    pitch = float(ext.max()) / (int(longest_axis_voxels) - 1)
    occ = voxelize(out, pitch)
    dx = float(ext[0]) / max(1, occ.shape[0] - 1)
    log(f"voxels: {occ.shape}, pitch {pitch:.6f}, dx {dx:.6f}, {int(occ.sum())} inside")
    lab = pad_z(label_end_planes(label_voxels(occ)))
    dx_m = dx * unit_m
    pe = pe_changes(lab, dx_m, left_m, right_m, outlet_radius_m, block)
    log(f"geometry: {pe.shape}, fluid-like {int((pe != 2).sum())}")
    return Result(Geometry(vtk_values(pe), dx_m, {"kind": "stl"}), Geometry(vtk_values(lab), dx_m, {"kind": "stl-pre-pe"}),
                  out, ends, pitch, dx_m, occ)


def main(argv=None) -> int:
    """Command line of scripts/generate_geometry.py: paths and sizes from the
    geometry section of configs/simulation.yaml, each overridable."""
    import yaml
    from .run_case import CONFIGS, ROOT
    ap = argparse.ArgumentParser(prog="generate_geometry",
                                 description="Voxelise the TCPC STL into the solver's wall-mask VTK.")
    ap.add_argument("--config", default=str(CONFIGS / "simulation.yaml"))
    ap.add_argument("--stl", default=None, help="input STL (default: the configured input_stl)")
    ap.add_argument("--longest-axis-voxels", type=int, default=None)
    ap.add_argument("--output", default=None, help="geometry VTK (default: the configured runtime_vtk)")
    ap.add_argument("--test-output", default=None, help="pre-PE label VTK (default: the configured test_vtk)")
    ap.add_argument("--tapered-stl", default=None, help="extended surface STL (default: the configured tapered_stl)")
    a = ap.parse_args(argv)
    cfg = yaml.safe_load(Path(a.config).read_text(encoding="utf-8"))["geometry"]

    def path(v, key):
        p = Path(v if v is not None else cfg[key])
        return p if p.is_absolute() else ROOT / p

    stl = path(a.stl, "input_stl")
    if not stl.exists():
        print(f"generate_geometry: {stl} not found", file=sys.stderr)
        return 1
    res = build(stl, a.longest_axis_voxels or int(cfg["longest_axis_voxels"]), left_m=float(cfg["pe_extension_left_m"]),
                right_m=float(cfg["pe_extension_right_m"]), outlet_radius_m=float(cfg["pe_outlet_radius_m"]),
                block=int(cfg["pe_block_size"]))
    outs = {"geometry_vtk": path(a.output, "runtime_vtk"), "test_vtk": path(a.test_output, "test_vtk"),
            "tapered_stl": path(a.tapered_stl, "tapered_stl")}
    for p in outs.values():
        p.parent.mkdir(parents=True, exist_ok=True)
    write_wall_vtk(outs["geometry_vtk"], res.geometry, coords="double")
    write_wall_vtk(outs["test_vtk"], res.test, coords="double")
    write_stl(outs["tapered_stl"], res.mesh)
    for k, p in outs.items():
        print(f"{k}={p}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
