"""The STL -> wall-mask pipeline (nonnewlbm_amd/stl_geometry.py).

Its reference (geometry/tcpc_geometry.py with trimesh and vtk) cannot run
here, so parity with the reference's geometry.vtk is unpinned; these tests
pin each step by what the mathematics fixes instead:

* the subdivision voxeliser: its lattice equals recursive midpoint
  subdivision (brute force on random triangles), and the voxel volume of
  closed meshes lies between the enclosed volume V and V + area x pitch;
* the labels: the shell and near-wall bands equal a cell-by-cell
  26-neighbour scan on a small blob; end tags sit on the extreme fluid planes;
* the PE changes: edge replication, the axis swap, the padding of the new y
  extent to a multiple of 32 (a whole block when it already is one) and the
  tanh narrowing of the outlets to the 2 mm radius;
* the whole pipeline on two crossing capped cylinders with closed-form cap
  areas, and on TCPC25_03.stl where the reference checkout is present: cap
  areas against the labelled planes, the volume bounds, the reader round trip
  (Python and the sim_tcpc log), and the exterior-as-fluid behaviour of the
  pre-PE file (P7: the solver's inflow area there includes exterior cells).
"""
from __future__ import annotations

import itertools
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from nonnewlbm_amd import _abi, lbm355
from nonnewlbm_amd import stl_geometry as sg
from nonnewlbm_amd.geometry import read_wall_vtk, write_wall_vtk

ROOT = Path(__file__).resolve().parents[1]
REF_STL = Path("/root/reference/geometry/input/TCPC25_03.stl")


def cylinder(center, axis: int, radius: float, length: float, segments: int = 96, hub: bool = True) -> sg.Mesh:
    """Closed, outward-oriented capped cylinder along `axis` (regular polygon
    section).  hub: caps fan from a centre vertex (an interior cap vertex),
    else from the first rim vertex."""
    u, v = [a for a in range(3) if a != axis]
    t = 2 * np.pi * np.arange(segments) / segments
    ring = np.zeros((segments, 3))
    ring[:, u] = radius * np.cos(t)
    ring[:, v] = radius * np.sin(t)
    lo, hi = ring.copy(), ring.copy()
    lo[:, axis] = -length / 2
    hi[:, axis] = length / 2
    V = [lo, hi]
    n = segments
    j = np.roll(np.arange(n), -1)
    F = [np.stack([np.arange(n), j, n + j], 1), np.stack([np.arange(n), n + j, n + np.arange(n)], 1)]
    if hub:
        c0 = np.zeros(3); c0[axis] = -length / 2
        c1 = np.zeros(3); c1[axis] = length / 2
        V += [c0[None], c1[None]]
        F += [np.stack([np.full(n, 2 * n), j, np.arange(n)], 1), np.stack([np.full(n, 2 * n + 1), n + np.arange(n), n + j], 1)]
    else:
        k = np.arange(1, n - 1)
        F += [np.stack([np.zeros(n - 2, int), k + 1, k], 1), np.stack([np.full(n - 2, n), n + k, n + k + 1], 1)]
    m = sg.Mesh(np.concatenate(V) + np.asarray(center, float), np.concatenate(F).astype(np.int64))
    # u x v = axis for (u, v) in cyclic order; flip when the pair is anti-cyclic
    if (u, v) == (0, 2):
        m.faces = m.faces[:, ::-1].copy()
    return m


def join(*meshes: sg.Mesh) -> sg.Mesh:
    V, F, off = [], [], 0
    for m in meshes:
        V.append(m.vertices)
        F.append(m.faces + off)
        off += len(m.vertices)
    return sg.Mesh(np.concatenate(V), np.concatenate(F))


# --- voxeliser ----------------------------------------------------------
def _midpoint_subdivide(tri: np.ndarray, max_edge: float) -> set:
    out = set()
    stack = [tri]
    while stack:
        a, b, c = stack.pop()
        if max(np.linalg.norm(b - a), np.linalg.norm(c - b), np.linalg.norm(a - c)) > max_edge:
            ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
            stack += [np.array(x) for x in ((a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca))]
        else:
            for p in (a, b, c):
                out.add(tuple(np.round(p, 9)))
    return out


def test_subdivision_lattice_equals_midpoint_recursion():
    rng = np.random.default_rng(3)
    for _ in range(6):
        tri = rng.uniform(-3, 3, (3, 3))
        max_edge = float(rng.uniform(0.4, 1.5))
        m = sg.Mesh(tri, np.array([[0, 1, 2]]))
        got = {tuple(np.round(p, 9)) for p in sg.surface_points(m, max_edge)}
        assert got == _midpoint_subdivide(tri, max_edge)


@pytest.mark.parametrize("hub", [True, False])
def test_voxel_volume_bounds(hub):
    m = cylinder((0.3, -0.2, 0.1), 2, radius=6.0, length=20.0, hub=hub)
    assert (m.edge_use() == 2).all() and m.volume() > 0
    pitch = 0.5
    occ = sg.voxelize(m, pitch)
    vox = occ.sum() * pitch ** 3
    assert m.volume() <= vox <= m.volume() + m.area() * pitch
    ext = m.bounds[1] - m.bounds[0]
    assert occ.shape == tuple(int(v) for v in np.ceil(ext / pitch) + 1)
    # solid: the cross-section through the middle has no holes
    mid = occ[:, :, occ.shape[2] // 2]
    assert mid[occ.shape[0] // 2, occ.shape[1] // 2]


# --- labels ---------------------------------------------------------------
def _brute_labels(occ: np.ndarray) -> np.ndarray:
    X, Y, Z = occ.shape
    lab = np.where(occ, 1, 0)

    def nbr(mask, x, y, z):
        for dx, dy, dz in itertools.product((-1, 0, 1), repeat=3):
            if (dx, dy, dz) == (0, 0, 0):
                continue
            a, b, c = x + dx, y + dy, z + dz
            if 0 <= a < X and 0 <= b < Y and 0 <= c < Z and mask[a, b, c]:
                return True
        return False

    shell = [(x, y, z) for x in range(X) for y in range(Y) for z in range(Z) if not occ[x, y, z] and nbr(occ, x, y, z)]
    for p in shell:
        lab[p] = 2
    for band, nxt in ((2, 3), (3, 4), (4, 5)):
        m = lab == band
        hits = [(x, y, z) for x in range(X) for y in range(Y) for z in range(Z) if lab[x, y, z] == 1 and nbr(m, x, y, z)]
        for p in hits:
            lab[p] = nxt
    return lab


def test_labels_equal_neighbour_scan():
    rng = np.random.default_rng(5)
    x, y, z = np.meshgrid(*(np.arange(14),) * 3, indexing="ij")
    occ = ((x - 6.5) ** 2 / 30 + (y - 7) ** 2 / 20 + (z - 6) ** 2 / 25 < 1) | (rng.random((14, 14, 14)) > 0.97)
    np.testing.assert_array_equal(sg.label_voxels(occ), _brute_labels(occ))


def test_end_planes_on_the_fluid_extent():
    occ = np.zeros((20, 24, 9), bool)
    occ[3:17, 5:19, 2:7] = True        # a box of fluid
    occ[9:11, 2:5, 4] = True           # a spur reaching y = 2
    lab = sg.label_end_planes(sg.label_voxels(occ))
    assert set(np.unique(lab[:, 2, :])) >= {sg.LABEL_Y_MIN}
    assert (lab[:, 2, :] == sg.LABEL_Y_MIN).sum() == 2
    # y-max: the box face, less the two x-plane edges that were tagged after it
    assert (lab[:, 18, :] == sg.LABEL_Y_MAX).sum() == 14 * 5 - 2 * 5
    assert (lab[:, 18, :] == sg.LABEL_X_MAX).sum() + (lab[:, 18, :] == sg.LABEL_X_MIN).sum() == 10
    assert (lab[3] == sg.LABEL_X_MIN).sum() == 14 * 5 and (lab[16] == sg.LABEL_X_MAX).sum() == 14 * 5
    assert not np.isin(lab[4:16, 6:18], (13, 14, 15, 16)).any()


# --- PE changes -------------------------------------------------------------
def _pe_input(ny0: int, nx0: int = 10, nz: int = 12, r: int = 3):
    """Pre-PE labels with a straight pulmonary tube along the first axis:
    square section of half-width r, end tags on both end planes."""
    lab = np.full((ny0, nx0, nz), 2, np.int32)
    xc, zc = nx0 // 2, nz // 2
    lab[:, xc - r:xc + r + 1, zc - r:zc + r + 1] = 1
    lab[0][lab[0] == 1] = sg.LABEL_X_MIN
    lab[-1][lab[-1] == 1] = sg.LABEL_X_MAX
    return lab, xc, zc


@pytest.mark.parametrize("ny0", [37, 64 - 2 * 10])
def test_pe_extension_and_padding(ny0):
    dx = 1e-3
    lab, xc, zc = _pe_input(ny0)
    pe = sg.pe_changes(lab, dx, left_m=10e-3, right_m=10e-3, outlet_radius_m=1.5e-3, block=32)
    n_left = 10
    rest = (ny0 + 20) % 32
    n_right = 10 + 32 - rest
    assert pe.shape == (lab.shape[1], ny0 + n_left + n_right, lab.shape[2])
    assert pe.shape[1] % 32 == 0
    # interior: the tube is fluid, end tags gone off the boundary planes
    np.testing.assert_array_equal(pe[:, n_left + 1:n_left + ny0 - 1, :] == 1, (lab[1:-1] != 2).transpose(1, 0, 2))
    # the outermost planes keep the tags inside the outlet radius
    x, z = np.meshgrid(np.arange(pe.shape[0]), np.arange(pe.shape[2]), indexing="ij")
    inside = ((x - xc) ** 2 + (z - zc) ** 2) * dx * dx <= 1.5e-3 ** 2
    np.testing.assert_array_equal(pe[:, 0, :] == sg.LABEL_X_MIN, inside)
    np.testing.assert_array_equal(pe[:, -1, :] == sg.LABEL_X_MAX, inside)
    # the radius shrinks monotonically from the end into the extension
    open_count = [(pe[:, y, :] != 2).sum() for y in range(n_left + 1)]
    assert all(a <= b for a, b in zip(open_count, open_count[1:]))
    # first extension cell: rfact (1 - tanh(-2)) / 2 of the way from the outlet radius to r_end
    r_end = 3 * dx
    rf = (1 - math.tanh(-2.0)) / 2
    r0 = 1.5e-3 + rf * (r_end - 1.5e-3)
    np.testing.assert_array_equal(pe[:, n_left - 1, :] != 2, ((x - xc) ** 2 + (z - zc) ** 2) * dx * dx <= r0 ** 2)
    assert not np.isin(pe, (0, 3, 4, 5)).any()


def test_near_wall_without_second_band_becomes_wall():
    lab = np.full((40, 9, 9), 2, np.int32)
    lab[:, 3:6, 3:6] = 3                # a thin tube: only near-wall cells
    lab[:, 4, 4] = 4                    # with a second-band core
    lab[0][lab[0] != 2] = sg.LABEL_X_MIN
    lab[-1][lab[-1] != 2] = sg.LABEL_X_MAX
    lab[10:20, 4, 4] = 3                # core missing over 10 planes
    pe = sg.pe_changes(lab, 1e-3, left_m=4e-3, right_m=4e-3, outlet_radius_m=5e-3, block=32)
    y0 = 4
    seg = pe[:, y0 + 12:y0 + 18, :]     # planes whose 26-neighbourhood has no 4
    assert (seg != 2).sum() == 0
    assert (pe[:, y0 + 2:y0 + 8, :] != 2).sum() == 6 * 9


# --- whole pipeline -----------------------------------------------------------
def _two_tubes(tmp_path: Path) -> Path:
    """Two capped cylinders: one along x (r = 6, x in [-30, 30]) and one along y
    (r = 9, y in [-45, 45]), offset in z so that they do not touch."""
    m = join(cylinder((0, 0, 0), 0, 6.0, 60.0, 128), cylinder((0, 0, 20.0), 1, 9.0, 90.0, 160, hub=False))
    p = tmp_path / "tubes.stl"
    sg.write_stl(p, m)
    return p


def _check_pipeline(res: sg.Result, cap_tol: float):
    geo, pre = res.geometry.labels, res.test.labels
    dx_mm = res.dx_m * 1e3
    assert geo.shape[1] % sg.PE_BLOCK_SIZE == 0
    assert set(np.unique(geo)) <= {0, 1, 3, 4, 5, 6}
    # the y-facing caps (ends 2, 3) are the new x planes, tagged 4 (y-min) and 3
    for end, plane in ((2, geo[0]), (3, geo[-1])):
        want = res.ends[end].area / dx_mm ** 2
        got = int((plane == (4 if end == 2 else 3)).sum())
        assert abs(got - want) < cap_tol * want, (end, got, want)
    assert (geo[:, 0] == 5).sum() > 0 and (geo[:, -1] == 6).sum() > 0
    # the outlets narrow to pi r_out^2
    want = math.pi * (sg.PE_OUTLET_RADIUS_M / res.dx_m) ** 2
    for v, plane in ((5, geo[:, 0]), (6, geo[:, -1])):
        assert abs((plane == v).sum() - want) < 0.3 * want + 4
    # volume bounds of the occupancy
    vox = res.occupancy.sum() * res.pitch ** 3
    assert res.mesh.volume() <= vox <= res.mesh.volume() + res.mesh.area() * res.pitch
    # z faces are walls in both files
    assert (geo[:, :, 0] == 1).all() and (geo[:, :, -1] == 1).all()
    assert (pre[:, :, 0] == 1).all() and (pre[:, :, -1] == 1).all()


def _labels_info(geom):
    kw = dict(storage=_abi.F32, nu_ref=1e-4)
    return lbm355.labels(geom.flat(), geom.dims, **kw)


def test_pipeline_on_two_tubes(tmp_path):
    res = sg.build(_two_tubes(tmp_path), longest_axis_voxels=160, log=lambda *a: None)
    assert [round(e.area, 3) for e in res.ends] == [pytest.approx(a, rel=2e-3) for a in
                                                      (math.pi * 36, math.pi * 36, math.pi * 81, math.pi * 81)]
    assert res.ends[0].centroid[0] < 0 < res.ends[1].centroid[0]
    assert res.ends[2].normal[1] < -0.99 and res.ends[3].normal[1] > 0.99
    _check_pipeline(res, cap_tol=0.15)


def _roundtrip(res: sg.Result, tmp_path: Path):
    write_wall_vtk(tmp_path / "geometry.vtk", res.geometry, coords="double")
    write_wall_vtk(tmp_path / "test.vtk", res.test, coords="double")
    back = read_wall_vtk(tmp_path / "geometry.vtk")
    np.testing.assert_array_equal(back.labels, res.geometry.labels)
    assert back.dl == pytest.approx(res.dx_m, rel=1e-12)
    pre = read_wall_vtk(tmp_path / "test.vtk")
    np.testing.assert_array_equal(pre.labels, res.test.labels)
    return back, pre


def _inflow_area(lab: np.ndarray) -> int:
    """x=0 inflow cells of the sim_tcpc layout when the plane x=1 is open: the
    open-boundary extrusion opens x=0 wherever x=1 is open
    (sim_tcpc.cu:549-586); the y edges are outlets, the z edges wall."""
    return int(((lab[0] != 1) | (lab[1] != 1))[1:-1, 1:-1].sum())


def test_pipeline_exterior_as_fluid_and_round_trip(tmp_path):
    res = sg.build(_two_tubes(tmp_path), longest_axis_voxels=160, log=lambda *a: None)
    geo, pre = _roundtrip(res, tmp_path)
    _, info = _labels_info(geo)
    _, info_pre = _labels_info(pre)
    assert info["area_ivc"] == _inflow_area(geo.labels)
    assert info["area_ivc"] >= int((geo.labels[0] == 4).sum()) > 0
    # test.vtk: exterior (0) cells on the x=0 plane are inflow cells too (P7)
    ext = int((pre.labels[0, 1:-1, 1:-1] == 0).sum())
    assert ext > 0
    assert info_pre["area_ivc"] == _inflow_area(pre.labels)
    assert info_pre["area_ivc"] >= info["area_ivc"] + ext - int((geo.labels[0] != 1).sum())


def test_command_line_writes_three_files(tmp_path):
    stl = _two_tubes(tmp_path)
    cfg = tmp_path / "sim.yaml"
    cfg.write_text("geometry:\n  input_stl: x.stl\n  runtime_vtk: rt/geometry.vtk\n  test_vtk: rt/test.vtk\n"
                   "  tapered_stl: rt/tapered.stl\n  longest_axis_voxels: 120\n  pe_extension_left_m: 0.025\n"
                   "  pe_extension_right_m: 0.025\n  pe_outlet_radius_m: 0.002\n  pe_block_size: 32\n")
    out = tmp_path / "o"
    rc = sg.main(["--config", str(cfg), "--stl", str(stl), "--output", str(out / "geometry.vtk"),
                  "--test-output", str(out / "test.vtk"), "--tapered-stl", str(out / "t.stl")])
    assert rc == 0
    g = read_wall_vtk(out / "geometry.vtk")
    assert g.dims[1] % 32 == 0
    t = sg.read_stl(out / "t.stl")
    assert (t.edge_use() == 2).all()
    # the extended surface is 10 mm longer on ends 0, 1 (x) and 2 (y-min)
    b0, b1 = sg.read_stl(stl).bounds, t.bounds
    np.testing.assert_allclose(b1[0, :2] - b0[0, :2], [-10, -10], atol=1e-4)
    np.testing.assert_allclose(b1[1, :2] - b0[1, :2], [10, 0], atol=1e-4)
    assert sg.main(["--config", str(cfg), "--stl", str(tmp_path / "missing.stl")]) == 1


@pytest.mark.skipif(not REF_STL.exists(), reason="TCPC25_03.stl of the reference checkout is not present")
def test_tcpc25_03(tmp_path):
    res = sg.build(REF_STL, log=lambda *a: None)
    assert res.geometry.dims == (520, 448, 80)
    _check_pipeline(res, cap_tol=0.10)
    # the caps are the extremes of the surface
    b = sg.read_stl(REF_STL).bounds
    for e, (axis, side) in zip(res.ends, ((0, 0), (0, 1), (1, 0), (1, 1))):
        assert e.centroid[axis] == pytest.approx(b[side, axis], abs=1e-3)
    geo, pre = _roundtrip(res, tmp_path)
    _, info = _labels_info(geo)
    assert info["terminate"] == 0
    assert info["area_ivc"] == _inflow_area(geo.labels)
    assert info["area_svc"] == _inflow_area(geo.labels[::-1])
    _, info_pre = _labels_info(pre)
    assert info_pre["area_ivc"] > info["area_ivc"]                  # exterior inflow cells (P7)
    # the solver reads the same labels from the file (host-only part of the run)
    exe = ROOT / "nonnewlbm_amd" / "bin" / "sim_tcpc"
    r = subprocess.run([str(exe), "--geometry", str(tmp_path / "geometry.vtk"), "--material", "newtonian", "--case-id",
                        "stl", "--ivc-mls", "20", "--svc-mls", "15", "--steps", "1", "--vtk-period", "-1",
                        "--print-period", "-1"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode in (0, 1), r.stderr
    log = (tmp_path / "results_stl" / "log_rank000").read_text()
    assert (f"IVC(label3)={info['area_ivc']} SVC(label4)={info['area_svc']} outlet5={info['area_out5']} "
            f"outlet6={info['area_out6']}") in log
