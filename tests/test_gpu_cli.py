"""sim_tcpc end to end on the GPU: the executable's fields equal an in-process
Case run on the same descriptor, its output tree and counters follow
execute() (core.h:1476-1628), the flow post-processor recovers the prescribed
inflows, and a zero-area geometry exits with code 2."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from nonnewlbm_amd import _abi, flows
from nonnewlbm_amd.build import EXE
from nonnewlbm_amd.geometry import read_output_vtk, tcpc_mask, write_wall_vtk
from nonnewlbm_amd.lbm355 import Case
from nonnewlbm_amd.materials import lattice_units, material_by_name

pytestmark = pytest.mark.gpu


def _run(args, cwd):
    print("sim_tcpc", " ".join(args), flush=True)
    return subprocess.run([str(EXE), *args], cwd=cwd, capture_output=True, text=True, timeout=300)


def test_cli_matches_library_and_writes_outputs(gpu_lib, tmp_path):
    g = tcpc_mask(0.25)
    write_wall_vtk(tmp_path / "geometry.vtk", g)
    mat, nu_lbm, ivc, svc = "sx", 0.01, 0.5, 0.3
    u = lattice_units(material_by_name(mat), np.float32(g.dl).item(), nu_lbm)
    steps = 60
    # vtk every 20 steps -> files 0 (t=0), 1, 2, 3
    args = ["--material", mat, "--case-id", "c1", "--ivc-mls", str(ivc), "--svc-mls", str(svc),
            "--lbm-viscosity", str(nu_lbm), "--steps", str(steps), "--vtk-period", str(20 * u.dt * 0.999999),
            "--print-period", str(30 * u.dt), "--precision", "fp32"]
    r = _run(args, tmp_path)
    assert r.returncode == 0, r.stderr + r.stdout
    res = tmp_path / "results_c1"
    log = (res / "log_rank000").read_text()
    assert "START: simulation LBM:CUM dimensions 64 x 48 x 32" in log
    assert "requested steps reached" in log and "MLUPS=" in log
    files = sorted((res / "vtk3D").glob("rank000_data_*.vtk"))
    assert [f.name for f in files] == [f"rank000_data_{i}.vtk" for i in range(4)]
    out = read_output_vtk(res / "vtk3D" / "rank000_data_3.vtk")
    assert out["field"]["CYCLE"][0] == 3
    # the same case through the library
    dl = float(np.float32(g.dl))
    dt = u.dt
    kw = dict(dl=dl, dt=dt, collision=_abi.CUM, storage=_abi.F32, nu_ref=nu_lbm, cy_enabled=1, cy_model=_abi.CY_NEQ,
              cy_nu0=u.nu0, cy_lambda=u.lam, cy_a=u.a, cy_n=u.n, profile=_abi.PROFILE_PLUG,
              ivc_flux=ivc * 1e-6 / (dl * dl * dl) * dt, svc_flux=svc * 1e-6 / (dl * dl * dl) * dt)
    c = Case(g.flat(), g.dims, **kw)
    c.step(steps)
    lab = c.field("label")
    vconv = dl / dt
    for k, name in enumerate(("vx", "vy", "vz")):
        ref = (c.field(name) * vconv).astype(np.float32)
        np.testing.assert_array_equal(out["velocity"][..., k], ref, err_msg=name)
    np.testing.assert_array_equal(out["wall"], lab.astype(np.float32))
    # compute_flows on the output: plug inflows are exact up to float32 output
    row = flows.compute_file(res / "vtk3D" / "rank000_data_3.vtk")
    assert row["IVC_l_min"] == pytest.approx(ivc * 0.06, rel=1e-5)
    assert row["SVC_l_min"] == pytest.approx(svc * 0.06, rel=1e-5)
    assert row["IVC_cells"] == c.info()["area_ivc"]


def test_cli_refuses_half_precision_it_cannot_serve(gpu_lib, tmp_path):
    """--precision bf16 always, and --precision fp16 below --lbm-viscosity
    1e-2 (the sim_tcpc studies run 1e-4), exit 1 with the reason; fp16 at
    1e-2 runs."""
    g = tcpc_mask(0.25)
    write_wall_vtk(tmp_path / "geometry.vtk", g)
    common = ["--material", "newtonian", "--ivc-mls", "0.4", "--svc-mls", "0.2", "--steps", "4", "--print-period", "-1"]
    r = _run(common + ["--case-id", "b", "--lbm-viscosity", "1e-2", "--precision", "bf16"], tmp_path)
    assert r.returncode == 1 and "bf16 is not admitted" in r.stderr + r.stdout
    r = _run(common + ["--case-id", "h", "--lbm-viscosity", "1e-4", "--precision", "fp16"], tmp_path)
    assert r.returncode == 1 and "fp16 is not admitted below --lbm-viscosity 1e-2" in r.stderr + r.stdout
    r = _run(common + ["--case-id", "h2", "--lbm-viscosity", "1e-2", "--precision", "fp16"], tmp_path)
    assert r.returncode == 0, r.stderr + r.stdout


def test_cli_zero_area_terminates(gpu_lib, tmp_path):
    g = tcpc_mask(0.25)
    g.labels[:] = 1
    write_wall_vtk(tmp_path / "geometry.vtk", g)
    r = _run(["--case-id", "z", "--steps", "10", "--vtk-period", "1e9"], tmp_path)
    assert r.returncode == 2, r.stderr + r.stdout
    log = (tmp_path / "results_z" / "log_rank000").read_text()
    assert "error: zero inlet area detected: left=0 right=0" in log
    assert "terminate flag triggered" in log
    assert (tmp_path / "results_z" / "vtk3D" / "rank000_data_0.vtk").exists()


def test_cli_restart_through_loadstate(gpu_lib, tmp_path):
    """--save-state-at-end writes results_<id>/current_state and the loadstate
    flag; the next run with the same id continues from there (core.h:1511-1517)
    and its output equals an uninterrupted run."""
    g = tcpc_mask(0.25)
    write_wall_vtk(tmp_path / "geometry.vtk", g)
    u = lattice_units(material_by_name("newtonian"), float(np.float32(g.dl)), 1e-4)
    common = ["--material", "newtonian", "--ivc-mls", "0.4", "--svc-mls", "0.2", "--lbm-viscosity", "1e-4",
              "--vtk-period", str(20 * u.dt * 0.999999), "--print-period", "-1"]
    r = _run(common + ["--case-id", "split", "--steps", "40", "--save-state-at-end"], tmp_path)
    assert r.returncode == 0, r.stderr + r.stdout
    assert (tmp_path / "results_split" / "loadstate").exists()
    assert (tmp_path / "results_split" / "current_state" / "df_0_rank000").exists()
    r = _run(common + ["--case-id", "split", "--steps", "60"], tmp_path)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "[loadState invoked]" in (tmp_path / "results_split" / "log_rank000").read_text()
    r = _run(common + ["--case-id", "whole", "--steps", "60"], tmp_path)
    assert r.returncode == 0, r.stderr + r.stdout
    names = sorted(p.name for p in (tmp_path / "results_split" / "vtk3D").glob("*.vtk"))
    assert names == [f"rank000_data_{i}.vtk" for i in range(4)]
    a = read_output_vtk(tmp_path / "results_split" / "vtk3D" / "rank000_data_3.vtk")
    b = read_output_vtk(tmp_path / "results_whole" / "vtk3D" / "rank000_data_3.vtk")
    for name in ("velocity", "lbm_rho", "mean_velocity"):
        np.testing.assert_allclose(a[name], b[name], rtol=1e-5, atol=1e-12, err_msg=name)
    np.testing.assert_array_equal(a["velocity"], b["velocity"])
