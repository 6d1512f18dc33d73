"""run_all on the GPU: the pool mode's flow summary (device-side plane fluxes)
equals the same cases stepped alone, and the process mode runs sim_tcpc per
case and harvests the flows of each case's last VTK snapshot."""
from __future__ import annotations

import csv
import glob
import math
from pathlib import Path

import pytest

from nonnewlbm_amd import _abi, flows, run_all
from nonnewlbm_amd.geometry import Geometry, read_wall_vtk, tcpc_mask, write_wall_vtk
from nonnewlbm_amd.lbm355 import Case, labels
from nonnewlbm_amd.materials import case_kwargs, material_by_name
from nonnewlbm_amd.run_case import study

pytestmark = pytest.mark.gpu


def _geometry(tmp_path: Path) -> Path:
    """The 128x96x64 junction at 1 mm spacing (vessel radius 10 mm): with
    nu_lbm = 1e-3 the prescribed caval flows of v1/v2 are lattice velocities
    of 0.01 to 0.05."""
    g = tcpc_mask(0.5)
    path = tmp_path / "geometry.vtk"
    write_wall_vtk(path, Geometry(g.labels, 1.0e-3, g.meta))
    return path


def _summary(tmp_path: Path) -> list[dict]:
    files = glob.glob(str(tmp_path / "runs" / "run-*" / "flow_summary.csv"))
    assert len(files) == 1, files
    with open(files[0], newline="", encoding="utf-8") as f:
        return list(csv.DictReader(f))


def test_run_all_pool_summary_equals_cases_alone(gpu_lib, tmp_path):
    geo = _geometry(tmp_path)
    rc = run_all.main(["--pool", "--geometry", str(geo), "--regimes", "v1", "v2", "--materials", "newtonian", "sx",
                       "--final-time", "0.02", "--lbm-viscosity", "1e-3", "--runs-root", "runs",
                       "--workdir", str(tmp_path)])
    assert rc == 0
    rows = _summary(tmp_path)
    assert [r["case_id"] for r in rows] == ["v1_newtonian", "v1_sx", "v2_newtonian", "v2_sx"]
    g = read_wall_vtk(geo)
    regimes = study().regimes
    for r in rows:
        assert r["state"] == "COMPLETED" and r["harvest_status"] == "ok" and r["source"] == "device_flux"
        kw = case_kwargs(material_by_name(r["material"]), g.dl, 1e-3)
        dt = kw["dt"]
        flow = regimes[r["regime"]]
        kw.update(storage=_abi.F32, collision=_abi.CUM, profile=_abi.PROFILE_PLUG, stats_every=1,
                  ivc_flux=flow["ivc_mls"] * 1e-6 / g.dl ** 3 * dt, svc_flux=flow["svc_mls"] * 1e-6 / g.dl ** 3 * dt)
        steps = run_all.steps_for(0.02, dt)
        assert int(r["steps"]) == steps
        c = Case(g.flat(), g.dims, **kw)
        c.step(steps)
        _, info = labels(g.flat(), g.dims, **kw)
        ref = run_all.flows_row(c.flux(mean=False), info, kw["dl"], dt)
        c.close()
        for k, v in ref.items():
            assert math.isfinite(v) and float(r[k]) == v, (r["case_id"], k, r[k], v)
        assert float(r["IVC_l_min"]) > 0 and float(r["lpa_split_percent"]) > 0


def test_run_all_processes_and_harvest(gpu_lib, tmp_path):
    geo = _geometry(tmp_path)
    rc = run_all.main(["--geometry", str(geo), "--regimes", "v1", "--materials", "newtonian", "--final-time", "0.02",
                       "--vtk-period", "0.01", "--print-period", "0.01", "--lbm-viscosity", "1e-3", "--harvest",
                       "--runs-root", "runs", "--workdir", str(tmp_path)])
    assert rc == 0
    (row,) = _summary(tmp_path)
    assert row["case_id"] == "v1_newtonian" and row["state"] == "COMPLETED" and row["harvest_status"] == "ok"
    vtks = sorted(glob.glob(str(tmp_path / "results_v1_newtonian" / "vtk3D" / "rank000_data_*.vtk")),
                  key=flows.snapshot_order)
    assert int(row["snapshot_count"]) == len(vtks) >= 2
    ref = flows.compute_file(Path(vtks[-1]))
    for k in ("IVC_l_min", "SVC_l_min", "RPA_l_min", "LPA_l_min", "mass_balance_l_min", "lpa_split_percent"):
        assert math.isfinite(ref[k]) and float(row[k]) == pytest.approx(ref[k], rel=1e-12), k
    # the prescribed inflow reaches the IVC plane: 8 ml/s = 0.48 l/min
    assert float(row["IVC_l_min"]) == pytest.approx(0.48, rel=0.05)
    assert (tmp_path / "results_v1_newtonian" / "flows.csv").exists()
