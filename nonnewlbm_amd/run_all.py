"""Run every configured TCPC case and collect the flow summary.

This module stands in for two reference scripts on one node:
- the run_all_cases.py driver (scripts/run_all_cases.py:13-52);
- the harvest step of slurm_run_cases.py (scripts/slurm_run_cases.py:421-483).

Modes
  default    one sim_tcpc process per (regime, material), in the order of
             configs/simulation.yaml, printing each command and stopping at
             the first non-zero exit code (run_all_cases.py:34-52).
  --harvest  also write <runs-root>/run-<timestamp>/flow_summary.csv after the
             runs. Each case contributes its row of the flows of its last VTK
             snapshot, in the row layout of slurm_run_cases.py:444-483
             (case_id, regime, material, state, the flows columns, flow_csv,
             snapshot_count, harvest_status).
  --pool     every case at once in one serving pool on this GPU, with no VTK
             output. The summary rows come from the device-side plane fluxes
             of each case's final step (the same cells, signs and l/min
             conversion as the flows module).

A case runs while step * dt <= final time, as the sim_tcpc loop does
(core.h:1476-1628): the first step index k with k * dt > final time ends it.
"""
from __future__ import annotations

import argparse
import csv
import dataclasses
import math
import shlex
import subprocess
import sys
import time
from pathlib import Path

from .run_case import build_command, repo_root, study

LABELS = (("IVC", "area_ivc"), ("SVC", "area_svc"), ("RPA", "area_out5"), ("LPA", "area_out6"))


def run_id() -> str:
    return time.strftime("run-%Y%m%d-%H%M%S")


def case_matrix(regimes: list[str] | None, materials: list[str] | None) -> list[tuple[str, str]]:
    """(regime, material) pairs, regime-major, in the configured order."""
    s = study()
    return [(r, m) for r in (regimes or s.regime_order) for m in (materials or s.materials)]


def steps_for(final_time: float, dt: float) -> int:
    """Number of steps the sim_tcpc loop runs: the first k with k * dt > T."""
    return int(math.floor(final_time / dt)) + 1


def flows_row(flux_lattice, cells: dict, dl: float, dt: float) -> dict:
    """Flows columns (l/min, cell counts, balance, splits) from the four
    device plane fluxes (IVC, SVC inflow; label-5 and label-6 outflow, in
    lattice units), through the flows module's own row builder."""
    from . import flows
    to_m3s = dl ** 3 / dt
    caps = [flows.Cap(name, 0, int(cells.get(key, 0)), 0, 1.0, 0.0) for name, key in LABELS]
    # a Cap's reported value is its negated outward flux for inlets: give
    # the inlets their outward (negative) flux so that they report inflow
    caps = [dataclasses.replace(c, flux=(-1.0 if c.name in flows.INLETS else 1.0) * float(f) * to_m3s)
            for c, f in zip(caps, flux_lattice)]
    return flows.summarise(caps)


class SummaryTable:
    """Rows of flow_summary.csv; the columns are the union of the rows' keys
    in first-seen order (cases that failed carry fewer columns)."""

    def __init__(self):
        self.rows: list[dict] = []

    def add(self, row: dict) -> None:
        self.rows.append(row)

    def columns(self) -> list[str]:
        return list(dict.fromkeys(key for row in self.rows for key in row))

    def write(self, path: Path) -> Path:
        path.parent.mkdir(parents=True, exist_ok=True)
        cols = self.columns()
        with path.open("w", newline="", encoding="utf-8") as fh:
            out = csv.writer(fh)
            out.writerow(cols)
            out.writerows([[row.get(c, "") for c in cols] for row in self.rows])
        return path


def snapshot_flows(result_dir: Path) -> dict:
    """Flows of every VTK snapshot of one case, written to
    <result_dir>/flows.csv (as the post-processing step of a finished job,
    slurm_run_cases.py:421-441); returns the summary columns of the last
    snapshot with the CSV path and the snapshot count."""
    from . import flows
    snaps = flows.expand_inputs([str(result_dir / "vtk3D" / "rank000_data_*.vtk")])
    if not snaps or not snaps[0].exists():
        raise FileNotFoundError(f"{result_dir}: no vtk3D/rank000_data_*.vtk snapshots")
    per_file = [flows.compute_file(p) for p in snaps]
    csv_path = result_dir / "flows.csv"
    flows.write_rows(str(csv_path), per_file)
    last = dict(per_file[-1])
    last.pop("file")
    return {**last, "flow_csv": str(csv_path), "snapshot_count": str(len(snaps))}


def run_processes(args, root: Path) -> int:
    work = Path(args.workdir).resolve() if args.workdir else root
    table = SummaryTable()
    rc_all = 0
    for regime, material in case_matrix(args.regimes, args.materials):
        case_id = f"{regime}_{material}"
        cmd = build_command(root, regime, material, case_id, geometry=args.geometry, final_time=args.final_time,
                            vtk_period=args.vtk_period, print_period=args.print_period,
                            stat_reset_period=args.stat_reset_period, lbm_viscosity=args.lbm_viscosity,
                            block_size=args.block_size, extra=args.extra)
        print(" ".join(shlex.quote(p) for p in cmd), flush=True)
        if args.dry_run:
            continue
        rc = subprocess.call(cmd, cwd=work)
        row = {"case_id": case_id, "regime": regime, "material": material,
               "state": "COMPLETED" if rc == 0 else f"FAILED({rc})"}
        if args.harvest and rc != 0:
            row["harvest_status"] = "skipped_non_completed"
        elif args.harvest:
            try:
                row.update(snapshot_flows(work / f"results_{case_id}"), harvest_status="ok")
            except (OSError, RuntimeError, ValueError) as exc:  # recorded per case in the summary
                row.update(harvest_status="failed", harvest_error=str(exc))
        table.add(row)
        if rc != 0:
            rc_all = rc
            break
    if args.harvest and table.rows:
        summary = table.write(work / args.runs_root / run_id() / "flow_summary.csv")
        print(f"[harvest] summary: {summary}", flush=True)
    return rc_all


def run_pool(args, root: Path) -> int:
    """All cases in one serving pool; the summary from device-side fluxes."""
    from . import _abi
    from .geometry import read_wall_vtk
    from .lbm355 import Pool, labels
    from .materials import case_kwargs, material_by_name

    s = study()
    geo_path = root / (args.geometry or s.geometry)
    g = read_wall_vtk(geo_path)
    flat = g.flat()
    final_time = args.final_time if args.final_time is not None else s.solver["final_time_s"]
    nu_lbm = args.lbm_viscosity if args.lbm_viscosity is not None else s.solver["lbm_viscosity"]
    jobs = []
    for regime, material in case_matrix(args.regimes, args.materials):
        mat = material_by_name(material)
        kw = case_kwargs(mat, g.dl, nu_lbm)
        dt = kw["dt"]
        flow = s.regimes[regime]
        kw.update(storage=_abi.STORAGES[args.precision], collision=_abi.COLLISIONS[args.collision],
                  profile=_abi.PROFILE_PLUG, stats_every=1,
                  ivc_flux=flow["ivc_mls"] * 1e-6 / g.dl ** 3 * dt, svc_flux=flow["svc_mls"] * 1e-6 / g.dl ** 3 * dt)
        _, info = labels(flat, g.dims, **kw)
        jobs.append(dict(case_id=f"{regime}_{material}", regime=regime, material=material, kw=kw, info=info,
                         steps=steps_for(final_time, dt)))
        print(f"[pool] {jobs[-1]['case_id']}: {jobs[-1]['steps']} steps, dt={dt:.6e} s", flush=True)
    if args.dry_run:
        return 0
    pool = Pool(budget_bytes=args.budget, steps_per_tick=64, graph=True)
    t0 = time.perf_counter()
    for j in jobs:
        j["id"] = pool.submit(flat, g.dims, j["steps"], report_every=0, **j["kw"])
    while not pool.drain(timeout=30.0):
        st = pool.stats()
        print(f"[pool] {st['done']} done, {st['running']} running, {st['queued']} queued, "
              f"{time.perf_counter() - t0:.0f} s", flush=True)
    table = SummaryTable()
    failures = 0
    for j in jobs:
        st = pool.poll(j["id"])
        row = {"case_id": j["case_id"], "regime": j["regime"], "material": j["material"],
               "state": "COMPLETED" if st["state"] == _abi.REQ_DONE else "FAILED", "steps": st["steps_done"]}
        if st["state"] == _abi.REQ_DONE:
            row.update(flows_row(st["flux"], j["info"], j["kw"]["dl"], j["kw"]["dt"]))
            row.update(source="device_flux", harvest_status="ok")
        else:
            row.update(harvest_status="failed", harvest_error=st["message"])
            failures += 1
        table.add(row)
    info = pool.stats()
    pool.close()
    work = Path(args.workdir).resolve() if args.workdir else root
    summary = table.write(work / args.runs_root / run_id() / "flow_summary.csv")
    print(f"[pool] {info['cell_updates'] / max(info['busy_seconds'], 1e-9) / 1e6:.0f} fluid MLUPS while stepping; "
          f"summary: {summary}", flush=True)
    return 1 if failures else 0


def main(argv: list[str] | None = None) -> int:
    p = argparse.ArgumentParser(description="Run or print all configured TCPC cases.")
    p.add_argument("--dry-run", action="store_true")
    p.add_argument("--materials", nargs="*", default=None)
    p.add_argument("--regimes", nargs="*", default=None)
    p.add_argument("--geometry", default=None)
    p.add_argument("--final-time", type=float, default=None)
    p.add_argument("--vtk-period", type=float, default=None)
    p.add_argument("--print-period", type=float, default=None)
    p.add_argument("--stat-reset-period", type=float, default=None)
    p.add_argument("--lbm-viscosity", type=float, default=None)
    p.add_argument("--block-size", type=int, default=None)
    p.add_argument("--harvest", action="store_true", help="write flow_summary.csv after the runs")
    p.add_argument("--pool", action="store_true", help="all cases at once in one serving pool on this GPU")
    p.add_argument("--runs-root", default="runtime/pool", help="summary directory, relative to the work directory")
    p.add_argument("--workdir", default=None, help="directory the solver runs in (default: the repository root)")
    p.add_argument("--budget", type=int, default=0, help="pool device-memory budget in bytes (0: 90%% of free)")
    p.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "fp16"])
    p.add_argument("--collision", default="cumulant", choices=["cumulant", "bgk", "mrt", "kbc"])
    args, extra = p.parse_known_args(argv)
    args.extra = extra
    root = repo_root()
    return run_pool(args, root) if args.pool else run_processes(args, root)


if __name__ == "__main__":
    sys.exit(main())
