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
import glob
import math
import shlex
import subprocess
import sys
import time
from pathlib import Path

from .run_case import CONFIGS, build_command, load_yaml, repo_root

LABELS = (("IVC", "area_ivc"), ("SVC", "area_svc"), ("RPA", "area_out5"), ("LPA", "area_out6"))


def run_id() -> str:
    return time.strftime("run-%Y%m%d-%H%M%S")


def case_matrix(regimes: list[str] | None, materials: list[str] | None) -> list[tuple[str, str]]:
    sim_cfg = load_yaml(CONFIGS / "simulation.yaml")
    regs = regimes or sim_cfg["cases"]["regimes"]
    mats = materials or sim_cfg["cases"]["materials"]
    return [(r, m) for r in regs for m in mats]


def steps_for(final_time: float, dt: float) -> int:
    """Number of steps the sim_tcpc loop runs: the first k with k * dt > T."""
    return int(math.floor(final_time / dt)) + 1


def flows_row(flux_lattice, cells: dict, dl: float, dt: float) -> dict:
    """Flows columns (l/min, cell counts, balance, splits) from the four
    device plane fluxes (IVC, SVC inflow; label-5 and label-6 outflow, in
    lattice units); the same arithmetic as flows.compute_arrays."""
    to_lmin = dl ** 3 / dt * 6.0e4
    row: dict = {}
    vals = {}
    for (name, key), f in zip(LABELS, flux_lattice):
        vals[name] = float(f) * to_lmin
        row[f"{name}_l_min"] = abs(vals[name])
        row[f"{name}_cells"] = int(cells.get(key, 0))
    inlet = abs(vals["IVC"]) + abs(vals["SVC"])
    outlet = abs(vals["RPA"]) + abs(vals["LPA"])
    row["mass_balance_l_min"] = inlet - outlet
    row["lpa_split_percent"] = 100.0 * abs(vals["LPA"]) / outlet if outlet > 0 else 0.0
    row["rpa_split_percent"] = 100.0 * abs(vals["RPA"]) / outlet if outlet > 0 else 0.0
    return row


def write_summary(path: Path, rows: list[dict]) -> None:
    fieldnames: list[str] = []
    for row in rows:
        for key in row:
            if key not in fieldnames:
                fieldnames.append(key)
    path.parent.mkdir(parents=True, exist_ok=True)
    with path.open("w", newline="", encoding="utf-8") as f:
        w = csv.DictWriter(f, fieldnames=fieldnames)
        w.writeheader()
        w.writerows(rows)


def harvest_case(root: Path, case_id: str) -> dict:
    """Flows of the last VTK snapshot of results_<case_id> (run_postprocess,
    slurm_run_cases.py:421-441): writes results_<case_id>/flows.csv."""
    from . import flows
    pattern = root / f"results_{case_id}" / "vtk3D" / "rank000_data_*.vtk"
    files = sorted(glob.glob(str(pattern)), key=flows.natural_key)
    if not files:
        raise FileNotFoundError(f"No VTK files found for {case_id}: {pattern}")
    out = root / f"results_{case_id}" / "flows.csv"
    rows = [flows.compute_file(Path(f)) for f in files]
    write_summary(out, rows)
    last = {k: v for k, v in rows[-1].items() if k != "file"}
    last.update({"flow_csv": str(out), "snapshot_count": str(len(files))})
    return last


def run_processes(args, root: Path) -> int:
    work = Path(args.workdir).resolve() if args.workdir else root
    rows = []
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
        if rc == 0 and args.harvest:
            try:
                row.update(harvest_case(work, case_id))
                row["harvest_status"] = "ok"
            except Exception as exc:  # the summary records the failure per case
                row.update(harvest_status="failed", harvest_error=str(exc))
        elif args.harvest:
            row["harvest_status"] = "skipped_non_completed"
        rows.append(row)
        if rc != 0:
            rc_all = rc
            break
    if args.harvest and rows:
        summary = work / args.runs_root / run_id() / "flow_summary.csv"
        write_summary(summary, rows)
        print(f"[harvest] summary: {summary}", flush=True)
    return rc_all


def run_pool(args, root: Path) -> int:
    """All cases in one serving pool; the summary from device-side fluxes."""
    from . import _abi
    from .geometry import read_wall_vtk
    from .lbm355 import Pool, labels
    from .materials import case_kwargs, material_by_name

    sim_cfg = load_yaml(CONFIGS / "simulation.yaml")
    regimes = load_yaml(CONFIGS / "regimes.yaml")["regimes"]
    solver = sim_cfg["solver"]
    geo_path = root / (args.geometry or sim_cfg["geometry"]["runtime_vtk"])
    g = read_wall_vtk(geo_path)
    flat = g.flat()
    final_time = args.final_time if args.final_time is not None else solver["final_time_s"]
    nu_lbm = args.lbm_viscosity if args.lbm_viscosity is not None else solver["lbm_viscosity"]
    jobs = []
    for regime, material in case_matrix(args.regimes, args.materials):
        mat = material_by_name(material)
        kw = case_kwargs(mat, g.dl, nu_lbm)
        dt = kw["dt"]
        flow = regimes[regime]
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
    rows = []
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
        rows.append(row)
    info = pool.stats()
    pool.close()
    work = Path(args.workdir).resolve() if args.workdir else root
    summary = work / args.runs_root / run_id() / "flow_summary.csv"
    write_summary(summary, rows)
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
    p.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "fp16", "bf16"])
    p.add_argument("--collision", default="cumulant", choices=["cumulant", "bgk", "mrt", "kbc"])
    args, extra = p.parse_known_args(argv)
    args.extra = extra
    root = repo_root()
    return run_pool(args, root) if args.pool else run_processes(args, root)


if __name__ == "__main__":
    sys.exit(main())
