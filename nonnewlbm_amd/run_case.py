"""Run or print one TCPC case: the run_case.py front end (scripts/run_case.py)
over this package's sim_tcpc executable.

``build_command`` maps (regime, material) and the solver defaults of
configs/simulation.yaml to the sim_tcpc argv exactly as
scripts/run_case.py:19-60 does; ``main`` prints the shell-quoted command and,
unless ``--dry-run``, runs it from the repository root and returns its exit
code (scripts/run_case.py:95).
"""
from __future__ import annotations

import argparse
import shlex
import subprocess
from pathlib import Path

import yaml

PKG = Path(__file__).resolve().parent
CONFIGS = PKG / "configs"


def repo_root() -> Path:
    return PKG.parent


def load_yaml(path: Path) -> dict:
    with path.open("r", encoding="utf-8") as f:
        return yaml.safe_load(f)


def build_command(root: Path, regime: str, material: str, case_id: str | None = None, *,
                  geometry: str | None = None, final_time: float | None = None, vtk_period: float | None = None,
                  print_period: float | None = None, stat_reset_period: float | None = None,
                  lbm_viscosity: float | None = None, block_size: int | None = None,
                  extra: list[str] | None = None) -> list[str]:
    sim_cfg = load_yaml(CONFIGS / "simulation.yaml")
    regimes = load_yaml(CONFIGS / "regimes.yaml")["regimes"]
    if regime not in regimes:
        raise ValueError(f"Unknown regime '{regime}'. Expected one of {sorted(regimes)}")
    if material not in sim_cfg["cases"]["materials"]:
        raise ValueError(f"Unknown material '{material}'. Expected one of {sim_cfg['cases']['materials']}")
    solver = sim_cfg["solver"]

    def pick(v, key):
        return v if v is not None else solver[key]

    flow = regimes[regime]
    cmd = [
        str(root / solver["executable"]),
        "--geometry", str(root / (geometry or sim_cfg["geometry"]["runtime_vtk"])),
        "--material", material,
        "--case-id", case_id or f"{regime}_{material}",
        "--ivc-mls", str(flow["ivc_mls"]),
        "--svc-mls", str(flow["svc_mls"]),
        "--final-time", str(pick(final_time, "final_time_s")),
        "--vtk-period", str(pick(vtk_period, "vtk_period_s")),
        "--print-period", str(pick(print_period, "print_period_s")),
        "--stat-reset-period", str(pick(stat_reset_period, "stat_reset_period_s")),
        "--lbm-viscosity", str(pick(lbm_viscosity, "lbm_viscosity")),
        "--block-size", str(pick(block_size, "block_size")),
    ]
    return cmd + list(extra or [])


def main(argv: list[str] | None = None) -> int:
    p = argparse.ArgumentParser(description="Run or print one TCPC non-Newtonian case.")
    p.add_argument("--regime", required=True)
    p.add_argument("--material", required=True, choices=["newtonian", "sx", "gx"])
    p.add_argument("--case-id", default=None)
    p.add_argument("--geometry", default=None)
    p.add_argument("--final-time", type=float, default=None)
    p.add_argument("--vtk-period", type=float, default=None)
    p.add_argument("--print-period", type=float, default=None)
    p.add_argument("--stat-reset-period", type=float, default=None)
    p.add_argument("--lbm-viscosity", type=float, default=None)
    p.add_argument("--block-size", type=int, default=None)
    p.add_argument("--dry-run", action="store_true")
    args, extra = p.parse_known_args(argv)
    root = repo_root()
    cmd = build_command(root, args.regime, args.material, args.case_id, geometry=args.geometry,
                        final_time=args.final_time, vtk_period=args.vtk_period, print_period=args.print_period,
                        stat_reset_period=args.stat_reset_period, lbm_viscosity=args.lbm_viscosity,
                        block_size=args.block_size, extra=extra)
    print(" ".join(shlex.quote(part) for part in cmd), flush=True)
    if args.dry_run:
        return 0
    return subprocess.call(cmd, cwd=root)


if __name__ == "__main__":
    raise SystemExit(main())
