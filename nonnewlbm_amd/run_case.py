"""One TCPC case on the command line: (regime, material) -> sim_tcpc argv.

Stands in for scripts/run_case.py.  Its contract is kept: the same options,
the argv it emits (same flags, values and order, scripts/run_case.py:46-60,
so that logs and job scripts that parse it keep working), printing the
shell-quoted command and running it from the repository root with the
solver's exit code (scripts/run_case.py:95).

Every sim_tcpc flag that the study sets comes from one table entry: the
regime table (caval flows) or the solver defaults of configs/simulation.yaml,
each overridable by the run_case option of the same name.  Unknown options
are passed through to sim_tcpc after the generated ones (this package's
extensions, e.g. --collision or --precision).
"""
from __future__ import annotations

import argparse
import functools
import shlex
import subprocess
import sys
from pathlib import Path

import yaml

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CONFIGS = PKG / "configs"

# sim_tcpc flag -> YAML key of its value in the regime entry
REGIME_FLAGS = (("--ivc-mls", "ivc_mls"), ("--svc-mls", "svc_mls"))
# sim_tcpc flag -> (key in the solver section, type of the override option)
SOLVER_FLAGS = (("--final-time", "final_time_s", float), ("--vtk-period", "vtk_period_s", float),
                ("--print-period", "print_period_s", float), ("--stat-reset-period", "stat_reset_period_s", float),
                ("--lbm-viscosity", "lbm_viscosity", float), ("--block-size", "block_size", int))


def repo_root() -> Path:
    return ROOT


def _read(path: Path) -> dict:
    return yaml.safe_load(path.read_text(encoding="utf-8"))


class Study:
    """The case matrix and solver defaults of the TCPC study."""

    def __init__(self, root: Path = ROOT, configs: Path = CONFIGS):
        self.root = Path(root)
        sim = _read(configs / "simulation.yaml")
        self.regimes: dict = _read(configs / "regimes.yaml")["regimes"]
        self.materials: list[str] = list(sim["cases"]["materials"])
        self.regime_order: list[str] = list(sim["cases"]["regimes"])
        self.solver: dict = sim["solver"]
        self.geometry: str = sim["geometry"]["runtime_vtk"]

    def check(self, regime: str, material: str) -> None:
        if regime not in self.regimes:
            raise ValueError(f"regime {regime!r} is not in the regime table (known: {', '.join(sorted(self.regimes))})")
        if material not in self.materials:
            raise ValueError(f"material {material!r} is not in the case matrix (known: {', '.join(self.materials)})")

    def argv(self, regime: str, material: str, case_id: str | None = None, overrides: dict | None = None,
             extra: list[str] | tuple = ()) -> list[str]:
        """The sim_tcpc command line of one case.  ``overrides`` maps a flag
        name ('--final-time', or 'geometry') to a value that replaces the
        configured one; None values are ignored."""
        self.check(regime, material)
        over = {k.lstrip("-").replace("_", "-"): v for k, v in (overrides or {}).items() if v is not None}
        flow = self.regimes[regime]
        cmd = [str(self.root / self.solver["executable"]),
               "--geometry", str(self.root / over.get("geometry", self.geometry)),
               "--material", material,
               "--case-id", case_id or f"{regime}_{material}"]
        for flag, key in REGIME_FLAGS:
            cmd += [flag, str(flow[key])]
        for flag, key, _ in SOLVER_FLAGS:
            cmd += [flag, str(over.get(flag[2:], self.solver[key]))]
        return cmd + list(extra)


@functools.lru_cache(maxsize=None)
def study(root: Path = ROOT) -> Study:
    return Study(root)


def build_command(root: Path, regime: str, material: str, case_id: str | None = None, extra=(), **overrides):
    """sim_tcpc argv; keyword overrides use the option names with
    underscores (final_time=..., geometry=...)."""
    return study(Path(root)).argv(regime, material, case_id, overrides, extra)


def parser(s: Study) -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(prog="run_case", description="Print and run the sim_tcpc command of one TCPC "
                                                              "case of the study (regime x material).")
    ap.add_argument("--regime", required=True, help="caval flow regime of the regime table")
    ap.add_argument("--material", required=True, choices=s.materials)
    ap.add_argument("--case-id", default=None, help="results_<case-id>/ (default <regime>_<material>)")
    ap.add_argument("--geometry", default=None, help="wall-mask VTK (default: the configured runtime geometry)")
    for flag, key, typ in SOLVER_FLAGS:
        ap.add_argument(flag, type=typ, default=None, help=f"override the configured {key}")
    ap.add_argument("--dry-run", action="store_true", help="print the command only")
    return ap


def main(argv: list[str] | None = None) -> int:
    s = study()
    args, passthrough = parser(s).parse_known_args(argv)
    opts = vars(args)
    overrides = {"geometry": opts["geometry"], **{flag: opts[flag[2:].replace("-", "_")] for flag, _, _ in SOLVER_FLAGS}}
    cmd = s.argv(args.regime, args.material, args.case_id, overrides, passthrough)
    print(shlex.join(cmd), flush=True)
    if args.dry_run:
        return 0
    return subprocess.call(cmd, cwd=s.root)


if __name__ == "__main__":
    sys.exit(main())
