"""Inlet/outlet flows and the LPA/RPA split from sim_tcpc output VTKs: the
compute_flows.py post-processor (postprocess/compute_flows.py:79-182) on
numpy arrays.

For each label L (3 IVC, 4 SVC, 5 RPA, 6 LPA by default) the cells whose
``wall`` value rounds to L form a cap; its axis is the one of least coordinate
spread, its outward sign is -1 if the cap centre lies below the domain middle
on that axis, else +1; flux = sum of velocity[axis] * sign * dx^2.  Inlets are
reported as positive inflow; l/min = m^3/s * 6e4; mass balance = inflow -
outflow; splits are percentages of the outlet total.
"""
from __future__ import annotations

import argparse
import csv
import glob
import re
from pathlib import Path

import numpy as np

from .geometry import read_output_vtk

DEFAULT_LABELS = {3: "IVC", 4: "SVC", 5: "RPA", 6: "LPA"}


def natural_key(path: str) -> list:
    return [int(p) if p.isdigit() else p for p in re.split(r"(\d+)", path)]


def _dx(coords: np.ndarray) -> float:
    d = np.abs(np.diff(np.asarray(coords, dtype=np.float64)))
    pos = d[d > 0]
    if pos.size == 0:
        raise RuntimeError("non-positive rectilinear x spacing")
    return float(pos[0])


def integrate_label(wall: np.ndarray, velocity: np.ndarray, dx: float, label: int,
                    rho: np.ndarray | None = None) -> tuple[float, int]:
    """(signed outward flux in m^3/s, cell count) of the cap with this label;
    wall is [x, y, z], velocity [x, y, z, 3] in m/s."""
    sel = np.rint(wall).astype(np.int64) == label
    coords = np.argwhere(sel)
    if coords.size == 0:
        return 0.0, 0
    spread = coords.max(axis=0) - coords.min(axis=0)
    axis = int(np.argmin(spread))
    centre = float(coords[:, axis].mean())
    sign = -1.0 if centre < 0.5 * (wall.shape[axis] - 1) else 1.0
    v = velocity[..., axis][sel].astype(np.float64)
    if rho is not None:
        v = v * rho[sel]
    return float(np.sum(v * sign * dx * dx)), int(coords.shape[0])


def liter_per_min(m3_s: float) -> float:
    return m3_s * 1000.0 * 60.0


def compute_arrays(wall, velocity, dx, label_names=None, file: str = "") -> dict:
    label_names = label_names or DEFAULT_LABELS
    row: dict = {"file": file}
    signed: dict[str, float] = {}
    for label, name in label_names.items():
        flux, count = integrate_label(wall, velocity, dx, label)
        reported = -flux if name in {"IVC", "SVC"} else flux
        signed[name] = reported
        row[f"{name}_l_min"] = abs(liter_per_min(reported))
        row[f"{name}_cells"] = count
    inlet = abs(signed.get("IVC", 0.0)) + abs(signed.get("SVC", 0.0))
    outlet = abs(signed.get("RPA", 0.0)) + abs(signed.get("LPA", 0.0))
    row["mass_balance_l_min"] = liter_per_min(inlet - outlet)
    row["lpa_split_percent"] = 100.0 * abs(signed.get("LPA", 0.0)) / outlet if outlet > 0 else 0.0
    row["rpa_split_percent"] = 100.0 * abs(signed.get("RPA", 0.0)) / outlet if outlet > 0 else 0.0
    return row


def compute_file(path: Path, label_names=None) -> dict:
    d = read_output_vtk(path)
    if "wall" not in d:
        raise RuntimeError(f"{path} has no 'wall' array")
    if "velocity" not in d:
        raise RuntimeError(f"{path} has no 'velocity' array")
    return compute_arrays(d["wall"], d["velocity"], _dx(d["coords"]["X"]), label_names, str(path))


def parse_label_map(text: str | None) -> dict[int, str]:
    if not text:
        return DEFAULT_LABELS.copy()
    out: dict[int, str] = {}
    for part in text.split(","):
        label, name = part.split(":", 1)
        out[int(label)] = name
    return out


def main(argv: list[str] | None = None) -> int:
    p = argparse.ArgumentParser(description="Compute TCPC inlet/outlet flows and LPA/RPA split.")
    p.add_argument("paths", nargs="+", help="VTK files or glob patterns.")
    p.add_argument("--labels", default=None, help="Comma list like 3:IVC,4:SVC,5:RPA,6:LPA.")
    p.add_argument("--output", default="flows.csv")
    args = p.parse_args(argv)
    files: list[Path] = []
    for item in args.paths:
        matches = sorted(glob.glob(item), key=natural_key)
        files.extend(Path(m) for m in (matches or [item]))
    names = parse_label_map(args.labels)
    rows = [compute_file(f, names) for f in files]
    if not rows:
        return 0
    with open(args.output, "w", newline="", encoding="utf-8") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0].keys()))
        w.writeheader()
        w.writerows(rows)
    print(args.output)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
