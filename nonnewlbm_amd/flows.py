"""Plane flows and the pulmonary flow split of sim_tcpc output VTKs.

Stands in for postprocess/compute_flows.py (the flux rule of
compute_flows.py:96-147 and its command line and CSV columns, which
downstream tools read) on numpy arrays, without VTK.

The rule (kept as the reference states it, so that a split means the same
thing): the cells whose ``wall`` value rounds to a label form a cap; its
normal is the axis along which the cap's cells spread least; the cap faces
out of the domain on the side of the domain middle it lies on (outward sign
-1 below the middle, +1 above); the outward flux is the sum over the cap of
velocity along the normal times that sign times dx^2.  Inlet caps (IVC, SVC)
are reported as inflow (the negated outward flux).  A row gives every cap's
flow in l/min (abs of m^3/s * 6e4) and cell count, the mass balance (inflow
of the inlets minus outflow of the RPA/LPA outlets) and the LPA/RPA shares of
the outlet total in percent.
"""
from __future__ import annotations

import argparse
import csv
import dataclasses
import glob
import re
import sys
from pathlib import Path

import numpy as np

from .geometry import read_output_vtk

# label -> cap name of the sim_tcpc layout (sim_tcpc.cu:194-203)
DEFAULT_LABELS = {3: "IVC", 4: "SVC", 5: "RPA", 6: "LPA"}
# which names are inlets and which outlets enter the balance and the split
INLETS = ("IVC", "SVC")
OUTLETS = ("RPA", "LPA")
M3S_TO_LMIN = 1.0e3 * 60.0


@dataclasses.dataclass(frozen=True)
class Cap:
    name: str
    label: int
    cells: int
    axis: int           # normal axis (0 x, 1 y, 2 z)
    outward: float      # +1 or -1
    flux: float         # outward flux in m^3/s

    @property
    def reported(self) -> float:
        """m^3/s as reported: inflow for inlets, outflow otherwise."""
        return -self.flux if self.name in INLETS else self.flux


def measure_cap(wall: np.ndarray, velocity: np.ndarray, dx: float, label: int, name: str = "") -> Cap:
    """The cap of one label; wall is [x, y, z], velocity [x, y, z, 3] in m/s."""
    members = np.rint(wall).astype(np.int64) == label
    idx = np.nonzero(members)
    if idx[0].size == 0:
        return Cap(name, label, 0, 0, 1.0, 0.0)
    pos = np.stack(idx, axis=1)
    axis = int(np.argmin(np.ptp(pos, axis=0)))
    middle = 0.5 * (wall.shape[axis] - 1)
    outward = -1.0 if pos[:, axis].mean() < middle else 1.0
    normal_v = np.asarray(velocity[..., axis], dtype=np.float64)[members]
    return Cap(name, label, int(pos.shape[0]), axis, outward, float(normal_v.sum() * outward * dx * dx))


def summarise(caps: list[Cap]) -> dict:
    """One CSV row (without the file column) from the caps, in cap order."""
    row: dict = {}
    by_name = {}
    for cap in caps:
        row[f"{cap.name}_l_min"] = abs(cap.reported * M3S_TO_LMIN)
        row[f"{cap.name}_cells"] = cap.cells
        by_name[cap.name] = abs(cap.reported)
    inflow = sum(by_name.get(n, 0.0) for n in INLETS)
    outflow = sum(by_name.get(n, 0.0) for n in OUTLETS)
    row["mass_balance_l_min"] = (inflow - outflow) * M3S_TO_LMIN
    for name in ("LPA", "RPA"):
        row[f"{name.lower()}_split_percent"] = 100.0 * by_name.get(name, 0.0) / outflow if outflow > 0 else 0.0
    return row


def compute_arrays(wall, velocity, dx, labels: dict[int, str] | None = None, file: str = "") -> dict:
    caps = [measure_cap(wall, velocity, dx, lab, name) for lab, name in (labels or DEFAULT_LABELS).items()]
    return {"file": file, **summarise(caps)}


def first_spacing(coords) -> float:
    """The first positive spacing of a rectilinear coordinate array (the grid
    step dl of a sim_tcpc VTK)."""
    steps = np.abs(np.diff(np.asarray(coords, dtype=np.float64)))
    positive = steps[steps > 0]
    if positive.size == 0:
        raise RuntimeError("the rectilinear x coordinates have no positive spacing")
    return float(positive[0])


def compute_file(path: Path, labels: dict[int, str] | None = None) -> dict:
    data = read_output_vtk(path)
    for needed in ("wall", "velocity"):
        if needed not in data:
            raise RuntimeError(f"{path}: no '{needed}' array")
    return compute_arrays(data["wall"], data["velocity"], first_spacing(data["coords"]["X"]), labels, str(path))


_LABEL_ITEM = re.compile(r"^\s*(-?\d+)\s*:\s*(\S+)\s*$")


def parse_labels(spec: str | None) -> dict[int, str]:
    """'3:IVC,4:SVC,...' -> {3: 'IVC', ...}; empty -> the sim_tcpc layout."""
    if not spec:
        return dict(DEFAULT_LABELS)
    out: dict[int, str] = {}
    for item in spec.split(","):
        m = _LABEL_ITEM.match(item)
        if not m:
            raise ValueError(f"label item {item!r} is not LABEL:NAME")
        out[int(m.group(1))] = m.group(2)
    return out


def snapshot_order(path: str) -> list:
    """Sort key that orders rank000_data_10 after rank000_data_9."""
    return [int(tok) if tok.isdigit() else tok for tok in re.split(r"(\d+)", path)]


def expand_inputs(items: list[str]) -> list[Path]:
    """Glob patterns in snapshot order; a pattern that matches nothing is
    taken as a literal path (and fails when read)."""
    out: list[Path] = []
    for item in items:
        hits = sorted(glob.glob(item), key=snapshot_order) or [item]
        out.extend(Path(h) for h in hits)
    return out


def write_rows(path: str, rows: list[dict]) -> None:
    header = list(rows[0])
    with open(path, "w", newline="", encoding="utf-8") as fh:
        out = csv.writer(fh)
        out.writerow(header)
        for row in rows:
            out.writerow([row.get(k, "") for k in header])


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(prog="flows", description="Cap flows (l/min), mass balance and LPA/RPA split of "
                                                           "sim_tcpc VTK snapshots, one CSV row per file.")
    ap.add_argument("paths", nargs="+", help="snapshot files or glob patterns")
    ap.add_argument("--labels", default=None, help="cap labels as LABEL:NAME,... (default 3:IVC,4:SVC,5:RPA,6:LPA)")
    ap.add_argument("--output", default="flows.csv", help="CSV to write (default flows.csv)")
    args = ap.parse_args(argv)
    labels = parse_labels(args.labels)
    rows = [compute_file(p, labels) for p in expand_inputs(args.paths)]
    if rows:
        write_rows(args.output, rows)
        print(args.output)
    return 0


if __name__ == "__main__":
    sys.exit(main())
