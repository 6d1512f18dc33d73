"""Writes tests/golden/tcpc25_03_280.npz: the wall mask that
nonnewlbm_amd.stl_geometry makes of the reference checkout's
geometry/input/TCPC25_03.stl at 280 voxels on the longest axis
(280x256x44 cells, 177,804 fluid; values 0 fluid, 1 wall, 3/4/5/6 end tags)
and its cell spacing in metres.  Input data only: the GPU test runs the solver
and the oracle on it.  Coarser settings (200-260 voxels) lose the x-max end
plane to the voxel rounding and are refused by the pipeline.

    python tests/golden/make_tcpc25_03_labels.py [path/to/TCPC25_03.stl]
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from nonnewlbm_amd import stl_geometry  # noqa: E402


def main() -> int:
    stl = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/geometry/input/TCPC25_03.stl")
    res = stl_geometry.build(stl, longest_axis_voxels=280)
    out = Path(__file__).resolve().parent / "tcpc25_03_280.npz"
    np.savez_compressed(out, labels=res.geometry.labels.astype(np.int8), dl=np.float64(res.dx_m))
    print(out, res.geometry.dims)
    return 0


if __name__ == "__main__":
    sys.exit(main())
