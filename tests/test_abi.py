"""The C ABI library without a GPU: it loads, exports every entry point that
include/lbm355.h declares, reports errors as status codes, and its host-side
label setup (the sim_tcpc setupBoundaries restatement that lbm355_case_create
runs before touching the device) equals the oracle's cell for cell.  No
compute call is made here."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from nonnewlbm_amd import _abi, lbm355
from nonnewlbm_amd.geometry import channel_mask, tcpc_mask
from oracle.oracle import OracleCase
from tests.helpers import channel_case, tcpc_case

HEADER = Path(__file__).resolve().parents[1] / "include" / "lbm355.h"


def declared() -> set[str]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return set(re.findall(r"\b(lbm355_\w+)\s*\(", text))


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 20
    assert names == set(_abi.EXPORTED)
    lib = _abi.load()
    for n in sorted(names):
        assert getattr(lib, n) is not None, n


def test_version_and_errors():
    lib = _abi.load()
    assert b"gfx950" in lib.lbm355_version()
    g = channel_mask()
    d, keep = _abi.make_desc(g.flat(), (2, 2, 2))
    out = C.c_void_p()
    assert lib.lbm355_case_create(C.byref(d), C.byref(out)) == -1
    assert b"at least 4" in lib.lbm355_last_error()
    assert not out.value
    assert lib.lbm355_case_footprint(None, C.byref(C.c_int64())) == -1
    # the null-handle paths answer EINVAL, never crash
    assert lib.lbm355_case_step(None, 1) == -1
    assert lib.lbm355_case_destroy(None) == 0      # like free(NULL)
    n = C.c_int(-1)
    assert lib.lbm355_device_count(C.byref(n)) == 0 and n.value >= 0
    del keep


def test_descriptor_validation():
    g, kw = channel_case()
    with pytest.raises(_abi.LbmError):
        lbm355.labels(g.flat(), g.dims, **dict(kw, nu_ref=0.0))
    with pytest.raises(_abi.LbmError):
        lbm355.labels(g.flat(), g.dims, **dict(kw, storage=7))
    with pytest.raises(_abi.LbmError):
        lbm355.labels(g.flat(), g.dims, **dict(kw, face_bc=(2, 2, 2, 1, 1, 1)))   # inflow on a y face
    with pytest.raises(TypeError):
        _abi.make_desc(g.flat(), g.dims, bogus=1)


CASES = [
    ("channel", lambda: channel_case()),
    ("tcpc-0.25", lambda: tcpc_case(0.25)),
    ("tcpc-0.3", lambda: tcpc_case(0.3)),
    ("tcpc-shifted", lambda: (tcpc_mask(0.25, dims=(66, 50, 34)), tcpc_case(0.25)[1])),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_labels_equal_oracle(name, make):
    g, kw = make()
    kw = dict(kw, storage=_abi.F64)
    lab, info = lbm355.labels(g.flat(), g.dims, **kw)
    o = OracleCase(g.flat(), g.dims, **kw)
    np.testing.assert_array_equal(lab, o.field("label").astype(np.int32))
    oi = o.info()
    for k in ("fluid", "wall", "area_ivc", "area_svc", "area_out5", "area_out6", "active_cells", "terminate"):
        assert info[k] == oi[k], k
    assert info["device_bytes"] == lbm355.footprint(g.flat(), g.dims, **kw)
    assert info["bricks_active"] * 64 >= info["active_cells"]


def test_zero_area_terminates():
    """No open cell anywhere: setupBoundaries logs zero inlet and outlet areas
    and the run exits with code 2 (sim_tcpc.cu:632-639, 837).  (Any open cell
    would be extruded out to every open face, sim_tcpc.cu:547-584.)"""
    dims = (16, 12, 10)
    mask = np.ones(dims, dtype=np.int32)
    lab, info = lbm355.labels(mask.transpose(2, 1, 0).ravel(), dims, storage=_abi.F64, nu_ref=0.05)
    assert info["terminate"] == 1 and info["area_ivc"] == 0 and info["area_out5"] == 0
    assert np.all(lab == 1)
    mask[:, 5:7, 4:6] = 0     # a tube along x reaches the y faces through extrusion
    lab, info = lbm355.labels(mask.transpose(2, 1, 0).ravel(), dims, storage=_abi.F64, nu_ref=0.05)
    # the whole y=0 row x = 0..15, z = 4,5 (the y planes overwrite the x planes)
    assert info["terminate"] == 0 and info["area_out5"] == 16 * 2


def test_case_above_4gib_is_refused_before_the_device():
    """Buffer-resource offsets are 32-bit: a case whose population buffer
    would reach 4 GiB (an all-fluid 288x288x256 grid at FP64: 21.2 M cells,
    about 4.6 GB) is refused at setup, before any device call."""
    lib = _abi.load()
    dims = (288, 288, 256)
    geom = np.zeros(int(np.prod(dims)), dtype=np.int32)
    desc, keep = _abi.make_desc(geom, dims, storage=_abi.F64, collision=_abi.CUM, nu_ref=0.02, dl=1e-3, dt=1e-4)
    h = C.c_void_p()
    rc = lib.lbm355_case_create(C.byref(desc), C.byref(h))
    assert rc == _abi.EINVAL and not h.value
    assert "4 GiB" in lib.lbm355_last_error().decode()
    del keep


def _min_active_bricks(lab: np.ndarray) -> int:
    """Fewest 8x8x1 bricks holding a non-wall cell over every brick-origin
    shift (ox, oy) that keeps the planes 0,1 and n-2,n-1 in one brick:
    0 <= o <= 6 and n - 1 + o not a multiple of 8 (counted by brute force
    on the label field, independently of the library's run-based count)."""
    X, Y, Z = lab.shape
    x, y, z = np.nonzero(lab != 1)
    best = None
    for ox in range(7):
        if (X - 1 + ox) % 8 == 0:
            continue
        for oy in range(7):
            if (Y - 1 + oy) % 8 == 0:
                continue
            keys = np.unique((z.astype(np.int64) * 4096 + (y + oy) // 8) * 4096 + (x + ox) // 8)
            best = len(keys) if best is None else min(best, len(keys))
    return best


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_origin_shift_minimises_active_bricks(name, make):
    """Case setup picks, among the brick-origin shifts that keep each BC plane
    pair in one brick, one with the fewest active bricks (one wavefront each)."""
    g, kw = make()
    lab, info = lbm355.labels(g.flat(), g.dims, **kw)
    assert lab.shape == tuple(g.dims)
    assert info["bricks_active"] == _min_active_bricks(lab)
