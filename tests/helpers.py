"""Shared cases for the tests: the 64x32x32 channel and small TCPC junctions."""
from __future__ import annotations

import numpy as np

from nonnewlbm_amd import _abi
from nonnewlbm_amd.geometry import channel_mask, tcpc_mask


def channel_case(storage=_abi.F64, collision=_abi.CUM, **extra):
    """The 64x32x32 channel: radius 12, parabolic inflow u_max = 0.05 at x-min,
    pressure outlet at x-max, tau = 0.6 (nu = 1/30)."""
    g = channel_mask()
    kw = dict(storage=storage, collision=collision, nu_ref=(0.6 - 0.5) / 3.0, profile=_abi.PROFILE_PARABOLIC,
              ivc_umax=0.05, ivc_radius=12.0, ivc_center=(15.5, 15.5), face_bc=_abi.CHANNEL_FACES,
              dl=1.0e-3, dt=1.0e-3)
    kw.update(extra)
    return g, kw


def tcpc_case(scale=0.25, storage=_abi.F32, collision=_abi.CUM, **extra):
    """Small synthetic TCPC junction with plug inflows in the sim_tcpc layout."""
    g = tcpc_mask(scale)
    kw = dict(storage=storage, collision=collision, nu_ref=0.02, profile=_abi.PROFILE_PLUG,
              ivc_flux=0.02 * (np.pi * (20 * scale) ** 2), svc_flux=0.015 * (np.pi * (16 * scale) ** 2),
              dl=g.dl, dt=1.0e-4)
    kw.update(extra)
    return g, kw


def stored_populations(act: np.ndarray, steps: int, balanced: bool = False) -> np.ndarray:
    """[27, n] mask over the populations of the active cells ``act`` (in the
    order of ``pops[:, act]``) that the AA layout holds after ``steps`` steps.

    After an odd number of steps (the last one even) population q of a cell
    sits in the cell's own slot opp(q): all are held.  After an even number
    (the last one odd) it sits in slot q of the downstream cell x + c_q;
    where that cell is outside the grid the population has no slot.  Those
    are exactly the outward directions of the cells on the grid faces, which
    no step reads (the clamped pull and the boundary partner reads only take
    inward and tangential directions), and the library reports them as 0.

    ``balanced``: the layout of the two-byte storage types, where pair (P, s)
    is stored at P - c_s for s < 13.  After any number of steps population q
    of cell x is then held unless q > 13 and x + c_q is outside the grid."""
    idx = np.nonzero(act)
    held = np.ones((27, idx[0].size), dtype=bool)
    if steps % 2 == 1 and not balanced:
        return held
    for q in range(27):
        if balanced and q <= 13:
            continue
        c = (q // 9 - 1, (q // 3) % 3 - 1, q % 3 - 1)
        for a in range(3):
            t = idx[a] + c[a]
            held[q] &= (t >= 0) & (t < act.shape[a])
    return held


def assert_populations_match(pg: np.ndarray, pc: np.ndarray, act: np.ndarray, steps: int, atol: float,
                             balanced: bool = False) -> None:
    """GPU populations ``pg`` against the oracle's ``pc`` ([27, X, Y, Z]) on
    every active cell: the held ones to ``atol``, the unheld ones exactly 0."""
    held = stored_populations(act, steps, balanced)
    a, b = pg[:, act], pc[:, act]
    assert np.all(a[~held] == 0.0), "a population without a slot is reported non-zero"
    worst = float(np.max(np.abs(a[held] - b[held])))
    assert worst < atol, worst


def rel_l2(a: np.ndarray, b: np.ndarray, mask: np.ndarray | None = None) -> float:
    if mask is not None:
        a = a[mask]
        b = b[mask]
    num = float(np.sqrt(np.sum((a - b) ** 2)))
    den = float(np.sqrt(np.sum(b ** 2)))
    return num / den if den > 0 else num
