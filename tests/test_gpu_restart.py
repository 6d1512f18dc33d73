"""Checkpoint / restart in the reference's current_state layout
(state.hpp:1552-1640): N steps, save, load into a fresh case, M steps equals
N + M steps without the break."""
from __future__ import annotations

import numpy as np
import pytest

from nonnewlbm_amd import _abi
from nonnewlbm_amd.lbm355 import Case
from tests.helpers import tcpc_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_first", [17, 20])
@pytest.mark.parametrize("storage", [_abi.F64, _abi.F32, _abi.F16_OFFSET])
def test_restart_equals_straight_run(gpu_lib, tmp_path, storage, n_first):
    g, kw = tcpc_case(0.25, storage=storage)
    straight = Case(g.flat(), g.dims, **kw)
    straight.step(n_first + 13)
    a = Case(g.flat(), g.dims, **kw)
    a.step(n_first)
    counters = [(k, 0.1 * (k + 1)) for k in range(11)]
    a.save_state(tmp_path / "current_state", final_time=2.5, counters=counters, stat_prev_iter=0)
    a.close()
    b = Case(g.flat(), g.dims, **kw)
    info = b.load_state(tmp_path / "current_state")
    assert info["iterations"] == n_first and info["final_time"] == 2.5
    assert [c for c, _ in info["counters"]] == list(range(11))
    np.testing.assert_allclose([p for _, p in info["counters"]], [0.1 * (k + 1) for k in range(11)], rtol=1e-6)
    b.step(13)
    pb, ps = b.populations(), straight.populations()
    act = straight.field("label") != 1
    if storage == _abi.F16_OFFSET:
        np.testing.assert_allclose(pb[:, act], ps[:, act], rtol=0, atol=1e-6)
    else:
        np.testing.assert_array_equal(pb, ps)
        for name in ("rho", "vx", "vy", "vz"):
            np.testing.assert_array_equal(b.field(name), straight.field(name), err_msg=name)
    # statistics sums are kept in the reference's real type (float for FP32)
    tol = 0 if storage == _abi.F64 else 1e-5
    for name in ("sum_rux", "sum_rruyy"):
        np.testing.assert_allclose(b.field(name), straight.field(name), rtol=tol, atol=1e-12, err_msg=name)
    np.testing.assert_allclose(b.flux(mean=True), straight.flux(mean=True), rtol=max(tol, 1e-12))


def test_restart_fd_force_and_files(gpu_lib, tmp_path):
    from nonnewlbm_amd.materials import case_kwargs, material_by_name
    g, kw = tcpc_case(0.25, storage=_abi.F32)
    kw.update(case_kwargs(material_by_name("gx"), g.dl, 0.02, cy_model=_abi.CY_FD_FORCE))
    straight = Case(g.flat(), g.dims, **kw)
    straight.step(25)
    a = Case(g.flat(), g.dims, **kw)
    a.step(11)
    d = tmp_path / "current_state"
    a.save_state(d)
    X, Y, Z = g.dims
    # file sizes of the reference layout (float dreal, int16 map, 20 macro fields)
    assert (d / "df_0_rank000").stat().st_size == 27 * X * Y * Z * 4
    assert (d / "df_1_rank000").stat().st_size == 27 * X * Y * Z * 4
    assert (d / "map_rank000").stat().st_size == X * Y * Z * 2
    assert (d / "macro_rank000").stat().st_size == 20 * X * Y * Z * 4
    assert (d / "config_rank000").read_text().split()[:4] == ["11", str(X), str(Y), str(Z)]
    # map[x][z][y] holds the labels; macro[x][id][z][y] id 0 = rho
    lab = a.field("label")
    m = np.fromfile(d / "map_rank000", dtype=np.int16).reshape(X, Z, Y).transpose(0, 2, 1)
    np.testing.assert_array_equal(m, lab.astype(np.int16))
    mac = np.fromfile(d / "macro_rank000", dtype=np.float32).reshape(X, 20, Z, Y).transpose(1, 0, 3, 2)
    np.testing.assert_array_equal(mac[0], a.field("rho").astype(np.float32))
    np.testing.assert_array_equal(mac[11], a.field("fx").astype(np.float32))
    b = Case(g.flat(), g.dims, **kw)
    b.load_state(d)
    b.step(14)
    for name in ("rho", "vx", "fx"):
        np.testing.assert_array_equal(b.field(name), straight.field(name), err_msg=name)
