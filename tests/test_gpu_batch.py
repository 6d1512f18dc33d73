"""Batched stepping: one fused launch per step over the concatenated brick
index of several cases (optionally replayed from a captured hipGraph) must
give every case exactly the result it gets alone."""
from __future__ import annotations

import numpy as np
import pytest

from nonnewlbm_amd import _abi
from nonnewlbm_amd.lbm355 import Batch, Case
from tests.helpers import tcpc_case

pytestmark = pytest.mark.gpu


def _cases(storage):
    out = []
    for scale, flux in ((0.25, 1.0), (0.3, 0.7), (0.25, 1.3)):
        g, kw = tcpc_case(scale, storage=storage)
        kw["ivc_flux"] *= flux
        out.append((g, kw))
    return out


@pytest.mark.parametrize("group_bytes", [None, 1])
@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("storage", [_abi.F32, _abi.F16_OFFSET])
def test_batch_equals_alone(gpu_lib, graph, storage, group_bytes):
    """group_bytes=1 steps every member as a group of its own (each group runs
    its chunk of steps before the next one starts)."""
    specs = _cases(storage)
    alone = [Case(g.flat(), g.dims, **kw) for g, kw in specs]
    mixed = [Case(g.flat(), g.dims, **kw) for g, kw in specs]
    for c in alone:
        c.step(37)
    # one case starts earlier: the batch must keep each case's own step index
    mixed[1].step(5)
    b = Batch(mixed, graph=graph)
    if group_bytes is not None:
        b.set_group_bytes(group_bytes)
    b.step(32)
    b.step(5)
    # case 1 is 5 steps ahead: step the others alone to 37 and compare
    b.close()
    for k in (0, 2):
        for name in ("rho", "vx", "vy", "vz"):
            np.testing.assert_array_equal(mixed[k].field(name), alone[k].field(name), err_msg=f"case {k} {name}")
    alone[1].step(5)
    for name in ("rho", "vx", "vy", "vz"):
        np.testing.assert_array_equal(mixed[1].field(name), alone[1].field(name), err_msg=f"case 1 {name}")


def test_batch_then_alone_continues(gpu_lib):
    """A case can leave a batch and keep stepping on its own (AA parity and
    the step index are carried over)."""
    (g, kw), = _cases(_abi.F32)[:1]
    a = Case(g.flat(), g.dims, **kw)
    m = Case(g.flat(), g.dims, **kw)
    a.step(40)
    b = Batch([m], graph=True)
    b.step(19)
    b.close()
    m.step(21)
    for name in ("rho", "vx"):
        np.testing.assert_array_equal(m.field(name), a.field(name))


def test_autotune_keeps_results(gpu_lib):
    """The autotuners time launch shapes, the cache policy and the wavefront
    cap on real steps; the trajectory is the one of plain stepping."""
    specs = _cases(_abi.F32)[:2]
    ref = [Case(g.flat(), g.dims, **kw) for g, kw in specs]
    pair = [Case(g.flat(), g.dims, **kw) for g, kw in specs]
    b = Batch(pair, graph=True)
    bpg, nt, cap, group = b.autotune()          # warm-up, 24 configurations, 3 group sizes
    assert bpg in (1, 2, 4, 8) and nt in (0, 1) and cap in (0, 16, 12) and group in (0, 256, 128)
    b.step(20)
    b.close()
    n = pair[0].info()["steps"]
    assert n > 24 * 10 and pair[1].info()["steps"] == n
    for c in ref:
        c.step(n)
    for k in range(2):
        np.testing.assert_array_equal(pair[k].field("vy"), ref[k].field("vy"))
    one = Case(specs[0][0].flat(), specs[0][0].dims, **specs[0][1])
    bpg, nt, cap = one.autotune()              # warm-up, 24 configurations, 3 passes over 4 finalists
    assert bpg in (1, 2, 4, 8) and nt in (0, 1) and cap in (0, 16, 12)
    m = one.info()["steps"]
    assert m > 24 * 10
    one.step(74)
    ref[0].step(m + 74 - n)
    for name in ("rho", "vx"):
        np.testing.assert_array_equal(one.field(name), ref[0].field(name))


@pytest.mark.parametrize("graph", [False, True])
def test_batch_follows_inflow_hook(gpu_lib, graph):
    """updateKernelVelocities (sim_tcpc.cu:657-663) on batched cases: the
    caval fluxes of one member change between batch calls (after 70 steps,
    which include a 64-step graph replay, then after 64 more); every member
    equals the same case stepped alone with the same changes."""
    specs = _cases(_abi.F32)
    alone = [Case(g.flat(), g.dims, **kw) for g, kw in specs]
    mixed = [Case(g.flat(), g.dims, **kw) for g, kw in specs]
    b = Batch(mixed, graph=graph)
    for n, scale in ((70, 1.7), (64, 0.4), (9, None)):
        b.step(n)
        for c in alone:
            c.step(n)
        if scale is not None:
            kw = specs[1][1]
            for c in (mixed[1], alone[1]):
                c.set_inflow(kw["ivc_flux"] * scale, kw["svc_flux"] * (2.0 - scale))
    b.close()
    for k in range(3):
        for name in ("rho", "vx", "vy", "vz"):
            np.testing.assert_array_equal(mixed[k].field(name), alone[k].field(name), err_msg=f"case {k} {name}")
    fl = mixed[1].flux(mean=False)
    assert fl[0] == pytest.approx(specs[1][1]["ivc_flux"] * 0.4, rel=1e-5)


def test_batch_waits_for_work_queued_on_a_case(gpu_lib):
    """Steps and a statistics reset queued on a member's own stream right
    before a batch call (no explicit synchronisation in between) complete
    before the batch touches the case, and the case's own later steps follow
    the batch's: the result equals the case stepped alone."""
    specs = _cases(_abi.F32)
    alone = [Case(g.flat(), g.dims, **dict(kw, stats_every=1)) for g, kw in specs]
    mixed = [Case(g.flat(), g.dims, **dict(kw, stats_every=1)) for g, kw in specs]
    b = Batch(mixed, graph=True)
    b.step(3)
    mixed[0].step(40)          # queued on case 0's stream, not waited for
    mixed[2].reset_stats()
    b.step(66)
    mixed[2].step(5)           # queued behind the batch's steps
    b.step(2)
    alone[0].step(3 + 40 + 66 + 2)
    alone[1].step(3 + 66 + 2)
    alone[2].step(3)
    alone[2].reset_stats()
    alone[2].step(66 + 5 + 2)
    b.close()
    for k in range(3):
        for name in ("rho", "vx", "vy", "sum_rux", "sum_rruyy"):
            np.testing.assert_array_equal(mixed[k].field(name), alone[k].field(name), err_msg=f"case {k} {name}")


@pytest.mark.parametrize("graph", [False, True])
def test_batch_membership_changes(gpu_lib, graph):
    """Members leave and join between batch calls (the device brick list and
    member constants are re-uploaded; the captured graph is kept unless the
    launch grid must grow): every case equals the same case stepped alone,
    bit for bit, whatever it did in between."""
    specs = _cases(_abi.F32)
    alone = [Case(g.flat(), g.dims, **kw) for g, kw in specs]
    mixed = [Case(g.flat(), g.dims, **kw) for g, kw in specs]
    b = Batch(mixed[:2], graph=graph)
    b.step(70)
    b.remove(mixed[0])
    assert b.size()[0] == 1
    mixed[0].step(9)            # on its own meanwhile
    b.add(mixed[2])             # a fresh case joins at its own step 0
    b.step(66)
    b.add(mixed[0])             # and the first one comes back
    b.step(65)
    cfg = b.config()
    b.close()
    steps = [70 + 9 + 65, 70 + 66 + 65, 66 + 65]
    for c, n in zip(alone, steps):
        c.step(n)
    for k in range(3):
        assert mixed[k].info()["steps"] == steps[k]
        for name in ("rho", "vx", "vy", "vz"):
            np.testing.assert_array_equal(mixed[k].field(name), alone[k].field(name), err_msg=f"case {k} {name}")
    assert cfg["uploads"] == 4
    if graph:
        assert 1 <= cfg["captures"] <= 3


@pytest.mark.parametrize("per_wave", [2, 4])
def test_batch_persistent_loop_equals_alone(gpu_lib, per_wave):
    """The persistent form (a grid of about 1/k of the bricks whose
    wavefronts loop over the group) gives the bits of one wavefront per
    brick."""
    specs = _cases(_abi.F16_OFFSET)
    alone = [Case(g.flat(), g.dims, **kw) for g, kw in specs]
    mixed = [Case(g.flat(), g.dims, **kw) for g, kw in specs]
    b = Batch(mixed, graph=True)
    b.set_per_wave(per_wave)
    b.set_group_bytes(1)        # every member a group of its own
    b.step(67)
    assert b.config()["per_wave"] == per_wave and b.config()["groups"] == 3
    b.close()
    for c in alone:
        c.step(67)
    for k in range(3):
        for name in ("rho", "vx", "vz"):
            np.testing.assert_array_equal(mixed[k].field(name), alone[k].field(name), err_msg=f"case {k} {name}")
