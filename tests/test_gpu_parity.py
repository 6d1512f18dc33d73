"""GPU parity: the HIP path against the CPU oracle on the same inputs.

Bar: relative L2 of velocity and density within 1e-5 at FP64 and
1e-3 at FP32 after the stated step count; the first steps are compared
population by population, which pins the AA addressing, the brick neighbour
tables and the boundary partner reads exactly.
"""
from __future__ import annotations

import numpy as np
import pytest

from nonnewlbm_amd import _abi
from nonnewlbm_amd.lbm355 import Case
from oracle.oracle import OracleCase
from tests.helpers import assert_populations_match, channel_case, rel_l2, stored_populations, tcpc_case

pytestmark = pytest.mark.gpu


def _pair(g, kw):
    gpu = Case(g.flat(), g.dims, **kw)
    cpu = OracleCase(g.flat(), g.dims, **kw)
    return gpu, cpu


def _active(cpu):
    return cpu.field("label") != 1


@pytest.mark.parametrize("steps", [1, 2, 3, 10])
def test_channel_populations_fp64_first_steps(gpu_lib, steps):
    g, kw = channel_case(_abi.F64)
    gpu, cpu = _pair(g, kw)
    gpu.step(steps)
    cpu.step(steps)
    act = _active(cpu)
    # every population of every active cell (wall cells hold no state in the
    # AA scheme); the outward directions of the grid-face cells have no slot
    # after an odd last step and must read 0
    assert_populations_match(gpu.populations(), cpu.populations(), act, steps, 1e-12)
    for name in ("rho", "vx", "vy", "vz"):
        np.testing.assert_allclose(gpu.field(name)[act], cpu.field(name)[act], rtol=0, atol=1e-12)


def test_channel_fp64_1000_steps(gpu_lib):
    """The 64x32x32 channel, cumulant, FP64, tau=0.6, 1000 steps."""
    g, kw = channel_case(_abi.F64)
    gpu, cpu = _pair(g, kw)
    gpu.step(1000)
    cpu.step(1000)
    act = _active(cpu)
    for name in ("rho", "vx", "vy", "vz"):
        assert rel_l2(gpu.field(name), cpu.field(name), act) < 1e-5, name


def test_channel_fp32_1000_steps(gpu_lib):
    g, kw = channel_case(_abi.F32)
    gpu = Case(g.flat(), g.dims, **kw)
    kw64 = dict(kw, storage=_abi.F64)
    cpu = OracleCase(g.flat(), g.dims, **kw64)
    gpu.step(1000)
    cpu.step(1000)
    act = _active(cpu)
    for name in ("rho", "vx", "vy", "vz"):
        assert rel_l2(gpu.field(name), cpu.field(name), act) < 1e-3, name


@pytest.mark.parametrize("collision", [_abi.CUM, _abi.BGK, _abi.MRT, _abi.KBC])
@pytest.mark.parametrize("storage", [_abi.F64, _abi.F32])
def test_tcpc_models(gpu_lib, collision, storage):
    g, kw = tcpc_case(0.25, storage=storage, collision=collision)
    gpu, cpu = _pair(g, kw)
    assert gpu.info()["area_ivc"] == cpu.info()["area_ivc"]
    np.testing.assert_array_equal(gpu.field("label"), cpu.field("label"))
    gpu.step(3)
    cpu.step(3)
    act = _active(cpu)
    tol3 = 1e-12 if storage == _abi.F64 else 2e-6
    for name in ("rho", "vx", "vy", "vz"):
        np.testing.assert_allclose(gpu.field(name)[act], cpu.field(name)[act], rtol=0, atol=tol3, err_msg=name)
    gpu.step(197)
    cpu.step(197)
    tol = 1e-5 if storage == _abi.F64 else 1e-3
    for name in ("rho", "vx", "vy", "vz"):
        assert rel_l2(gpu.field(name), cpu.field(name), act) < tol, name


@pytest.mark.parametrize("dims", [(65, 49, 33), (66, 50, 34), (67, 51, 35)])
def test_brick_origin_shift(gpu_lib, dims):
    """Grids whose last x or y plane would start a new brick (X - 1 and Y - 1
    multiples of the brick edge, as 65 x 49 for 8- and 4-cell bricks) get a
    shifted brick origin; the BC partner reads must still match the oracle
    exactly."""
    from nonnewlbm_amd.geometry import tcpc_mask
    g = tcpc_mask(0.25, dims=dims)
    kw = dict(storage=_abi.F64, collision=_abi.CUM, nu_ref=0.02, profile=_abi.PROFILE_PLUG,
              ivc_flux=5.0, svc_flux=3.0, dl=g.dl, dt=1e-4)
    gpu, cpu = _pair(g, kw)
    act = _active(cpu)
    done = 0
    for steps in (1, 1, 2):
        gpu.step(steps)
        cpu.step(steps)
        done += steps
        assert_populations_match(gpu.populations(), cpu.populations(), act, done, 1e-12)
    gpu.step(60)
    cpu.step(60)
    for name in ("rho", "vx", "vy", "vz"):
        assert rel_l2(gpu.field(name), cpu.field(name), act) < 1e-9, name


@pytest.mark.parametrize("every", [1, 3])
@pytest.mark.parametrize("storage", [_abi.F64, _abi.F32])
def test_statistics_sums(gpu_lib, storage, every):
    """Mean/rms sums (MacroTcpc e_vm_*, sim_tcpc.cu:347-353) against the
    oracle: read while a group of samples is still open (37 steps), after a
    call that closes it (48: a multiple of the 16-sample group, also at
    stats_every 3), and after a statistics reset, where the sums of steps
    48..67 are the oracle's sums at 68 minus those at 48."""
    g, kw = channel_case(storage, stats_every=every)
    gpu = Case(g.flat(), g.dims, **kw)
    cpu = OracleCase(g.flat(), g.dims, **dict(kw, storage=_abi.F64))
    act = _active(cpu)
    tol = 1e-9 if storage == _abi.F64 else 2e-5
    names = ("sum_rux", "sum_ruy", "sum_ruz", "sum_rruxx", "sum_rruyy", "sum_rruzz")
    gpu.step(37)
    cpu.step(37)
    for name in names:
        assert rel_l2(gpu.field(name), cpu.field(name), act) < tol, name
    # instantaneous inlet flux = the prescribed profile sum; the mean flux is
    # the density-weighted mean momentum of the plane cells (mean_velocity)
    lab = gpu.field("label")
    inst = gpu.flux(mean=False)
    np.testing.assert_allclose(inst[0], cpu.field("vx")[lab == 3].sum(), rtol=1e-6)
    n = 37 // every
    mean = gpu.flux(mean=True)
    np.testing.assert_allclose(mean[0], gpu.field("sum_rux")[lab == 3].sum() / n, rtol=1e-9)
    np.testing.assert_allclose(mean[1], gpu.field("sum_rux")[lab == 7].sum() / n, rtol=1e-9)
    gpu.step(11)
    cpu.step(11)
    at48 = {name: cpu.field(name) for name in names}
    for name in names:
        assert rel_l2(gpu.field(name), at48[name], act) < tol, name
    gpu.reset_stats()
    gpu.step(20)
    cpu.step(20)
    for name in names:
        assert rel_l2(gpu.field(name), cpu.field(name) - at48[name], act) < tol, name


@pytest.mark.parametrize("storage", [_abi.F16_OFFSET, _abi.BF16_OFFSET])
def test_statistics_ring_half_storage(gpu_lib, storage):
    """With 2-byte population storage the statistics go through the same FP32
    sample ring: the sums equal FP64 sums of the FP32 samples rho*u that the
    step's own macro output yields, stepped one step at a time, read with a
    group open (40 samples) and closed (48)."""
    g, kw = channel_case(storage, stats_every=1)
    gpu = Case(g.flat(), g.dims, **kw)
    act = gpu.field("label") != 1
    acc = {a: np.zeros(act.sum()) for a in ("x", "y", "z")}
    sq = {a: np.zeros(act.sum()) for a in ("x", "y", "z")}
    for step in range(1, 49):
        gpu.step(1)
        rho = gpu.field("rho")[act].astype(np.float32)
        for a in ("x", "y", "z"):
            v = (rho * gpu.field("v" + a)[act].astype(np.float32)).astype(np.float64)
            acc[a] += v
            sq[a] += v * v
        if step in (40, 48):
            for a in ("x", "y", "z"):
                np.testing.assert_allclose(gpu.field("sum_ru" + a)[act], acc[a], rtol=1e-12, atol=1e-15)
                np.testing.assert_allclose(gpu.field("sum_rru" + a * 2)[act], sq[a], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("material", ["cy-a", "cy-b"])
@pytest.mark.parametrize("storage", [_abi.F64, _abi.F32])
@pytest.mark.parametrize("collision", [_abi.CUM, _abi.KBC])
def test_carreau_yasuda_neq(gpu_lib, storage, collision, material):
    """Local Carreau-Yasuda viscosity from the non-equilibrium strain, sets A
    (a = 2, n = 0.3568) and B (a = 0.64, n = 0.2128: the strongest shear
    thinning, where the FP32 kernel's two Newton steps carry the most
    weight)."""
    from nonnewlbm_amd.materials import case_kwargs, material_by_name
    g, kw = tcpc_case(0.25, storage=storage, collision=collision)
    kw.update(case_kwargs(material_by_name(material), g.dl, 0.01))
    gpu, cpu = _pair(g, kw)
    act = _active(cpu)
    gpu.step(2)
    cpu.step(2)
    if storage == _abi.F64:
        assert_populations_match(gpu.populations(), cpu.populations(), act, 2, 1e-12)
    gpu.step(150)
    cpu.step(150)
    tol = 1e-8 if storage == _abi.F64 else 1e-3
    for name in ("rho", "vx", "vy", "vz"):
        assert rel_l2(gpu.field(name), cpu.field(name), act) < tol, name


@pytest.mark.parametrize("profile", [_abi.PROFILE_PLUG, _abi.PROFILE_PARABOLIC])
def test_pulsatile_inflow(gpu_lib, profile):
    """u_in(t) = u (1 + A sin(2 pi t / T)), t the step index (the per-step
    inflow hook, sim_tcpc.cu:657-663, driven inside the kernel)."""
    g, kw = tcpc_case(0.25, storage=_abi.F64, pulse_amp=0.3, pulse_period=40.0, profile=profile,
                      ivc_umax=0.04, svc_umax=0.03, ivc_center=(-1.0, -1.0), svc_center=(-1.0, -1.0))
    gpu, cpu = _pair(g, kw)
    act = _active(cpu)
    lab = cpu.field("label")
    done = 0
    for steps in (1, 2, 7):
        gpu.step(steps)
        cpu.step(steps)
        done += steps
        assert_populations_match(gpu.populations(), cpu.populations(), act, done, 1e-12)
    # the inlet velocity follows the pulse: step 10 (index 9) of period 40
    vx = gpu.field("vx")[lab == 3]
    gpu2 = Case(g.flat(), g.dims, **dict(kw, pulse_amp=0.0))
    gpu2.step(10)
    np.testing.assert_allclose(vx, gpu2.field("vx")[lab == 3] * (1 + 0.3 * np.sin(2 * np.pi * 9 / 40)), rtol=1e-12)
    gpu.step(90)
    cpu.step(90)
    for name in ("rho", "vx", "vy", "vz"):
        assert rel_l2(gpu.field(name), cpu.field(name), act) < 1e-9, name


@pytest.mark.parametrize("storage", [_abi.F64, _abi.F32])
def test_carreau_yasuda_fd_force(gpu_lib, storage):
    """The reference's multi-rank Carreau-Yasuda algorithm (core.h:293-384,
    407-576, 663-781): velocity pre-pass, FD strain, body force.  nu is
    evaluated in single precision as the reference's powf, on both sides as the
    correctly rounded float of the double pow."""
    from nonnewlbm_amd.materials import case_kwargs, material_by_name
    g, kw = tcpc_case(0.25, storage=storage)
    kw.update(case_kwargs(material_by_name("sx"), g.dl, 0.02, cy_model=_abi.CY_FD_FORCE))
    gpu = Case(g.flat(), g.dims, **kw)
    cpu = OracleCase(g.flat(), g.dims, **dict(kw, storage=_abi.F64))
    act = _active(cpu)
    fl = cpu.field("label") == 0
    gpu.step(2)
    cpu.step(2)
    tol = 1e-12 if storage == _abi.F64 else 1e-3
    for name in ("s11", "s12", "s33", "fx", "fy", "fz"):
        a, b = gpu.field(name)[fl], cpu.field(name)[fl]
        assert np.max(np.abs(a - b)) <= tol * max(np.max(np.abs(b)), 1e-30), name
    assert np.abs(cpu.field("fx")[fl]).max() > 0
    gpu.step(100)
    cpu.step(100)
    tol = 1e-10 if storage == _abi.F64 else 1e-3
    for name in ("rho", "vx", "vy", "vz", "fx"):
        assert rel_l2(gpu.field(name), cpu.field(name), act if name != "fx" else fl) < tol, name


def test_output_strain_matches_oracle(gpu_lib):
    """The output-path strain (cudaLBMKernelStress stencil, core.h:407-576)
    of the final velocity field."""
    g, kw = tcpc_case(0.25, storage=_abi.F64)
    gpu, cpu = _pair(g, kw)
    gpu.step(30)
    cpu.step(30)
    gpu.compute_strain()
    cpu.compute_strain()
    fl = cpu.field("label") == 0
    for name in ("s11", "s12", "s13", "s22", "s32", "s33"):
        np.testing.assert_allclose(gpu.field(name), cpu.field(name), rtol=0, atol=1e-13, err_msg=name)
    assert np.abs(cpu.field("s12")[fl]).max() > 1e-6


@pytest.mark.parametrize("storage", [_abi.F32, _abi.F16_OFFSET])
def test_nontemporal_policy_changes_nothing(gpu_lib, storage):
    """The junction x2 (above the working set where the even step streams with
    the non-temporal hint) gives the same bits with the hint forced on and
    off, and with and without a cap of 12 wavefronts per CU."""
    from nonnewlbm_amd.geometry import tcpc_mask
    g = tcpc_mask(2.0)
    kw = dict(storage=storage, collision=_abi.CUM, nu_ref=0.01, profile=_abi.PROFILE_PARABOLIC, ivc_umax=0.04,
              svc_umax=0.03, ivc_center=g.meta["ivc"]["center"], svc_center=g.meta["svc"]["center"],
              ivc_radius=g.meta["ivc"]["radius"], svc_radius=g.meta["svc"]["radius"], dl=g.dl, dt=1e-5)
    a = Case(g.flat(), g.dims, **kw)
    a.set_nontemporal(1)
    a.set_occupancy(12)
    a.step(6)
    va = a.field("vx")
    a.close()
    b = Case(g.flat(), g.dims, **kw)
    b.set_nontemporal(0)
    b.step(6)
    np.testing.assert_array_equal(va, b.field("vx"))
    assert np.abs(va).max() > 0


def test_near_4gib_case_matches_oracle(gpu_lib):
    """The largest case class that 32-bit buffer offsets allow: an all-fluid
    272x272x256 channel at FP64 stores 295,937 bricks, a 4.09 GB population
    buffer (3.81 GiB, just below the limit). Two steps, both AA parities,
    match the oracle to 1e-12 in every population of every active cell."""
    from nonnewlbm_amd.geometry import Geometry
    dims = (272, 272, 256)
    g = Geometry(np.zeros(dims, dtype=np.int32), 1.0e-3)
    kw = dict(storage=_abi.F64, collision=_abi.CUM, nu_ref=0.02, profile=_abi.PROFILE_PLUG, ivc_flux=2000.0,
              face_bc=_abi.CHANNEL_FACES, dl=1.0e-3, dt=1.0e-3)
    gpu, cpu = _pair(g, kw)
    info = gpu.info()
    assert info["bricks_stored"] * 27 * 64 * 8 > 3.9e9
    gpu.step(2)
    cpu.step(2)
    act = _active(cpu)
    pg, pc = gpu.populations(), cpu.populations()
    # as in the channel tests, direction by direction (the arrays are 4 GB
    # each): held populations to 1e-12, the unheld outward ones exactly 0
    held = stored_populations(act, 2)
    worst = 0.0
    for q in range(27):
        a, b = pg[q][act], pc[q][act]
        assert np.all(a[~held[q]] == 0.0), q
        worst = max(worst, float(np.max(np.abs(a[held[q]] - b[held[q]]))))
    assert worst < 1e-12, worst
    assert float(np.max(np.abs(pg[13][act] - 8.0 / 27.0))) > 1e-6   # the inflow has moved the state


@pytest.mark.parametrize("storage", [_abi.F64, _abi.F32])
def test_inflow_hook_mid_run(gpu_lib, storage):
    """updateKernelVelocities (sim_tcpc.cu:657-663) between steps: the caval
    fluxes change after 7 steps (odd, so the next step is even) and again
    after 12; the GPU follows the oracle through both changes."""
    g, kw = tcpc_case(0.25, storage=storage)
    gpu, cpu = _pair(g, kw)
    for n, scale in ((7, 1.7), (5, 0.4), (9, None)):
        gpu.step(n)
        cpu.step(n)
        if scale is not None:
            a, b = kw["ivc_flux"] * scale, kw["svc_flux"] * (2.0 - scale)
            gpu.set_inflow(a, b)
            cpu.set_inflow(a, b)
    act = _active(cpu)
    tol = 1e-12 if storage == _abi.F64 else 1e-3   # FP32: hardware rcp/sqrt, as in test_tcpc_models
    for name in ("rho", "vx", "vy", "vz"):
        assert rel_l2(gpu.field(name), cpu.field(name), act) < tol, name
    fl = gpu.flux(mean=False)
    # plug inflow: the IVC plane carries exactly the prescribed lattice flux
    assert fl[0] == pytest.approx(kw["ivc_flux"] * 0.4, rel=1e-5)


@pytest.mark.parametrize("collision", [_abi.CUM, _abi.BGK, _abi.MRT, _abi.KBC])
def test_tcpc_models_fp32_full_size_against_fp64(gpu_lib, collision):
    """The junction workload at its stated size and length: 256x192x128,
    20,000 steps of each collision model, FP32 storage against the FP64 path
    (the one held to the CPU oracle to 1e-12 on the smaller cases here; the
    oracle would need about 40 minutes per model at this size).  Steady
    parabolic inflows 0.04/0.03, Carreau-Yasuda set A (neq), nu_lbm = 0.01.
    Bar: 1e-3 relative L2 of velocity and of rho - 1 over the non-wall cells,
    flow split within 0.5 share points."""
    import bench
    r = bench.precision_study(1.0, 20000, collision, (_abi.F64, _abi.F32), pulse_amp=0.0)
    f32 = r["storages"]["f32"]
    print(f"model {collision}: {f32}")
    assert np.isfinite(f32["u_rel_l2"]) and f32["u_rel_l2"] < 1e-3
    assert f32["rho_minus_1_rel_l2"] < 1e-3
    assert f32["split_abs_diff_percent"] < 0.5


@pytest.mark.parametrize("collision", [_abi.CUM, _abi.BGK, _abi.MRT, _abi.KBC])
def test_tcpc_models_fp32_20000_steps(gpu_lib, collision):
    """The junction workload's stated length: 20,000 steps of each collision
    model at FP32 storage against the FP64 oracle, on the junction at scale
    0.25 (64x48x32) with parabolic inflows 0.04/0.03, Carreau-Yasuda set A
    (neq) and nu_lbm = 0.01.  Bar: 1e-3 relative L2 of velocity and density."""
    import bench
    g, kw = bench.make_case(0.25, _abi.F32, collision, 0.01)
    gpu = Case(g.flat(), g.dims, **kw)
    cpu = OracleCase(g.flat(), g.dims, **dict(kw, storage=_abi.F64))
    gpu.step(20000)
    cpu.step(20000)
    act = _active(cpu)
    vx = cpu.field("vx")
    assert np.isfinite(vx[act]).all() and np.abs(vx[act]).max() > 0.03
    for name in ("rho", "vx", "vy", "vz"):
        err = rel_l2(gpu.field(name), cpu.field(name), act)
        print(f"model {collision} {name}: {err:.3e}")
        assert err < 1e-3, (name, err)
