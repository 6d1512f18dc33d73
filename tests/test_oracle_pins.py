"""Pins of the CPU oracle (oracle/lbm_oracle.cpp) against mathematics that does
not share its code.

The reference cannot be built here (it needs TNL, VTK C++ and libpng, none of
which is in the image) and ships no golden vectors, so each part of the oracle
is checked against a property that fixes it independently:

* equilibrium: the product-form f_eq is the distribution whose central
  moments factorise as 1, 0, 1/3 per axis (lbm_eq_inv_cum.h:19-49), rebuilt
  here by inverting a 27x27 central-moment matrix; the pmm quirk (q=18 carries
  one more factor rho, lbm_eq_inv_cum.h:49) as an explicit special case;
* cumulant (lbm_col_cum.h, every rate but omega1 equal to 1): all cumulants of
  order >= 3 vanish after collision, so the post-collision central moments are
  rho times the moments of a Gaussian with the relaxed covariance (Isserlis'
  theorem), with the second-order part relaxed as Geier 2015 eqs 33-35;
* BGK / MRT_LES / KBC_N4: conservation, moment-level relaxation rates,
  the regularised Hermite rebuild, KBC's entropic orthogonality condition
  (Karlin, Boesch, Chikatamarla 2014) and the nu = 1/6 special cases;
* non-equilibrium Carreau-Yasuda viscosity: populations built from the
  Chapman-Enskog stress -2 rho c_s^2 tau S of a chosen shear rate must give
  back nu(shear rate);
* streaming / bounce-back: single-population transport on a tiny lattice;
* FD strain: exact on linear velocity fields;
* statistics and boundary planes: closed forms from the step history.
A mistake of the kind that is easy to make (a dropped term, a wrong sign or
index, a transposed operand) fails at least one of these.
"""
from __future__ import annotations

import itertools
import zlib

import numpy as np
import pytest

from nonnewlbm_amd import _abi
from nonnewlbm_amd.materials import lattice_units, material_by_name
from oracle import oracle as O

Q = np.arange(27)
CV = np.stack([Q // 9 - 1, (Q // 3) % 3 - 1, Q % 3 - 1], axis=1)   # defs.h:215-247
OPP = 26 - Q
WQ = np.array([{0: 8 / 27, 1: 2 / 27, 2: 1 / 54, 3: 1 / 216}[int((c * c).sum())] for c in CV])
RNG = np.random.default_rng(20261203)
# FP64 rounding of O(1) quantities that go through sums of 27 terms (27 ulp of
# 0.3, twice): the floor of the moment-level comparisons below
SUM27 = 4e-15


@pytest.fixture(autouse=True)
def _own_stream(request):
    """Each test draws from its own seeded stream, whatever ran before it in
    the same process (the helpers below read the module-level RNG)."""
    global RNG
    RNG = np.random.default_rng([20261203, zlib.crc32(request.node.name.encode())])


def central_matrix(u) -> np.ndarray:
    """M[9a+3b+c, q] = (cx-ux)^a (cy-uy)^b (cz-uz)^c."""
    d = CV - np.asarray(u, dtype=np.float64)
    M = np.empty((27, 27))
    for a, b, c in itertools.product(range(3), repeat=3):
        M[9 * a + 3 * b + c] = d[:, 0] ** a * d[:, 1] ** b * d[:, 2] ** c
    return M


def _pairings(items):
    if not items:
        yield []
        return
    first = items[0]
    for i in range(1, len(items)):
        rest = items[1:i] + items[i + 1:]
        for p in _pairings(rest):
            yield [(first, items[i])] + p


def gauss_moments(S: np.ndarray) -> np.ndarray:
    """E[X^a Y^b Z^c] (a,b,c <= 2) of a zero-mean Gaussian with covariance S."""
    out = np.zeros(27)
    for a, b, c in itertools.product(range(3), repeat=3):
        idx = [0] * a + [1] * b + [2] * c
        if len(idx) % 2:
            continue
        out[9 * a + 3 * b + c] = sum(np.prod([S[i, j] for i, j in p]) for p in _pairings(idx))
    return out


def feq_product(rho: float, u) -> np.ndarray:
    """Product-form equilibrium from its central moments rho * (1, 0, 1/3)^3."""
    return np.linalg.solve(central_matrix(u), rho * gauss_moments(np.eye(3) / 3.0))


def random_state(eps=0.03, umax=0.08):
    rho = 1.0 + 0.2 * (RNG.random() - 0.5)
    u = umax * (2 * RNG.random(3) - 1)
    f = feq_product(rho, u) * (1.0 + eps * RNG.standard_normal(27))
    return f


def moments2(f):
    return np.einsum("qa,qb,q->ab", CV, CV, f)


def third7(f):
    """xyy, xzz, xxy, yzz, xxz, yyz, xyz raw moments."""
    x, y, z = CV[:, 0], CV[:, 1], CV[:, 2]
    return np.array([(x * y * y * f).sum(), (x * z * z * f).sum(), (x * x * y * f).sum(), (y * z * z * f).sum(),
                     (x * x * z * f).sum(), (y * y * z * f).sum(), (x * y * z * f).sum()])


# ---------------------------------------------------------------------------
def test_weights_and_lattice():
    assert np.isclose(WQ.sum(), 1.0)
    np.testing.assert_array_equal(CV[OPP], -CV)
    np.testing.assert_allclose(O.feq(1.0, 0.0, 0.0, 0.0), WQ, rtol=0, atol=1e-15)


def test_feq_central_moments_and_pmm_quirk():
    for _ in range(20):
        u = 0.1 * (2 * RNG.random(3) - 1)
        rho = 0.8 + 0.4 * RNG.random()
        f1 = O.feq(1.0, *u)
        np.testing.assert_allclose(f1, feq_product(1.0, u), rtol=0, atol=1e-15)
        fr = O.feq(rho, *u)
        expect = rho * f1
        expect[18] *= rho   # pmm (cx,cy,cz) = (+1,-1,-1) carries rho^2
        np.testing.assert_allclose(fr, expect, rtol=1e-15, atol=0)
    assert tuple(CV[18]) == (1, -1, -1)


def test_density_velocity():
    for _ in range(20):
        f = random_state()
        F = 1e-3 * RNG.standard_normal(3)
        out = O.density_velocity(f, F)
        rho = f.sum()
        assert abs(out[0] - rho) < 1e-15 * rho * 4
        np.testing.assert_allclose(out[1:], (CV.T @ f + F / 2) / rho, rtol=0, atol=1e-15)


def _relaxed_covariance(K2, k000, om):
    tr = np.trace(K2)
    return (1 - om) * (K2 - tr / 3 * np.eye(3)) + k000 / 3 * np.eye(3)


@pytest.mark.parametrize("with_force", [False, True])
def test_cumulant_against_gaussian_closure(with_force):
    for _ in range(25):
        f = random_state()
        F = 1e-3 * RNG.standard_normal(3) if with_force else np.zeros(3)
        rho, *u = O.density_velocity(f, F)
        nu = float(RNG.choice([1.0 / 6.0, 0.003, 0.02, 0.1, 0.4]))
        om = 1.0 / (3 * nu + 0.5)
        out = O.collide(_abi.CUM, f, rho, u, F, nu)
        M = central_matrix(u)
        k = M @ f
        K2 = np.array([[k[18], k[12], k[10]], [k[12], k[6], k[4]], [k[10], k[4], k[2]]])
        K2p = _relaxed_covariance(K2, k[0], om)
        kp = rho * gauss_moments(K2p / rho)
        kp[0] = k[0]
        # forcing: first central moments change sign (lbm_col_cum.h:313-315)
        kp[9], kp[3], kp[1] = -k[9], -k[3], -k[1]
        np.testing.assert_allclose(out, np.linalg.solve(M, kp), rtol=0, atol=5e-15)
        # mass and momentum: + F exactly
        assert abs(out.sum() - f.sum()) < 1e-14
        np.testing.assert_allclose(CV.T @ out, CV.T @ f + F, rtol=0, atol=1e-15)


def test_cumulant_equilibrium_special_cases():
    for _ in range(10):
        rho = 0.9 + 0.2 * RNG.random()
        u = 0.08 * (2 * RNG.random(3) - 1)
        feq = feq_product(rho, u)
        for nu in (0.004, 1.0 / 6.0, 0.3):
            np.testing.assert_allclose(O.collide(_abi.CUM, feq, rho, u, nu=nu), feq, rtol=0, atol=SUM27)
        f = random_state()
        r, *v = O.density_velocity(f)
        np.testing.assert_allclose(O.collide(_abi.CUM, f, r, v, nu=1.0 / 6.0), feq_product(r, v), rtol=0,
                                   atol=SUM27)


def _orth_to_conserved(d):
    """Part of d with zero mass and momentum (project out 1, cx, cy, cz)."""
    B = np.column_stack([np.ones(27), CV]).astype(np.float64)
    coef, *_ = np.linalg.lstsq(B, d, rcond=None)
    return d - B @ coef


def test_bgk():
    for _ in range(20):
        f = random_state()
        F = 1e-3 * RNG.standard_normal(3)
        rho, *u = O.density_velocity(f, F)
        nu = float(RNG.choice([0.005, 0.05, 0.2]))
        om = 1.0 / (3 * nu + 0.5)
        out = O.collide(_abi.BGK, f, rho, u, F, nu)
        assert abs(out.sum() - f.sum()) < 1e-14
        # improved BGK with the Guo-like source: momentum gains F exactly
        np.testing.assert_allclose(CV.T @ out, CV.T @ f + F, rtol=0, atol=1e-15)
        # relaxation rate: a perturbation that keeps rho and u relaxes by 1 - omega
        d = _orth_to_conserved(1e-3 * WQ * RNG.standard_normal(27))
        out2 = O.collide(_abi.BGK, f + d, rho, u, F, nu)
        np.testing.assert_allclose(out2 - out, (1 - om) * d, rtol=0, atol=1e-15)
        # omega = 1 and no force: f_eq
        r0, *u0 = O.density_velocity(f)
        np.testing.assert_allclose(O.collide(_abi.BGK, f, r0, u0, nu=1.0 / 6.0), feq_product(r0, u0), atol=1e-15)


def test_mrt_les():
    C = 0.0342
    B = np.column_stack([np.ones(27), CV, *[CV[:, a] * CV[:, b] for a in range(3) for b in range(a, 3)]])
    for _ in range(20):
        f = random_state(eps=0.05)
        F = 1e-3 * RNG.standard_normal(3)
        rho, *u = O.density_velocity(f, F)
        u = np.array(u)
        nu = float(RNG.choice([0.001, 0.01, 0.1]))
        tau = 3 * nu + 0.5
        out = O.collide(_abi.MRT, f, rho, u, F, nu)
        P = moments2(f)
        Pn = P - rho * (np.eye(3) / 3 + np.outer(u, u))
        Qs = 2 * (Pn * Pn).sum()
        om = 2.0 / (np.sqrt(tau * tau + 18 * C * np.sqrt(Qs) / rho) + tau)
        assert abs(out.sum() - rho) < 1e-15
        # no force term: momentum is rho*u of the (half-force) velocity it is given
        np.testing.assert_allclose(CV.T @ out, rho * u, rtol=0, atol=1e-15)
        np.testing.assert_allclose(moments2(out), P - om * Pn, rtol=0, atol=1e-15)
        # regularised: out / w is a polynomial of degree 2 in c
        coef, *_ = np.linalg.lstsq(B, out / WQ, rcond=None)
        assert np.abs(B @ coef - out / WQ).max() < 1e-13


def test_kbc_n4():
    for _ in range(20):
        f = random_state(eps=0.05)
        rho, *u = O.density_velocity(f)
        u = np.array(u)
        nu = float(RNG.choice([0.002, 0.02, 0.1]))
        beta = 1.0 / (6 * nu + 1)
        out = O.collide(_abi.KBC, f, rho, u, nu=nu)
        feq = feq_product(rho, u)
        assert abs(out.sum() - f.sum()) < 1e-14
        np.testing.assert_allclose(CV.T @ out, CV.T @ f, rtol=0, atol=1e-15)
        # the shear part carries the whole second moment and the seven third
        # moments, which relax by 1 - 2 beta whatever the entropic gamma
        np.testing.assert_allclose(moments2(out) - moments2(feq), (1 - 2 * beta) * (moments2(f) - moments2(feq)),
                                   rtol=0, atol=SUM27)
        np.testing.assert_allclose(third7(out) - third7(feq), (1 - 2 * beta) * (third7(f) - third7(feq)),
                                   rtol=0, atol=SUM27)
        # entropic gamma (quadratic entropy): the post-collision deviation is
        # orthogonal, in the 1/f_eq product, to r = out - feq - (1-2b)(f - feq)
        r = out - feq - (1 - 2 * beta) * (f - feq)
        dev = out - feq
        # the floor is FP64 rounding of out (about 64 ulp in the same norm),
        # for draws where r itself is only a few ulp of out
        num = abs(np.sum(dev * r / feq))
        nd, nr, no = (np.sqrt(np.sum(v * v / feq)) for v in (dev, r, out))
        assert num < 1e-12 * nd * nr + 64 * np.finfo(float).eps * nd * no
        # nu = 1/6 (beta = 1/2): equilibrium; f_eq is a fixed point
        np.testing.assert_allclose(O.collide(_abi.KBC, f, rho, u, nu=1.0 / 6.0), feq, rtol=0, atol=SUM27)
        np.testing.assert_allclose(O.collide(_abi.KBC, feq, rho, u, nu=nu), feq, rtol=0, atol=SUM27)


# ---------------------------------------------------------------------------
def cy_nu(g, nu_ref, nu0, lam, a, n):
    """Carreau-Yasuda: nu = nu_inf + (nu0 - nu_inf)(1 + (g lam)^a)^((n-1)/a)."""
    return nu_ref + (nu0 - nu_ref) * (1.0 + (g * lam) ** a) ** ((n - 1.0) / a)


CY_SETS = [("cy-a", 2.5e-4, 0.01), ("cy-b", 2.5e-4, 0.01), ("sx", 1e-3, 0.05), ("gx", 1e-3, 0.05)]


@pytest.mark.parametrize("name,dl,nu_lbm", CY_SETS)
def test_neq_viscosity_recovers_shear_rate(name, dl, nu_lbm):
    lu = lattice_units(material_by_name(name), dl, nu_lbm)
    par = (nu_lbm, lu.nu0, lu.lam, lu.a, lu.n)
    E = np.array([[0.3, 0.5, -0.2], [0.5, -0.7, 0.1], [-0.2, 0.1, 0.4]])
    E -= np.trace(E) / 3 * np.eye(3)
    E /= np.sqrt((E * E).sum())
    Qt = np.einsum("qa,qb->qab", CV, CV) - np.eye(3) / 3
    for g in np.logspace(-9, -1.5, 16):
        rho = 1.02
        u = np.array([0.03, -0.01, 0.02])
        S = g / np.sqrt(2.0) * E                        # sqrt(2 S:S) = g
        tau = 3 * cy_nu(g, *par) + 0.5
        Pi = -2.0 * rho * tau / 3.0 * S                 # Chapman-Enskog stress
        f = feq_product(rho, u) + 4.5 * WQ * np.einsum("qab,ab->q", Qt, Pi)
        nu = O.neq_nu(f, rho, u, *par)
        assert nu == pytest.approx(cy_nu(g, *par), rel=2e-9), (name, g)


@pytest.mark.parametrize("name,dl,nu_lbm", CY_SETS)
def test_neq_solver_against_bisection(name, dl, nu_lbm):
    lu = lattice_units(material_by_name(name), dl, nu_lbm)
    par = (nu_lbm, lu.nu0, lu.lam, lu.a, lu.n)
    for G in np.logspace(-9, 0, 19):
        lo, hi = -80.0, 10.0      # log2(gamma)
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            g = 2.0 ** mid
            if g * (3 * cy_nu(g, *par) + 0.5) < G:
                lo = mid
            else:
                hi = mid
        expect = cy_nu(2.0 ** (0.5 * (lo + hi)), *par)
        assert O.neq_nu_from_G(G, *par) == pytest.approx(expect, rel=2e-9), (name, G)
    assert O.neq_nu_from_G(0.0, *par) == pytest.approx(cy_nu(0.0, *par), rel=1e-15)


# ---------------------------------------------------------------------------
def _box_case(dims, walls=(), storage=_abi.F64, **kw):
    X, Y, Z = dims
    mask = np.zeros(dims, dtype=np.int32)
    for sl in walls:
        mask[sl] = 1
    kw = dict(dict(storage=storage, collision=_abi.CUM, nu_ref=0.05, face_bc=(1,) * 6, dl=1e-3, dt=1e-3), **kw)
    return O.OracleCase(mask.transpose(2, 1, 0).ravel(), dims, **kw), mask


def test_streaming_single_population_transport():
    dims = (9, 8, 7)
    case, _ = _box_case(dims, walls=[np.s_[5:7, 2:5, 2:4]])
    lab = case.field("label")
    N = np.prod(dims)
    ids = (1.0 + np.arange(27 * N, dtype=np.float64)).reshape(27, *dims)
    case.set_populations(ids)
    case.stream_only(1)
    p1 = case.populations()
    case.stream_only(1)
    p2 = case.populations()
    n_fluid = n_wall = 0
    for x, y, z in zip(*np.nonzero(lab == 0)):
        for q in range(27):
            t = (x + CV[q, 0], y + CV[q, 1], z + CV[q, 2])
            if lab[t] == 0:
                assert p1[(q,) + t] == ids[q, x, y, z]
                n_fluid += 1
            else:  # full-way bounce-back: reversed inside the wall, back after 2 steps
                assert p1[(OPP[q],) + t] == ids[q, x, y, z]
                assert p2[OPP[q], x, y, z] == ids[q, x, y, z]
                n_wall += 1
    assert n_fluid > 1000 and n_wall > 500


def test_fd_strain_exact_on_linear_fields():
    dims = (10, 9, 8)
    # a one-cell sheet at x=7 has no fluid neighbour along x
    case, _ = _box_case(dims, walls=[np.s_[6, :, :], np.s_[8, :, :], np.s_[3, 4, 3]])
    lab = case.field("label")
    A = RNG.standard_normal((3, 3))
    b = RNG.standard_normal(3)
    xs = np.stack(np.meshgrid(*[np.arange(n) for n in dims], indexing="ij"), axis=-1).astype(np.float64)
    u = xs @ A.T + b
    case.set_velocity(u[..., 0], u[..., 1], u[..., 2])
    case.compute_strain()
    S = {k: case.field(k) for k in ("s11", "s12", "s13", "s22", "s32", "s33")}
    fl = lab == 0
    pad = np.pad(fl, 1, constant_values=False)
    X, Y, Z = dims
    checked = no_x = 0
    for x, y, z in zip(*np.nonzero(fl)):
        has = [pad[x, y + 1, z + 1] or pad[x + 2, y + 1, z + 1], pad[x + 1, y, z + 1] or pad[x + 1, y + 2, z + 1],
               pad[x + 1, y + 1, z] or pad[x + 1, y + 1, z + 2]]
        D = np.array([A[:, a] if has[a] else np.zeros(3) for a in range(3)])   # D[a, c] = d u_c / d x_a
        E = 0.5 * (D + D.T)
        got = np.array([[S["s11"][x, y, z], S["s12"][x, y, z], S["s13"][x, y, z]],
                        [S["s12"][x, y, z], S["s22"][x, y, z], S["s32"][x, y, z]],
                        [S["s13"][x, y, z], S["s32"][x, y, z], S["s33"][x, y, z]]])
        np.testing.assert_allclose(got, E, rtol=0, atol=1e-12)
        checked += 1
        no_x += not has[0]
    assert checked > 200 and no_x > 20
    for k in S:
        assert np.all(S[k][~fl] == 0.0)


def test_fd_force_closed_form():
    """fd-force body force (core.h:663-781) from a state with u_x = a (y - y0)^2:
    central differences are exact on it, so S12 = a (y - y0), div S = (a, 0, 0),
    gamma = 2 |a (y - y0)| and F = 2 (nu(gamma) - nu_ref) a with the textbook
    Carreau-Yasuda nu, rho = 1, away from the walls."""
    dims = (8, 16, 8)
    mat = material_by_name("cy-a")
    u = lattice_units(mat, 0.25e-3, 0.01)
    kw = dict(dl=0.25e-3, dt=u.dt, nu_ref=0.01, cy_enabled=1, cy_model=_abi.CY_FD_FORCE, cy_nu0=u.nu0,
              cy_lambda=u.lam, cy_a=u.a, cy_n=u.n)
    case, _ = _box_case(dims, **kw)
    lab = case.field("label")
    a, y0 = 1.0e-3, 7.25
    pops = np.empty((27,) + dims)
    for y in range(dims[1]):
        pops[:, :, y, :] = O.feq(1.0, a * (y - y0) ** 2, 0.0, 0.0)[:, None, None]
    case.set_populations(pops)
    case.step(1)
    fx, fy, fz = (case.field(k) for k in ("fx", "fy", "fz"))
    checked = 0
    for x, y, z in zip(*np.nonzero(lab == 0)):
        if not (2 <= x < dims[0] - 2 and 3 <= y < dims[1] - 3 and 2 <= z < dims[2] - 2):
            continue
        assert np.all(lab[x - 1:x + 2, y - 2:y + 3, z - 1:z + 2] == 0)
        gam = 2.0 * abs(a * (y - y0))
        nu = 0.01 + (u.nu0 - 0.01) * (1.0 + (gam * u.lam) ** u.a) ** ((u.n - 1.0) / u.a)
        np.testing.assert_allclose(fx[x, y, z], 2.0 * (nu - 0.01) * a, rtol=1e-5, atol=0)
        assert abs(fy[x, y, z]) < 1e-15 and abs(fz[x, y, z]) < 1e-15
        checked += 1
    assert checked >= 4 * 10 * 4
    # the shear-thinning part is what the force carries here
    assert 0.0 < fx[4, 3, 4] < fx[4, 7, 4]


def test_label_precedence_and_extrusion():
    dims = (8, 7, 6)
    case, _ = _box_case(dims, face_bc=(0,) * 6)
    lab = case.field("label")
    # sim_tcpc.cu:522-527: x planes, then y planes, then z planes overwrite
    assert lab[0, 3, 3] == 3 and lab[7, 3, 3] == 4
    assert lab[3, 0, 3] == 5 and lab[3, 6, 3] == 6 and lab[0, 0, 3] == 5 and lab[7, 6, 3] == 6
    assert lab[3, 3, 0] == 1 and lab[0, 3, 5] == 1
    # extrusion (sim_tcpc.cu:547-584): x = 0,1 walled, x = 2 open in a disc
    mask = np.ones(dims, dtype=np.int32)
    mask[2:, 2:5, 2:4] = 0
    mask[4, :, 2:4] = 0   # pulmonary opening through y at x = 4
    kw = dict(storage=_abi.F64, nu_ref=0.05, dl=1e-3, dt=1e-3)
    o = O.OracleCase(mask.transpose(2, 1, 0).ravel(), dims, **kw)
    lab = o.field("label")
    np.testing.assert_array_equal(lab[0, 2:5, 2:4], 3)
    assert o.info()["area_ivc"] == 6 and not o.info()["terminate"]


def test_statistics_sums_closed_form():
    from tests.helpers import tcpc_case
    for every in (1, 3):
        g, kw = tcpc_case(0.25, storage=_abi.F64, stats_every=every)
        o = O.OracleCase(g.flat(), g.dims, **kw)
        acc = {k: 0.0 for k in ("sum_rux", "sum_ruy", "sum_ruz", "sum_rruxx", "sum_rruyy", "sum_rruzz")}
        for it in range(1, 8):
            o.step(1)
            if it % every:
                continue
            r = o.field("rho")
            for c, name in zip("xyz", ("vx", "vy", "vz")):
                ru = r * o.field(name)
                acc[f"sum_ru{c}"] = acc[f"sum_ru{c}"] + ru
                acc[f"sum_rru{c * 2}"] = acc[f"sum_rru{c * 2}"] + ru * ru
        for k, v in acc.items():
            np.testing.assert_allclose(o.field(k), v, rtol=1e-13, atol=1e-18, err_msg=f"{k} every={every}")


def test_boundary_planes_closed_form():
    from tests.helpers import tcpc_case
    g, kw = tcpc_case(0.25, storage=_abi.F64)
    o = O.OracleCase(g.flat(), g.dims, **kw)
    o.step(3)
    info = o.info()
    lab, vx, vy, vz, rho = (o.field(k) for k in ("label", "vx", "vy", "vz", "rho"))
    # plug inflow u = flux / area (sim_tcpc.cu:295-309), outlets rho = 1
    np.testing.assert_allclose(vx[lab == 3], kw["ivc_flux"] / info["area_ivc"], rtol=1e-15)
    np.testing.assert_allclose(vx[lab == 4], -kw["svc_flux"] / info["area_svc"], rtol=1e-15)
    assert np.all(vy[(lab == 3) | (lab == 4)] == 0) and np.all(vz[(lab == 3) | (lab == 4)] == 0)
    assert np.all(rho[(lab == 5) | (lab == 6)] == 1.0)
    # parabolic profile: u_max (1 - r^2 / R^2) about the given centre
    c, R = g.meta["ivc"]["center"], g.meta["ivc"]["radius"]
    g2, kw2 = tcpc_case(0.25, storage=_abi.F64, profile=_abi.PROFILE_PARABOLIC, ivc_umax=0.04, svc_umax=0.03,
                        ivc_center=c, ivc_radius=R, svc_center=g.meta["svc"]["center"],
                        svc_radius=g.meta["svc"]["radius"])
    o2 = O.OracleCase(g2.flat(), g2.dims, **kw2)
    o2.step(1)
    lab2, vx2 = o2.field("label"), o2.field("vx")
    ys, zs = np.nonzero(lab2[0] == 3)
    r2 = ((ys - c[0]) ** 2 + (zs - c[1]) ** 2) / R ** 2
    np.testing.assert_allclose(vx2[0, ys, zs], 0.04 * np.maximum(0.0, 1 - r2), rtol=1e-14, atol=1e-18)


# inward normal of each boundary-plane label: the partner cell whose incoming
# populations a BC cell pulls sits one plane inwards (sim_tcpc.cu:242-273)
_INWARD = {3: (1, 0, 0), 4: (-1, 0, 0), 5: (0, 1, 0), 6: (0, -1, 0), 7: (-1, 0, 0), 8: (1, 0, 0)}
_OUTFLOW_FACES = (_abi.FACE_OUTFLOW,) * 4 + (_abi.FACE_WALL,) * 2


@pytest.mark.parametrize("faces", [(0,) * 6, _OUTFLOW_FACES, _abi.CHANNEL_FACES], ids=["tcpc", "outflows", "channel"])
def test_boundary_partner_reads_tagged(faces):
    """The non-local reads of the boundary planes, pinned by tagged populations.

    Restated independently of the oracle's index arrays: a cell C of label L
    pulls, for each direction q, the population f_q that streams into its
    partner P = C + n_L (n_L the plane's inward normal) from P - c_q, every
    coordinate clamped into the grid (sim_tcpc.cu:246,256,263,269 with the
    clamped neighbour ids of core.h:651-661).  Inflow cells (3/4) then carry
    rho = sum_q f_q; outflow cells (5/6 and the x outflows 7/8) carry
    u = sum_q c_q f_q / rho and rho = 1.  Every population holds a distinct
    random integer below 2^20, so the sums are exact and a read from any
    other cell shows as a different value; the edge cells where the clamp
    changes a source are counted and must be present."""
    dims = (9, 8, 7)
    X, Y, Z = dims
    mask = np.zeros(dims, dtype=np.int32)
    mask[4:6, 3:5, 2:4] = 1   # an interior wall block: some sources hold bounce-back slots
    kw = dict(storage=_abi.F64, collision=_abi.CUM, nu_ref=0.05, face_bc=faces, dl=1e-3, dt=1e-3,
              ivc_flux=1.0, svc_flux=1.0)
    case = O.OracleCase(mask.transpose(2, 1, 0).ravel(), dims, **kw)
    lab = case.field("label").astype(int)
    ids = RNG.choice(np.arange(1, 1 << 20), size=27 * X * Y * Z, replace=False).astype(np.float64)
    ids = ids.reshape(27, *dims)
    case.set_populations(ids)
    case.step(1)
    rho, ux, uy, uz = (case.field(k) for k in ("rho", "vx", "vy", "vz"))
    hi = np.array(dims) - 1
    checked = {L: 0 for L in _INWARD}
    clamped = 0
    for C in zip(*np.nonzero(np.isin(lab, list(_INWARD)))):
        L = int(lab[C])
        P = np.array(C) + np.array(_INWARD[L])
        src = P[None, :] - CV
        clamped += bool(np.any((src < 0) | (src > hi)))
        src = np.clip(src, 0, hi)
        f = ids[Q, src[:, 0], src[:, 1], src[:, 2]]
        r, J = f.sum(), CV.T @ f
        if L in (3, 4):
            assert rho[C] == r, (C, L)
        else:
            assert rho[C] == 1.0, (C, L)
            assert (ux[C], uy[C], uz[C]) == tuple(J / r), (C, L)
        checked[L] += 1
    present = {L for L, n in checked.items() if n}
    expect = {(0,) * 6: {3, 4, 5, 6}, _OUTFLOW_FACES: {5, 6, 7, 8}, _abi.CHANNEL_FACES: {3, 7}}[faces]
    assert present == expect
    # y-plane cells on the x faces (tcpc, outflows) have an x-clamped source
    assert clamped > 0 or faces == _abi.CHANNEL_FACES


@pytest.mark.slow
def test_steady_mass_balance():
    """IVC + SVC inflow equals the label-5/6 outflow once the flow is steady."""
    from tests.helpers import tcpc_case
    g, kw = tcpc_case(0.25, storage=_abi.F64, nu_ref=0.05)
    o = O.OracleCase(g.flat(), g.dims, **kw)
    o.step(2500)
    lab, rho, vx, vy = (o.field(k) for k in ("label", "rho", "vx", "vy"))
    # momentum through the planes one cell inside each boundary
    inflow = (rho * vx)[1][lab[0] == 3].sum() - (rho * vx)[-2][lab[-1] == 4].sum()
    outflow = -(rho * vy)[:, 1][lab[:, 0] == 5].sum() + (rho * vy)[:, -2][lab[:, -1] == 6].sum()
    assert inflow > 0
    assert outflow == pytest.approx(inflow, rel=0.03)
