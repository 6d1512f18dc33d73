// lbm_oracle.cpp - CPU oracle for the lbm3d D3Q27 step behind solver/sim_tcpc.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: a plain restatement of
// the reference algorithm (dense A/B pull over every cell, one OpenMP loop per
// step, exactly as the reference's Host build runs it at
// solver/lbm3d/core.h:1313-1320).  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load it.  The product (nonnewlbm_amd/) never does.
//
// Each function names the reference lines it restates.  The reference itself
// cannot be compiled here (TNL, VTK C++ and libpng are absent) and ships no
// golden vectors, so this restatement is pinned by the numpy checks in
// tests/test_oracle_pins.py: central moments by 27x27 matrices, the Gaussian
// (Isserlis) closure of the cumulant collision, closed forms, conservation,
// fixed points, the Chapman-Enskog stress of a known shear rate, the fd-force
// body force of a quadratic velocity field in closed form, and brute-force
// transport on tiny lattices.
//
// Parity unpinned at the bit level: the rounding sequence of the reference's
// own cumulant and KBC code (the closed forms pin them to rounding only).
//
// Deviations from the reference, all deliberate:
//  - statistics sums are kept in double for every storage type (the reference
//    keeps them in dreal, sim_tcpc.cu:347-353);
//  - e_vm_plus_x and e_vm_minus_x are kept as their sum (only the sum is ever
//    read, sim_tcpc.cu:724-725);
//  - extensions the reference does not have: parabolic / pulsatile inflow,
//    the CY "neq" local-viscosity mode, BGK/MRT/KBC selectable at run time.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <omp.h>

#include "../include/lbm355.h"

namespace {

thread_local std::string g_err;

enum : int { GEO_FLUID = 0, GEO_WALL = 1, GEO_IN_L = 3, GEO_IN_R = 4, GEO_OUT_F = 5, GEO_OUT_B = 6 };

// D3Q27 ordering, defs.h:215-247: q = 9(cx+1) + 3(cy+1) + (cz+1)
inline int CX(int q) { return q / 9 - 1; }
inline int CY(int q) { return (q / 3) % 3 - 1; }
inline int CZ(int q) { return q % 3 - 1; }

template <class T>
struct Cell {
    T f[27];
    T rho = 1, vx = 0, vy = 0, vz = 0;
    T fx = 0, fy = 0, fz = 0;
    T nu = 0;  // viscosity used by the collision
};

// ---------------------------------------------------------------------------
// moments: computeDensityAndVelocity, lbm_common.h:16-39.  The reference sums
// rho in Geier's paired order; the pairing below is the same set of pairs.
template <class T>
void density_velocity(Cell<T>& k)
{
    const T* f = k.f;
    auto P = [&](int a, int b) { return f[a] + f[b]; };
    // pairs of opposite directions (q, 26-q)
    T corners = (P(26, 0) + P(20, 6)) + (P(24, 2) + P(8, 18));
    T edges = ((P(17, 9) + P(15, 11)) + (P(23, 3) + P(21, 5))) + (P(25, 1) + P(19, 7));
    T faces = (P(22, 4) + P(16, 10)) + P(14, 12);
    k.rho = ((corners + edges) + faces) + f[13];
    T mx = 0, my = 0, mz = 0;
    for (int q = 0; q < 27; ++q) {
        mx += CX(q) * f[q];
        my += CY(q) * f[q];
        mz += CZ(q) * f[q];
    }
    k.vx = (mx + k.fx * T(0.5)) / k.rho;
    k.vy = (my + k.fy * T(0.5)) / k.rho;
    k.vz = (mz + k.fz * T(0.5)) / k.rho;
}

// product-form ("inverse cumulant") equilibrium, lbm_eq_inv_cum.h:19-49, with
// the extra rho factor of feq_pmm (lbm_eq_inv_cum.h:49) reproduced.
template <class T>
T phi(int c, T u)
{
    // 1-D factors: c=0: (2 - 3u^2)/3 ; c=+-1: (3u^2 +- 3u + 1)/6
    if (c == 0) return (T(2) - T(3) * u * u) / T(3);
    return (T(3) * u * u + T(3) * c * u + T(1)) / T(6);
}
template <class T>
void feq_invcum(T rho, T ux, T uy, T uz, T* out)
{
    for (int q = 0; q < 27; ++q)
        out[q] = rho * phi(CX(q), ux) * phi(CY(q), uy) * phi(CZ(q), uz);
    out[18] *= rho;  // pmm = q 18 (cx=+1, cy=-1, cz=-1)
}

// ---------------------------------------------------------------------------
// Cumulant collision, lbm_col_cum.h:12-455, default build (no 2017 limiter,
// no antialiasing, defs.h:163-166): omega1 = 1/(3 nu + 1/2), every other
// relaxation rate 1, A=B=0.  Chimera transform z, y, x (Geier 2015 eqs 6-14).
template <class T>
void chimera_fwd(T* m, int stride, int outer_stride, int inner_stride, T u)
{
    // transforms the triples m[b], m[b+stride], m[b+2 stride] for 9 bases
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            int base = a * outer_stride + b * inner_stride;
            T fm = m[base], f0 = m[base + stride], fp = m[base + 2 * stride];
            T k0 = (fp + fm) + f0;
            T k1 = (fp - fm) - u * k0;
            T k2 = (fp + fm) - T(2) * u * (fp - fm) + u * u * k0;
            m[base] = k0;
            m[base + stride] = k1;
            m[base + 2 * stride] = k2;
        }
}
template <class T>
void chimera_bwd(T* m, int stride, int outer_stride, int inner_stride, T u)
{
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            int base = a * outer_stride + b * inner_stride;
            T k0 = m[base], k1 = m[base + stride], k2 = m[base + 2 * stride];
            T f0 = k0 * (T(1) - u * u) - T(2) * u * k1 - k2;
            T fm = (k0 * (u * u - u) + k1 * (T(2) * u - T(1)) + k2) * T(0.5);
            T fp = (k0 * (u * u + u) + k1 * (T(2) * u + T(1)) + k2) * T(0.5);
            m[base] = fm;
            m[base + stride] = f0;
            m[base + 2 * stride] = fp;
        }
}

// central-moment index: K(a,b,c) with a along x, b along y, c along z
inline int K(int a, int b, int c) { return 9 * a + 3 * b + c; }

template <class T>
void collide_cum(Cell<T>& k)
{
    T m[27];
    for (int q = 0; q < 27; ++q) m[q] = k.f[q];
    // along z (stride 1), then y (stride 3), then x (stride 9)
    chimera_fwd(m, 1, 9, 3, k.vz);
    chimera_fwd(m, 3, 9, 1, k.vy);
    chimera_fwd(m, 9, 3, 1, k.vx);
    const T rho = k.rho;
    auto k_ = [&](int a, int b, int c) { return m[K(a, b, c)]; };
    // 4th..6th order cumulants are relaxed with rate 1, i.e. to zero; the
    // 2nd order ones as Geier 2015 eqs 33-35 with omega2 = 1.
    const T om = T(1) / (T(3) * k.nu + T(0.5));
    const T k000 = k_(0, 0, 0);
    const T c110 = (T(1) - om) * k_(1, 1, 0);
    const T c101 = (T(1) - om) * k_(1, 0, 1);
    const T c011 = (T(1) - om) * k_(0, 1, 1);
    const T dxy = (T(1) - om) * (k_(2, 0, 0) - k_(0, 2, 0));
    const T dxz = (T(1) - om) * (k_(2, 0, 0) - k_(0, 0, 2));
    const T c200 = (dxy + dxz + k000) / T(3);
    const T c020 = (-T(2) * dxy + dxz + k000) / T(3);
    const T c002 = (dxy - T(2) * dxz + k000) / T(3);
    // post-collision central moments (Geier 2015 eqs 81-84, 3rd-order = 0)
    T n[27];
    for (int i = 0; i < 27; ++i) n[i] = T(0);
    n[K(0, 0, 0)] = k000;
    n[K(1, 0, 0)] = -k_(1, 0, 0);  // forcing: lbm_col_cum.h:313-315
    n[K(0, 1, 0)] = -k_(0, 1, 0);
    n[K(0, 0, 1)] = -k_(0, 0, 1);
    n[K(1, 1, 0)] = c110;
    n[K(1, 0, 1)] = c101;
    n[K(0, 1, 1)] = c011;
    n[K(2, 0, 0)] = c200;
    n[K(0, 2, 0)] = c020;
    n[K(0, 0, 2)] = c002;
    n[K(2, 1, 1)] = (c200 * c011 + T(2) * c101 * c110) / rho;
    n[K(1, 2, 1)] = (c020 * c101 + T(2) * c110 * c011) / rho;
    n[K(1, 1, 2)] = (c002 * c110 + T(2) * c011 * c101) / rho;
    n[K(2, 2, 0)] = (c020 * c200 + T(2) * c110 * c110) / rho;
    n[K(0, 2, 2)] = (c002 * c020 + T(2) * c011 * c011) / rho;
    n[K(2, 0, 2)] = (c200 * c002 + T(2) * c101 * c101) / rho;
    n[K(2, 2, 2)] = (c200 * n[K(0, 2, 2)] + c020 * n[K(2, 0, 2)] + c002 * n[K(2, 2, 0)] +
                     T(4) * (c011 * n[K(2, 1, 1)] + c101 * n[K(1, 2, 1)] + c110 * n[K(1, 1, 2)])) / rho -
                    (T(16) * c110 * c101 * c011 +
                     T(4) * (c101 * c101 * c020 + c011 * c011 * c200 + c110 * c110 * c002) +
                     T(2) * c200 * c020 * c002) / (rho * rho);
    // backward: x, then y, then z (Geier 2015 eqs 88-96)
    chimera_bwd(n, 9, 3, 1, k.vx);
    chimera_bwd(n, 3, 9, 1, k.vy);
    chimera_bwd(n, 1, 9, 3, k.vz);
    for (int q = 0; q < 27; ++q) k.f[q] = n[q];
}

// improved BGK with the product-form equilibrium and Guo-like source,
// lbm_col_bgk.h:14-140 (no Galilean correction)
template <class T>
void collide_bgk(Cell<T>& k)
{
    const T om = T(1) / (T(3) * k.nu + T(0.5));
    for (int q = 0; q < 27; ++q) {
        T feq = k.rho * phi(CX(q), k.vx) * phi(CY(q), k.vy) * phi(CZ(q), k.vz);
        T S = T(3) * ((CX(q) - k.vx) * k.fx + (CY(q) - k.vy) * k.fy + (CZ(q) - k.vz) * k.fz) / k.rho;
        k.f[q] += (feq - k.f[q]) * om + (T(1) - T(0.5) * om) * S * feq;
    }
}

inline double wq(int q)
{
    int c2 = CX(q) * CX(q) + CY(q) * CY(q) + CZ(q) * CZ(q);
    return c2 == 0 ? 8.0 / 27.0 : c2 == 1 ? 2.0 / 27.0 : c2 == 2 ? 1.0 / 54.0 : 1.0 / 216.0;
}

// regularised collision with a Smagorinsky rate ("MRT_LES"), lbm_col_mrt.h:14-77
template <class T>
void collide_mrt(Cell<T>& k)
{
    T P[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int q = 0; q < 27; ++q) {
        int c[3] = {CX(q), CY(q), CZ(q)};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) P[a][b] += T(c[a] * c[b]) * k.f[q];
    }
    T u[3] = {k.vx, k.vy, k.vz};
    T Pn[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Pn[a][b] = P[a][b] - k.rho * ((a == b ? T(1) / T(3) : T(0)) + u[a] * u[b]);
    T Q = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Q += Pn[a][b] * Pn[a][b];
    Q *= T(2);
    const T tau = T(3) * k.nu + T(0.5);
    const T C = T(0.0342);
    const T om = T(2) / (std::sqrt(tau * tau + T(18) * C * std::sqrt(Q) / k.rho) + tau);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) P[a][b] -= om * Pn[a][b];
    const T trP = P[0][0] + P[1][1] + P[2][2];
    for (int q = 0; q < 27; ++q) {
        int c[3] = {CX(q), CY(q), CZ(q)};
        int c2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
        T cc = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) cc += T(c[a] * c[b]) * P[a][b];
        T cu = c[0] * u[0] + c[1] * u[1] + c[2] * u[2];
        k.f[q] = T(wq(q)) * (k.rho * (T(2.5) - T(1.5) * c2 + T(3) * cu) + T(4.5) * cc - T(1.5) * trP);
    }
}

// KBC N4 (raw moments, shear part D+T+Q), lbm_col_kbc_n.h:21-289,1024-1259
template <class T>
void kbc_shear(const T M[3][3], const T Q3[10], T* s)
{
    // Q3: xyy, xzz, xxy, yzz, xxz, yyz, xyz (indices 0..6)
    const T Nxz = M[0][0] - M[2][2], Nyz = M[1][1] - M[2][2];
    const T Tr = M[0][0] + M[1][1] + M[2][2];
    for (int q = 0; q < 27; ++q) {
        int c[3] = {CX(q), CY(q), CZ(q)};
        int nz = (c[0] != 0) + (c[1] != 0) + (c[2] != 0);
        T v = 0;
        if (nz == 0) {
            v = -Tr;
        } else if (nz == 1) {
            T D;
            if (c[0]) D = (T(2) * Nxz - Nyz) / T(6);
            else if (c[1]) D = (-Nxz + T(2) * Nyz) / T(6);
            else D = (-Nxz - Nyz) / T(6);
            T Qv;
            if (c[0]) Qv = -T(c[0]) * (Q3[0] + Q3[1]) / T(2);
            else if (c[1]) Qv = -T(c[1]) * (Q3[2] + Q3[3]) / T(2);
            else Qv = -T(c[2]) * (Q3[4] + Q3[5]) / T(2);
            v = D + Tr / T(6) + Qv;
        } else if (nz == 2) {
            T D, Qv;
            if (!c[2]) {
                D = T(c[0] * c[1]) * M[0][1] / T(4);
                Qv = (T(c[0]) * Q3[0] + T(c[1]) * Q3[2]) / T(4);
            } else if (!c[1]) {
                D = T(c[0] * c[2]) * M[0][2] / T(4);
                Qv = (T(c[0]) * Q3[1] + T(c[2]) * Q3[4]) / T(4);
            } else {
                D = T(c[1] * c[2]) * M[1][2] / T(4);
                Qv = (T(c[1]) * Q3[3] + T(c[2]) * Q3[5]) / T(4);
            }
            v = D + Qv;
        } else {
            v = T(c[0] * c[1] * c[2]) * Q3[6] / T(8);
        }
        s[q] = v;
    }
}

template <class T>
void collide_kbc(Cell<T>& k)
{
    T feq[27];
    for (int q = 0; q < 27; ++q) feq[q] = k.rho * phi(CX(q), k.vx) * phi(CY(q), k.vy) * phi(CZ(q), k.vz);
    T M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    T Q3[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < 27; ++q) {
        int c[3] = {CX(q), CY(q), CZ(q)};
        T f = k.f[q];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) M[a][b] += T(c[a] * c[b]) * f;
        Q3[0] += T(c[0] * c[1] * c[1]) * f;
        Q3[1] += T(c[0] * c[2] * c[2]) * f;
        Q3[2] += T(c[0] * c[0] * c[1]) * f;
        Q3[3] += T(c[1] * c[2] * c[2]) * f;
        Q3[4] += T(c[0] * c[0] * c[2]) * f;
        Q3[5] += T(c[1] * c[1] * c[2]) * f;
        Q3[6] += T(c[0] * c[1] * c[2]) * f;
    }
    // equilibrium moments (lbm_col_kbc_n.h:38-50)
    const T r = k.rho, ux = k.vx, uy = k.vy, uz = k.vz, cs = T(1) / T(3);
    T Me[3][3] = {{r * (cs + ux * ux), r * ux * uy, r * ux * uz},
                  {r * ux * uy, r * (cs + uy * uy), r * uy * uz},
                  {r * ux * uz, r * uy * uz, r * (cs + uz * uz)}};
    T Qe[10] = {r * ux * (cs + uy * uy), r * ux * (cs + uz * uz), r * uy * (cs + ux * ux),
                r * uy * (cs + uz * uz), r * uz * (cs + ux * ux), r * uz * (cs + uy * uy),
                r * ux * uy * uz, 0, 0, 0};
    T s[27], se[27];
    kbc_shear(M, Q3, s);
    kbc_shear(Me, Qe, se);
    T ds[27], dh[27];
    T num = 0, den = 0;
    for (int q = 0; q < 27; ++q) {
        ds[q] = s[q] - se[q];
        dh[q] = k.f[q] - feq[q] - ds[q];
        num += ds[q] * dh[q] / feq[q];
        den += dh[q] * dh[q] / feq[q];
    }
    const T beta = T(1) / (T(2) * k.nu / cs + T(1));
    // 0/0 when dh == 0 exactly; any finite gamma gives the same update then
    const T gamma = den != T(0) ? T(1) / beta - (T(2) - T(1) / beta) * num / den : T(2);
    for (int q = 0; q < 27; ++q) {
        T S = T(3) * ((CX(q) - ux) * k.fx + (CY(q) - uy) * k.fy + (CZ(q) - uz) * k.fz) / r;
        k.f[q] -= beta * (T(2) * ds[q] + gamma * dh[q]) - (T(1) - beta) * S * feq[q];
    }
}

// Carreau-Yasuda lattice viscosity (core.h:766, sim_tcpc.cu:697-707)
inline double cy_nu(double gamma, double nu_ref, double nu0, double lam, double a, double n)
{
    return nu_ref + (nu0 - nu_ref) * std::pow(1.0 + std::pow(gamma * lam, a), (n - 1.0) / a);
}

// "neq" mode: gamma from the deviatoric non-equilibrium second moment P,
// gamma * tau(gamma) = G = 3 sqrt(2) |P_dev| / (2 rho) with tau = 3 nu(gamma) + 1/2.
// Solved for t = log2(gamma) by 3 Newton steps on H(t) = t + log2(tau) - log2(G)
// from the middle of the bracket [log2 G/tau_max, log2 G/tau_min]; gamma*tau is
// close to a power law, so H is nearly linear in t (converged to ~1e-14 for the
// Carreau-Yasuda sets used here).  G = 0 gives gamma = 0, nu = nu0.
inline double neq_nu_from_G(double G, double nu_ref, double nu0, double lam, double a, double n)
{
    if (!(G > 1e-30)) return cy_nu(0.0, nu_ref, nu0, lam, a, n);
    const double e = (n - 1.0) / a;
    const double lG = std::log2(G), L = std::log2(lam);
    const double tr = 3.0 * nu_ref + 0.5, t0 = 3.0 * nu0 + 0.5;
    const double lo = lG - std::log2(std::max(tr, t0)), hi = lG - std::log2(std::min(tr, t0));
    double t = 0.5 * (lo + hi);
    for (int it = 0; it < 3; ++it) {
        const double x = std::exp2(a * (L + t));
        const double b = 1.0 + x;
        const double pw = std::exp2(e * std::log2(b));
        const double tau = 3.0 * (nu_ref + (nu0 - nu_ref) * pw) + 0.5;
        const double H = t + std::log2(tau) - lG;
        const double dnu = (nu0 - nu_ref) * pw * (n - 1.0) * x / b;  // d nu / d ln(gamma)
        const double Hp = 1.0 + 3.0 * dnu / tau;
        t = std::min(std::max(t - H / Hp, lo), hi);
    }
    const double x = std::exp2(a * (L + t));
    return nu_ref + (nu0 - nu_ref) * std::exp2(e * std::log2(1.0 + x));
}

template <class T>
T neq_local_nu(const Cell<T>& k, double nu_ref, double nu0, double lam, double a, double n)
{
    double P[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int q = 0; q < 27; ++q) {
        int c[3] = {CX(q), CY(q), CZ(q)};
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) P[x][y] += double(c[x] * c[y]) * double(k.f[q]);
    }
    double u[3] = {double(k.vx), double(k.vy), double(k.vz)};
    double rho = double(k.rho);
    for (int x = 0; x < 3; ++x)
        for (int y = 0; y < 3; ++y) P[x][y] -= rho * u[x] * u[y];
    double tr = (P[0][0] + P[1][1] + P[2][2]) / 3.0;
    double s2 = 0;
    for (int x = 0; x < 3; ++x)
        for (int y = 0; y < 3; ++y) {
            double d = P[x][y] - (x == y ? tr : 0.0);
            s2 += d * d;
        }
    const double G = 3.0 * std::sqrt(2.0) * std::sqrt(s2) / (2.0 * rho);
    return T(neq_nu_from_G(G, nu_ref, nu0, lam, a, n));
}

// ---------------------------------------------------------------------------
struct OracleBase {
    virtual ~OracleBase() {}
    virtual int step(int64_t n) = 0;
    virtual int get_field(int field, double* out) = 0;
    virtual int get_populations(double* out) = 0;
    virtual int compute_strain() = 0;
    virtual int info(lbm355_case_info* i) = 0;
    virtual int set_inflow(double a, double b) = 0;
    virtual int set_populations(const double* in) = 0;
    virtual int stream_only(int64_t n) = 0;
    virtual int set_velocity(const double* ux, const double* uy, const double* uz) = 0;
};

template <class T>
struct Oracle : OracleBase {
    lbm355_case_desc d;
    int X, Y, Z;
    int64_t N;
    std::vector<int8_t> map;
    std::vector<uint8_t> wall;  // wall mask (lbm.wall)
    std::vector<T> df[2];       // [q][x][z][y] as lat_permutation x,q,z,y
    int cur = 0;
    int64_t iterations = 0;
    std::vector<T> rho, vx, vy, vz, fx, fy, fz, S[6];
    std::vector<double> acc[6];
    std::vector<double> inflow_factor;  // per cell profile factor on planes 3/4
    long long area_l = 0, area_r = 0, out_f = 0, out_b = 0, n_wall = 0, n_fluid = 0;
    long long in3 = 0, in4 = 0, in5 = 0, in6 = 0;
    bool terminate = false;
    double ivc_val, svc_val;

    int64_t I(int x, int y, int z) const { return x + (int64_t)X * (y + (int64_t)Y * z); }
    int64_t DF(int q, int x, int y, int z) const { return y + (int64_t)Y * (z + (int64_t)Z * (q + 27 * (int64_t)x)); }

    explicit Oracle(const lbm355_case_desc& desc) : d(desc)
    {
        X = d.nx; Y = d.ny; Z = d.nz;
        N = (int64_t)X * Y * Z;
        map.assign(N, GEO_FLUID);
        wall.assign(N, 0);
        setup_boundaries();
        for (int k = 0; k < 2; ++k) df[k].assign(27 * N, T(0));
        for (auto* v : {&rho, &vx, &vy, &vz, &fx, &fy, &fz}) v->assign(N, T(0));
        for (int i = 0; i < 6; ++i) { S[i].assign(N, T(0)); acc[i].assign(N, 0.0); }
        std::fill(rho.begin(), rho.end(), T(1));
        setup_profile();
        ivc_val = d.profile == LBM355_PROFILE_PARABOLIC ? d.ivc_umax : d.ivc_flux;
        svc_val = d.profile == LBM355_PROFILE_PARABOLIC ? d.svc_umax : d.svc_flux;
        // resetLattice: f = f_eq(1,0) in both buffers (sim_tcpc.cu:647-655, lbm.hpp:2-39)
        T feq[27];
        feq_invcum<T>(T(1), T(0), T(0), T(0), feq);
        for (int k = 0; k < 2; ++k)
            for (int x = 0; x < X; ++x)
                for (int q = 0; q < 27; ++q)
                    for (int z = 0; z < Z; ++z)
                        for (int y = 0; y < Y; ++y) df[k][DF(q, x, y, z)] = feq[q];
        // LBMKernelInit (core.h:839-863): one statistics term of the initial state
        // (zero for every cell since u = 0 and F = 0)
        if (d.stats_every > 0) {
            for (int64_t i = 0; i < N; ++i) {
                // rho*u = 0 for u = 0; kept explicit for clarity
                for (int a = 0; a < 6; ++a) acc[a][i] += 0.0;
            }
        }
    }

    // setupBoundaries (sim_tcpc.cu:494-645) + projectWall (lbm.hpp:145-155)
    void setup_boundaries()
    {
        for (int x = 0; x < X; ++x)
            for (int y = 0; y < Y; ++y)
                for (int z = 0; z < Z; ++z) {
                    int lab = d.geometry[I(x, y, z)];
                    in3 += lab == 3; in4 += lab == 4; in5 += lab == 5; in6 += lab == 6;
                    map[I(x, y, z)] = GEO_FLUID;
                    wall[I(x, y, z)] = lab == GEO_WALL;
                }
        // face labels (sim_tcpc.cu:524-529 for the default layout)
        int fk[6];
        bool tcpc = true;
        for (int i = 0; i < 6; ++i) tcpc = tcpc && d.face_bc[i] == LBM355_FACE_TCPC;
        const int tc[6] = {LBM355_FACE_INFLOW, LBM355_FACE_INFLOW, LBM355_FACE_OUTFLOW,
                           LBM355_FACE_OUTFLOW, LBM355_FACE_WALL, LBM355_FACE_WALL};
        for (int i = 0; i < 6; ++i) fk[i] = tcpc ? tc[i] : d.face_bc[i];
        auto face_label = [&](int face) {
            int k = fk[face];
            if (k == LBM355_FACE_WALL) return (int)GEO_WALL;
            if (face == 0) return k == LBM355_FACE_INFLOW ? 3 : 8;
            if (face == 1) return k == LBM355_FACE_INFLOW ? 4 : 7;
            if (face == 2) return 5;
            if (face == 3) return 6;
            return (int)GEO_WALL;
        };
        for (int y = 0; y < Y; ++y)
            for (int z = 0; z < Z; ++z) { map[I(0, y, z)] = face_label(0); map[I(X - 1, y, z)] = face_label(1); }
        for (int x = 0; x < X; ++x)
            for (int z = 0; z < Z; ++z) { map[I(x, 0, z)] = face_label(2); map[I(x, Y - 1, z)] = face_label(3); }
        for (int x = 0; x < X; ++x)
            for (int y = 0; y < Y; ++y) { map[I(x, y, 0)] = face_label(4); map[I(x, y, Z - 1)] = face_label(5); }
        auto openX = [&](int x) {
            long long c = 0;
            for (int y = 0; y < Y; ++y)
                for (int z = 0; z < Z; ++z) c += !wall[I(x, y, z)];
            return c;
        };
        auto openY = [&](int y) {
            long long c = 0;
            for (int x = 0; x < X; ++x)
                for (int z = 0; z < Z; ++z) c += !wall[I(x, y, z)];
            return c;
        };
        auto extrudeX = [&](int bx, int step) {
            int sx = bx + step;
            while (sx > 0 && sx < X - 1 && openX(sx) == 0) sx += step;
            if (sx <= 0 || sx >= X - 1) return;
            for (int y = 0; y < Y; ++y)
                for (int z = 0; z < Z; ++z)
                    if (!wall[I(sx, y, z)])
                        for (int x = bx; x != sx + step; x += step) wall[I(x, y, z)] = 0;
        };
        auto extrudeY = [&](int by, int step) {
            int sy = by + step;
            while (sy > 0 && sy < Y - 1 && openY(sy) == 0) sy += step;
            if (sy <= 0 || sy >= Y - 1) return;
            for (int x = 0; x < X; ++x)
                for (int z = 0; z < Z; ++z)
                    if (!wall[I(x, sy, z)])
                        for (int y = by; y != sy + step; y += step) wall[I(x, y, z)] = 0;
        };
        if (fk[0] != LBM355_FACE_WALL) extrudeX(0, 1);
        if (fk[1] != LBM355_FACE_WALL) extrudeX(X - 1, -1);
        if (fk[2] != LBM355_FACE_WALL) extrudeY(0, 1);
        if (fk[3] != LBM355_FACE_WALL) extrudeY(Y - 1, -1);
        for (int64_t i = 0; i < N; ++i) {
            int lab = map[i];
            bool w = wall[i] || lab == GEO_WALL;
            area_l += (lab == GEO_IN_L || lab == 8) && !w;
            area_r += (lab == GEO_IN_R || lab == 7) && !w;
            out_f += lab == GEO_OUT_F && !w;
            out_b += lab == GEO_OUT_B && !w;
            n_wall += w;
            n_fluid += lab == GEO_FLUID && !w;
        }
        if ((fk[0] != LBM355_FACE_WALL && area_l <= 0) || (fk[1] != LBM355_FACE_WALL && area_r <= 0) ||
            (fk[2] != LBM355_FACE_WALL && out_f <= 0) || (fk[3] != LBM355_FACE_WALL && out_b <= 0))
            terminate = true;
        for (int64_t i = 0; i < N; ++i)
            if (wall[i]) map[i] = GEO_WALL;
    }

    // inflow profile factors on the IVC (x=0) and SVC (x=X-1) planes
    void setup_profile()
    {
        inflow_factor.assign(N, 0.0);
        for (int side = 0; side < 2; ++side) {
            int x = side == 0 ? 0 : X - 1;
            int lab = side == 0 ? GEO_IN_L : GEO_IN_R;
            double cy = side == 0 ? d.ivc_center[0] : d.svc_center[0];
            double cz = side == 0 ? d.ivc_center[1] : d.svc_center[1];
            double R = side == 0 ? d.ivc_radius : d.svc_radius;
            double sy = 0, sz = 0;
            long long cnt = 0;
            for (int y = 0; y < Y; ++y)
                for (int z = 0; z < Z; ++z)
                    if (map[I(x, y, z)] == lab) { sy += y; sz += z; ++cnt; }
            if (cy < 0 && cnt) cy = sy / cnt;
            if (cz < 0 && cnt) cz = sz / cnt;
            if (R <= 0) {
                double r2max = 0;
                for (int y = 0; y < Y; ++y)
                    for (int z = 0; z < Z; ++z)
                        if (map[I(x, y, z)] == lab) r2max = std::max(r2max, (y - cy) * (y - cy) + (z - cz) * (z - cz));
                R = std::sqrt(r2max) + 0.5;
            }
            for (int y = 0; y < Y; ++y)
                for (int z = 0; z < Z; ++z) {
                    if (map[I(x, y, z)] != lab) continue;
                    double fac = 1.0;
                    if (d.profile == LBM355_PROFILE_PARABOLIC) {
                        double r2 = ((y - cy) * (y - cy) + (z - cz) * (z - cz)) / (R * R);
                        fac = std::max(0.0, 1.0 - r2);
                    }
                    inflow_factor[I(x, y, z)] = fac;
                }
        }
    }

    int set_inflow(double a, double b) override { ivc_val = a; svc_val = b; return 0; }

    // pull streaming, lbm_streaming.h:10-40: f_q <- df(q, x - c_q) with the
    // caller's (possibly shifted / clamped) neighbour indices
    void stream(const std::vector<T>& src, T* f, int xm, int x, int xp, int ym, int y, int yp, int zm, int z, int zp) const
    {
        int xs[3] = {xp, x, xm}, ys[3] = {yp, y, ym}, zs[3] = {zp, z, zm};
        for (int q = 0; q < 27; ++q) f[q] = src[DF(q, xs[CX(q) + 1], ys[CY(q) + 1], zs[CZ(q) + 1])];
    }
    // streamingBounceBack, lbm_streaming.h:78-107: f_opp(q) <- df(q, x - c_q)
    void stream_bb(const std::vector<T>& src, T* f, int xm, int x, int xp, int ym, int y, int yp, int zm, int z, int zp) const
    {
        int xs[3] = {xp, x, xm}, ys[3] = {yp, y, ym}, zs[3] = {zp, z, zm};
        for (int q = 0; q < 27; ++q) f[26 - q] = src[DF(q, xs[CX(q) + 1], ys[CY(q) + 1], zs[CZ(q) + 1])];
    }

    void collide(Cell<T>& k) const
    {
        switch (d.collision) {
        case LBM355_BGK: collide_bgk(k); break;
        case LBM355_MRT: collide_mrt(k); break;
        case LBM355_KBC: collide_kbc(k); break;
        default: collide_cum(k); break;
        }
    }

    T pulse() const
    {
        if (d.pulse_amp == 0.0 || d.pulse_period <= 0.0) return T(1);
        return T(1.0 + d.pulse_amp * std::sin(2.0 * M_PI * double(iterations) / d.pulse_period));
    }

    // velocity pre-pass of the multi-rank CY path (core.h:293-384): fluid
    // cells only, since the strain stencil reads fluid neighbours only
    void velocity_prepass()
    {
        const std::vector<T>& src = df[cur];
#pragma omp parallel for schedule(static) collapse(2)
        for (int x = 0; x < X; ++x)
            for (int z = 0; z < Z; ++z)
                for (int y = 0; y < Y; ++y) {
                    int64_t i = I(x, y, z);
                    if (map[i] != GEO_FLUID) continue;
                    Cell<T> k;
                    k.fx = fx[i]; k.fy = fy[i]; k.fz = fz[i];
                    stream(src, k.f, std::max(x - 1, 0), x, std::min(x + 1, X - 1), std::max(y - 1, 0), y,
                           std::min(y + 1, Y - 1), std::max(z - 1, 0), z, std::min(z + 1, Z - 1));
                    density_velocity(k);
                    rho[i] = k.rho; vx[i] = k.vx; vy[i] = k.vy; vz[i] = k.vz;
                }
    }

    // strain from macro velocities, cudaLBMKernelStress (core.h:407-576)
    int compute_strain() override
    {
#pragma omp parallel for schedule(static) collapse(2)
        for (int x = 0; x < X; ++x)
            for (int z = 0; z < Z; ++z)
                for (int y = 0; y < Y; ++y) {
                    int64_t i = I(x, y, z);
                    T s11 = 0, s12 = 0, s13 = 0, s22 = 0, s32 = 0, s33 = 0;
                    if (map[i] == GEO_FLUID) {
                        int xp = std::min(x + 1, X - 1), xm = std::max(x - 1, 0);
                        int yp = std::min(y + 1, Y - 1), ym = std::max(y - 1, 0);
                        int zp = std::min(z + 1, Z - 1), zm = std::max(z - 1, 0);
                        const std::vector<T>* V[3] = {&vx, &vy, &vz};
                        auto fl = [&](int64_t j) { return map[j] == GEO_FLUID; };
                        // derivative d/da of the three velocity components
                        auto deriv = [&](int64_t jm, int64_t jp, T out[3], bool& any) {
                            bool m = fl(jm), p = fl(jp);
                            any = m || p;
                            for (int c = 0; c < 3; ++c) {
                                const std::vector<T>& v = *V[c];
                                if (!m && !p) out[c] = 0;
                                else if (!m) out[c] = v[jp] - v[i];
                                else if (!p) out[c] = v[i] - v[jm];
                                else out[c] = T(0.5) * (v[jp] - v[jm]);
                            }
                        };
                        T dx[3], dy[3], dz[3];
                        bool ax, ay, az;
                        deriv(I(xm, y, z), I(xp, y, z), dx, ax);
                        deriv(I(x, ym, z), I(x, yp, z), dy, ay);
                        deriv(I(x, y, zm), I(x, y, zp), dz, az);
                        s11 = dx[0]; s22 = dy[1]; s33 = dz[2];
                        s12 = T(0.5) * dx[1] + T(0.5) * dy[0];
                        s13 = T(0.5) * dx[2] + T(0.5) * dz[0];
                        s32 = T(0.5) * dy[2] + T(0.5) * dz[1];
                    }
                    S[0][i] = s11; S[1][i] = s12; S[2][i] = s13; S[3][i] = s22; S[4][i] = s32; S[5][i] = s33;
                }
        return 0;
    }

    // one SimUpdate: cudaLBMKernel / LBMKernel over all cells (core.h:598-800)
    void step_once()
    {
        if (d.cy_enabled && d.cy_model == LBM355_CY_FD_FORCE) {
            velocity_prepass();
            compute_strain();
        }
        const std::vector<T>& src = df[cur];
        std::vector<T>& dst = df[1 - cur];
        const T pl = pulse();
        const T in_l = T(ivc_val), in_r = T(svc_val);
        const bool stats = d.stats_every > 0 && ((iterations + 1) % d.stats_every == 0);
#pragma omp parallel for schedule(static) collapse(2)
        for (int x = 0; x < X; ++x)
            for (int z = 0; z < Z; ++z)
                for (int y = 0; y < Y; ++y) {
                    const int64_t i = I(x, y, z);
                    const int lab = map[i];
                    const int xp = std::min(x + 1, X - 1), xm = std::max(x - 1, 0);
                    const int yp = std::min(y + 1, Y - 1), ym = std::max(y - 1, 0);
                    const int zp = std::min(z + 1, Z - 1), zm = std::max(z - 1, 0);
                    Cell<T> k;
                    k.nu = T(d.nu_ref);
                    if (d.cy_enabled && d.cy_model == LBM355_CY_FD_FORCE) {
                        // core.h:663-781
                        T F[3] = {0, 0, 0};
                        if (lab == GEO_FLUID) {
                            auto fl = [&](int64_t j) { return map[j] == GEO_FLUID; };
                            // rows of S used by d/dx, d/dy, d/dz: (11,12,13), (12,22,32), (13,32,33)
                            const int rows[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
                            int64_t jm[3] = {I(xm, y, z), I(x, ym, z), I(x, y, zm)};
                            int64_t jp[3] = {I(xp, y, z), I(x, yp, z), I(x, y, zp)};
                            for (int a = 0; a < 3; ++a) {
                                bool m = fl(jm[a]), p = fl(jp[a]);
                                for (int c = 0; c < 3; ++c) {
                                    const std::vector<T>& s = S[rows[a][c]];
                                    if (!m && !p) continue;
                                    if (!m) F[c] += s[jp[a]] - s[i];
                                    else if (!p) F[c] += s[i] - s[jm[a]];
                                    else F[c] += T(0.5) * (s[jp[a]] - s[jm[a]]);
                                }
                            }
                        }
                        T s11 = S[0][i], s12 = S[1][i], s13 = S[2][i], s22 = S[3][i], s32 = S[4][i], s33 = S[5][i];
                        T gam = std::sqrt(T(2)) * std::sqrt(s11 * s11 + s22 * s22 + s33 * s33 +
                                                            T(2) * (s12 * s12 + s13 * s13 + s32 * s32));
                        // the reference evaluates nu in single precision with powf
                        // (core.h:766); powf here is the correctly rounded float of
                        // the double pow, as on the device (the reference's own
                        // builds differ: glibc powf on the host, fast-math __powf
                        // in CUDA), so that the two agree bit for bit
                        auto powf_cr = [](float x, float y) { return float(std::pow(double(x), double(y))); };
                        T nu = T(d.nu_ref) + (T(d.cy_nu0) - T(d.nu_ref)) *
                               T(powf_cr(float(1.0 + powf_cr(float(gam * T(d.cy_lambda)), float(d.cy_a))),
                                         float((d.cy_n - 1.0) / d.cy_a)));
                        k.fx += T(2) * (nu - T(d.nu_ref)) * F[0] * k.rho;
                        k.fy += T(2) * (nu - T(d.nu_ref)) * F[1] * k.rho;
                        k.fz += T(2) * (nu - T(d.nu_ref)) * F[2] * k.rho;
                    }
                    if (lab == GEO_FLUID) {
                        stream(src, k.f, xm, x, xp, ym, y, yp, zm, z, zp);
                        density_velocity(k);
                        if (d.cy_enabled && d.cy_model == LBM355_CY_NEQ)
                            k.nu = neq_local_nu(k, d.nu_ref, d.cy_nu0, d.cy_lambda, d.cy_a, d.cy_n);
                        collide(k);
                    } else if (lab == GEO_WALL) {
                        stream_bb(src, k.f, xm, x, xp, ym, y, yp, zm, z, zp);
                        k.rho = 1; k.vx = 0; k.vy = 0; k.vz = 0;
                    } else if (lab == GEO_IN_L || lab == GEO_IN_R) {
                        // sim_tcpc.cu:242-261: rho from the populations entering the
                        // cell one step inwards, u prescribed
                        Cell<T> kl;
                        if (lab == GEO_IN_L) stream(src, kl.f, x, xp, xp + 1, ym, y, yp, zm, z, zp);
                        else stream(src, kl.f, xm - 1, xm, x, ym, y, yp, zm, z, zp);
                        density_velocity(kl);
                        // plug: u = flux/area (sim_tcpc.cu:298,306); parabolic: u_max * profile
                        T u;
                        if (d.profile == LBM355_PROFILE_PARABOLIC) {
                            u = (lab == GEO_IN_L ? in_l : -in_r) * T(inflow_factor[i]) * pl;
                        } else {
                            long long area = lab == GEO_IN_L ? area_l : area_r;
                            u = area > 0 ? (lab == GEO_IN_L ? in_l : -in_r) / T(area) * pl : T(0);
                        }
                        k.rho = kl.rho; k.vx = u; k.vy = 0; k.vz = 0;
                        feq_invcum(k.rho, k.vx, k.vy, k.vz, k.f);
                    } else {
                        // sim_tcpc.cu:262-273: u from the populations entering the
                        // cell one step inwards, rho = 1
                        if (lab == GEO_OUT_F) stream(src, k.f, xm, x, xp, y, yp, yp + 1, zm, z, zp);
                        else if (lab == GEO_OUT_B) stream(src, k.f, xm, x, xp, ym - 1, ym, y, zm, z, zp);
                        else if (lab == 7) stream(src, k.f, xm - 1, xm, x, ym, y, yp, zm, z, zp);
                        else stream(src, k.f, x, xp, xp + 1, ym, y, yp, zm, z, zp);
                        density_velocity(k);
                        k.rho = 1;
                        feq_invcum(k.rho, k.vx, k.vy, k.vz, k.f);
                    }
                    for (int q = 0; q < 27; ++q) dst[DF(q, x, y, z)] = k.f[q];
                    // MacroTcpc::outputMacro, sim_tcpc.cu:337-354
                    rho[i] = k.rho; vx[i] = k.vx; vy[i] = k.vy; vz[i] = k.vz;
                    fx[i] = k.fx; fy[i] = k.fy; fz[i] = k.fz;
                    if (stats) {
                        double r = k.rho, ux = k.vx, uy = k.vy, uz = k.vz;
                        acc[0][i] += r * ux; acc[1][i] += r * uy; acc[2][i] += r * uz;
                        acc[3][i] += r * r * ux * ux; acc[4][i] += r * r * uy * uy; acc[5][i] += r * r * uz * uz;
                    }
                }
        cur = 1 - cur;
        ++iterations;
    }

    int step(int64_t n) override
    {
        for (int64_t s = 0; s < n; ++s) step_once();
        return 0;
    }

    int get_field(int field, double* out) override
    {
        const std::vector<T>* src = nullptr;
        switch (field) {
        case LBM355_FIELD_RHO: src = &rho; break;
        case LBM355_FIELD_VX: src = &vx; break;
        case LBM355_FIELD_VY: src = &vy; break;
        case LBM355_FIELD_VZ: src = &vz; break;
        case LBM355_FIELD_FX: src = &fx; break;
        case LBM355_FIELD_FY: src = &fy; break;
        case LBM355_FIELD_FZ: src = &fz; break;
        default: break;
        }
        if (src) { for (int64_t i = 0; i < N; ++i) out[i] = double((*src)[i]); return 0; }
        if (field >= LBM355_FIELD_SUM_RUX && field <= LBM355_FIELD_SUM_RRUZZ) {
            const std::vector<double>& a = acc[field - LBM355_FIELD_SUM_RUX];
            std::copy(a.begin(), a.end(), out);
            return 0;
        }
        if (field >= LBM355_FIELD_S11 && field <= LBM355_FIELD_S33) {
            const std::vector<T>& s = S[field - LBM355_FIELD_S11];
            for (int64_t i = 0; i < N; ++i) out[i] = double(s[i]);
            return 0;
        }
        if (field == LBM355_FIELD_LABEL) { for (int64_t i = 0; i < N; ++i) out[i] = map[i]; return 0; }
        if (field == LBM355_FIELD_WALLMASK) { for (int64_t i = 0; i < N; ++i) out[i] = wall[i]; return 0; }
        g_err = "unknown field";
        return LBM355_EINVAL;
    }

    // test hooks: overwrite the current populations (layout of get_populations)
    int set_populations(const double* in) override
    {
        std::vector<T>& s = df[cur];
        for (int q = 0; q < 27; ++q)
            for (int z = 0; z < Z; ++z)
                for (int y = 0; y < Y; ++y)
                    for (int x = 0; x < X; ++x) s[DF(q, x, y, z)] = T(in[q * N + I(x, y, z)]);
        return 0;
    }

    // test hook: n steps of streaming alone (pull for fluid, full-way
    // bounce-back for wall cells, no collision), for lattices with fluid and
    // wall labels only
    int stream_only(int64_t n) override
    {
        for (int64_t i = 0; i < N; ++i)
            if (map[i] != GEO_FLUID && map[i] != GEO_WALL) { g_err = "stream_only: boundary labels present"; return LBM355_EINVAL; }
        for (int64_t s = 0; s < n; ++s) {
            const std::vector<T>& src = df[cur];
            std::vector<T>& dst = df[1 - cur];
            for (int x = 0; x < X; ++x)
                for (int z = 0; z < Z; ++z)
                    for (int y = 0; y < Y; ++y) {
                        T f[27];
                        const int xp = std::min(x + 1, X - 1), xm = std::max(x - 1, 0);
                        const int yp = std::min(y + 1, Y - 1), ym = std::max(y - 1, 0);
                        const int zp = std::min(z + 1, Z - 1), zm = std::max(z - 1, 0);
                        if (map[I(x, y, z)] == GEO_FLUID) stream(src, f, xm, x, xp, ym, y, yp, zm, z, zp);
                        else stream_bb(src, f, xm, x, xp, ym, y, yp, zm, z, zp);
                        for (int q = 0; q < 27; ++q) dst[DF(q, x, y, z)] = f[q];
                    }
            cur = 1 - cur;
        }
        return 0;
    }

    // test hook: overwrite the macroscopic velocity (input of compute_strain)
    int set_velocity(const double* ux, const double* uy, const double* uz) override
    {
        for (int64_t i = 0; i < N; ++i) { vx[i] = T(ux[i]); vy[i] = T(uy[i]); vz[i] = T(uz[i]); }
        return 0;
    }

    int get_populations(double* out) override
    {
        const std::vector<T>& s = df[cur];
        for (int q = 0; q < 27; ++q)
            for (int z = 0; z < Z; ++z)
                for (int y = 0; y < Y; ++y)
                    for (int x = 0; x < X; ++x) out[q * N + I(x, y, z)] = double(s[DF(q, x, y, z)]);
        return 0;
    }

    int info(lbm355_case_info* o) override
    {
        std::memset(o, 0, sizeof(*o));
        o->fluid = n_fluid; o->wall = n_wall;
        o->area_ivc = area_l; o->area_svc = area_r; o->area_out5 = out_f; o->area_out6 = out_b;
        o->input_label3 = in3; o->input_label4 = in4; o->input_label5 = in5; o->input_label6 = in6;
        o->active_cells = N - n_wall;
        o->steps = iterations;
        o->terminate = terminate ? 1 : 0;
        return 0;
    }
};

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

// OpenMP threads that a step uses (OMP_NUM_THREADS, else every online CPU)
int oracle_threads(void) { return omp_get_max_threads(); }

void* oracle_create(const lbm355_case_desc* d)
{
    if (!d || d->nx < 3 || d->ny < 3 || d->nz < 3 || !d->geometry) { g_err = "bad descriptor"; return nullptr; }
    if (d->storage == LBM355_F64) return static_cast<OracleBase*>(new Oracle<double>(*d));
    return static_cast<OracleBase*>(new Oracle<float>(*d));
}
void oracle_destroy(void* o) { delete static_cast<OracleBase*>(o); }
int oracle_step(void* o, int64_t n) { return static_cast<OracleBase*>(o)->step(n); }
int oracle_set_inflow(void* o, double a, double b) { return static_cast<OracleBase*>(o)->set_inflow(a, b); }
int oracle_get_field(void* o, int field, double* out) { return static_cast<OracleBase*>(o)->get_field(field, out); }
int oracle_get_populations(void* o, double* out) { return static_cast<OracleBase*>(o)->get_populations(out); }
int oracle_compute_strain(void* o) { return static_cast<OracleBase*>(o)->compute_strain(); }
int oracle_info(void* o, lbm355_case_info* i) { return static_cast<OracleBase*>(o)->info(i); }
int oracle_set_populations(void* o, const double* in) { return static_cast<OracleBase*>(o)->set_populations(in); }
int oracle_stream_only(void* o, int64_t n) { return static_cast<OracleBase*>(o)->stream_only(n); }
int oracle_set_velocity(void* o, const double* ux, const double* uy, const double* uz)
{
    return static_cast<OracleBase*>(o)->set_velocity(ux, uy, uz);
}

// single-cell entry points for the numpy pins
int oracle_collide(int model, int f64, const double* fin, double rho, const double* u, const double* F,
                   double nu, double* fout)
{
    if (f64) {
        Cell<double> k;
        for (int q = 0; q < 27; ++q) k.f[q] = fin[q];
        k.rho = rho; k.vx = u[0]; k.vy = u[1]; k.vz = u[2]; k.fx = F[0]; k.fy = F[1]; k.fz = F[2]; k.nu = nu;
        switch (model) {
        case LBM355_BGK: collide_bgk(k); break;
        case LBM355_MRT: collide_mrt(k); break;
        case LBM355_KBC: collide_kbc(k); break;
        default: collide_cum(k); break;
        }
        for (int q = 0; q < 27; ++q) fout[q] = k.f[q];
    } else {
        Cell<float> k;
        for (int q = 0; q < 27; ++q) k.f[q] = float(fin[q]);
        k.rho = float(rho); k.vx = float(u[0]); k.vy = float(u[1]); k.vz = float(u[2]);
        k.fx = float(F[0]); k.fy = float(F[1]); k.fz = float(F[2]); k.nu = float(nu);
        switch (model) {
        case LBM355_BGK: collide_bgk(k); break;
        case LBM355_MRT: collide_mrt(k); break;
        case LBM355_KBC: collide_kbc(k); break;
        default: collide_cum(k); break;
        }
        for (int q = 0; q < 27; ++q) fout[q] = k.f[q];
    }
    return 0;
}

int oracle_feq(double rho, double ux, double uy, double uz, double* out)
{
    feq_invcum<double>(rho, ux, uy, uz, out);
    return 0;
}

int oracle_density_velocity(const double* f, const double* F, double* out4)
{
    Cell<double> k;
    for (int q = 0; q < 27; ++q) k.f[q] = f[q];
    k.fx = F[0]; k.fy = F[1]; k.fz = F[2];
    density_velocity(k);
    out4[0] = k.rho; out4[1] = k.vx; out4[2] = k.vy; out4[3] = k.vz;
    return 0;
}

double oracle_neq_nu_from_G(double G, double nu_ref, double nu0, double lam, double a, double n)
{
    return neq_nu_from_G(G, nu_ref, nu0, lam, a, n);
}

double oracle_neq_nu(const double* f, double rho, const double* u, double nu_ref, double nu0, double lam,
                     double a, double n)
{
    Cell<double> k;
    for (int q = 0; q < 27; ++q) k.f[q] = f[q];
    k.rho = rho; k.vx = u[0]; k.vy = u[1]; k.vz = u[2];
    return neq_local_nu(k, nu_ref, nu0, lam, a, n);
}

}  // extern "C"
