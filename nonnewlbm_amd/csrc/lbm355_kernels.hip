// lbm355_kernels.hip - gfx950 kernels of the D3Q27 step behind sim_tcpc.
//
// One wavefront = one brick of 64 cells.  The fused kernel reads the 27
// incoming populations of each non-wall cell with the AA access pattern,
// computes rho/u, applies the collision operator (or the boundary equilibrium
// on BC cells) and writes the 27 outgoing populations back into the same
// buffer.  It replaces cudaLBMKernel (solver/lbm3d/core.h:598-800) with the
// policies LBM_STREAMING, LBM_CUM/BGK/MRT/KBC_N4, LBM_BC_Tcpc and MacroTcpc
// (solver/sim_tcpc/sim_tcpc.cu:187-418).
//
// HBM traffic per non-wall cell and step: 27 loads + 27 stores of the storage
// type (the algorithmic 2*27*s bytes), 1 label byte, and on request 4 macro
// values and 6 FP64 statistics read-modify-writes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "lbm355_internal.h"

namespace lbm355 {

// ---------------------------------------------------------------------------
// lattice helpers (q = 9(cx+1) + 3(cy+1) + (cz+1))
__device__ __forceinline__ constexpr int qx(int q) { return q / 9 - 1; }
__device__ __forceinline__ constexpr int qy(int q) { return (q / 3) % 3 - 1; }
__device__ __forceinline__ constexpr int qz(int q) { return q % 3 - 1; }

template <class R>
__device__ __forceinline__ R wgt(int q)
{
    return R(lattice_weight(q));
}

// ---------------------------------------------------------------------------
// buffer resources: populations and statistics are addressed through an SGPR
// resource descriptor and a 32-bit per-lane byte offset (buffer_load/store
// "offen"), which keeps the per-lane address math to one 32-bit multiply-add
// and turns any offset past the end of the buffer into a zero read / dropped
// write instead of a fault.
using rsrc_t = __amdgpu_buffer_rsrc_t;

// a wave-uniform copy of a pointer (readfirstlane of both halves), so that
// descriptors built from it stay in SGPRs even inside divergent code
template <class T>
__device__ __forceinline__ T* uniform(T* p)
{
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
    return reinterpret_cast<T*>((uint64_t(hi) << 32) | lo);
}

// descriptor word 3 of a raw buffer on the gfx9 family: DATA_FORMAT = 32
// bits (a zero format makes every access out of range)
constexpr int kRsrcWord3 = 0x00020000;

__device__ __forceinline__ rsrc_t make_rsrc(const void* p, int64_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(uniform(p)), (short)0,
                                             __builtin_amdgcn_readfirstlane(int(uint32_t(bytes))), kRsrcWord3);
}
// AUX is the cache-policy operand of the buffer instructions (gfx950: bit 1 =
// nt, the non-temporal hint)
template <int AUX = 0>
__device__ __forceinline__ float ld_f32(rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, AUX));
}
template <int AUX = 0>
__device__ __forceinline__ void st_f32(rsrc_t r, uint32_t off, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, off, 0, AUX);
}
template <int AUX = 0>
__device__ __forceinline__ double ld_f64(rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, AUX));
}
template <int AUX = 0>
__device__ __forceinline__ void st_f64(rsrc_t r, uint32_t off, double v)
{
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, off, 0, AUX);
}
template <int AUX = 0>
__device__ __forceinline__ uint16_t ld_u16(rsrc_t r, uint32_t off)
{
    return __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, AUX);
}
template <int AUX = 0>
__device__ __forceinline__ void st_u16(rsrc_t r, uint32_t off, uint16_t v)
{
    __builtin_amdgcn_raw_buffer_store_b16(v, r, off, 0, AUX);
}

// cache policy of the even step's population accesses: with nt (non-temporal)
// the full-line reads and writes of the even step stream past the caches.
// That pays when the working set is larger than the Infinity Cache (the odd
// step then finds the caches free for its partial-line traffic) and costs
// when it fits (the next step would have hit there); the host chooses per
// launch.  The statistics ring and sums follow the same choice on both step
// parities.  The odd step's population accesses keep the default policy:
// their partial lines need L2.
constexpr int kAuxNt = 2;  // gfx950 buffer-instruction aux: bit 1 = nt

// ---------------------------------------------------------------------------
// storage policies: T is what lives in HBM, R the arithmetic type; ld/st take
// the element index i (converted to a byte offset here) and the direction q
template <int S>
struct Pop;

template <>
struct Pop<S_F64> {
    using T = double;
    using R = double;
    template <int A = 0>
    __device__ static R ld(rsrc_t r, uint32_t i, int) { return ld_f64<A>(r, i * 8u); }
    template <int A = 0>
    __device__ static void st(rsrc_t r, uint32_t i, int, R v, uint32_t) { st_f64<A>(r, i * 8u, v); }
    // plain-pointer variants for the fill kernel
    __device__ static void put(T* p, int64_t i, int, R v) { p[i] = v; }
};

// FP32 holds f_q - w_q as well.  At lattice viscosity 1e-4 the
// non-equilibrium stress that carries the viscosity is about 1e-7 of a
// population, which FP32 rounds to 3e-8 near w_q: stored in full, the
// cumulant's velocity on TCPC25_03 (material sx, regime v1) drifted 3.2e-3
// from FP64 in 20,000 steps.  The deviation keeps 24 bits of its own
template <>
struct Pop<S_F32> {
    using T = float;
    using R = float;
    template <int A = 0>
    __device__ static R ld(rsrc_t r, uint32_t i, int q) { return ld_f32<A>(r, i * 4u) + wgt<float>(q); }
    template <int A = 0>
    __device__ static R ld_dev(rsrc_t r, uint32_t i, int) { return ld_f32<A>(r, i * 4u); }
    template <int A = 0>
    __device__ static void st(rsrc_t r, uint32_t i, int q, R v, uint32_t) { st_f32<A>(r, i * 4u, v - wgt<float>(q)); }
    template <int A = 0>
    __device__ static void st_dev(rsrc_t r, uint32_t i, int, R d, uint32_t) { st_f32<A>(r, i * 4u, d); }
    __device__ static void put(T* p, int64_t i, int q, R v) { p[i] = v - wgt<float>(q); }
};

// Two-byte storage rounds stochastically: a value is rounded up with the
// probability of its distance from the lower neighbour (IEEE half: gfx950's
// v_cvt_sr_f16_f32 with a random word; bfloat16: uniform random bits added
// below the kept mantissa, then truncation).  The rounding error then has zero
// mean.  With round-to-nearest the same populations of a steady or periodic
// flow are rounded the same way step after step, and that bias piles up in
// the slowly damped modes: on the 512x384x256 junction with pulsatile inflow
// (cumulant, CY-A, 40,000 steps) velocity missed FP64 by 5.6e-3 (IEEE half)
// and 0.21 (bfloat16) with round-to-nearest, against 9.2e-4 and 4.3e-3 with
// stochastic rounding of full-population arithmetic, and 3.1e-4 (half) once
// the collisions work on the deviations f - w from load to store (Croci et
// al. 2022, "Stochastic rounding: implementation, error analysis and
// applications", for the method).
//
// The random bits are a function of (step, brick, lane, q) only, so a case
// gives the same bits alone, in a batch and across a restart.

// f_q - w_q in IEEE half, scaled by 2^e_q (half_scale_exp): the deviation from
// the rest state is what needs the mantissa (prior art for the offset:
// solver/lbm3d/lbm_eq_well.h:7-46)
template <>
struct Pop<S_F16> {
    using T = _Float16;
    using R = float;
    template <int A = 0>
    __device__ static R ld(rsrc_t r, uint32_t i, int q)
    {
        const float h = float(__builtin_bit_cast(_Float16, ld_u16<A>(r, i * 2u)));
        return fmaf(h, 1.0f / float(1 << half_scale_exp(q)), wgt<float>(q));
    }
    // the stored deviation f_q - w_q itself (exact in FP32)
    template <int A = 0>
    __device__ static R ld_dev(rsrc_t r, uint32_t i, int q)
    {
        return float(__builtin_bit_cast(_Float16, ld_u16<A>(r, i * 2u))) * (1.0f / float(1 << half_scale_exp(q)));
    }
    template <int A = 0>
    __device__ static void st(rsrc_t r, uint32_t i, int q, R v, uint32_t rnd)
    {
        st_dev<A>(r, i, q, v - wgt<float>(q), rnd);
    }
    // stores the deviation d = f_q - w_q, scaled, rounded stochastically by
    // gfx950's v_cvt_sr_f16_f32 with the word rnd (half subnormals included,
    // FP32 denormals being kept).  |x| < 2 in practice (x = (f/w - 1) 16/27),
    // far below the half range (65504)
    template <int A = 0>
    __device__ static void st_dev(rsrc_t r, uint32_t i, int q, R d, uint32_t rnd)
    {
        typedef __fp16 half2_t __attribute__((ext_vector_type(2)));
        const float x = d * float(1 << half_scale_exp(q));
        const half2_t h = __builtin_amdgcn_cvt_sr_f16_f32(half2_t{0, 0}, x, rnd, false);
        st_u16<A>(r, i * 2u, uint16_t(__builtin_bit_cast(uint32_t, h) & 0xFFFFu));
    }
    __device__ static void put(T* p, int64_t i, int q, R v)
    {
        p[i] = _Float16((v - wgt<float>(q)) * float(1 << half_scale_exp(q)));
    }
};

// f_q - w_q in bfloat16 (the upper half of its FP32 image, rounded
// stochastically with 16 random bits)
__device__ __forceinline__ uint16_t to_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
template <>
struct Pop<S_BF16> {
    using T = uint16_t;
    using R = float;
    template <int A = 0>
    __device__ static R ld(rsrc_t r, uint32_t i, int q)
    {
        return __uint_as_float(uint32_t(ld_u16<A>(r, i * 2u)) << 16) + wgt<float>(q);
    }
    template <int A = 0>
    __device__ static R ld_dev(rsrc_t r, uint32_t i, int)
    {
        return __uint_as_float(uint32_t(ld_u16<A>(r, i * 2u)) << 16);
    }
    template <int A = 0>
    __device__ static void st(rsrc_t r, uint32_t i, int q, R v, uint32_t rnd)
    {
        st_dev<A>(r, i, q, v - wgt<float>(q), rnd);
    }
    template <int A = 0>
    __device__ static void st_dev(rsrc_t r, uint32_t i, int, R d, uint32_t rnd)
    {
        st_u16<A>(r, i * 2u, uint16_t((__float_as_uint(d) + (rnd & 0xFFFFu)) >> 16));
    }
    __device__ static void put(T* p, int64_t i, int q, R v) { p[i] = to_bf16(v - wgt<float>(q)); }
};

// the random bits of a cell's 27 stores in one step: a hash of the cell
// (brick * 64 + lane) and of the step (the step part is wave-uniform), then a
// Weyl sequence over q (Wellons' lowbias32 mixer)
__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ uint32_t rnd_q(uint32_t cell_hash, int q) { return cell_hash + uint32_t(q) * 0x9E3779B9u; }

// ---------------------------------------------------------------------------
// arithmetic policy: FP64 (the parity path) uses correctly rounded division,
// sqrt and pow; FP32 uses the hardware reciprocal, square root, log2 and exp2
// (about 1 ulp), which keeps the collision and the Carreau-Yasuda solve a few
// instructions each.
template <class R>
struct Fm;
template <>
struct Fm<double> {
    __device__ static double rcp(double x) { return 1.0 / x; }
    __device__ static double div(double a, double b) { return a / b; }
    __device__ static double sqrt_(double x) { return sqrt(x); }
    __device__ static double pow_(double b, double e) { return pow(b, e); }
};
template <>
struct Fm<float> {
    __device__ static float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    __device__ static float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
    // b >= 0: exp2(e log2 b); b = 0 gives exp2(-inf) = 0 for e > 0
    __device__ static float pow_(float b, float e) { return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(b)); }
};

// ---------------------------------------------------------------------------
// 1-D central-moment transforms of a velocity-ordered triple (f-, f0, f+)
// (Geier et al. 2015, the chimera transform used by lbm_col_cum.h:50-146 and
// its inverse, lbm_col_cum.h:319-415)
template <class R>
__device__ __forceinline__ void to_central(R& a, R& b, R& c, R u)
{
    const R s = c + a, d = c - a;
    const R k0 = s + b;
    const R k1 = d - u * k0;
    const R k2 = s - R(2) * u * d + u * u * k0;
    a = k0; b = k1; c = k2;
}
// f- = ((u^2 - u) k0 + (2u - 1) k1 + k2) / 2 with the halving folded into
// the coefficients (scaling by 1/2 is exact): one multiply less per call.
// The compiler contracts the two forms differently, so the last bits differ;
// the 40,000-step velocity errors against FP64 stay put (FP32 2.03e-5 vs
// 2.02e-5, FP16-offset 9.35e-4 vs 9.36e-4)
template <class R>
__device__ __forceinline__ void from_central(R& a, R& b, R& c, R u)
{
    const R k0 = a, k1 = b, k2 = c;
    const R fm = R(0.5) * (u * u - u) * k0 + (u - R(0.5)) * k1 + R(0.5) * k2;
    const R fp = fm + u * k0 + k1;
    a = fm;
    b = k0 - fm - fp;
    c = fp;
}

// the same step on deviations from the rest state: the column's k0 is
// K0 + a, its k2 K0/3 + c, its output (K0/6, 2 K0/3, K0/6) + (a, b, c).  K0
// enters only through the velocity terms, and b = a - fm - fp keeps the
// column's mass deviation exactly
template <class R>
__device__ __forceinline__ void from_central_dev(R& a, R& b, R& c, R u, R K0)
{
    const R k0 = a, k1 = b, k2 = c;
    const R k0f = k0 + K0;
    const R fm = R(0.5) * (u * u - u) * k0f + (u - R(0.5)) * k1 + R(0.5) * k2;
    const R fp = fm + u * k0f + k1;
    a = fm;
    b = k0 - fm - fp;
    c = fp;
}

// Moments of a population set up to second order: density, raw momentum
// J = sum c f and raw second moments P_ab = sum c_a c_b f, summed over the 13
// opposite pairs (q, 26-q).  The integer lattice coefficients are resolved at
// compile time (after unrolling), so no multiply by 0 or +-1 is issued.
// Differences of opposite populations (equal weights) are taken first, so the
// momentum sums small numbers.  Summing one axis at a time instead (66
// additions, the chimera order) was 0.4 % faster but less accurate: on the
// 512x384x256 junction at 40,000 steps the velocity error against FP64 grew
// from 2.02e-5 to 2.09e-5 (FP32) and from 9.4e-4 to 1.13e-3 (FP16-offset).
template <class R>
struct Moments {
    R rho, jx, jy, jz, pxx, pyy, pzz, pxy, pxz, pyz;
};

template <class R>
__device__ __forceinline__ Moments<R> moments2(const R (&f)[27])
{
    Moments<R> m;
    R rho = f[13], jx = R(0), jy = R(0), jz = R(0);
    R pxx = R(0), pyy = R(0), pzz = R(0), pxy = R(0), pxz = R(0), pyz = R(0);
#pragma unroll
    for (int q = 0; q < 13; ++q) {
        const int cx = qx(q), cy = qy(q), cz = qz(q);
        const R sum = f[q] + f[26 - q];
        const R dif = f[q] - f[26 - q];  // direction q minus direction -c_q
        rho += sum;
        if (cx > 0) jx += dif; else if (cx < 0) jx -= dif;
        if (cy > 0) jy += dif; else if (cy < 0) jy -= dif;
        if (cz > 0) jz += dif; else if (cz < 0) jz -= dif;
        if (cx != 0) pxx += sum;
        if (cy != 0) pyy += sum;
        if (cz != 0) pzz += sum;
        if (cx * cy > 0) pxy += sum; else if (cx * cy < 0) pxy -= sum;
        if (cx * cz > 0) pxz += sum; else if (cx * cz < 0) pxz -= sum;
        if (cy * cz > 0) pyz += sum; else if (cy * cz < 0) pyz -= sum;
    }
    m.rho = rho; m.jx = jx; m.jy = jy; m.jz = jz;
    m.pxx = pxx; m.pyy = pyy; m.pzz = pzz; m.pxy = pxy; m.pxz = pxz; m.pyz = pyz;
    return m;
}

// The same moments summed one axis at a time, as the chimera transform does:
// per (cy, cz) column the x sums f+ + f- and differences f+ - f-, then per cz
// the y sums and differences of those, then the z ones; 68 additions instead
// of 111.  Over full populations (values near w_q) this order rounds worse
// than the opposite-pair differences; over the deviations f - w of the
// two-byte storage types the sums are small and it costs no accuracy
template <class R>
__device__ __forceinline__ Moments<R> moments2_axis(const R (&f)[27])
{
    // x: column (b, c) = (cy + 1, cz + 1) holds q = 3b + c + 9a, a = cx + 1
    R m0[9], m1[9], m2[9];  // sum f, sum cx f, sum cx^2 f over the column
#pragma unroll
    for (int col = 0; col < 9; ++col) {
        const R fm = f[col], f0 = f[9 + col], fp = f[18 + col];
        m2[col] = fp + fm;
        m1[col] = fp - fm;
        m0[col] = m2[col] + f0;
    }
    // y, per cz column c: sums over cy of m0, cy m0, cy^2 m0, m1, cy m1, m2
    R t0[3], b0[3], a0[3], t1[3], b1[3], t2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        a0[c] = m0[6 + c] + m0[c];
        b0[c] = m0[6 + c] - m0[c];
        t0[c] = a0[c] + m0[3 + c];
        const R s1 = m1[6 + c] + m1[c];
        b1[c] = m1[6 + c] - m1[c];
        t1[c] = s1 + m1[3 + c];
        t2[c] = (m2[6 + c] + m2[c]) + m2[3 + c];
    }
    // z
    Moments<R> m;
    const R pz = t0[2] + t0[0];
    m.rho = pz + t0[1];
    m.jz = t0[2] - t0[0];
    m.pzz = pz;
    m.jy = (b0[2] + b0[0]) + b0[1];
    m.pyz = b0[2] - b0[0];
    m.pyy = (a0[2] + a0[0]) + a0[1];
    m.jx = (t1[2] + t1[0]) + t1[1];
    m.pxz = t1[2] - t1[0];
    m.pxy = (b1[2] + b1[0]) + b1[1];
    m.pxx = (t2[2] + t2[0]) + t2[1];
    return m;
}

// index of the central moment kappa_{abc} (orders along x, y, z)
__device__ __forceinline__ constexpr int K(int a, int b, int c) { return 9 * a + 3 * b + c; }

// Cumulant collision of lbm_col_cum.h (default build: omega1 = 1/(3 nu + 1/2),
// every other rate 1).  With those rates the post-collision state depends on
// the pre-collision populations only through rho, the first central moments
// (mirrored, which applies the force, lbm_col_cum.h:313-315) and the second
// central moments: 3rd-order cumulants go to 0 and the 4th/6th-order central
// moments become products of the relaxed 2nd-order ones (lbm_col_cum.h:289-307).
// The second central moments about u follow from the raw moments as
// kappa_ab = P_ab - u_a J_b - u_b J_a + rho u_a u_b, which replaces the
// forward chimera transform; the backward chimera (lbm_col_cum.h:319-415)
// rebuilds the 27 populations.
// With DEV the populations are the deviations f_q - w_q from the rest state
// (the two-byte storage types keep exactly those): M then holds the full rho
// but the raw second moments of the deviations (P_aa less 1/3), dr = rho - 1,
// and the diagonal 2nd, 4th and 6th-order central moments are carried as
// their deviations from 1/3, 1/9 and 1/27.  This is the same collision with
// no quantity near 1 rounded in FP32 on the way: a post-collision population
// of a cell at rest is w_q = 8/27 ... 1/216 + a small part, and rounding the
// whole value each step is a systematic error in the conserved moments
template <class R, bool DEV = false>
__device__ __forceinline__ void collide_cumulant(R (&m)[27], const Moments<R>& M, R dr, R ux, R uy, R uz, R om)
{
    const R rho = M.rho;
    const R inv = Fm<R>::rcp(rho);
    const R om1 = R(1) - om;
    const R k100 = M.jx - ux * rho, k010 = M.jy - uy * rho, k001 = M.jz - uz * rho;
    const R k200 = M.pxx - ux * (R(2) * M.jx - ux * rho);
    const R k020 = M.pyy - uy * (R(2) * M.jy - uy * rho);
    const R k002 = M.pzz - uz * (R(2) * M.jz - uz * rho);
    const R k110 = M.pxy - ux * M.jy - uy * M.jx + rho * ux * uy;
    const R k101 = M.pxz - ux * M.jz - uz * M.jx + rho * ux * uz;
    const R k011 = M.pyz - uy * M.jz - uz * M.jy + rho * uy * uz;
    const R c110 = om1 * k110, c101 = om1 * k101, c011 = om1 * k011;
    const R dxy = om1 * (k200 - k020);
    const R dxz = om1 * (k200 - k002);
    const R third = R(1) / R(3);
    const R r0 = DEV ? dr : rho;
    const R c200 = (dxy + dxz + r0) * third;
    const R c020 = (dxz - R(2) * dxy + r0) * third;
    const R c002 = (dxy - R(2) * dxz + r0) * third;
    const R C200 = DEV ? c200 + third : c200, C020 = DEV ? c020 + third : c020, C002 = DEV ? c002 + third : c002;
    const R k211 = (C200 * c011 + R(2) * c101 * c110) * inv;
    const R k121 = (C020 * c101 + R(2) * c110 * c011) * inv;
    const R k112 = (C002 * c110 + R(2) * c011 * c101) * inv;
    R k220, k022, k202, k222;
    if constexpr (DEV) {
        // k_220 - 1/9 = ((c020 + c200)/3 + c020 c200 + 2 c110^2 - dr/9) / rho
        // with c_aa the deviations; k_222 - 1/27 likewise, its constant parts
        // summing to dr (1 - dr) / 27 / rho^2 (c200 + c020 + c002 = dr)
        const R n9 = dr * (R(1) / R(9));
        k220 = ((c020 + c200) * third + c020 * c200 + R(2) * c110 * c110 - n9) * inv;
        k022 = ((c002 + c020) * third + c002 * c020 + R(2) * c011 * c011 - n9) * inv;
        k202 = ((c200 + c002) * third + c200 * c002 + R(2) * c101 * c101 - n9) * inv;
        const R s1 = n9 + (k022 + k202 + k220) * third + (c200 * k022 + c020 * k202 + c002 * k220) +
                     R(4) * (c011 * k211 + c101 * k121 + c110 * k112);
        const R s3 = R(2) * n9 + R(2) * third * (c200 * c020 + c020 * c002 + c002 * c200) + R(2) * c200 * c020 * c002 +
                     R(16) * c110 * c101 * c011 + R(4) * (c101 * c101 * C020 + c011 * c011 * C200 + c110 * c110 * C002);
        k222 = s1 * inv + (dr * (R(1) - dr) * (R(1) / R(27)) - s3) * (inv * inv);
    } else {
        k220 = (c020 * c200 + R(2) * c110 * c110) * inv;
        k022 = (c002 * c020 + R(2) * c011 * c011) * inv;
        k202 = (c200 * c002 + R(2) * c101 * c101) * inv;
        k222 = (c200 * k022 + c020 * k202 + c002 * k220 + R(4) * (c011 * k211 + c101 * k121 + c110 * k112)) * inv -
               (R(16) * c110 * c101 * c011 + R(4) * (c101 * c101 * c020 + c011 * c011 * c200 + c110 * c110 * c002) +
                R(2) * c200 * c020 * c002) * (inv * inv);
    }

    // backward along x for the 9 (y,z)-moment columns; the 3rd and 5th order
    // central moments are zero, which the per-column code below drops.  k0f
    // is the column's full zeroth moment, k0 the one whose mass f- + f0 + f+
    // takes (the same but for the rest part under DEV)
    const R a0 = R(0.5) * (ux * ux - ux), a1 = ux - R(0.5);  // f- = a0 k0 + a1 k1 + k2/2
    auto bx = [&](int col, R k0f, R k0, R k1, R k2) {
        const R fm = a0 * k0f + a1 * k1 + R(0.5) * k2;
        const R fp = fm + ux * k0f + k1;
        m[col] = fm;
        m[9 + col] = k0 - fm - fp;
        m[18 + col] = fp;
    };
    auto bx0 = [&](int col, R k0f, R k0, R k2) {  // k1 = 0
        const R fm = a0 * k0f + R(0.5) * k2;
        const R fp = fm + ux * k0f;
        m[col] = fm;
        m[9 + col] = k0 - fm - fp;
        m[18 + col] = fp;
    };
    auto bx1 = [&](int col, R k1) {  // k0 = k2 = 0
        const R fm = a1 * k1;
        const R fp = fm + k1;
        m[col] = fm;
        m[9 + col] = -fm - fp;
        m[18 + col] = fp;
    };
    // columns indexed by (b,c) = y,z orders: col = 3b + c
    bx(K(0, 0, 0), rho, r0, -k100, c200);
    bx(K(0, 1, 0), -k010, -k010, c110, R(0));
    bx0(K(0, 2, 0), C020, c020, k220);
    bx(K(0, 0, 1), -k001, -k001, c101, R(0));
    bx0(K(0, 1, 1), c011, c011, k211);
    bx1(K(0, 2, 1), k121);
    bx0(K(0, 0, 2), C002, c002, k202);
    bx1(K(0, 1, 2), k112);
    bx0(K(0, 2, 2), DEV ? k022 + R(1) / R(9) : k022, k022, k222);
    // then along y and z (general triples).  At rest the x pass leaves
    // w1(a) Z(c) in the y column (a, c), w1 = (1/6, 2/3, 1/6) the 1-D weights
    // and Z = (1, 0, 1/3) the 1-D rest moments, and the y pass w1(a) w1(b) in
    // the z column (a, b)
    constexpr double w1[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}, z1[3] = {1.0, 0.0, 1.0 / 3.0};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (DEV && c != 1) from_central_dev(m[9 * a + c], m[9 * a + 3 + c], m[9 * a + 6 + c], uy, R(w1[a] * z1[c]));
            else from_central(m[9 * a + c], m[9 * a + 3 + c], m[9 * a + 6 + c], uy);
        }
#pragma unroll
    for (int a = 0; a < 9; ++a) {
        if (DEV) from_central_dev(m[3 * a], m[3 * a + 1], m[3 * a + 2], uz, R(w1[a / 3] * w1[a % 3]));
        else from_central(m[3 * a], m[3 * a + 1], m[3 * a + 2], uz);
    }
}

// 1-D factors of the product-form equilibrium: phi(0) = 2/3 - u^2,
// phi(+-1) = (u^2 +- u + 1/3)/2 (lbm_col_bgk.h:43-53, lbm_eq_inv_cum.h:19-49).
// The three sum to 1, so that sum_q f_eq = rho.  The rounded 2/3 and 1/3
// must keep that: in FP32 both round up, their sum is 1 + 3e-8, and with it
// every BGK or KBC step added 9e-8 rho to each cell (40,000-step junction:
// velocity 5e-3 off FP64).  1/3 is therefore taken as 1 - (2/3 rounded),
// which is exact (Sterbenz) and makes the pair sum to exactly 1
template <class R>
__device__ __forceinline__ void phi3(R u, R (&p)[3])
{
    constexpr R k23 = R(2) / R(3), k13 = R(1) - k23;
    const R uu = u * u;
    p[1] = k23 - uu;
    p[0] = R(0.5) * (uu - u + k13);
    p[2] = R(0.5) * (uu + u + k13);
}

// boundary equilibrium f_eq^INV_CUM(rho, u) including the extra rho of feq_pmm
// (lbm_eq_inv_cum.h:49), which matters at inflow cells (rho != 1)
template <class R>
__device__ __forceinline__ void feq_boundary(R (&f)[27], R rho, R ux, R uy, R uz)
{
    R px[3], py[3], pz[3];
    phi3(ux, px); phi3(uy, py); phi3(uz, pz);
#pragma unroll
    for (int q = 0; q < 27; ++q) f[q] = rho * px[q / 9] * py[(q / 3) % 3] * pz[q % 3];
    f[18] *= rho;
}

// c_q . F with the lattice coefficients resolved at compile time
template <class R>
__device__ __forceinline__ R dot_c(int q, R Fx, R Fy, R Fz)
{
    R v = R(0);
    if (qx(q) > 0) v += Fx; else if (qx(q) < 0) v -= Fx;
    if (qy(q) > 0) v += Fy; else if (qy(q) < 0) v -= Fy;
    if (qz(q) > 0) v += Fz; else if (qz(q) < 0) v -= Fz;
    return v;
}

// The product-form equilibrium as deviations from the rest state, from small
// quantities only: phi_a = p0_a (1 + e_a) with p0 = (1/6, 2/3, 1/6) the 1-D
// rest weights, e(0) = -3/2 u^2 and e(+-1) = 3 (u^2 +- u), so that
// f_eq,q - w_q = w_q (rho - 1 + rho E_q), E_q = prod_a (1 + e_a) - 1.  The
// 1-D corrections p0_a e_a sum to zero, so sum_q w_q E_q = 0 and the mass of
// the deviations is rho - 1 up to rounding of small numbers
template <class R>
__device__ __forceinline__ void eq_dev3(R u, R (&e)[3])
{
    const R uu = u * u;
    e[1] = R(-1.5) * uu;
    e[0] = R(3) * (uu - u);
    e[2] = R(3) * (uu + u);
}
// sum_q w_q - 1 (a = 0) and 3 sum_q w_q c_qx^2 - 1 (a = 1) over the lattice
// weights as rounded to R: 2^-27 each in FP32
template <class R>
__host__ __device__ constexpr double weight_excess(int a)
{
    double s = 0.0;
    for (int q = 0; q < 27; ++q) {
        const int cx = q / 9 - 1;
        s += double(R(lattice_weight(q))) * (a == 0 ? 1.0 : 3.0 * cx * cx);
    }
    return s - 1.0;
}

// With e_xy = (1 + e_x)(1 + e_y) - 1 per (x, y) pair, rho - 1 + rho E_q =
// A + B e_z with A = rho - 1 + rho e_xy and B = rho (1 + e_xy): two
// operations per direction once the nine pairs are formed
template <class R>
struct EqDev {
    R A[9], B[9], ez[3];
    __device__ EqDev(R rho, R dr, R ux, R uy, R uz)
    {
        R ex[3], ey[3];
        eq_dev3(ux, ex); eq_dev3(uy, ey); eq_dev3(uz, ez);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const R exy = ex[a] + ey[b] + ex[a] * ey[b];
                A[3 * a + b] = dr + rho * exy;
                B[3 * a + b] = rho + rho * exy;
            }
    }
    // f_eq,q - w_q
    __device__ R operator()(int q) const { return wgt<R>(q) * (A[q / 3] + B[q / 3] * ez[q % 3]); }
    // the same over the rounded weights, whose sum and momentum sums exceed 1
    // by em and 2 ej (weight_excess): the rest direction gives back em times
    // the mass it rests on (base: rho for populations stored in full over the
    // rounded weights, rho - 1 for deviations), each face ej rho (c.u).  The
    // amounts (7e-9 rho at FP32) are far below the spacing of FP32 near w_q,
    // so this is only possible on the deviation
    __device__ R fixed(int q, R base, R rho, R cu) const
    {
        constexpr R em = R(weight_excess<R>(0)), ej = R(0.5 * weight_excess<R>(1));
        const int c2 = qx(q) * qx(q) + qy(q) * qy(q) + qz(q) * qz(q);
        R de = (*this)(q);
        if (c2 == 0) de -= em * base;
        else if (c2 == 1) de -= ej * rho * cu;
        return de;
    }
};

// improved BGK with product-form equilibrium and source term
// (lbm_col_bgk.h:14-140).  DEV: f holds f - w (in and out), dr = rho - 1.
// FP32 populations stored in full take the equilibrium as w + EqDev too,
// with f - w exact (Sterbenz), so that no value near w_q is rounded before
// the final update (FP32, 40,000-step junction: velocity 1.4e-4 off FP64 with
// the product form, whose 2/3 - u^2 alone rounds by up to 3e-8); FP64 keeps
// the product form of the reference
// FORCE: the case carries a body force (the fd-force Carreau-Yasuda mode);
// without one the source term is zero and is not formed
template <class R, bool DEV = false, bool FORCE = true>
__device__ __forceinline__ void collide_bgk(R (&f)[27], R rho, R dr, R ux, R uy, R uz, R Fx, R Fy, R Fz, R om)
{
    const R src = (R(1) - R(0.5) * om) * R(3) * Fm<R>::rcp(rho);
    const R uF = ux * Fx + uy * Fy + uz * Fz;
    if constexpr (DEV || sizeof(R) == 4) {
        const EqDev<R> E(rho, dr, ux, uy, uz);
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            const R w = wgt<R>(q);
            const R de = E.fixed(q, DEV ? dr : rho, rho, dot_c(q, ux, uy, uz));
            const R d = DEV ? f[q] : f[q] - w;
            f[q] += (de - d) * om;
            if constexpr (FORCE) f[q] += src * (dot_c(q, Fx, Fy, Fz) - uF) * (de + w);
        }
    } else {
        R px[3], py[3], pz[3];
        phi3(ux, px); phi3(uy, py); phi3(uz, pz);
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            const R feq = rho * px[q / 9] * py[(q / 3) % 3] * pz[q % 3];
            if constexpr (FORCE) f[q] += (feq - f[q]) * om + src * (dot_c(q, Fx, Fy, Fz) - uF) * feq;
            else f[q] += (feq - f[q]) * om;
        }
    }
}

// regularised collision with a Smagorinsky rate, "MRT_LES"
// (lbm_col_mrt.h:14-77); the force is not applied, as in the reference
// sum_q w_q (5/2 - 3/2 c_q^2) - 1 (a = 0) and 3 sum_q w_q c_qx^2 - 1 (a = 1)
// over the lattice weights as rounded to R: the mass and momentum that one
// MRT reconstruction gains per unit of rho and rho u (2^-27 each in FP32)
template <class R>
__host__ __device__ constexpr double mrt_excess(int a)
{
    double s = 0.0;
    for (int q = 0; q < 27; ++q) {
        const int cx = q / 9 - 1, c2 = cx * cx + ((q / 3) % 3 - 1) * ((q / 3) % 3 - 1) + (q % 3 - 1) * (q % 3 - 1);
        s += double(R(lattice_weight(q))) * (a == 0 ? 2.5 - 1.5 * c2 : 3.0 * cx * cx);
    }
    return s - 1.0;
}

// DEV: f holds f - w (in and out) and M the moments of the deviations (rho
// full, P_aa less 1/3), dr = rho - 1.  With P~ = P - rho/3 I the relaxed
// moments less their isotropic part, f_q - w_q = w_q (dr + 3 rho c.u +
// 9/2 c.P~.c - 3/2 tr P~): the same populations, from small numbers only
template <class R, bool DEV = false>
__device__ __forceinline__ void collide_mrt(R (&f)[27], const Moments<R>& M, R dr, R ux, R uy, R uz, R nu)
{
    const R rho = M.rho;
    R Pxx = M.pxx, Pyy = M.pyy, Pzz = M.pzz, Pxy = M.pxy, Pxz = M.pxz, Pyz = M.pyz;
    const R third = R(1) / R(3);
    R nxx, nyy, nzz;
    if constexpr (DEV) {
        const R r3 = dr * third;
        nxx = Pxx - r3 - rho * ux * ux; nyy = Pyy - r3 - rho * uy * uy; nzz = Pzz - r3 - rho * uz * uz;
    } else {
        nxx = Pxx - rho * (third + ux * ux); nyy = Pyy - rho * (third + uy * uy); nzz = Pzz - rho * (third + uz * uz);
    }
    const R nxy = Pxy - rho * ux * uy, nxz = Pxz - rho * ux * uz, nyz = Pyz - rho * uy * uz;
    const R Q = R(2) * (nxx * nxx + nyy * nyy + nzz * nzz + R(2) * (nxy * nxy + nxz * nxz + nyz * nyz));
    const R tau = R(3) * nu + R(0.5);
    const R om = Fm<R>::div(R(2), Fm<R>::sqrt_(tau * tau + R(18) * R(0.0342) * Fm<R>::sqrt_(Q) * Fm<R>::rcp(rho)) + tau);
    Pxx -= om * nxx; Pyy -= om * nyy; Pzz -= om * nzz;
    Pxy -= om * nxy; Pxz -= om * nxz; Pyz -= om * nyz;
    if constexpr (DEV) {
        const R r3 = dr * third;  // P~_aa = P_aa - 1/3 - dr/3
        Pxx -= r3; Pyy -= r3; Pzz -= r3;
    }
    const R tr = Pxx + Pyy + Pzz;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const int cx = qx(q), cy = qy(q), cz = qz(q);
        const int c2 = cx * cx + cy * cy + cz * cz;
        R cPc = R(0);
        if (cx) cPc += Pxx;
        if (cy) cPc += Pyy;
        if (cz) cPc += Pzz;
        if (cx * cy > 0) cPc += R(2) * Pxy; else if (cx * cy < 0) cPc -= R(2) * Pxy;
        if (cx * cz > 0) cPc += R(2) * Pxz; else if (cx * cz < 0) cPc -= R(2) * Pxz;
        if (cy * cz > 0) cPc += R(2) * Pyz; else if (cy * cz < 0) cPc -= R(2) * Pyz;
        const R cu = dot_c(q, ux, uy, uz);
        if constexpr (DEV) {
            // the rounded weights' momentum excess (2^-27 in FP32) off the
            // faces; their mass excess here is 2^-27 (rho - 1), negligible
            constexpr R ej = R(0.5 * mrt_excess<R>(1));
            f[q] = wgt<R>(q) * (dr + R(3) * rho * cu + R(4.5) * cPc - R(1.5) * tr);
            if (c2 == 1) f[q] -= ej * rho * cu;
        } else {
            // full populations: the rounded weights' excess mass (rest
            // direction) and momentum (the six faces, half each) is taken out
            // inside the one rounding of the fused multiply-add; outside it,
            // 7e-9 rho is below the spacing of FP32 near w_q and the excess
            // would accumulate, step after step (40,000-step junction:
            // velocity 2.6e-4 off FP64 before)
            constexpr R em = R(mrt_excess<R>(0)), ej = R(0.5 * mrt_excess<R>(1));
            const R br = rho * (R(2.5) - R(1.5) * R(c2) + R(3) * cu) + R(4.5) * cPc - R(1.5) * tr;
            R corr = R(0);
            if (c2 == 0) corr = -em * rho;
            else if (c2 == 1) corr = -ej * rho * cu;
            f[q] = fma(wgt<R>(q), br, corr);
        }
    }
}

// shear part s_q(D + T + Q) of KBC N4 from the non-equilibrium raw moments
// (lbm_col_kbc_n.h:21-219): M2 = {xx,yy,zz,xy,xz,yz}, M3 = {xyy,xzz,xxy,yzz,xxz,yyz,xyz}
template <class R>
__device__ __forceinline__ R kbc_shear(int q, const R (&M2)[6], const R (&M3)[7])
{
    const int cx = qx(q), cy = qy(q), cz = qz(q);
    const int nz = (cx != 0) + (cy != 0) + (cz != 0);
    const R Nxz = M2[0] - M2[2], Nyz = M2[1] - M2[2], Tr = M2[0] + M2[1] + M2[2];
    const R sixth = R(1) / R(6);
    if (nz == 0) return -Tr;
    if (nz == 1) {
        if (cx) return (R(2) * Nxz - Nyz) * sixth + Tr * sixth - R(cx) * (M3[0] + M3[1]) * R(0.5);
        if (cy) return (R(2) * Nyz - Nxz) * sixth + Tr * sixth - R(cy) * (M3[2] + M3[3]) * R(0.5);
        return -(Nxz + Nyz) * sixth + Tr * sixth - R(cz) * (M3[4] + M3[5]) * R(0.5);
    }
    if (nz == 2) {
        if (!cz) return (R(cx * cy) * M2[3] + R(cx) * M3[0] + R(cy) * M3[2]) * R(0.25);
        if (!cy) return (R(cx * cz) * M2[4] + R(cx) * M3[1] + R(cz) * M3[4]) * R(0.25);
        return (R(cy * cz) * M2[5] + R(cy) * M3[3] + R(cz) * M3[5]) * R(0.25);
    }
    return R(cx * cy * cz) * M3[6] * R(0.125);
}

// entropic KBC N4 (lbm_col_kbc_n.h:1024-1259): f = k + s + h, shear part
// relaxed with beta = 1/(6 nu + 1), higher part with the entropic gamma
// an empty asm statement that hides a register value's origin from the
// optimiser: no instruction is emitted, it only blocks the reuse of values
// computed earlier from the same inputs
__device__ __forceinline__ void opaque(double& v) { asm volatile("" : "+v"(v)); }

// DEV (FP32 only): f holds f - w (in and out), M the moments of the
// deviations (rho full, P_aa less 1/3), dr = rho - 1; the equilibrium enters
// as EqDev's deviations, and in full only where it weighs (1/f_eq, source).
// FP32 populations stored in full take the equilibrium from EqDev as well (as
// collide_bgk does); FP64 keeps the product form
template <class R, bool DEV = false, bool FORCE = true>
__device__ __forceinline__ void collide_kbc(R (&f)[27], const Moments<R>& M, R dr, R ux, R uy, R uz, R Fx, R Fy, R Fz,
                                            R nu)
{
    static_assert(!DEV || sizeof(R) == 4, "deviation form at FP32 only");
    const R rho = M.rho;
    R px[3], py[3], pz[3];
    phi3(ux, px); phi3(uy, py); phi3(uz, pz);
    // third-order raw moments from the opposite-pair differences
    R M3[7] = {R(0), R(0), R(0), R(0), R(0), R(0), R(0)};
#pragma unroll
    for (int q = 0; q < 13; ++q) {
        const int cx = qx(q), cy = qy(q), cz = qz(q);
        const R d = f[q] - f[26 - q];
        if (cx * cy * cy > 0) M3[0] += d; else if (cx * cy * cy < 0) M3[0] -= d;
        if (cx * cz * cz > 0) M3[1] += d; else if (cx * cz * cz < 0) M3[1] -= d;
        if (cx * cx * cy > 0) M3[2] += d; else if (cx * cx * cy < 0) M3[2] -= d;
        if (cy * cz * cz > 0) M3[3] += d; else if (cy * cz * cz < 0) M3[3] -= d;
        if (cx * cx * cz > 0) M3[4] += d; else if (cx * cx * cz < 0) M3[4] -= d;
        if (cy * cy * cz > 0) M3[5] += d; else if (cy * cy * cz < 0) M3[5] -= d;
        if (cx * cy * cz > 0) M3[6] += d; else if (cx * cy * cz < 0) M3[6] -= d;
    }
    const R third = R(1) / R(3);
    // non-equilibrium parts of the moments: s is linear in them
    const R r3 = DEV ? dr * third : R(0);
    R M2[6] = {DEV ? M.pxx - r3 - rho * ux * ux : M.pxx - rho * (third + ux * ux),
               DEV ? M.pyy - r3 - rho * uy * uy : M.pyy - rho * (third + uy * uy),
               DEV ? M.pzz - r3 - rho * uz * uz : M.pzz - rho * (third + uz * uz),
               M.pxy - rho * ux * uy, M.pxz - rho * ux * uz, M.pyz - rho * uy * uz};
    M3[0] -= rho * ux * (third + uy * uy); M3[1] -= rho * ux * (third + uz * uz);
    M3[2] -= rho * uy * (third + ux * ux); M3[3] -= rho * uy * (third + uz * uz);
    M3[4] -= rho * uz * (third + ux * ux); M3[5] -= rho * uz * (third + uy * uy);
    M3[6] -= rho * ux * uy * uz;
    constexpr bool kEq = DEV || sizeof(R) == 4;  // equilibrium from EqDev
    const EqDev<R> E(rho, dr, ux, uy, uz);
    R num = 0, den = 0;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const R de = kEq ? E.fixed(q, DEV ? dr : rho, rho, dot_c(q, ux, uy, uz)) : R(0);
        const R feq = kEq ? de + wgt<R>(q) : rho * px[q / 9] * py[(q / 3) % 3] * pz[q % 3];
        const R ds = kbc_shear<R>(q, M2, M3);
        const R dh = DEV ? f[q] - de - ds : kEq ? (f[q] - wgt<R>(q)) - de - ds : f[q] - feq - ds;
        const R ifeq = Fm<R>::rcp(feq);
        num += ds * dh * ifeq;
        den += dh * dh * ifeq;
    }
    const R beta = Fm<R>::rcp(R(6) * nu + R(1));
    const R ib = R(6) * nu + R(1);
    // 0/0 only when dh == 0 everywhere, where any finite gamma gives the same
    // f.  FP32 divides through the hardware reciprocal, which overflows to
    // infinity for a subnormal den (below 2^-126): with the exact deviations
    // of FP32 storage, a quiescent cell's dh can fall under 1e-19, and the
    // infinite gamma times a zero dh made NaNs that spread over the whole case
    // (KBC, 512x384x256 pulsatile junction, within 40,000 steps).  Such a cell
    // takes gamma = 2 as well; its dh term is then below 1e-19
    const bool den_ok = sizeof(R) == 4 ? fabs(den) >= R(1.17549435e-38f) : den != R(0);
    const R gamma = den_ok ? ib - (R(2) - ib) * Fm<R>::div(num, den) : R(2);
    const R src = (R(1) - beta) * R(3) * Fm<R>::rcp(rho);
    const R uF = ux * Fx + uy * Fy + uz * Fz;
    if constexpr (sizeof(R) == 8) {
        // FP64: recompute f_eq and the shear part from opaque copies of their
        // inputs, so that the 27 + 27 values of the first pass are not kept
        // live across the gamma reduction.  Without this the FP64 kernel
        // spills to scratch (256 VGPRs + 208 B per lane): 85 -> 56 us per step
        // on the 256x192x128 junction.  At FP32 the reuse is cheaper.
        R r2 = rho, x2 = ux, y2 = uy, z2 = uz;
        R N2[6], N3[7];
#pragma unroll
        for (int a = 0; a < 6; ++a) N2[a] = M2[a];
#pragma unroll
        for (int a = 0; a < 7; ++a) N3[a] = M3[a];
        opaque(r2); opaque(x2); opaque(y2); opaque(z2);
#pragma unroll
        for (int a = 0; a < 6; ++a) opaque(N2[a]);
#pragma unroll
        for (int a = 0; a < 7; ++a) opaque(N3[a]);
        R qx3[3], qy3[3], qz3[3];
        phi3(x2, qx3); phi3(y2, qy3); phi3(z2, qz3);
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            const R feq = r2 * qx3[q / 9] * qy3[(q / 3) % 3] * qz3[q % 3];
            const R ds = kbc_shear<R>(q, N2, N3);
            const R dh = f[q] - feq - ds;
            if constexpr (FORCE) f[q] -= beta * (R(2) * ds + gamma * dh) - src * (dot_c(q, Fx, Fy, Fz) - uF) * feq;
            else f[q] -= beta * (R(2) * ds + gamma * dh);
        }
    } else {
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            const R de = E.fixed(q, DEV ? dr : rho, rho, dot_c(q, ux, uy, uz));
            const R feq = de + wgt<R>(q);
            const R ds = kbc_shear<R>(q, M2, M3);
            const R dh = DEV ? f[q] - de - ds : (f[q] - wgt<R>(q)) - de - ds;
            if constexpr (FORCE) f[q] -= beta * (R(2) * ds + gamma * dh) - src * (dot_c(q, Fx, Fy, Fz) - uF) * feq;
            else f[q] -= beta * (R(2) * ds + gamma * dh);
        }
    }
}

// Carreau-Yasuda: nu = nu_ref + (nu0 - nu_ref) (1 + (lambda gamma)^a)^((n-1)/a)
// (solver/lbm3d/core.h:766, sim_tcpc.cu:697-707)
template <class R>
struct Lg;
template <>
struct Lg<double> {
    __device__ static double log2_(double x) { return log2(x); }
    __device__ static double exp2_(double x) { return exp2(x); }
};
template <>
struct Lg<float> {
    __device__ static float log2_(float x) { return __builtin_amdgcn_logf(x); }
    __device__ static float exp2_(float x) { return __builtin_amdgcn_exp2f(x); }
};

// the wave-uniform Carreau-Yasuda constants of a case in the compute type
// (precomputed on the host, so no conversion, division or logarithm of
// uniform values runs per wavefront)
template <class R>
struct CyK;
template <>
struct CyK<double> {
    template <class CT>
    __device__ static double nu_ref(const CT& C) { return C.nu_ref; }
    template <class CT>
    __device__ static double nu0(const CT& C) { return C.cy_nu0; }
    template <class CT>
    __device__ static double a(const CT& C) { return C.cy_a; }
    template <class CT>
    __device__ static double n(const CT& C) { return C.cy_n; }
    template <class CT>
    __device__ static double e(const CT& C) { return C.cy_e; }
    template <class CT>
    __device__ static double log2_lambda(const CT& C) { return C.cy_log2_lambda; }
    template <class CT>
    __device__ static double log2_tau_max(const CT& C) { return C.cy_log2_tau_max; }
    template <class CT>
    __device__ static double log2_tau_min(const CT& C) { return C.cy_log2_tau_min; }
};
template <>
struct CyK<float> {
    template <class CT>
    __device__ static float nu_ref(const CT& C) { return C.nu_ref_f; }
    template <class CT>
    __device__ static float nu0(const CT& C) { return C.cy_nu0_f; }
    template <class CT>
    __device__ static float a(const CT& C) { return C.cy_a_f; }
    template <class CT>
    __device__ static float n(const CT& C) { return C.cy_n_f; }
    template <class CT>
    __device__ static float e(const CT& C) { return C.cy_e_f; }
    template <class CT>
    __device__ static float log2_lambda(const CT& C) { return C.cy_log2_lambda_f; }
    template <class CT>
    __device__ static float log2_tau_max(const CT& C) { return C.cy_log2_tau_max_f; }
    template <class CT>
    __device__ static float log2_tau_min(const CT& C) { return C.cy_log2_tau_min_f; }
};

// local viscosity of the "neq" coupling: gamma from the deviatoric
// non-equilibrium second moment P, gamma * tau(gamma) = G = 3 sqrt(2)|P|/(2 rho).
// Solved for t = log2(gamma) by Newton steps on H(t) = t + log2(tau) - log2(G)
// from the middle of the bracket (gamma*tau is close to a power law, so H is
// nearly linear in t); the oracle runs the same iteration (3 steps).
template <class R, class CT>
__device__ __forceinline__ R neq_viscosity(const Moments<R>& M, R ux, R uy, R uz, const CT& C)
{
    const R rho = M.rho;
    const R Pxx = M.pxx - rho * ux * ux, Pyy = M.pyy - rho * uy * uy, Pzz = M.pzz - rho * uz * uz;
    const R Pxy = M.pxy - rho * ux * uy, Pxz = M.pxz - rho * ux * uz, Pyz = M.pyz - rho * uy * uz;
    const R tr = (Pxx + Pyy + Pzz) * (R(1) / R(3));
    const R dxx = Pxx - tr, dyy = Pyy - tr, dzz = Pzz - tr;
    const R s2 = dxx * dxx + dyy * dyy + dzz * dzz + R(2) * (Pxy * Pxy + Pxz * Pxz + Pyz * Pyz);
    const R G = R(2.1213203435596424) * Fm<R>::sqrt_(s2) * Fm<R>::rcp(rho);  // 3 sqrt(2) / 2
    using K = CyK<R>;
    const R nr = K::nu_ref(C), n0 = K::nu0(C), a = K::a(C), n = K::n(C);
    const R e = K::e(C);
    if (!(G > R(1e-30))) return n0;
    const R lG = Lg<R>::log2_(G), L = K::log2_lambda(C);
    const R lo = lG - K::log2_tau_max(C), hi = lG - K::log2_tau_min(C);
    R t = R(0.5) * (lo + hi);
    // 3 Newton steps reach 1e-14 relative in nu (FP64); 2 reach 8e-7 for the
    // Carreau-Yasuda sets sx, gx, cy-a and cy-b over gamma = 1e-9..1 (lattice),
    // the accuracy of the FP32 evaluation itself, so FP32 kernels take 2
    constexpr int kNewton = sizeof(R) == 4 ? 2 : 3;
#pragma unroll
    for (int it = 0; it < kNewton; ++it) {
        const R x = Lg<R>::exp2_(a * (L + t));
        const R b = R(1) + x;
        const R pw = Lg<R>::exp2_(e * Lg<R>::log2_(b));
        const R tau = R(3) * (nr + (n0 - nr) * pw) + R(0.5);
        const R H = t + Lg<R>::log2_(tau) - lG;
        const R dnu = (n0 - nr) * pw * (n - R(1)) * Fm<R>::div(x, b);
        const R Hp = R(1) + R(3) * Fm<R>::div(dnu, tau);
        t = fmin(fmax(t - Fm<R>::div(H, Hp), lo), hi);
    }
    const R x = Lg<R>::exp2_(a * (L + t));
    return nr + (n0 - nr) * Lg<R>::exp2_(e * Lg<R>::log2_(R(1) + x));
}

// ---------------------------------------------------------------------------
// neighbour brick of the wavefront for a brick offset in {-1,0,1}^3 selected
// per lane: t[] holds the 27 storage indices (wave-uniform, in SGPRs), the
// booleans say whether the lane's shifted cell left the brick along x/y/z
// in the direction (dx,dy,dz).  Compile-time offsets fold the select tree.
__device__ __forceinline__ int pick_brick(const int (&t)[27], int dx, int dy, int dz, bool ex, bool ey, bool ez)
{
    const int a000 = t[13], a100 = t[13 + 9 * dx], a010 = t[13 + 3 * dy], a001 = t[13 + dz];
    const int a110 = t[13 + 9 * dx + 3 * dy], a101 = t[13 + 9 * dx + dz], a011 = t[13 + 3 * dy + dz];
    const int a111 = t[13 + 9 * dx + 3 * dy + dz];
    const int r00 = ex ? a100 : a000, r10 = ex ? a110 : a010;
    const int r01 = ex ? a101 : a001, r11 = ex ? a111 : a011;
    const int r0 = ey ? r10 : r00, r1 = ey ? r11 : r01;
    return ez ? r1 : r0;
}

__device__ __forceinline__ uint64_t lanes_below(int lane) { return (uint64_t(1) << lane) - 1; }

// balanced AA addressing of the two-byte storage types: the pair (cell P,
// slot s) of the AA scheme is stored at cell P - c_s for s < 13 and at P
// otherwise (host: lbm355_case::pair_position).  The AA steps then become
// alike: each reads the 13 populations q < 13 one cell upstream and writes the
// 13 populations q > 13 one cell downstream, instead of the odd step shifting
// all 26 (512x384x256 FP16: 93.7 -> 84.1 us per step).  The loads of a
// step: even steps read slot q, odd steps slot 26 - q
template <int S, bool DEV = false>
__device__ __forceinline__ void balanced_loads(const rsrc_t& F, const int (&t)[27], uint32_t own, int i, int j, int k,
                                               int lane, int parity, typename Pop<S>::R (&f)[27])
{
    using P = Pop<S>;
    auto ld = [&](uint32_t e, int q) { if constexpr (DEV) return P::ld_dev(F, e, q); else return P::ld(F, e, q); };
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const uint32_t sl = uint32_t(parity == 0 ? q : 26 - q) * kLanes;
        if (q < 13) {
            const int cx = qx(q), cy = qy(q), cz = qz(q);
            const int si = i - cx, sj = j - cy, sk = k - cz;
            const int nb = pick_brick(t, -cx, -cy, -cz, (si & ~(kBX - 1)) != 0, (sj & ~(kBY - 1)) != 0,
                                      (sk & ~(kBZ - 1)) != 0);
            f[q] = ld(uint32_t(nb + brick_lane(si, sj, sk)) + sl, q);
        } else {
            f[q] = ld(own + sl, q);
        }
    }
}

// ---------------------------------------------------------------------------
// the fused step
// one time step of brick b of case C at launch index `launch` (the case's
// step index is launch - C.start)
template <int S, int MODEL, bool NT, class CT>
__device__ __forceinline__ void step_brick(const CT& C, int64_t launch, int b, int lane)
{
    constexpr int kEvenAux = NT ? kAuxNt : 0;
    using P = Pop<S>;
    using R = typename P::R;
    const rsrc_t F = make_rsrc(C.f, C.f_bytes);
    const int64_t step = launch - C.start;
    const int parity = int(step & 1);
    // per-step output decisions, wave-uniform (32-bit step arithmetic)
    const uint32_t next = uint32_t(step) + 1u;
    // statistics are sampled at the steps whose index + 1 is a multiple of
    // stats_every (as MacroTcpc every step when it is 1).  Samples are counted
    // in groups of kStatBlock from the epoch (the last reset or restart): the
    // first kStatBlock - 1 of a group store their rho u in a ring of FP32
    // slots, the last adds the whole group to the FP64 sums
    const uint32_t kev = uint32_t(C.stats_every > 0 ? C.stats_every : 1);
    const uint32_t s32 = uint32_t(step), e32 = uint32_t(C.stats_epoch);
    const bool do_stats = C.stats && C.stats_every > 0 && (kev == 1u || (s32 + 1u) % kev == 0u);
    const uint32_t nsamp = kev == 1u ? s32 - e32 + 1u : (s32 + 1u) / kev - e32 / kev;  // samples since the epoch
    const uint32_t spos = (nsamp - 1u) % kStatBlock;
    const int64_t ring_slot_bytes = int64_t(C.n_active) * kRingFields * kLanes * 4;
    const bool do_macro = C.macro && (step == C.macro_step ||
                                      (C.macro_every > 0 && next % uint32_t(C.macro_every) == 0u));
    const rsrc_t SR = make_rsrc(C.stats, C.stats_bytes);
    R* const mo = uniform(static_cast<R*>(C.macro));
    // pulsatile inflow factor of this step (wave-uniform)
    double pulse = 1.0;
    if (C.pulse_amp != 0.0 && C.pulse_period > 0.0)
        pulse = 1.0 + C.pulse_amp * sin(6.283185307179586 * double(step) / C.pulse_period);
    const int lab = ((const __attribute__((address_space(1))) uint8_t*)C.label)[b * kLanes + lane];
    const bool fluid = lab == L_FLUID;
    const bool bc = is_bc_label(lab);
    const uint64_t bc_lanes = __ballot(bc);

    // the 27 neighbour bricks are wave-uniform: scalar loads into SGPRs
    const __attribute__((address_space(4))) int32_t* nbr_c =
        (const __attribute__((address_space(4))) int32_t*)C.nbr + b * 27;
    int t[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) t[q] = nbr_c[q];

    const int i = lane & (kBX - 1), j = (lane >> kSX) & (kBY - 1), k = lane >> (kSX + kSY);
    const uint32_t own = uint32_t(t[13]) + uint32_t(lane);

    // FP32 and two-byte storage run every collision on the stored deviations
    // f_q - w_q from load to store (the DEV forms of the collide_* functions);
    // FP64, the parity path, keeps full populations
    constexpr bool kDev = kBalanced<S> || S == S_F32;
    R f[27];
    R rho = R(1), ux = R(0), uy = R(0), uz = R(0);
    if (fluid) {
        if constexpr (kBalanced<S>) {
            balanced_loads<S, kDev>(F, t, own, i, j, k, lane, parity, f);
        } else if (parity == 0) {
#pragma unroll
            for (int q = 0; q < 27; ++q) {
                if constexpr (kDev) f[q] = P::template ld_dev<kEvenAux>(F, own + uint32_t(q * kLanes), q);
                else f[q] = P::template ld<kEvenAux>(F, own + uint32_t(q * kLanes), q);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 27; ++q) {
                const int cx = qx(q), cy = qy(q), cz = qz(q);
                const int si = i - cx, sj = j - cy, sk = k - cz;
                const int nb = pick_brick(t, -cx, -cy, -cz, (si & ~(kBX - 1)) != 0, (sj & ~(kBY - 1)) != 0,
                                          (sk & ~(kBZ - 1)) != 0);
                const int l2 = brick_lane(si, sj, sk);
                if constexpr (kDev) f[q] = P::ld_dev(F, uint32_t(nb + (26 - q) * kLanes + l2), q);
                else f[q] = P::ld(F, uint32_t(nb + (26 - q) * kLanes + l2), q);
            }
        }
        R Fx = R(0), Fy = R(0), Fz = R(0);
        if (C.force) {
            const R* fo = static_cast<const R*>(C.force);
            Fx = fo[(b * 3 + 0) * kLanes + lane];
            Fy = fo[(b * 3 + 1) * kLanes + lane];
            Fz = fo[(b * 3 + 2) * kLanes + lane];
        }
        // density and momentum (lbm_common.h:16-39) with the second moments
        // (with kDev: the sums of the deviations, so M.rho is rho - 1 and
        // M.p_aa is P_aa - 1/3; the other moments of w vanish)
        Moments<R> M;
        if constexpr (kDev) M = moments2_axis<R>(f);
        else M = moments2<R>(f);
        const R dr = kDev ? M.rho : M.rho - R(1);
        if constexpr (kDev) M.rho += R(1);
        rho = M.rho;
        const R ir = Fm<R>::rcp(rho);
        ux = (M.jx + R(0.5) * Fx) * ir;
        uy = (M.jy + R(0.5) * Fy) * ir;
        uz = (M.jz + R(0.5) * Fz) * ir;
        R nu = CyK<R>::nu_ref(C);
        if (C.cy_mode == 1) nu = neq_viscosity<R>(M, ux, uy, uz, C);
        const R om = Fm<R>::rcp(R(3) * nu + R(0.5));
        if constexpr (MODEL == M_CUM) collide_cumulant<R, kDev>(f, M, dr, ux, uy, uz, om);
        // BGK and KBC without a body force drop the source term at compile
        // time (FP32 arithmetic only: a second copy of the FP64 KBC, at 256
        // VGPRs with scratch, halves its speed)
        else if constexpr (MODEL == M_BGK) {
            if (sizeof(R) == 8 || C.force) collide_bgk<R, kDev, true>(f, rho, dr, ux, uy, uz, Fx, Fy, Fz, om);
            else if constexpr (sizeof(R) == 4) collide_bgk<R, kDev, false>(f, rho, dr, ux, uy, uz, Fx, Fy, Fz, om);
        }
        else if constexpr (MODEL == M_MRT) collide_mrt<R, kDev>(f, M, dr, ux, uy, uz, nu);
        else {
            if (sizeof(R) == 8 || C.force) collide_kbc<R, kDev, true>(f, M, dr, ux, uy, uz, Fx, Fy, Fz, nu);
            else if constexpr (sizeof(R) == 4) collide_kbc<R, kDev, false>(f, M, dr, ux, uy, uz, Fx, Fy, Fz, nu);
        }
    } else if (bc) {
        // sim_tcpc.cu:242-273: the populations entering the partner cell one
        // plane inwards give rho (inflow) or u (outflow); the partner and any
        // clamped-edge writer of those slots are lanes of this wavefront (see
        // the brick origin shift), so these loads precede their stores
        const int slot = C.bc_base[b] + __popcll(bc_lanes & lanes_below(lane));
        const int32_t* rd = C.bc_read + (slot * 2 + parity) * 27;
        R g[27];
#pragma unroll
        for (int q = 0; q < 27; ++q) g[q] = P::ld(F, uint32_t(rd[q]), q);
        const Moments<R> G = moments2<R>(g);
        if (is_inflow_label(lab)) {
            const R base = lab == L_IN_XM ? R(C.inflow_xm) : -R(C.inflow_xp);
            const R fac = R(C.bc_factor[slot]);
            ux = (C.plug ? base / fac : base * fac) * R(pulse);
            rho = G.rho;
        } else {
            ux = G.jx / G.rho; uy = G.jy / G.rho; uz = G.jz / G.rho;
            rho = R(1);
        }
        feq_boundary<R>(f, rho, ux, uy, uz);
        if constexpr (kDev) {
#pragma unroll
            for (int q = 0; q < 27; ++q) f[q] -= wgt<R>(q);
        }
    }
    if (fluid || bc) {
        // random bits of the two-byte stochastic rounding (unused at FP32/FP64)
        const uint32_t H = (S == S_F16 || S == S_BF16)
                               ? mix32(uint32_t(b * kLanes + lane) ^ mix32(uint32_t(step) * 0x85ebca6bu + 0x632be5abu))
                               : 0u;
        if constexpr (kBalanced<S>) {
            // the writes of q > 13 land one cell downstream, the others in the
            // cell's own slots: even steps write slot 26 - q, odd steps slot q
#pragma unroll
            for (int q = 0; q < 27; ++q) {
                const uint32_t sl = uint32_t(parity == 0 ? 26 - q : q) * kLanes;
                if (q > 13) {
                    const int cx = qx(q), cy = qy(q), cz = qz(q);
                    const int di = i + cx, dj = j + cy, dk = k + cz;
                    const int nb = pick_brick(t, cx, cy, cz, (di & ~(kBX - 1)) != 0, (dj & ~(kBY - 1)) != 0,
                                              (dk & ~(kBZ - 1)) != 0);
                    const uint32_t e = uint32_t(nb + brick_lane(di, dj, dk)) + sl;
                    if constexpr (kDev) P::st_dev(F, e, q, f[q], rnd_q(H, q));
                    else P::st(F, e, q, f[q], rnd_q(H, q));
                } else {
                    if constexpr (kDev) P::st_dev(F, own + sl, q, f[q], rnd_q(H, q));
                    else P::st(F, own + sl, q, f[q], rnd_q(H, q));
                }
            }
        } else if (parity == 0) {
#pragma unroll
            for (int q = 0; q < 27; ++q)
                if constexpr (kDev) P::template st_dev<kEvenAux>(F, own + uint32_t((26 - q) * kLanes), q, f[q], rnd_q(H, q));
                else P::template st<kEvenAux>(F, own + uint32_t((26 - q) * kLanes), q, f[q], rnd_q(H, q));
        } else {
#pragma unroll
            for (int q = 0; q < 27; ++q) {
                const int cx = qx(q), cy = qy(q), cz = qz(q);
                const int di = i + cx, dj = j + cy, dk = k + cz;
                const int nb = pick_brick(t, cx, cy, cz, (di & ~(kBX - 1)) != 0, (dj & ~(kBY - 1)) != 0,
                                          (dk & ~(kBZ - 1)) != 0);
                const int l2 = brick_lane(di, dj, dk);
                // unconditional: a BC-plane cell's writes that leave the grid
                // land in an outside lane of its brick or in the sink brick,
                // neither of which is ever read
                if constexpr (kDev) P::st_dev(F, uint32_t(nb + q * kLanes + l2), q, f[q], rnd_q(H, q));
                else P::st(F, uint32_t(nb + q * kLanes + l2), q, f[q], rnd_q(H, q));
            }
        }
        // macroscopic output (MacroTcpc::outputMacro, sim_tcpc.cu:337-354)
        if (do_macro) {
            mo[(b * 4 + 0) * kLanes + lane] = rho;
            mo[(b * 4 + 1) * kLanes + lane] = ux;
            mo[(b * 4 + 2) * kLanes + lane] = uy;
            mo[(b * 4 + 3) * kLanes + lane] = uz;
        }
        if (do_stats) {
            const uint32_t idx = uint32_t(b * kStatFields) * kLanes + uint32_t(lane);
            if constexpr (S == S_F64) {
                // FP64 runs add every sample to the FP64 sums directly
                const double rx = rho * ux, ry = rho * uy, rz = rho * uz;
                const double add[6] = {rx, ry, rz, rx * rx, ry * ry, rz * rz};
                double cur[6];
#pragma unroll
                for (int a = 0; a < 6; ++a) cur[a] = ld_f64(SR, (idx + uint32_t(a * kLanes)) * 8u);
#pragma unroll
                for (int a = 0; a < 6; ++a) st_f64(SR, (idx + uint32_t(a * kLanes)) * 8u, cur[a] + add[a]);
            } else {
                // 12 B per sample into ring slot spos; the group's last sample
                // reads the kStatBlock - 1 slots and adds the group to the FP64
                // sums (squares of the FP32 samples taken in FP64).  In the
                // non-temporal variant (cases larger than the Infinity Cache)
                // these accesses carry the hint as well: a slot is read back
                // up to 15 steps later, long after a larger case has swept the
                // cache, so allocating it only evicts population lines that
                // the odd step shares (512x384x256: 150.9 -> 144.9 us per step)
                const float r[3] = {float(rho * ux), float(rho * uy), float(rho * uz)};
                const uint32_t ridx = uint32_t(b * kRingFields) * kLanes + uint32_t(lane);
                if (spos < uint32_t(kStatBlock - 1)) {
                    const rsrc_t RS = make_rsrc(C.stats_part + int64_t(spos) * (ring_slot_bytes / 4), ring_slot_bytes);
#pragma unroll
                    for (int a = 0; a < 3; ++a) st_f32<kEvenAux>(RS, (ridx + uint32_t(a * kLanes)) * 4u, r[a]);
                } else {
                    float ring[kStatBlock - 1][3];
#pragma unroll
                    for (int sl = 0; sl < kStatBlock - 1; ++sl) {
                        const rsrc_t RS = make_rsrc(C.stats_part + int64_t(sl) * (ring_slot_bytes / 4), ring_slot_bytes);
#pragma unroll
                        for (int a = 0; a < 3; ++a) ring[sl][a] = ld_f32<kEvenAux>(RS, (ridx + uint32_t(a * kLanes)) * 4u);
                    }
                    double cur[6];
#pragma unroll
                    for (int a = 0; a < 6; ++a) cur[a] = ld_f64<kEvenAux>(SR, (idx + uint32_t(a * kLanes)) * 8u);
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        double sum = 0.0, sq = 0.0;
#pragma unroll
                        for (int sl = 0; sl < kStatBlock - 1; ++sl) {
                            const double v = double(ring[sl][a]);
                            sum += v;
                            sq += v * v;
                        }
                        const double v = double(r[a]);
                        st_f64<kEvenAux>(SR, (idx + uint32_t(a * kLanes)) * 8u, cur[a] + (sum + v));
                        st_f64<kEvenAux>(SR, (idx + uint32_t((a + 3) * kLanes)) * 8u, cur[a + 3] + (sq + v * v));
                    }
                }
            }
        }
    }
}

// velocity pre-pass of the fd-force Carreau-Yasuda mode (core.h:293-384):
// rho and u = (sum c f + F/2) / rho of every fluid cell from the populations
// its step is about to read (same AA addressing as step_brick) and the force
// of the previous step, into the macro buffer
template <int S>
__device__ __forceinline__ void prepass_brick(const StepLaunch& L, const BrickRef br, int lane)
{
    using P = Pop<S>;
    using R = typename P::R;
    const CaseDev& C = L.cases[br.case_id];
    const int b = int(br.brick);
    const rsrc_t F = make_rsrc(C.f, C.f_bytes);
    const int64_t step = *L.base + L.rel - C.start;
    const int parity = int(step & 1);
    const int lab = ((const __attribute__((address_space(1))) uint8_t*)C.label)[b * kLanes + lane];
    if (lab != L_FLUID) return;
    const __attribute__((address_space(4))) int32_t* nbr_c =
        (const __attribute__((address_space(4))) int32_t*)C.nbr + b * 27;
    int t[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) t[q] = nbr_c[q];
    const int i = lane & (kBX - 1), j = (lane >> kSX) & (kBY - 1), k = lane >> (kSX + kSY);
    R f[27];
    if constexpr (kBalanced<S>) {
        balanced_loads<S>(F, t, uint32_t(t[13]) + uint32_t(lane), i, j, k, lane, parity, f);
    } else if (parity == 0) {
        const uint32_t own = uint32_t(t[13]) + uint32_t(lane);
#pragma unroll
        for (int q = 0; q < 27; ++q) f[q] = P::ld(F, own + uint32_t(q * kLanes), q);
    } else {
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            const int cx = qx(q), cy = qy(q), cz = qz(q);
            const int si = i - cx, sj = j - cy, sk = k - cz;
            const int nb = pick_brick(t, -cx, -cy, -cz, (si & ~(kBX - 1)) != 0, (sj & ~(kBY - 1)) != 0,
                                      (sk & ~(kBZ - 1)) != 0);
            const int l2 = brick_lane(si, sj, sk);
            f[q] = P::ld(F, uint32_t(nb + (26 - q) * kLanes + l2), q);
        }
    }
    const R* fo = static_cast<const R*>(C.force);
    const R Fx = fo[(b * 3 + 0) * kLanes + lane], Fy = fo[(b * 3 + 1) * kLanes + lane];
    const R Fz = fo[(b * 3 + 2) * kLanes + lane];
    const Moments<R> M = moments2<R>(f);
    R* mo = static_cast<R*>(C.macro);
    mo[(b * 4 + 0) * kLanes + lane] = M.rho;
    mo[(b * 4 + 1) * kLanes + lane] = (M.jx + R(0.5) * Fx) / M.rho;
    mo[(b * 4 + 2) * kLanes + lane] = (M.jy + R(0.5) * Fy) / M.rho;
    mo[(b * 4 + 3) * kLanes + lane] = (M.jz + R(0.5) * Fz) / M.rho;
}

// XCD-aware mapping: workgroups go round-robin to the 8 XCDs; remap them so
// that each XCD walks one contiguous (slab-ordered, hence spatially compact)
// chunk of the brick list and its L2 sees the neighbour bricks.  Odd launches
// walk each chunk backwards, so a step starts on the bricks the previous step
// finished with, whose lines are still in L2 and the Infinity Cache (bricks
// of one step are independent, so the order does not change results).
// 256x192x128 junction: 19.2 -> 18.6 us per step.
__device__ __forceinline__ int xcd_block(int bid, int grid, bool backwards)
{
    const int per = grid >> 3;
    const int pos = backwards ? per - 1 - (bid >> 3) : (bid >> 3);
    return (bid & 7) * per + pos;
}

// the step index of a launch is *base + rel: host-driven launches pass it in
// rel (base stays put), a captured graph of K steps bakes rel = 0..K-1 and
// ends with k_advance, so replays need no host update and no device-wide
// atomics or fences inside the step kernels
__global__ void k_advance(int64_t* base, int64_t k) { *base += k; }

template <int S, int MODEL, bool NT>
__global__ void __launch_bounds__(512) k_step(StepLaunch L)
{
    const int wid = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
    const int64_t launch = *L.base + L.rel;
    const bool backwards = (launch & 1) != 0;
    const int blk = xcd_block(int(blockIdx.x), int(gridDim.x), backwards);
    const int w = blk * L.bricks_per_group + wid;
    if (w < L.n_list) {
        const BrickRef br = L.list[w];
        step_brick<S, MODEL, NT>(L.cases[br.case_id], launch, int(br.brick), int(threadIdx.x) & 63);
    }
}

// the batch descriptor, the brick list and the member constants are not
// written while a step runs, so the batch kernels read them through the
// constant address space (scalar loads into SGPRs)
template <class T>
using const_as = const __attribute__((address_space(4))) T;

// Batches step through a device-resident descriptor (brick list, member
// constants, the brick range of the group being stepped), so that a captured
// graph of these launches stays valid when the membership changes between
// replays, and one graph serves every group of the batch (the range is set
// before each replay).  k_step_range keeps one brick per wavefront: the grid
// (n_list workgroups, fixed at capture, at least the largest group) is split
// into 8 XCD chunks over the range's ceil(n / bpg) workgroups as k_step does,
// and the workgroups past the range exit at once.
template <int S, int MODEL, bool NT>
__global__ void __launch_bounds__(512) k_step_range(StepLaunch L)
{
    const_as<BatchDev>& B = *(const_as<BatchDev>*)L.batch;
    const int n = B.end - B.begin;
    const int64_t launch = *L.base + L.rel;
    const bool backwards = (launch & 1) != 0;
    const int bpg = L.bricks_per_group;
    const int chunk = ((n + bpg - 1) / bpg + 7) >> 3;
    const int pos = int(blockIdx.x) >> 3;
    if (pos >= chunk) return;
    const int wid = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
    const int w = ((int(blockIdx.x) & 7) * chunk + (backwards ? chunk - 1 - pos : pos)) * bpg + wid;
    if (w < n) {
        const_as<BrickRef>& br = ((const_as<BrickRef>*)B.list)[B.begin + w];
        step_brick<S, MODEL, NT>(((const_as<CaseDev>*)B.cases)[br.case_id], launch, int(br.brick),
                                 int(threadIdx.x) & 63);
    }
}

// The persistent form for batches: a grid of a fixed size, multiple of 8,
// whose wavefronts stride over the brick range [begin, end) of the batch
// descriptor.  The brick list, the member constants and the range are read
// from device memory, so a captured graph of these launches stays valid when
// the batch's membership changes between replays, and one graph serves every
// group of the batch (the range is set before each replay).  XCD x (blockIdx
// % 8) walks the x-th eighth of the range, backwards on odd launches, as
// k_step does; every wavefront leaves the loop when its next index passes
// the end of its eighth.
template <int S, int MODEL, bool NT>
__global__ void __launch_bounds__(512) k_step_loop(StepLaunch L)
{
    const_as<BatchDev>& B = *(const_as<BatchDev>*)L.batch;
    const_as<BrickRef>* const list = (const_as<BrickRef>*)B.list;
    const_as<CaseDev>* const cases = (const_as<CaseDev>*)B.cases;
    const int begin = B.begin, n = B.end - B.begin;
    const int64_t launch = *L.base + L.rel;
    const bool backwards = (launch & 1) != 0;
    const int wid = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
    const int xcd = int(blockIdx.x) & 7;
    const int chunk = (n + 7) >> 3;
    const int c0 = xcd * chunk, c1 = min(c0 + chunk, n);
    const int stride = (int(gridDim.x) >> 3) * L.bricks_per_group;
#pragma nounroll
    for (int k = (int(blockIdx.x) >> 3) * L.bricks_per_group + wid; c0 + k < c1; k += stride) {
        const_as<BrickRef>& br = list[begin + (backwards ? c1 - 1 - k : c0 + k)];
        // the lane's 54 neighbour offsets and edge masks depend on the lane
        // only; hoisted out of the loop they would stay live across
        // iterations (134 instead of 63 VGPRs, 3 instead of 8 wavefronts per
        // SIMD), so the lane index is made opaque to loop-invariant motion
        int lane = int(threadIdx.x) & 63;
        asm volatile("" : "+v"(lane));
        step_brick<S, MODEL, NT>(cases[br.case_id], launch, int(br.brick), lane);
    }
}

template <int S>
__global__ void __launch_bounds__(64) k_prepass(StepLaunch L)
{
    if (int(blockIdx.x) < L.n_list) prepass_brick<S>(L, L.list[blockIdx.x], int(threadIdx.x));
}

// a brick and its one-cell halo, staged in LDS by the strain and force kernels
constexpr int kHX = kBX + 2, kHY = kBY + 2, kHZ = kBZ + 2, kHalo = kHX * kHY * kHZ;

// body force of the fd-force mode (core.h:663-781) from the strain of the
// pre-pass velocities: F_c = sum_a d_a S_{a c} with the one-sided / central /
// zero rule of the fluid neighbours, nu from the cell's own strain in single
// precision as the reference's powf (core.h:766), and
// F = 2 (nu - nu_ref) F_c rho with rho still 1 at that point of the reference
// kernel.  Non-fluid cells get no force.
template <class R>
__global__ void __launch_bounds__(64) k_force(const CaseDev* Cp)
{
    const CaseDev& C = *Cp;
    const int b = int(blockIdx.x);
    const int lane = int(threadIdx.x);
    __shared__ R ss[6][kHalo];
    __shared__ uint8_t sf[kHalo];
    const R* so = static_cast<const R*>(C.strain);
    for (int c = lane; c < kHalo; c += 64) {
        const int hx = c % kHX, hy = (c / kHX) % kHY, hz = c / (kHX * kHY);
        const int gx = hx - 1, gy = hy - 1, gz = hz - 1;
        const int ox = (gx < 0) ? -1 : (gx > kBX - 1 ? 1 : 0);
        const int oy = (gy < 0) ? -1 : (gy > kBY - 1 ? 1 : 0);
        const int oz = (gz < 0) ? -1 : (gz > kBZ - 1 ? 1 : 0);
        const int nb = C.nbr[b * 27 + 13 + 9 * ox + 3 * oy + oz] / (27 * kLanes);
        const int l = brick_lane(gx, gy, gz);
        const bool fl = C.label[nb * kLanes + l] == L_FLUID && nb < C.n_active;
        sf[c] = fl ? 1 : 0;
#pragma unroll
        for (int m = 0; m < 6; ++m) ss[m][c] = fl ? so[(nb * 6 + m) * kLanes + l] : R(0);
    }
    __syncthreads();
    const int i = lane & (kBX - 1), j = (lane >> kSX) & (kBY - 1), k = lane >> (kSX + kSY);
    const int c0 = (i + 1) + kHX * ((j + 1) + kHY * (k + 1));
    R Fo[3] = {R(0), R(0), R(0)};
    if (sf[c0]) {
        // S components: 0 S11, 1 S12, 2 S13, 3 S22, 4 S32, 5 S33; rows of the
        // derivative along x, y, z
        const int rows[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
        const int st[3] = {1, kHX, kHX * kHY};
        R Fc[3] = {R(0), R(0), R(0)};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int cm = c0 - st[a], cp = c0 + st[a];
            const bool m = sf[cm], p = sf[cp];
            if (!m && !p) continue;
#pragma unroll
            for (int comp = 0; comp < 3; ++comp) {
                const R* v = ss[rows[a][comp]];
                if (!m) Fc[comp] += v[cp] - v[c0];
                else if (!p) Fc[comp] += v[c0] - v[cm];
                else Fc[comp] += R(0.5) * (v[cp] - v[cm]);
            }
        }
        const R s11 = ss[0][c0], s12 = ss[1][c0], s13 = ss[2][c0], s22 = ss[3][c0], s32 = ss[4][c0], s33 = ss[5][c0];
        const R gam = sqrt(R(2)) * sqrt(s11 * s11 + s22 * s22 + s33 * s33 + R(2) * (s12 * s12 + s13 * s13 + s32 * s32));
        // single-precision pow as the correctly rounded float of the double
        // pow (the oracle evaluates the same expression)
        auto powf_cr = [](float x, float y) { return float(pow(double(x), double(y))); };
        const float inner = powf_cr(float(gam * R(C.cy_lambda)), float(C.cy_a));
        const R nu = R(C.nu_ref) + (R(C.cy_nu0) - R(C.nu_ref)) *
                                       R(powf_cr(float(1.0 + double(inner)), float((C.cy_n - 1.0) / C.cy_a)));
#pragma unroll
        for (int comp = 0; comp < 3; ++comp) Fo[comp] = R(2) * (nu - R(C.nu_ref)) * Fc[comp] * R(1);
    }
    R* fo = static_cast<R*>(C.force);
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) fo[(b * 3 + comp) * kLanes + lane] = Fo[comp];
}

template <int S>
__global__ void k_fill_rest(typename Pop<S>::T* f, int64_t n)
{
    using P = Pop<S>;
    using R = typename P::R;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int q = int((i / kLanes) % kQ);
    // rest state f = f_eq(1, 0) = w_q (resetLattice, sim_tcpc.cu:647-655)
    P::put(f, i, q, R(lattice_weight(q)));
}

__global__ void k_zero(uint32_t* p, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// single-value device updates (kernel arguments are copied at launch, so the
// host value can change right after the call; no memset/memcpy nodes)
__global__ void k_set_u64(uint64_t* p, uint64_t v) { *p = v; }
__global__ void k_set_batch(BatchDev* p, BatchDev v) { *p = v; }

// ---------------------------------------------------------------------------
// strain S of the velocity field with the reference stencil
// (cudaLBMKernelStress, solver/lbm3d/core.h:407-576): per axis, central
// difference when both face neighbours are fluid, one-sided when one is, zero
// when none is; S = 0 on non-fluid cells.  The brick and its one-cell halo are
// staged in LDS (its (kBX+2) x (kBY+2) x (kBZ+2) cells x (fluid flag, u)).
template <class R>
__global__ void __launch_bounds__(64) k_strain(const CaseDev* Cp)
{
    const CaseDev& C = *Cp;
    const int b = int(blockIdx.x);
    const int lane = int(threadIdx.x);
    __shared__ R su[3][kHalo];
    __shared__ uint8_t sf[kHalo];
    const R* mo = static_cast<const R*>(C.macro);
    for (int c = lane; c < kHalo; c += 64) {
        const int hx = c % kHX, hy = (c / kHX) % kHY, hz = c / (kHX * kHY);
        const int gx = hx - 1, gy = hy - 1, gz = hz - 1;  // cell offset inside the brick (-1..4)
        const int ox = (gx < 0) ? -1 : (gx > kBX - 1 ? 1 : 0);
        const int oy = (gy < 0) ? -1 : (gy > kBY - 1 ? 1 : 0);
        const int oz = (gz < 0) ? -1 : (gz > kBZ - 1 ? 1 : 0);
        const int nb = C.nbr[b * 27 + 13 + 9 * ox + 3 * oy + oz] / (27 * kLanes);
        const int l = brick_lane(gx, gy, gz);
        const bool fl = C.label[nb * kLanes + l] == L_FLUID;
        sf[c] = fl ? 1 : 0;
        R v0 = R(0), v1 = R(0), v2 = R(0);
        if (fl && nb < C.n_active) {
            v0 = mo[(nb * 4 + 1) * kLanes + l];
            v1 = mo[(nb * 4 + 2) * kLanes + l];
            v2 = mo[(nb * 4 + 3) * kLanes + l];
        }
        su[0][c] = v0; su[1][c] = v1; su[2][c] = v2;
    }
    __syncthreads();
    const int i = lane & (kBX - 1), j = (lane >> kSX) & (kBY - 1), k = lane >> (kSX + kSY);
    const int c0 = (i + 1) + kHX * ((j + 1) + kHY * (k + 1));
    R S11 = 0, S12 = 0, S13 = 0, S22 = 0, S32 = 0, S33 = 0;
    if (sf[c0]) {
        const int st[3] = {1, kHX, kHX * kHY};
        R d[3][3];  // d[axis][component]
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int cm = c0 - st[a], cp = c0 + st[a];
            const bool m = sf[cm], p = sf[cp];
#pragma unroll
            for (int comp = 0; comp < 3; ++comp) {
                R v;
                if (!m && !p) v = R(0);
                else if (!m) v = su[comp][cp] - su[comp][c0];
                else if (!p) v = su[comp][c0] - su[comp][cm];
                else v = R(0.5) * (su[comp][cp] - su[comp][cm]);
                d[a][comp] = v;
            }
        }
        S11 = d[0][0]; S22 = d[1][1]; S33 = d[2][2];
        S12 = R(0.5) * d[0][1] + R(0.5) * d[1][0];
        S13 = R(0.5) * d[0][2] + R(0.5) * d[2][0];
        S32 = R(0.5) * d[1][2] + R(0.5) * d[2][1];
    }
    R* so = static_cast<R*>(C.strain);
    so[(b * 6 + 0) * kLanes + lane] = S11;
    so[(b * 6 + 1) * kLanes + lane] = S12;
    so[(b * 6 + 2) * kLanes + lane] = S13;
    so[(b * 6 + 3) * kLanes + lane] = S22;
    so[(b * 6 + 4) * kLanes + lane] = S32;
    so[(b * 6 + 5) * kLanes + lane] = S33;
}

// ---------------------------------------------------------------------------
// plane fluxes: out[0..3] += sum over BC cells of u_n * sign for the IVC,
// SVC/x+ plane, label-5 and label-6 planes (compute_flows.py:107-120, with
// inlets reported as positive inflow, compute_flows.py:133-137); wavefront
// shuffle reduction then one FP64 atomic per wavefront and plane.
//   mean = 0: u of the BC cell from the macro output of the last step;
//   mean = 1: the density-weighted mean momentum sum rho u (the mean_velocity
//             output, sim_tcpc.cu:724-727) from the FP64 statistics sums plus
//             the first `open` FP32 samples of the statistics ring (the open
//             group); the caller divides by the sample count.
template <class R>
__global__ void __launch_bounds__(256) k_flux(const CaseDev* Cp, double* out, int mean, int open)
{
    const CaseDev& C = *Cp;
    const int s = int(blockIdx.x * blockDim.x + threadIdx.x);
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (s < C.n_bc) {
        const int lab = C.bc_label[s];
        int comp = 1, plane = 0;
        double sign = 1.0;
        switch (lab) {
        case L_IN_XM: plane = 0; comp = 1; sign = 1.0; break;   // inflow along +x
        case L_OUT_XM: plane = 0; comp = 1; sign = -1.0; break;
        case L_IN_XP: plane = 1; comp = 1; sign = -1.0; break;  // inflow along -x
        case L_OUT_XP: plane = 1; comp = 1; sign = 1.0; break;
        case L_OUT_YM: plane = 2; comp = 2; sign = -1.0; break; // outward = -y
        default: plane = 3; comp = 2; sign = 1.0; break;        // outward = +y
        }
        const int cell = C.bc_cell[s];
        const int b = cell >> 6, l = cell & 63;
        double u;
        if (!mean) {
            u = double(static_cast<const R*>(C.macro)[(b * kMacroFields + comp) * kLanes + l]);
        } else {
            // rho u_comp: field comp - 1 of the sums and of each ring slot
            u = C.stats[(int64_t(b) * kStatFields + comp - 1) * kLanes + l];
            const int64_t slot = int64_t(C.n_active) * kRingFields * kLanes;
            for (int k = 0; k < open; ++k)
                u += double(C.stats_part[k * slot + (int64_t(b) * kRingFields + comp - 1) * kLanes + l]);
        }
        v[plane] = u * sign;
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        double x = v[p];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if ((threadIdx.x & 63) == 0 && x != 0.0) atomicAdd(out + p, x);
    }
}

// ---------------------------------------------------------------------------
// difference norms of two cases with the same brick index (the same geometry
// and grid): per non-wall cell |u_a - u_b|^2, |u_b|^2, (rho_a - rho_b)^2 and
// (rho_b - 1)^2 from the macro output of their last steps.  Pass 1: each
// workgroup of kCompareWaves wavefronts takes kCompareWaves consecutive
// bricks and writes its 4 sums to partials[4 * blockIdx + k] (every
// workgroup writes all 4, so nothing has to start at zero); pass 2: one
// workgroup adds the partials in a fixed order into partials[4 * nblocks + k].
// The summation order depends on the brick count only, so the result is the
// same bits whatever buffer holds the partials.
constexpr int kCompareWaves = 4;

template <class Ra, class Rb>
__global__ void __launch_bounds__(64 * kCompareWaves) k_compare(const CaseDev* A, const CaseDev* B, double* partials)
{
    __shared__ double red[kCompareWaves][4];
    const int wid = int(threadIdx.x) >> 6, lane = int(threadIdx.x) & 63;
    const int b = int(blockIdx.x) * kCompareWaves + wid;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (b < B->n_active) {
        const int lab = B->label[b * kLanes + lane];
        if (lab != L_WALL && lab != L_OUTSIDE) {
            const Ra* ma = static_cast<const Ra*>(A->macro);
            const Rb* mb = static_cast<const Rb*>(B->macro);
            double d2 = 0.0, r2 = 0.0;
            for (int c = 1; c <= 3; ++c) {
                const double ua = double(ma[(b * kMacroFields + c) * kLanes + lane]);
                const double ub = double(mb[(b * kMacroFields + c) * kLanes + lane]);
                d2 += (ua - ub) * (ua - ub);
                r2 += ub * ub;
            }
            const double ra = double(ma[b * kMacroFields * kLanes + lane]);
            const double rb = double(mb[b * kMacroFields * kLanes + lane]);
            v[0] = d2;
            v[1] = r2;
            v[2] = (ra - rb) * (ra - rb);
            v[3] = (rb - 1.0) * (rb - 1.0);
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) red[wid][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double x = 0.0;
        for (int w = 0; w < kCompareWaves; ++w) x += red[w][threadIdx.x];
        partials[4 * blockIdx.x + threadIdx.x] = x;
    }
}

__global__ void __launch_bounds__(256) k_sum_partials(double* partials, int nblocks)
{
    __shared__ double red[4][4];
    const int lane = int(threadIdx.x) & 63, wid = int(threadIdx.x) >> 6;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = int(threadIdx.x); i < nblocks; i += int(blockDim.x))
        for (int k = 0; k < 4; ++k) v[k] += partials[4 * i + k];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) red[wid][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < 4) partials[4 * nblocks + threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x] +
                                                               red[2][threadIdx.x] + red[3][threadIdx.x];
}

// ---------------------------------------------------------------------------
static thread_local const char* g_kerr = "";
const char* kernel_error() { return g_kerr; }

static int check_launch()
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_kerr = hipGetErrorString(e);
        return -3;
    }
    return 0;
}

template <int S, int MODEL>
static int launch_step_t(const StepLaunch& L, hipStream_t st)
{
    const int bpg = L.bricks_per_group;
    const unsigned lds = unsigned(L.lds_bytes);
    if (L.batch) {
        // batch forms: L.n_list is the grid size in workgroups
        const int groups = std::max(8, (L.n_list + 7) / 8 * 8);
        if (L.per_wave > 1) {
            if (L.nontemporal) hipLaunchKernelGGL((k_step_loop<S, MODEL, true>), dim3(groups), dim3(64 * bpg), lds, st, L);
            else hipLaunchKernelGGL((k_step_loop<S, MODEL, false>), dim3(groups), dim3(64 * bpg), lds, st, L);
        } else {
            if (L.nontemporal) hipLaunchKernelGGL((k_step_range<S, MODEL, true>), dim3(groups), dim3(64 * bpg), lds, st, L);
            else hipLaunchKernelGGL((k_step_range<S, MODEL, false>), dim3(groups), dim3(64 * bpg), lds, st, L);
        }
        return check_launch();
    }
    int groups = (L.n_list + bpg - 1) / bpg;
    groups = ((groups + 7) / 8) * 8;
    if (groups == 0) groups = 8;
    if (L.nontemporal) hipLaunchKernelGGL((k_step<S, MODEL, true>), dim3(groups), dim3(64 * bpg), lds, st, L);
    else hipLaunchKernelGGL((k_step<S, MODEL, false>), dim3(groups), dim3(64 * bpg), lds, st, L);
    return check_launch();
}

template <int S>
static int launch_step_s(int model, const StepLaunch& L, hipStream_t st)
{
    switch (model) {
    case M_CUM: return launch_step_t<S, M_CUM>(L, st);
    case M_BGK: return launch_step_t<S, M_BGK>(L, st);
    case M_MRT: return launch_step_t<S, M_MRT>(L, st);
    case M_KBC: return launch_step_t<S, M_KBC>(L, st);
    default: g_kerr = "unknown collision"; return -1;
    }
}

int launch_step(int storage, int model, const StepLaunch& L, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (storage) {
    case S_F64: return launch_step_s<S_F64>(model, L, st);
    case S_F32: return launch_step_s<S_F32>(model, L, st);
    case S_F16: return launch_step_s<S_F16>(model, L, st);
    case S_BF16: return launch_step_s<S_BF16>(model, L, st);
    default: g_kerr = "unknown storage"; return -1;
    }
}

int launch_fill_populations(int storage, void* f, int64_t n_bricks, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t n = n_bricks * kQ * kLanes;
    const int nb = int((n + 255) / 256);
    switch (storage) {
    case S_F64: hipLaunchKernelGGL(k_fill_rest<S_F64>, dim3(nb), dim3(256), 0, st, static_cast<double*>(f), n); break;
    case S_F32: hipLaunchKernelGGL(k_fill_rest<S_F32>, dim3(nb), dim3(256), 0, st, static_cast<float*>(f), n); break;
    case S_F16: hipLaunchKernelGGL(k_fill_rest<S_F16>, dim3(nb), dim3(256), 0, st, static_cast<_Float16*>(f), n); break;
    case S_BF16: hipLaunchKernelGGL(k_fill_rest<S_BF16>, dim3(nb), dim3(256), 0, st, static_cast<uint16_t*>(f), n); break;
    default: return -1;
    }
    return check_launch();
}

int launch_strain(int storage, const CaseDev* dev_case, int n_active, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (n_active <= 0) return 0;
    if (storage == S_F64) hipLaunchKernelGGL(k_strain<double>, dim3(n_active), dim3(64), 0, st, dev_case);
    else hipLaunchKernelGGL(k_strain<float>, dim3(n_active), dim3(64), 0, st, dev_case);
    return check_launch();
}

int launch_flux(int storage, const CaseDev* dev_case, int n_bc, double* dev_out, int mean, int open, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (n_bc <= 0) return 0;
    const int nb = (n_bc + 255) / 256;
    if (storage == S_F64) hipLaunchKernelGGL(k_flux<double>, dim3(nb), dim3(256), 0, st, dev_case, dev_out, mean, open);
    else hipLaunchKernelGGL(k_flux<float>, dim3(nb), dim3(256), 0, st, dev_case, dev_out, mean, open);
    return check_launch();
}

int compare_blocks(int n_active) { return std::max(1, (n_active + kCompareWaves - 1) / kCompareWaves); }

int launch_compare(int storage_a, int storage_b, const CaseDev* a, const CaseDev* b, int n_active, double* partials,
                   void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int nb = compare_blocks(n_active);
    const bool da = storage_a == S_F64, db = storage_b == S_F64;
    const dim3 grid(nb), block(64 * kCompareWaves);
    if (da && db) hipLaunchKernelGGL((k_compare<double, double>), grid, block, 0, st, a, b, partials);
    else if (da) hipLaunchKernelGGL((k_compare<double, float>), grid, block, 0, st, a, b, partials);
    else if (db) hipLaunchKernelGGL((k_compare<float, double>), grid, block, 0, st, a, b, partials);
    else hipLaunchKernelGGL((k_compare<float, float>), grid, block, 0, st, a, b, partials);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, partials, nb);
    return check_launch();
}

// the partials of lbm355_case_compare_noalloc: a file-scope device array, so
// that the call allocates nothing.  Each GPU holds its own copy for the life
// of the process; every call on one GPU shares it (the kernels write every
// partial before reading it, so nothing has to start at zero), and calls
// that overlap on one GPU race.  In host code `scratch` is only a stand-in
// for the symbol: kernels get the device address that hipGetSymbolAddress
// returns for each device.
constexpr int kCompareScratchBlocks = 16384;  // workgroups of 4 bricks: up to 65,536 active bricks
static __device__ __align__(256) double scratch[4 * kCompareScratchBlocks + 4];

int compare_scratch(int device, double** ptr, int* max_blocks)
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static double* per_device[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) { g_kerr = "device index out of range"; return -1; }
    std::lock_guard<std::mutex> lock(mu);
    if (!per_device[device]) {
        // looked up once per device, with that device current
        int prev = 0;
        if (hipGetDevice(&prev) != hipSuccess) { g_kerr = "hipGetDevice failed"; return -3; }
        if (prev != device && hipSetDevice(device) != hipSuccess) { g_kerr = "hipSetDevice failed"; return -3; }
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != device) (void)hipSetDevice(prev);
        if (e != hipSuccess || !p) { g_kerr = e != hipSuccess ? hipGetErrorString(e) : "null symbol address"; return -3; }
        per_device[device] = static_cast<double*>(p);
    }
    *ptr = per_device[device];
    *max_blocks = kCompareScratchBlocks;
    return 0;
}

int launch_zero(void* ptr, int64_t bytes, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t n = bytes / 4;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_zero, dim3(int((n + 255) / 256)), dim3(256), 0, st, static_cast<uint32_t*>(ptr), n);
    return check_launch();
}

int launch_set_batch(BatchDev* dev_ptr, const BatchDev& value, void* stream)
{
    hipLaunchKernelGGL(k_set_batch, dim3(1), dim3(1), 0, static_cast<hipStream_t>(stream), dev_ptr, value);
    return check_launch();
}

int launch_set_u64(void* dev_ptr, uint64_t value, void* stream)
{
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, static_cast<hipStream_t>(stream),
                       static_cast<uint64_t*>(dev_ptr), value);
    return check_launch();
}

// plain vectorised copy, the on-box bandwidth reference
// the on-box bandwidth reference: one 16-byte element per work-item with
// non-temporal loads and stores (the fastest plain copy measured on MI355X:
// 6.6 TB/s against 4.8 TB/s for a grid-stride loop)
typedef float f32x4_t __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_copy(const f32x4_t* __restrict__ src, f32x4_t* __restrict__ dst, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
}

int launch_copy(void* dst, const void* src, int64_t bytes, void* stream)
{
    const int64_t n = bytes / 16;
    const int64_t groups = std::max<int64_t>(1, (n + 255) / 256);
    if (groups > int64_t(0x7FFFFFFF)) { g_kerr = "copy too large"; return -1; }
    hipLaunchKernelGGL(k_copy, dim3(uint32_t(groups)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const f32x4_t*>(src), static_cast<f32x4_t*>(dst), n);
    return check_launch();
}

int launch_advance(int64_t* base, int64_t k, void* stream)
{
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, static_cast<hipStream_t>(stream), base, k);
    return check_launch();
}

int launch_fd_force_prepass(int storage, const StepLaunch& L, const CaseDev* dev_case, int n_active, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (n_active <= 0) return 0;
    switch (storage) {
    case S_F64: hipLaunchKernelGGL(k_prepass<S_F64>, dim3(n_active), dim3(64), 0, st, L); break;
    case S_F32: hipLaunchKernelGGL(k_prepass<S_F32>, dim3(n_active), dim3(64), 0, st, L); break;
    case S_F16: hipLaunchKernelGGL(k_prepass<S_F16>, dim3(n_active), dim3(64), 0, st, L); break;
    case S_BF16: hipLaunchKernelGGL(k_prepass<S_BF16>, dim3(n_active), dim3(64), 0, st, L); break;
    default: g_kerr = "unknown storage"; return -1;
    }
    int rc = check_launch();
    if (rc) return rc;
    if ((rc = launch_strain(storage, dev_case, n_active, stream))) return rc;
    if (storage == S_F64) hipLaunchKernelGGL(k_force<double>, dim3(n_active), dim3(64), 0, st, dev_case);
    else hipLaunchKernelGGL(k_force<float>, dim3(n_active), dim3(64), 0, st, dev_case);
    return check_launch();
}

}  // namespace lbm355
