// frg_device.h -- device primitives for the CDNA4 (gfx950) path tracker.
//
//  * cx<R>       complex number over R = double or float ({re, im}, bit-compatible with the
//                reference's Cplx for R = double, classes.cuh:26-33).
//  * group ops   8- or 16-lane groups inside a wave64 (one path per group): butterfly sums,
//                maxima, arg-max and broadcasts built on DPP (quad_perm, row_half_mirror and, for
//                16 lanes, row_mirror), so every lane of a group ends with bitwise identical
//                results; pair ops between lanes 2i and 2i+1.
//  * dd          double-double arithmetic (error-free transformations) for extended residuals.
//  * Philox4x32-10 (Salmon et al., SC'11) counter-based RNG.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace frg {

// ------------------------------------------------------------------------------------------
// complex arithmetic
// ------------------------------------------------------------------------------------------
template <typename R>
struct cx {
    R re, im;
};

template <typename R>
__host__ __device__ __forceinline__ cx<R> mk(R a, R b) {
    cx<R> z;
    z.re = a;
    z.im = b;
    return z;
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator-(cx<R> a) { return mk(-a.re, -a.im); }
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) {
    return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator*(R s, cx<R> a) { return mk(s * a.re, s * a.im); }
template <typename R>
__host__ __device__ __forceinline__ R abs2(cx<R> a) { return a.re * a.re + a.im * a.im; }
template <typename R>
__host__ __device__ __forceinline__ cx<R> conj(cx<R> a) { return mk(a.re, -a.im); }
// Real reciprocal.  On the device: the hardware estimate (v_rcp_f64 / v_rcp_f32) refined by
// Newton-Raphson steps r <- r + r(1 - x r), two for FP64 and one for FP32, instead of the full
// IEEE division sequence; the operands here are finite, nonzero and far from the exponent
// limits.  Generic R (used by the host-side operation counter): 1/x.
template <typename R>
__host__ __device__ __forceinline__ R recip(R x) { return R(1) / x; }
__device__ __forceinline__ double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    return r;
}
__device__ __forceinline__ float recip(float x) {
    float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}
// 1/a = conj(a)/|a|^2 (one real reciprocal); values here stay far from over/underflow.
template <typename R>
__host__ __device__ __forceinline__ cx<R> inv(cx<R> a) {
    R r = recip(abs2(a));
    return mk(a.re * r, -a.im * r);
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> cdiv(cx<R> a, cx<R> b) { return a * inv(b); }
// acc - a*b
template <typename R>
__host__ __device__ __forceinline__ cx<R> fms(cx<R> acc, cx<R> a, cx<R> b) {
    return mk(acc.re - a.re * b.re + a.im * b.im, acc.im - a.re * b.im - a.im * b.re);
}
// acc + a*b
template <typename R>
__host__ __device__ __forceinline__ cx<R> fma(cx<R> acc, cx<R> a, cx<R> b) {
    return mk(acc.re + a.re * b.re - a.im * b.im, acc.im + a.re * b.im + a.im * b.re);
}
template <typename R, typename S>
__host__ __device__ __forceinline__ cx<R> cvt(cx<S> a) { return mk(R(a.re), R(a.im)); }
template <typename R>
__host__ __device__ __forceinline__ cx<R> sel(bool c, cx<R> a, cx<R> b) { return mk(c ? a.re : b.re, c ? a.im : b.im); }

// ------------------------------------------------------------------------------------------
// Double-double arithmetic (value = hi + lo, |lo| <= ulp(hi)/2): Knuth's TwoSum, Dekker's
// product through a fused multiply-add, and the accurate sum and product of Hida, Li and Bailey
// ("Algorithms for quad-double precision floating point arithmetic", 2001).  About 106 bits;
// used for the residual of the final root refinement and for pow_sixth.  Every function carries
// `#pragma clang fp contract(off)`: the transformations are exact only if each product and sum
// is rounded by itself, and hipcc fuses a product with a later sum into an FMA by default (after
// inlining, across these functions), which turns TwoSum and QuickTwoSum into plain sums.
// ------------------------------------------------------------------------------------------
struct dd {
    double hi, lo;
    __host__ __device__ dd() : hi(0.0), lo(0.0) {}
    __host__ __device__ dd(double a) : hi(a), lo(0.0) {}
    __host__ __device__ dd(double a, double b) : hi(a), lo(b) {}
};
__host__ __device__ __forceinline__ dd two_sum(double a, double b) {
#pragma clang fp contract(off)
    const double s = a + b, bb = s - a;
    return dd(s, (a - (s - bb)) + (b - bb));
}
__host__ __device__ __forceinline__ dd quick_two_sum(double a, double b) {  // |a| >= |b|
#pragma clang fp contract(off)
    const double s = a + b;
    return dd(s, b - (s - a));
}
__host__ __device__ __forceinline__ dd two_prod(double a, double b) {
#pragma clang fp contract(off)
    const double p = a * b;
    return dd(p, __builtin_fma(a, b, -p));
}
__host__ __device__ __forceinline__ dd operator+(dd a, dd b) {
#pragma clang fp contract(off)
    dd s = two_sum(a.hi, b.hi);
    const dd t = two_sum(a.lo, b.lo);
    s = quick_two_sum(s.hi, s.lo + t.hi);
    return quick_two_sum(s.hi, s.lo + t.lo);
}
__host__ __device__ __forceinline__ dd operator-(dd a) { return dd(-a.hi, -a.lo); }
__host__ __device__ __forceinline__ dd operator-(dd a, dd b) {
#pragma clang fp contract(off)
    return a + (-b);
}
__host__ __device__ __forceinline__ dd operator*(dd a, dd b) {
#pragma clang fp contract(off)
    const dd p = two_prod(a.hi, b.hi);
    return quick_two_sum(p.hi, p.lo + (a.hi * b.lo + a.lo * b.hi));
}
// 1/a: a double estimate x, then one Newton step x + x (1 - a x) in double-double
__host__ __device__ __forceinline__ dd recip(dd a) {
#pragma clang fp contract(off)
    const double x = 1.0 / a.hi;
    const dd e = dd(1.0) - a * dd(x);
    return dd(x) + dd(x) * e;
}

// pow(x, 1./6.), correctly rounded, for the RK step-size factor (rungekutta.cu:107-112 compute
// pow((tol/err), 1./6.); the exponent is the double nearest 1/6, 1/6 - kSixthDelta).
//   x = y 2^(6k) with y in [0.5, 32), so x^(1/6) = 2^k y^(1/6);
//   r: y^(1/6) from the hardware log2 and exp2 in single precision (relative error below 1e-6),
//      then one Newton step on r^6 = y in FP64 (relative error below 1e-11);
//   c = y / r^6 - 1 with r^6 in double-double (error-free products), so that
//      y^(1/6) = r (1 + c)^(1/6) = r (1 + c/6 - 5c^2/72 + ...) to far below an ulp;
//   x^(1/6 - kSixthDelta) = x^(1/6) (1 - kSixthDelta ln x + ...), ln x from the same log2;
//   one rounding: r + r t with t = c/6 - 5c^2/72 - kSixthDelta ln x, then the exact 2^k.
// The result is the double nearest x^(1./6.) unless that value lies within about 1e-20 relative
// of a rounding boundary.  The library pow gets there through a long double-double sequence whose
// coefficients do not stay in registers inside the tracker loop.  x = 0 gives 0, +inf gives +inf,
// NaN gives NaN.
constexpr double kSixthDelta = 9.25185853854297e-18;  // 1/6 - (double)(1./6.)
__device__ __forceinline__ double pow_sixth(double x) {
    if (!(x > 0.0) || x == __builtin_inf()) return x == 0.0 ? 0.0 : x;
    const int e = __builtin_amdgcn_frexp_exp(x);
    const int k = e >= 0 ? e / 6 : -((5 - e) / 6);
    const double y = __builtin_amdgcn_ldexp(__builtin_amdgcn_frexp_mant(x), e - 6 * k);
    const float ly = __builtin_amdgcn_logf((float)y);
    double r = (double)__builtin_amdgcn_exp2f(ly * (1.0f / 6.0f));
    const double r5 = (r * r) * (r * r) * r;
    r = (5.0 * r + y * recip(r5)) * (1.0 / 6.0);
    const dd r1(r);
    const dd r3 = (r1 * r1) * r1;
    const dd r6 = r3 * r3;
    const double c = (dd(y) - r6).hi * recip(y);
    const double lnx = (double)((ly + (float)(6 * k)) * 0.6931471805599453f);
    const double t = c * (1.0 / 6.0) - (5.0 / 72.0) * (c * c) - kSixthDelta * lnx;
    return __builtin_amdgcn_ldexp(__builtin_fma(r, t, r), k);
}

// Principal square root by the half-angle formulas of classes.cuh:39-45.
__host__ __device__ __forceinline__ cx<double> csqrt_ref(cx<double> x) {
    double r = sqrt(x.re * x.re + x.im * x.im);
    double cosA = x.re / r;
    cx<double> o = mk(sqrt(r * (cosA + 1.0) / 2.0), sqrt(r * (1.0 - cosA) / 2.0));
    if (signbit(x.im)) o.im = -o.im;
    return o;
}

// ------------------------------------------------------------------------------------------
// Group primitives (DPP) for groups of W = 8 or 16 lanes inside a wave64 (one path per group).
// DPP controls (gfx9): quad_perm [1,0,3,2] = 0xB1, quad_perm [2,3,0,1] = 0x4E, row_half_mirror
// = 0x141 (lane i <-> 7-i within each 8 lanes), row_mirror = 0x140 (lane i <-> 15-i within each
// 16-lane row).  Each butterfly step combines a lane with its mirror image, so every lane of a
// group ends with bitwise identical results (x + y = y + x, max is symmetric).
// ------------------------------------------------------------------------------------------
// Every control used here reads a lane of the same group, so no lane falls outside the row and
// the old value and bound_ctrl never matter; old = 0 with bound_ctrl set lets the compiler fold
// the move into the instruction that consumes it (v_max_u32_dpp and the like).
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
    int lo = __double2loint(x), hi = __double2hiint(x);
    lo = dpp_i<CTRL>(lo);
    hi = dpp_i<CTRL>(hi);
    return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float x) {
    return __int_as_float(dpp_i<CTRL>(__float_as_int(x)));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double x) { return dpp_d<CTRL>(x); }
template <int CTRL>
__device__ __forceinline__ float dpp(float x) { return dpp_f<CTRL>(x); }
template <int CTRL, typename R>
__device__ __forceinline__ cx<R> dpp(cx<R> z) { return mk(dpp<CTRL>(z.re), dpp<CTRL>(z.im)); }

template <int W = 8, typename R>
__device__ __forceinline__ R grp_sum(R x) {
    x = x + dpp<0xB1>(x);
    x = x + dpp<0x4E>(x);
    x = x + dpp<0x141>(x);
    if (W == 16) x = x + dpp<0x140>(x);
    return x;
}
template <int W = 8, typename R>
__device__ __forceinline__ R grp_max(R x) {
    x = fmax(x, dpp<0xB1>(x));
    x = fmax(x, dpp<0x4E>(x));
    x = fmax(x, dpp<0x141>(x));
    if (W == 16) x = fmax(x, dpp<0x140>(x));
    return x;
}
template <int W = 8, typename R>
__device__ __forceinline__ cx<R> grp_sum(cx<R> z) { return mk(grp_sum<W>(z.re), grp_sum<W>(z.im)); }

// Value of lane `src` of the group, delivered to every lane (sum with zeros elsewhere).
template <int W = 8, typename R>
__device__ __forceinline__ R grp_bcast(R x, int role, int src) { return grp_sum<W>(role == src ? x : R(0)); }
template <int W = 8, typename R>
__device__ __forceinline__ cx<R> grp_bcast(cx<R> z, int role, int src) {
    return mk(grp_bcast<W>(z.re, role, src), grp_bcast<W>(z.im, role, src));
}
template <int W = 8>
__device__ __forceinline__ int grp_bcast_i(int v, int role, int src) {
    int x = (role == src) ? v : 0;
    x = x + dpp_i<0xB1>(x);
    x = x + dpp_i<0x4E>(x);
    x = x + dpp_i<0x141>(x);
    if (W == 16) x = x + dpp_i<0x140>(x);
    return x;
}

// Group maximum of an unsigned word (DPP).
template <int W = 8>
__device__ __forceinline__ uint32_t grp_max_u32(uint32_t x) {
    x = max(x, (uint32_t)dpp_i<0xB1>((int)x));
    x = max(x, (uint32_t)dpp_i<0x4E>((int)x));
    x = max(x, (uint32_t)dpp_i<0x141>((int)x));
    if (W == 16) x = max(x, (uint32_t)dpp_i<0x140>((int)x));
    return x;
}

// Sum and maximum over a lane pair (2i, 2i+1): bitwise identical in both lanes.
template <typename R>
__device__ __forceinline__ R pair_sum(R x) { return x + dpp<0xB1>(x); }
template <typename R>
__device__ __forceinline__ cx<R> pair_sum(cx<R> z) { return mk(pair_sum(z.re), pair_sum(z.im)); }
template <typename R>
__device__ __forceinline__ R pair_max(R x) { return fmax(x, dpp<0xB1>(x)); }
// The other lane of the pair's value.
template <typename R>
__device__ __forceinline__ cx<R> pair_swap(cx<R> z) { return dpp<0xB1>(z); }

// Pivot search key of a row with pivot magnitude |a|^2 = m and id (0..15) in its group: larger
// m wins, then the smaller id.  Finite m (compared in FP32, to 2^-19 relative) beats NaN, which
// beats a row already used as a pivot.
template <typename R>
__device__ __forceinline__ uint32_t pivot_key(R m, bool used, int id) {
    const float f = (float)m;
    const uint32_t mag = (f == f) ? ((uint32_t)__float_as_int(f) >> 4) + 2u : 1u;
    return ((used ? 0u : mag) << 4) | (uint32_t)(15 - id);
}

// Orders this lane's LDS accesses at compiler level only (no wave-wide code motion barrier).
__device__ __forceinline__ void lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Orders this lane's LDS accesses before and after the call at compiler level.  LDS requests of
// one wave are processed in issue order, so a write by one lane is seen by a later read of
// another lane of the same wave.
__device__ __forceinline__ void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------------------------------
// Philox4x32-10.  Draw k of a path is half (k & 1) of block k >> 1; a half is the 53 high bits
// of a 64-bit word plus one, times 2^-53, giving a double in (0, 1].
// ------------------------------------------------------------------------------------------
struct u4 {
    uint32_t x, y, z, w;
};
__device__ __forceinline__ u4 philox(uint64_t blk, uint64_t pid, uint64_t seed) {
    uint32_t c0 = (uint32_t)blk, c1 = (uint32_t)(blk >> 32), c2 = (uint32_t)pid, c3 = (uint32_t)(pid >> 32);
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    u4 o;
    o.x = c0;
    o.y = c1;
    o.z = c2;
    o.w = c3;
    return o;
}
__device__ __forceinline__ double u53(uint32_t lo, uint32_t hi) {
    uint64_t w = ((uint64_t)hi << 32) | lo;
    return (double)((w >> 11) + 1) * 0x1.0p-53;
}
// The two uniforms of block `blk`: draws 2*blk (first) and 2*blk+1 (second).
__device__ __forceinline__ void uniform_pair(uint64_t blk, uint64_t pid, uint64_t seed, double &a, double &b) {
    u4 o = philox(blk, pid, seed);
    a = u53(o.x, o.y);
    b = u53(o.z, o.w);
}

}  // namespace frg
