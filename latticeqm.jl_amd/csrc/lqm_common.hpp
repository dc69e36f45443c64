// Shared device helpers for the gfx950 k-space band path: a minimal complex type,
// wave64 reductions and the phase factor e^{2 pi i x}.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lqm {

constexpr int kWave = 64;   // CDNA wavefront width
constexpr int kTPB = 256;   // threads per block of the per-k kernels (4 waves)

template <typename R>
struct cx {
    R re, im;
};

template <typename R>
__host__ __device__ __forceinline__ cx<R> mkc(R a, R b) { return cx<R>{a, b}; }
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator*(R s, cx<R> a) { return {s * a.re, s * a.im}; }
template <typename R>
__host__ __device__ __forceinline__ cx<R> conj(cx<R> a) { return {a.re, -a.im}; }
template <typename R>
__host__ __device__ __forceinline__ R abs2(cx<R> a) { return a.re * a.re + a.im * a.im; }
// acc + a*b with fused multiply-adds
template <typename R>
__device__ __forceinline__ cx<R> cmad(cx<R> acc, cx<R> a, cx<R> b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
    return acc;
}
// acc + conj(a)*b
template <typename R>
__device__ __forceinline__ cx<R> cmadc(cx<R> acc, cx<R> a, cx<R> b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(-a.im, b.re, acc.im);
    return acc;
}
// acc + a*conj(b)
template <typename R>
__device__ __forceinline__ cx<R> cmadcb(cx<R> acc, cx<R> a, cx<R> b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(a.im, b.im, acc.re);
    acc.im = fma(a.im, b.re, acc.im);
    acc.im = fma(-a.re, b.im, acc.im);
    return acc;
}
// Reciprocal from the hardware approximation: FP32 as is (1 ulp), FP64 with one Newton step
// (for normal, finite x).
template <typename R>
__device__ __forceinline__ R rcp_nr(R x);
template <>
__device__ __forceinline__ float rcp_nr(float x) {
    return __builtin_amdgcn_rcpf(x);
}
template <>
__device__ __forceinline__ double rcp_nr(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// 1/z
template <typename R>
__device__ __forceinline__ cx<R> crecip(cx<R> z) {
    R den = z.re * z.re + z.im * z.im;
    return {z.re / den, -z.im / den};
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
template <typename R>
__device__ __forceinline__ cx<R> wave_sum(cx<R> v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        v.re += __shfl_xor(v.re, o);
        v.im += __shfl_xor(v.im, o);
    }
    return v;
}

// Sixteen wave sums at once by a halving exchange: each xor step trades half of the values
// with the partner lane, so 17 exchanges replace 16 x 6.  Lane l ends with the sum over the
// wave of x[u], u = 8 b5 + 4 b4 + 2 b3 + b2 (b = bits of l; lanes differing only in bits
// 0 and 1 hold the same sum).
template <typename R>
__device__ __forceinline__ cx<R> xor_cx(cx<R> v, int m) {
    return {__shfl_xor(v.re, m), __shfl_xor(v.im, m)};
}
__device__ __forceinline__ int wave_sum16_index(int lane) {
    return ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 + ((lane >> 2) & 1);
}
template <typename R>
__device__ __forceinline__ cx<R> wave_sum16(const cx<R> (&x)[16], int lane) {
    cx<R> a[8], b[4], c[2];
    const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8, h2 = lane & 4;
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = (h5 ? x[8 + u] : x[u]) + xor_cx(h5 ? x[u] : x[8 + u], 32);
#pragma unroll
    for (int u = 0; u < 4; ++u) b[u] = (h4 ? a[4 + u] : a[u]) + xor_cx(h4 ? a[u] : a[4 + u], 16);
#pragma unroll
    for (int u = 0; u < 2; ++u) c[u] = (h3 ? b[2 + u] : b[u]) + xor_cx(h3 ? b[u] : b[2 + u], 8);
    cx<R> d = (h2 ? c[1] : c[0]) + xor_cx(h2 ? c[0] : c[1], 4);
    d = d + xor_cx(d, 2);
    return d + xor_cx(d, 1);
}

// e^{2 pi i x} in FP64 (sincospi keeps the argument reduction exact).
__device__ __forceinline__ cx<double> cis2pi(double x) {
    double s, c;
    sincospi(2.0 * x, &s, &c);
    return {c, s};
}

// Numerical-type constants used by the eigensolver.
template <typename R> struct num;
// rot_lo, rot_hi: the range of f^2 + g^2 in which a Givens rotation is formed from the
// sum of squares directly (no overflow, no loss to subnormals).
template <> struct num<double> {
    static constexpr double half_eps = 1.1102230246251565e-16;  // 2^-53
    static constexpr double tiny = 2.2250738585072014e-308;     // smallest normal
    static constexpr double huge = 1.7976931348623157e308;      // largest finite
    static constexpr double rot_lo = 1e-290, rot_hi = 1e290;
};
template <> struct num<float> {
    static constexpr float half_eps = 5.9604645e-08f;  // 2^-24
    static constexpr float tiny = 1.17549435e-38f;     // smallest normal
    static constexpr float huge = 3.40282347e+38f;     // largest finite
    static constexpr float rot_lo = 1e-30f, rot_hi = 1e30f;
};

// Householder scalars (zlarfg: H^H x = beta e_1, H = I - tau v v^H, v = [1, x(1:) scale])
// from alpha = x(0) and xn2 = ||x(1:)||^2: beta = -sign(Re alpha) ||x||, tau = (beta -
// alpha) / beta, scale = 1 / (alpha - beta); tau = 0 when x(1:) = 0 and alpha is real.
// Returns false, leaving the outputs undefined, when |alpha - beta|^2 falls below the normal
// range of R (a column of tiny entries): the caller then rescales x by its largest entry,
// calls hh_scalars_scaled on the rescaled column and scales beta and scale back.
template <typename R>
__device__ __forceinline__ bool hh_scalars(cx<R> alpha, R xn2, cx<R>& tau, cx<R>& scale, R& beta) {
    tau = mkc<R>(0, 0);
    scale = mkc<R>(0, 0);
    beta = alpha.re;
    if (xn2 == (R)0 && alpha.im == (R)0) return true;
    const R nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
    beta = alpha.re >= (R)0 ? -nrm : nrm;
    const R zr = alpha.re - beta, den = zr * zr + alpha.im * alpha.im;
    if (!(den >= num<R>::rot_lo)) return false;
    // |alpha - beta|^2 lies in [beta^2, 4 beta^2]: reciprocals from the hardware approximation
    const R rb = rcp_nr(beta), rz = rcp_nr(den);
    tau = mkc<R>((beta - alpha.re) * rb, -alpha.im * rb);
    scale = mkc<R>(zr * rz, -alpha.im * rz);
    return true;
}
// The same for a column rescaled so that its largest component has magnitude 1 (IEEE
// divisions; |beta| >= 1 there).
template <typename R>
__device__ __forceinline__ void hh_scalars_scaled(cx<R> alpha, R xn2, cx<R>& tau, cx<R>& scale, R& beta) {
    tau = mkc<R>(0, 0);
    scale = mkc<R>(0, 0);
    beta = alpha.re;
    if (xn2 == (R)0 && alpha.im == (R)0) return;
    const R nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
    beta = alpha.re >= (R)0 ? -nrm : nrm;
    const R zr = alpha.re - beta, den = zr * zr + alpha.im * alpha.im;
    tau = mkc<R>((beta - alpha.re) / beta, -alpha.im / beta);
    scale = mkc<R>(zr / den, -alpha.im / den);
}

// Device-side status bits, OR-ed into a per-call flag word.
enum : int {
    kStatNoConv = 1,       // QL exceeded its iteration budget
    kStatLogFull = 2,      // rotation log capacity exceeded
    kStatNonFinite = 4,    // NaN/Inf met
};

}  // namespace lqm
