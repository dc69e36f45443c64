// Register-tile Householder tridiagonalization for N <= 64 (one 256-thread block per
// k-point).
//
// The N x N Hermitian matrix is held entirely in registers: thread (ti, tj) (ti =
// tid & 15, tj = tid >> 4) owns the T x T tile A[off + T ti + u][off + T tj + v] of the
// trailing matrix from row/column off, both triangles.  T starts at the smallest size with
// 16 T >= N and drops by one whenever the trailing matrix fits the next smaller size
// (N - j <= 48, 32, 16), so the lanes stay busy with live entries as the matrix shrinks.
// The 16 threads of one tile column tj are one 16-lane row of a wave, so every
// reduction over rows is a lane butterfly inside that row; LDS only carries the
// Householder vector, p = tau A v and 16 partial dot products per step (double
// buffered, two barriers per step).
//
// Step j (zhetd2, lower; the same outputs as k_hk_tridiag):
//   x = A[j+1:N, j]; beta = -sign(Re x_0) ||x||; tau = (beta - x_0)/beta;
//   v = [1, x[1:]/(x_0 - beta)]                       (zlarfg: H^H x = beta e_1)
//   y = A22 v, computed column-wise as y_c = sum_r conj(A_rc) v_r (A Hermitian)
//   p = tau y; s = v^H y (real); alpha2 = -1/2 |tau|^2 s; w = p + alpha2 v
//   A22 <- A22 - v w^H - w v^H
// H(k) comes in materialised (FP64, full Hermitian) and is rounded to the working
// precision R on load.
#pragma once
#include "lqm_common.hpp"

namespace lqm {

// Lane-row (16 lanes) reductions by DPP row rotations: plain VALU moves, no LDS
// traffic.  After rotations by 1, 2, 4 and 8 every lane of the row holds the row sum
// (each lane adds in its own rotation order, so lanes may differ in the last bit;
// row_bcast_sum returns one lane's value to every lane of a wave-uniform row).
// (Row rotations read a valid lane everywhere, so the DPP move needs no old value: mov_dpp
// leaves the destination undefined where no lane is read, and no zero-fill move is issued.)
template <int CTRL>
__device__ __forceinline__ float dpp_r(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_r(double v) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)x, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(x >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float rl(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double rl(double v, int lane) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)x, lane);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <typename R>
__device__ __forceinline__ R row16_sum(R v) {
    v += dpp_r<0x121>(v);  // row_ror:1
    v += dpp_r<0x122>(v);  // row_ror:2
    v += dpp_r<0x124>(v);  // row_ror:4
    v += dpp_r<0x128>(v);  // row_ror:8
    return v;
}

// Shared memory of k_tridiag_r (double-buffered per step parity).  The vectors are read
// T entries per lane (rows off + T ti .. off + T ti + T - 1), so entry r sits at
// rpad(r) = r + r/4: one spare slot per four entries spreads the 16 lanes of a row over
// distinct banks.  L carries the trailing matrix from one tile size to the next.
__device__ __forceinline__ int rpad(int r) { return r + (r >> 2); }
constexpr int kRegLd = 49;
template <typename R>
struct RegTriSmem {
    cx<R> sv[2][80];  // Householder vector v (zero above row j+1), entry r at rpad(r)
    cx<R> sp[2][80];  // p = tau A22 v, entry r at rpad(r)
    R ss[2][4];       // per wave partial of v^H A22 v
    cx<R> stau[2];
    cx<R> L[48 * kRegLd];  // trailing matrix (<= 48 x 48) between tile sizes, column-major
};

// One Householder step j = off + T jt + JV on T x T tiles (thread (ti, tj) owns rows
// off + T ti + u and columns off + T tj + v of the trailing matrix).  T and JV are
// compile-time constants so that every register of the tile is addressed statically.
template <typename R, int T, int JV>
__device__ __forceinline__ void regtri_step(cx<R> (&a)[T][T], RegTriSmem<R>& S, int N, int Nkc, int kk,
                                            int off, int jend, int jt, int ti, int tj,
                                            cx<R>* __restrict__ rf, R* __restrict__ dd, R* __restrict__ ee,
                                            cx<R>* __restrict__ tau) {
    const int j = off + T * jt + JV;
    if (j + 1 >= N || j >= jend) return;
    const int buf = j & 1;
    // (a) Householder vector of column j, by the 16 lanes of tile column jt.  Every other
    //     wave waits for it at the barrier, so its wave is issued ahead of the waves of other
    //     blocks on the same SIMD while it runs (s_setprio).
    if (tj == jt) {
        __builtin_amdgcn_s_setprio(3);
        R xn2 = 0;
#pragma unroll
        for (int u = 0; u < T; ++u) {
            const int r = off + T * ti + u;
            if (r >= j + 2 && r < N) xn2 += abs2(a[u][JV]);
        }
        // the row of lanes is wave-uniform here (tile column jt), so one lane's sum and
        // the owner of A[j+1][j] are read back with readlane: bitwise the same everywhere
        const int base = 16 * (jt & 3);
        xn2 = rl(row16_sum(xn2), base);
        constexpr int UA = (JV + 1) % T;  // row j+1 is row UA of tile row (j+1-off)/T
        const int src = base + (j + 1 - off) / T;
        const cx<R> alpha = mkc<R>(rl(a[UA][JV].re, src), rl(a[UA][JV].im, src));
        cx<R> tj_, scale;
        R beta;
        if (!hh_scalars(alpha, xn2, tj_, scale, beta)) {
            // a column of tiny entries: the scalars of x / max|x|, then beta and scale back
            R mx = 0;
#pragma unroll
            for (int u = 0; u < T; ++u) {
                const int r = off + T * ti + u;
                if (r > j && r < N) mx = fmax(mx, fmax(fabs(a[u][JV].re), fabs(a[u][JV].im)));
            }
            mx = fmax(mx, dpp_r<0x121>(mx));
            mx = fmax(mx, dpp_r<0x122>(mx));
            mx = fmax(mx, dpp_r<0x124>(mx));
            mx = fmax(mx, dpp_r<0x128>(mx));
            mx = rl(mx, base);
            const R inv = (R)1 / mx;
            R xs2 = 0;
#pragma unroll
            for (int u = 0; u < T; ++u) {
                const int r = off + T * ti + u;
                if (r >= j + 2 && r < N) xs2 += abs2(inv * a[u][JV]);
            }
            xs2 = rl(row16_sum(xs2), base);
            hh_scalars_scaled(inv * alpha, xs2, tj_, scale, beta);
            beta *= mx;
            scale = inv * scale;
            if (!(inv <= num<R>::huge)) {
                // entries below 1 / huge (subnormal): left in place, H_j = I
                tj_ = mkc<R>(0, 0);
                scale = mkc<R>(0, 0);
                beta = alpha.re;
            }
        }
        const bool triv = (tj_.re == (R)0 && tj_.im == (R)0);
#pragma unroll
        for (int u = 0; u < T; ++u) {
            const int r = off + T * ti + u;
            cx<R> vr = mkc<R>(0, 0);
            if (r == j + 1) vr = mkc<R>(1, 0);
            else if (r > j + 1 && r < N && !triv) vr = a[u][JV] * scale;
            S.sv[buf][rpad(r)] = vr;
            if (r > j && r < N) rf[(size_t)j * N + r] = vr;
        }
        if (ti == jt) dd[(size_t)j * Nkc + kk] = a[JV][JV].re;  // A[j][j]
        if (ti == 0) {
            S.stau[buf] = tj_;
            tau[(size_t)kk * N + j] = tj_;
            ee[(size_t)j * Nkc + kk] = beta;
        }
        __builtin_amdgcn_s_setprio(0);
    }
    __syncthreads();
    const cx<R> t = S.stau[buf];
    if (t.re == (R)0 && t.im == (R)0) return;  // H_j = I (uniform branch)

    // (b) y_c = sum_r conj(A_rc) v_r for this thread's columns, summed over the lane row.
    //     A wave whose four tile columns all lie at or left of column j has no part in the
    //     rest of the step (wave-uniform test): it only posts a zero partial of v^H A22 v.
    const bool wact = off + T * (4 * (tj >> 2) + 4) - 1 > j;
    cx<R> vr[T], vc[T], p[T];
    R spart = 0;
    if (wact) {
#pragma unroll
        for (int u = 0; u < T; ++u) vr[u] = S.sv[buf][rpad(off + T * ti + u)];
#pragma unroll
        for (int v = 0; v < T; ++v) vc[v] = S.sv[buf][rpad(off + T * tj + v)];
#pragma unroll
        for (int v = 0; v < T; ++v) {
            cx<R> y = mkc<R>(0, 0);
#pragma unroll
            for (int u = 0; u < T; ++u) y = cmadc(y, a[u][v], vr[u]);
            y.re = row16_sum(y.re);
            y.im = row16_sum(y.im);
            p[v] = y;
        }
#pragma unroll
        for (int v = 0; v < T; ++v) {
            const int c = off + T * tj + v;
            if (c > j && c < N) spart += p[v].re * vc[v].re + p[v].im * vc[v].im;  // Re(conj(y_c) v_c)
            p[v] = t * p[v];
        }
        if (ti == 0) {
#pragma unroll
            for (int v = 0; v < T; ++v) S.sp[buf][rpad(off + T * tj + v)] = p[v];
        }
    }
    {
        // the wave's four tile columns (its rows of 16 lanes, lane ti = 0 of each) summed with
        // readlane: one partial of v^H A22 v per wave
        const R s4 = ((rl(spart, 0) + rl(spart, 16)) + rl(spart, 32)) + rl(spart, 48);
        if (ti == 0 && (tj & 3) == 0) S.ss[buf][tj >> 2] = s4;
    }
    __syncthreads();

    // (c) A22 -= v w^H + w v^H, w = p + alpha2 v, alpha2 = -1/2 |tau|^2 v^H A22 v
    const R sum = (S.ss[buf][0] + S.ss[buf][1]) + (S.ss[buf][2] + S.ss[buf][3]);
    const R alpha2 = (R)-0.5 * (t.re * t.re + t.im * t.im) * sum;
    if (wact && off + T * ti + T - 1 > j && off + T * tj + T - 1 > j) {
        cx<R> wr[T], wc[T];
#pragma unroll
        for (int u = 0; u < T; ++u) {
            const int r = off + T * ti + u;
            const cx<R> pr = S.sp[buf][rpad(r)];
            wr[u] = (r > j && r < N) ? cmad(pr, mkc<R>(alpha2, 0), vr[u]) : mkc<R>(0, 0);
        }
#pragma unroll
        for (int v = 0; v < T; ++v) {
            const int c = off + T * tj + v;
            wc[v] = (c > j && c < N) ? cmad(p[v], mkc<R>(alpha2, 0), vc[v]) : mkc<R>(0, 0);
        }
        // straight into the tile: a - v conj(w) - w conj(v), the negations folded into the
        // multiply-adds' source modifiers
#pragma unroll
        for (int u = 0; u < T; ++u)
#pragma unroll
            for (int v = 0; v < T; ++v) {
                a[u][v] = cmadcb(a[u][v], mkc<R>(-vr[u].re, -vr[u].im), wc[v]);
                a[u][v] = cmadcb(a[u][v], mkc<R>(-wr[u].re, -wr[u].im), vc[v]);
            }
    }
}

// The trailing matrix from row/column off on T x T tiles: steps j in [off, jend), then the
// part from row/column off' = max(0, N - 16 (T - 1)) moves through LDS onto (T-1) x (T-1)
// tiles for the steps that follow (the active part never needs more tiles than it has rows).
template <typename R, int T>
__device__ __forceinline__ void regtri_run(cx<R> (&a)[T][T], RegTriSmem<R>& S, int N, int Nkc, int kk,
                                           int off, int ti, int tj, cx<R>* __restrict__ rf,
                                           R* __restrict__ dd, R* __restrict__ ee, cx<R>* __restrict__ tau) {
    const int offn = T > 1 ? max(0, N - 16 * (T - 1)) : N;
    const int jend = T > 1 ? offn : N - 1;
    for (int jt = 0; off + T * jt < jend; ++jt) {
        regtri_step<R, T, 0>(a, S, N, Nkc, kk, off, jend, jt, ti, tj, rf, dd, ee, tau);
        if constexpr (T > 1) regtri_step<R, T, 1>(a, S, N, Nkc, kk, off, jend, jt, ti, tj, rf, dd, ee, tau);
        if constexpr (T > 2) regtri_step<R, T, 2>(a, S, N, Nkc, kk, off, jend, jt, ti, tj, rf, dd, ee, tau);
        if constexpr (T > 3) regtri_step<R, T, 3>(a, S, N, Nkc, kk, off, jend, jt, ti, tj, rf, dd, ee, tau);
    }
    if constexpr (T > 1) {
        // rows/columns [offn, N) of the trailing matrix into L (column-major, offset offn)
#pragma unroll
        for (int u = 0; u < T; ++u)
#pragma unroll
            for (int v = 0; v < T; ++v) {
                const int r = off + T * ti + u, c = off + T * tj + v;
                if (r >= offn && r < N && c >= offn && c < N) S.L[(c - offn) * kRegLd + (r - offn)] = a[u][v];
            }
        __syncthreads();
        constexpr int T1 = T - 1;
        cx<R> b[T1][T1];
#pragma unroll
        for (int u = 0; u < T1; ++u)
#pragma unroll
            for (int v = 0; v < T1; ++v) {
                const int r = offn + T1 * ti + u, c = offn + T1 * tj + v;
                b[u][v] = (r < N && c < N) ? S.L[(c - offn) * kRegLd + (r - offn)] : mkc<R>(0, 0);
            }
        regtri_run<R, T1>(b, S, N, Nkc, kk, offn, ti, tj, rf, dd, ee, tau);
    } else {
        // last diagonal element
        const int n1 = N - 1;
        if (off + ti == n1 && off + tj == n1) {
            dd[(size_t)n1 * Nkc + kk] = a[0][0].re;
            ee[(size_t)n1 * Nkc + kk] = (R)0;
        }
    }
}

template <typename R, int T>
__device__ __forceinline__ void regtri_start(const cx<double>* __restrict__ H, RegTriSmem<R>& S, int N, int Nkc,
                                             int kk, int ti, int tj, cx<R>* __restrict__ rf,
                                             R* __restrict__ dd, R* __restrict__ ee, cx<R>* __restrict__ tau) {
    // all T x T loads in flight together (clamped addresses: every load is in bounds; the
    // padding rows and columns are zeroed after)
    cx<double> h[T][T];
#pragma unroll
    for (int u = 0; u < T; ++u)
#pragma unroll
        for (int v = 0; v < T; ++v) h[u][v] = H[(size_t)min(T * tj + v, N - 1) * N + min(T * ti + u, N - 1)];
    cx<R> a[T][T];
#pragma unroll
    for (int u = 0; u < T; ++u)
#pragma unroll
        for (int v = 0; v < T; ++v) {
            const int r = T * ti + u, c = T * tj + v;
            a[u][v] = (r < N && c < N) ? mkc<R>((R)h[u][v].re, (R)h[u][v].im) : mkc<R>(0, 0);
        }
    regtri_run<R, T>(a, S, N, Nkc, kk, 0, ti, tj, rf, dd, ee, tau);
}

template <typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, sizeof(R) == 4 ? 4 : 3))) void k_tridiag_r(
    int N, int Nkc, const cx<double>* __restrict__ Hk, R* __restrict__ dd, R* __restrict__ ee,
    cx<R>* __restrict__ refl, cx<R>* __restrict__ tau) {
    __shared__ RegTriSmem<R> S;
    const int tid = threadIdx.x;
    const int ti = tid & 15, tj = tid >> 4;
    const int kk = blockIdx.x;
    const size_t NN = (size_t)N * N;
    const cx<double>* H = Hk + (size_t)kk * NN;
    cx<R>* rf = refl + (size_t)kk * NN;
    // the smallest tile holding the whole matrix: 16 T >= N (N <= 64)
    if (N > 48) regtri_start<R, 4>(H, S, N, Nkc, kk, ti, tj, rf, dd, ee, tau);
    else if (N > 32) regtri_start<R, 3>(H, S, N, Nkc, kk, ti, tj, rf, dd, ee, tau);
    else if (N > 16) regtri_start<R, 2>(H, S, N, Nkc, kk, ti, tj, rf, dd, ee, tau);
    else regtri_start<R, 1>(H, S, N, Nkc, kk, ti, tj, rf, dd, ee, tau);
}

}  // namespace lqm
