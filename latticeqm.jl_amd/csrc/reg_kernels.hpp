// Register-tile Householder tridiagonalization for N <= 64 (one 256-thread block per
// k-point).
//
// The N x N Hermitian matrix is held entirely in registers: thread (ti, tj) (ti =
// tid & 15, tj = tid >> 4) owns the 4 x 4 tile A[4ti + u][4tj + v], both triangles.
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
template <int CTRL>
__device__ __forceinline__ float dpp_r(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_r(double v) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), CTRL, 0xf, 0xf, false);
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
// four entries per lane (rows 4 ti .. 4 ti + 3), so entry r sits at rpad(r) = r + r/4: one
// spare slot per four entries spreads the 16 lanes of a row over distinct banks.
__device__ __forceinline__ int rpad(int r) { return r + (r >> 2); }
template <typename R>
struct RegTriSmem {
    cx<R> sv[2][80];  // Householder vector v (zero above row j+1), entry r at rpad(r)
    cx<R> sp[2][80];  // p = tau A22 v, entry r at rpad(r)
    R ss[2][16];      // per tile column partial of v^H A22 v
    cx<R> stau[2];
};

// One Householder step j = 4 jt + JV.  JV is a compile-time constant so that every
// register of the tile is addressed statically (no private-memory array).
template <typename R, int JV>
__device__ __forceinline__ void regtri_step(cx<R> (&a)[4][4], RegTriSmem<R>& S, int N, int Nkc, int kk,
                                            int jt, int ti, int tj, cx<R>* __restrict__ rf,
                                            R* __restrict__ dd, R* __restrict__ ee,
                                            cx<R>* __restrict__ tau) {
    const int j = 4 * jt + JV;
    if (j + 1 >= N) return;
    const int buf = j & 1;
    // (a) Householder vector of column j, by the 16 lanes of tile column jt.  Every other
    //     wave waits for it at the barrier, so its wave is issued ahead of the waves of other
    //     blocks on the same SIMD while it runs (s_setprio).
    if (tj == jt) {
        __builtin_amdgcn_s_setprio(3);
        R xn2 = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = 4 * ti + u;
            if (r >= j + 2 && r < N) xn2 += abs2(a[u][JV]);
        }
        // the row of lanes is wave-uniform here (tile column jt), so one lane's sum and
        // the owner of A[j+1][j] are read back with readlane: bitwise the same everywhere
        const int base = 16 * (jt & 3);
        xn2 = rl(row16_sum(xn2), base);
        constexpr int UA = (JV + 1) & 3;  // row j+1 is row UA of tile row (j+1)>>2
        const int src = base + ((j + 1) >> 2);
        const cx<R> alpha = mkc<R>(rl(a[UA][JV].re, src), rl(a[UA][JV].im, src));
        cx<R> tj_ = mkc<R>(0, 0), scale = mkc<R>(0, 0);
        R beta = alpha.re;
        if (!(xn2 == (R)0 && alpha.im == (R)0)) {
            const R nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
            beta = alpha.re >= (R)0 ? -nrm : nrm;
            // reciprocals from the hardware approximation while |alpha - beta|^2 (<= 4 beta^2,
            // >= beta^2) sits in the normal range of R, IEEE division below it (wave-uniform)
            const R zr = alpha.re - beta, den = zr * zr + alpha.im * alpha.im;
            R rb, rz;
            if (den >= num<R>::rot_lo) {
                rb = rcp_nr(beta);
                rz = rcp_nr(den);
            } else {
                rb = (R)1 / beta;
                rz = (R)1 / den;
            }
            tj_ = mkc<R>((beta - alpha.re) * rb, -alpha.im * rb);
            scale = mkc<R>(zr * rz, -alpha.im * rz);
        }
        const bool triv = (tj_.re == (R)0 && tj_.im == (R)0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = 4 * ti + u;
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
    const bool wact = 16 * (tj >> 2) + 15 > j;
    cx<R> vr[4], vc[4], p[4];
    R spart = 0;
    if (wact) {
#pragma unroll
        for (int u = 0; u < 4; ++u) vr[u] = S.sv[buf][5 * ti + u];
#pragma unroll
        for (int v = 0; v < 4; ++v) vc[v] = S.sv[buf][5 * tj + v];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            cx<R> y = mkc<R>(0, 0);
#pragma unroll
            for (int u = 0; u < 4; ++u) y = cmadc(y, a[u][v], vr[u]);
            y.re = row16_sum(y.re);
            y.im = row16_sum(y.im);
            p[v] = y;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = 4 * tj + v;
            if (c > j && c < N) spart += p[v].re * vc[v].re + p[v].im * vc[v].im;  // Re(conj(y_c) v_c)
            p[v] = t * p[v];
        }
        if (ti == 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) S.sp[buf][5 * tj + v] = p[v];
        }
    }
    if (ti == 0) S.ss[buf][tj] = spart;
    __syncthreads();

    // (c) A22 -= v w^H + w v^H, w = p + alpha2 v, alpha2 = -1/2 |tau|^2 v^H A22 v
    R sum = 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) sum += S.ss[buf][q];
    const R alpha2 = (R)-0.5 * (t.re * t.re + t.im * t.im) * sum;
    if (wact && 4 * ti + 3 > j && 4 * tj + 3 > j) {
        cx<R> wr[4], wc[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = 4 * ti + u;
            const cx<R> pr = S.sp[buf][5 * ti + u];
            wr[u] = (r > j && r < N) ? cmad(pr, mkc<R>(alpha2, 0), vr[u]) : mkc<R>(0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = 4 * tj + v;
            wc[v] = (c > j && c < N) ? cmad(p[v], mkc<R>(alpha2, 0), vc[v]) : mkc<R>(0, 0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                cx<R> d = mkc<R>(0, 0);
                d = cmadcb(d, vr[u], wc[v]);
                d = cmadcb(d, wr[u], vc[v]);
                a[u][v] = a[u][v] - d;
            }
    }
}

template <typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 3))) void k_tridiag_r(
    int N, int Nkc, const cx<double>* __restrict__ Hk, R* __restrict__ dd, R* __restrict__ ee,
    cx<R>* __restrict__ refl, cx<R>* __restrict__ tau) {
    __shared__ RegTriSmem<R> S;
    const int tid = threadIdx.x;
    const int ti = tid & 15, tj = tid >> 4;
    const int kk = blockIdx.x;
    const size_t NN = (size_t)N * N;
    const cx<double>* H = Hk + (size_t)kk * NN;

    cx<R> a[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = 4 * ti + u, c = 4 * tj + v;
            cx<R> x = mkc<R>(0, 0);
            if (r < N && c < N) {
                const cx<double> h = H[(size_t)c * N + r];
                x = mkc<R>((R)h.re, (R)h.im);
            }
            a[u][v] = x;
        }

    cx<R>* rf = refl + (size_t)kk * NN;
    const int ntiles = (N + 2) / 4;  // steps j = 0 .. N-2 in groups of four
    for (int jt = 0; jt < ntiles; ++jt) {
        regtri_step<R, 0>(a, S, N, Nkc, kk, jt, ti, tj, rf, dd, ee, tau);
        regtri_step<R, 1>(a, S, N, Nkc, kk, jt, ti, tj, rf, dd, ee, tau);
        regtri_step<R, 2>(a, S, N, Nkc, kk, jt, ti, tj, rf, dd, ee, tau);
        regtri_step<R, 3>(a, S, N, Nkc, kk, jt, ti, tj, rf, dd, ee, tau);
    }
    // last diagonal element
    const int n1 = N - 1;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (4 * ti + u == n1 && 4 * tj + v == n1) {
                dd[(size_t)n1 * Nkc + kk] = a[u][v].re;
                ee[(size_t)n1 * Nkc + kk] = (R)0;
            }
}

}  // namespace lqm
