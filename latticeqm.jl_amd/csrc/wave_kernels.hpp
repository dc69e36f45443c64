// One-wave-per-k-point kernels for N <= 64*NQ (NQ rows per lane).
//
// The Hermitian matrix of a k-point lives in LDS as its packed lower triangle
// (column-major, column c holds rows c..N-1), N(N+1)/2 complex entries, i.e. half of
// the full tile, so twice as many k-points share a CU.  A single wave owns the
// whole matrix: lane l owns rows l, l+64, ...  Every cross-lane step is a DPP
// row-rotate reduction or an LDS broadcast; no workgroup barrier joins several waves.
//
//   k_hk_tridiag_w  H(k) assembly + Householder tridiagonalization (same outputs as
//                   k_hk_tridiag: d, e, reflectors, tau)
//   k_zfz           M = Z diag(f) Z^T, the tridiagonal-basis density matrix (real)
//   k_rho_w         rho_k = Q M Q^H by applying the reflectors from both sides:
//                   M <- H_j M H_j^H for j = N-2 .. 0, written out as a full matrix
#pragma once
#include "lqm_common.hpp"

namespace lqm {

template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const long long x = __double_as_longlong(v);
    // row rotations only: every lane reads a valid lane, so no old value (no zero-fill move)
    const int lo = __builtin_amdgcn_mov_dpp((int)x, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(x >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double lane_read(double v, int lane) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)x, lane);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over all 64 lanes, returned in every lane.  Must be called by the full wave.
// Cyclic rotations by 1, 2, 4, 8 inside each 16-lane row give every lane its row
// sum; the four row sums are combined in a fixed order.
__device__ __forceinline__ double wave_allsum(double v) {
    v += dpp_d<0x121>(v);  // row_ror:1
    v += dpp_d<0x122>(v);  // row_ror:2
    v += dpp_d<0x124>(v);  // row_ror:4
    v += dpp_d<0x128>(v);  // row_ror:8
    return (lane_read(v, 0) + lane_read(v, 16)) + (lane_read(v, 32) + lane_read(v, 48));
}

// Start of packed lower column c of an N x N matrix.
__device__ __forceinline__ int pcol(int N, int c) { return c * N - ((c * (c - 1)) >> 1); }

// Hermitian element (r, c) from the packed lower triangle.
template <typename R>
__device__ __forceinline__ cx<R> herm_at(const cx<R>* A, int N, int r, int c) {
    const bool lo = r >= c;
    const int idx = lo ? pcol(N, c) + r - c : pcol(N, r) + c - r;
    cx<R> x = A[idx];
    if (!lo) x.im = -x.im;
    return x;
}

// sum_{c = c0}^{N-1} Herm(r, c) x_c, four independent accumulators so the LDS
// loads of one group are in flight together.
template <typename R>
__device__ __forceinline__ cx<R> herm_dot(const cx<R>* A, const cx<R>* x, int N, int r, int c0) {
    cx<R> a0 = {0, 0}, a1 = {0, 0}, a2 = {0, 0}, a3 = {0, 0};
    int c = c0;
    for (; c + 4 <= N; c += 4) {
        const cx<R> h0 = herm_at(A, N, r, c), h1 = herm_at(A, N, r, c + 1);
        const cx<R> h2 = herm_at(A, N, r, c + 2), h3 = herm_at(A, N, r, c + 3);
        const cx<R> x0 = x[c], x1 = x[c + 1], x2 = x[c + 2], x3 = x[c + 3];
        a0 = cmad(a0, h0, x0);
        a1 = cmad(a1, h1, x1);
        a2 = cmad(a2, h2, x2);
        a3 = cmad(a3, h3, x3);
    }
    for (; c < N; ++c) a0 = cmad(a0, herm_at(A, N, r, c), x[c]);
    return (a0 + a1) + (a2 + a3);
}

template <typename R>
__device__ __forceinline__ cx<R> wave_allsum(cx<R> v) {
    return {(R)wave_allsum((double)v.re), (R)wave_allsum((double)v.im)};
}

// ---------------------------------------------------------------------------------
// H(k) + Householder tridiagonalization, one wave per k-point.
// LDS: packed A [N(N+1)/2] | sv [N] | sw [N] | phases [nR]
// ---------------------------------------------------------------------------------
template <typename R, int NQ>
__global__ __launch_bounds__(kWave) void k_hk_tridiag_w(
    int N, int d, int nR, const int* __restrict__ Rk, const cx<double>* __restrict__ T,
    const double* __restrict__ ks, int Nkc, R* __restrict__ dd, R* __restrict__ ee,
    cx<R>* __restrict__ refl, cx<R>* __restrict__ tau) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int kk = blockIdx.x;
    const int lane = threadIdx.x;
    const int np = (N * (N + 1)) / 2;
    cx<R>* A = reinterpret_cast<cx<R>*>(smem);
    cx<R>* sv = A + np;
    cx<R>* sw = sv + N;
    cx<double>* sph = reinterpret_cast<cx<double>*>(sw + N);

    const double* k = ks + (size_t)kk * d;
    for (int r = lane; r < nR; r += kWave) {
        double x = 0.0;
        for (int a = 0; a < d; ++a) x += k[a] * (double)Rk[r * d + a];
        sph[r] = cis2pi(x);
    }
    __syncthreads();

    // Assembly: packed upper element e <-> (i, j), i <= j, column-major; the lower
    // entry (j, i) of Hermitian(H) is conj(H[i, j]) (Hermitian(M) reads the upper part).
    const size_t NN = (size_t)N * N;
    for (int e0 = 0; e0 < np; e0 += 4 * kWave) {
        int ii[4], jj[4];
        cx<double> h[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = e0 + u * kWave + lane;
            int j = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
            if (((j + 1) * (j + 2)) / 2 <= e) ++j;
            if ((j * (j + 1)) / 2 > e) --j;
            ii[u] = e - (j * (j + 1)) / 2;
            jj[u] = e < np ? j : -1;
            h[u] = {0.0, 0.0};
        }
        for (int r = 0; r < nR; ++r) {
            const cx<double> ph = sph[r];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (jj[u] >= 0) h[u] = cmad(h[u], T[r * NN + (size_t)jj[u] * N + ii[u]], ph);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (jj[u] < 0) continue;
            const int i = ii[u], j = jj[u];
            A[pcol(N, i) + j - i] = (i == j) ? mkc<R>((R)h[u].re, R(0)) : mkc<R>((R)h[u].re, (R)-h[u].im);
        }
    }
    __syncthreads();

    cx<R>* rf = refl + (size_t)kk * NN;
    for (int j = 0; j + 1 < N; ++j) {
        // (a) reflector of column j (packed column j, rows j+1..N-1, is contiguous)
        const int pj = pcol(N, j);
        R xn2 = 0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int r = lane + q * kWave;
            if (r >= j + 2 && r < N) xn2 += abs2(A[pj + r - j]);
        }
        xn2 = (R)wave_allsum((double)xn2);
        const cx<R> alpha = A[pj + 1];
        cx<R> tj, scale;
        R beta;
        if (!hh_scalars(alpha, xn2, tj, scale, beta)) {
            // a column of tiny entries: the scalars of x / max|x|, then beta and scale back
            R mx = 0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int r = lane + q * kWave;
                if (r > j && r < N) mx = fmax(mx, fmax(fabs(A[pj + r - j].re), fabs(A[pj + r - j].im)));
            }
#pragma unroll
            for (int o = kWave / 2; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
            const R inv = R(1) / mx;
            R xs2 = 0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int r = lane + q * kWave;
                if (r >= j + 2 && r < N) xs2 += abs2(inv * A[pj + r - j]);
            }
            xs2 = (R)wave_allsum((double)xs2);
            hh_scalars_scaled(inv * alpha, xs2, tj, scale, beta);
            beta *= mx;
            scale = inv * scale;
            if (!(inv <= num<R>::huge)) {
                tj = mkc<R>(0, 0);
                scale = mkc<R>(0, 0);
                beta = alpha.re;
            }
        }
        const bool triv = (tj.re == R(0) && tj.im == R(0));
        cx<R> vq[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int r = lane + q * kWave;
            vq[q] = mkc<R>(0, 0);
            if (r > j && r < N) {
                const cx<R> v = (r == j + 1) ? mkc<R>(1, 0) : (triv ? mkc<R>(0, 0) : A[pj + r - j] * scale);
                vq[q] = v;
                sv[r] = v;
                rf[(size_t)j * N + r] = v;
            }
        }
        if (lane == 0) {
            tau[(size_t)kk * N + j] = tj;
            dd[(size_t)j * Nkc + kk] = A[pj].re;
            ee[(size_t)j * Nkc + kk] = beta;
        }
        __syncthreads();
        if (triv) continue;  // uniform

        // (b) p = tau A22 v for the owned rows, then w = p - (tau/2)(p^H v) v
        cx<R> pq[NQ];
        cx<R> dotp = {0, 0};
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int r = lane + q * kWave;
            pq[q] = mkc<R>(0, 0);
            if (r > j && r < N) {
                pq[q] = tj * herm_dot(A, sv, N, r, j + 1);
                dotp = cmadc(dotp, pq[q], vq[q]);
            }
        }
        const cx<R> dot = wave_allsum(dotp);
        const cx<R> alpha2 = R(-0.5) * (tj * dot);
        cx<R> wq[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int r = lane + q * kWave;
            wq[q] = cmad(pq[q], alpha2, vq[q]);
            if (r > j && r < N) sw[r] = wq[q];
        }
        __syncthreads();
        // (c) A22 -= v w^H + w v^H on the owned rows (lower part c <= r)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int r = lane + q * kWave;
            if (r > j && r < N) {
                const cx<R> vr = vq[q], wr = wq[q];
                int pc = pcol(N, j + 1);
                int c = j + 1;
                for (; c + 4 <= r + 1; c += 4) {
                    const int p0 = pc + r - c, p1 = p0 + N - c - 1, p2 = p1 + N - c - 2, p3 = p2 + N - c - 3;
                    cx<R> a0 = A[p0], a1 = A[p1], a2 = A[p2], a3 = A[p3];
                    const cx<R> w0 = sw[c], w1 = sw[c + 1], w2 = sw[c + 2], w3 = sw[c + 3];
                    const cx<R> v0 = sv[c], v1 = sv[c + 1], v2 = sv[c + 2], v3 = sv[c + 3];
                    cx<R> u0 = {0, 0}, u1 = {0, 0}, u2 = {0, 0}, u3 = {0, 0};
                    u0 = cmadcb(cmadcb(u0, vr, w0), wr, v0);
                    u1 = cmadcb(cmadcb(u1, vr, w1), wr, v1);
                    u2 = cmadcb(cmadcb(u2, vr, w2), wr, v2);
                    u3 = cmadcb(cmadcb(u3, vr, w3), wr, v3);
                    A[p0] = a0 - u0;
                    A[p1] = a1 - u1;
                    A[p2] = a2 - u2;
                    A[p3] = a3 - u3;
                    pc += 4 * N - 4 * c - 6;
                }
                for (; c <= r; ++c) {
                    cx<R> u = {0, 0};
                    u = cmadcb(u, vr, sw[c]);
                    u = cmadcb(u, wr, sv[c]);
                    A[pc + r - c] = A[pc + r - c] - u;
                    pc += N - c;
                }
            }
        }
        __syncthreads();
    }
    if (lane == 0) {
        dd[(size_t)(N - 1) * Nkc + kk] = A[pcol(N, N - 1)].re;
        ee[(size_t)(N - 1) * Nkc + kk] = R(0);
    }
}

// ---------------------------------------------------------------------------------
// M = Z diag(f) Z^T (real), f_c = fermidirac(e_c - mu; T) for Z column c (or the given
// occupation occ[kk*N + n] of the n-th smallest eigenvalue), and the
// band-energy term eb[kk] = w_k sum_n e_n f_n.  One 256-thread block per k-point,
// Z staged in LDS, 4x4 register tiles.
// ---------------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kTPB) void k_zfz(int N, const R* __restrict__ Z,
                                              const R* __restrict__ evals,
                                              const int* __restrict__ perm,
                                              const double* __restrict__ kw, double wuni,
                                              double mu, double Temp,
                                              const double* __restrict__ occ, R* __restrict__ M,
                                              double* __restrict__ eb) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int kk = blockIdx.x;
    const int tid = threadIdx.x;
    const int ld = N + 1;
    R* Zs = reinterpret_cast<R*>(smem);  // [N][ld], Zs[c*ld + i] = Z(i, c)
    double* sf = reinterpret_cast<double*>(Zs + (size_t)N * ld + (((size_t)N * ld) & 1));
    double* sred = sf + N;
    const size_t NN = (size_t)N * N;
    const R* zk = Z + (size_t)kk * NN;
    for (int e = tid; e < N * N; e += kTPB) Zs[(size_t)(e / N) * ld + (e % N)] = zk[e];
    const R* ev = evals + (size_t)kk * N;
    const int* pm = perm + (size_t)kk * N;
    double ebp = 0.0;
    for (int n = tid; n < N; n += kTPB) {
        const double en = (double)ev[n];
        const double x = en - mu;
        const double f = occ ? occ[(size_t)kk * N + n]
                             : (Temp > 0.0 ? 1.0 / (exp(x / Temp) + 1.0) : (x < 0.0 ? 1.0 : (x > 0.0 ? 0.0 : 0.5)));
        sf[min(max(pm[n], 0), N - 1)] = f;
        ebp += en * f;
    }
    ebp = wave_sum(ebp);
    if ((tid & (kWave - 1)) == 0) sred[tid / kWave] = ebp;
    __syncthreads();
    if (tid == 0) eb[kk] = (kw ? kw[kk] : wuni) * (sred[0] + sred[1] + sred[2] + sred[3]);
    const int nb = (N + 3) / 4;
    R* Mk = M + (size_t)kk * NN;
    for (int b = tid; b < nb * nb; b += kTPB) {
        const int bi = b % nb, bj = b / nb;
        if (bj > bi) continue;  // lower block triangle, mirrored on store
        R acc[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] = R(0);
        for (int c = 0; c < N; ++c) {
            const R f = (R)sf[c];
            if (f == R(0)) continue;
            const R* zc = Zs + (size_t)c * ld;
            R a[4], bb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = 4 * bi + u, jx = 4 * bj + u;
                a[u] = i < N ? zc[i] : R(0);
                bb[u] = jx < N ? f * zc[jx] : R(0);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] = fma(a[u], bb[v], acc[u][v]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = 4 * bi + u, jx = 4 * bj + v;
                if (i < N && jx < N) {
                    Mk[(size_t)jx * N + i] = acc[u][v];
                    Mk[(size_t)i * N + jx] = acc[u][v];
                }
            }
    }
}

// ---------------------------------------------------------------------------------
// rho_k = H_0 ... H_{N-2} M H_{N-2}^H ... H_0^H, one wave per k-point.
// For each reflector (v supported on j+1..N-1): y = M v, w = conj(tau) y -
// (|tau|^2/2)(v^H y) v, M -= v w^H + w v^H (rows r > j of the packed lower part).
// LDS: packed M [N(N+1)/2] | sv [N] | sw [N]
// ---------------------------------------------------------------------------------
template <typename R, int NQ>
__global__ __launch_bounds__(kWave) void k_rho_w(int N, const cx<R>* __restrict__ refl,
                                                 const cx<R>* __restrict__ tau,
                                                 const R* __restrict__ M,
                                                 cx<double>* __restrict__ rhok) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int kk = blockIdx.x;
    const int lane = threadIdx.x;
    const int np = (N * (N + 1)) / 2;
    cx<R>* A = reinterpret_cast<cx<R>*>(smem);
    cx<R>* sv = A + np;
    cx<R>* sw = sv + N;
    const size_t NN = (size_t)N * N;
    const R* Mk = M + (size_t)kk * NN;
    // packed lower load: column c, rows c..N-1 (coalesced along the column)
    for (int c = 0; c < N; ++c) {
        const int pc = pcol(N, c);
        for (int r = c + lane; r < N; r += kWave) A[pc + r - c] = mkc<R>(Mk[(size_t)c * N + r], R(0));
    }
    const cx<R>* rfk = refl + (size_t)kk * NN;
    const cx<R>* tk = tau + (size_t)kk * N;
    // reflector j-1 is fetched while j is applied
    cx<R> vq[NQ], vnext[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int r = lane + q * kWave, j = N - 2;
        vq[q] = (j >= 0 && r > j && r < N) ? rfk[(size_t)j * N + r] : mkc<R>(0, 0);
    }
    __syncthreads();
    for (int j = N - 2; j >= 0; --j) {
        const cx<R> tj = tk[j];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int r = lane + q * kWave;
            vnext[q] = (j >= 1 && r > j - 1 && r < N) ? rfk[(size_t)(j - 1) * N + r] : mkc<R>(0, 0);
            if (r > j && r < N) sv[r] = vq[q];
        }
        const bool triv = (tj.re == R(0) && tj.im == R(0));
        if (!triv) {
            __syncthreads();
            // y = M v over all rows; dot = v^H y
            cx<R> yq[NQ];
            cx<R> dotp = {0, 0};
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int r = lane + q * kWave;
                const cx<R> acc = r < N ? herm_dot(A, sv, N, r, j + 1) : mkc<R>(0, 0);
                yq[q] = acc;
                dotp = cmadc(dotp, vq[q], acc);
            }
            const R dot = (R)wave_allsum((double)dotp.re);  // v^H M v is real
            const R h = R(0.5) * abs2(tj) * dot;
            cx<R> wq[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int r = lane + q * kWave;
                cx<R> w = conj(tj) * yq[q];
                w.re -= h * vq[q].re;
                w.im -= h * vq[q].im;
                wq[q] = w;
                if (r < N) sw[r] = w;
            }
            __syncthreads();
            // rows r > j: M(r, c) -= v_r conj(w_c) + w_r conj(v_c), c = 0..r (v_c = 0 for c <= j)
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int r = lane + q * kWave;
                if (r > j && r < N) {
                    const cx<R> vr = vq[q], wr = wq[q];
                    int pc = 0;
                    int c = 0;
                    for (; c + 4 <= j + 1; c += 4) {
                        const int p0 = pc + r - c, p1 = p0 + N - c - 1, p2 = p1 + N - c - 2, p3 = p2 + N - c - 3;
                        cx<R> a0 = A[p0], a1 = A[p1], a2 = A[p2], a3 = A[p3];
                        const cx<R> w0 = sw[c], w1 = sw[c + 1], w2 = sw[c + 2], w3 = sw[c + 3];
                        cx<R> u0 = {0, 0}, u1 = {0, 0}, u2 = {0, 0}, u3 = {0, 0};
                        u0 = cmadcb(u0, vr, w0);
                        u1 = cmadcb(u1, vr, w1);
                        u2 = cmadcb(u2, vr, w2);
                        u3 = cmadcb(u3, vr, w3);
                        A[p0] = a0 - u0;
                        A[p1] = a1 - u1;
                        A[p2] = a2 - u2;
                        A[p3] = a3 - u3;
                        pc += 4 * N - 4 * c - 6;
                    }
                    for (; c <= j; ++c) {
                        cx<R> u = {0, 0};
                        u = cmadcb(u, vr, sw[c]);
                        A[pc + r - c] = A[pc + r - c] - u;
                        pc += N - c;
                    }
                    c = j + 1;  // pc is pcol(N, j + 1) here
                    for (; c + 4 <= r + 1; c += 4) {
                        const int p0 = pc + r - c, p1 = p0 + N - c - 1, p2 = p1 + N - c - 2, p3 = p2 + N - c - 3;
                        cx<R> a0 = A[p0], a1 = A[p1], a2 = A[p2], a3 = A[p3];
                        const cx<R> w0 = sw[c], w1 = sw[c + 1], w2 = sw[c + 2], w3 = sw[c + 3];
                        const cx<R> v0 = sv[c], v1 = sv[c + 1], v2 = sv[c + 2], v3 = sv[c + 3];
                        cx<R> u0 = {0, 0}, u1 = {0, 0}, u2 = {0, 0}, u3 = {0, 0};
                        u0 = cmadcb(cmadcb(u0, vr, w0), wr, v0);
                        u1 = cmadcb(cmadcb(u1, vr, w1), wr, v1);
                        u2 = cmadcb(cmadcb(u2, vr, w2), wr, v2);
                        u3 = cmadcb(cmadcb(u3, vr, w3), wr, v3);
                        A[p0] = a0 - u0;
                        A[p1] = a1 - u1;
                        A[p2] = a2 - u2;
                        A[p3] = a3 - u3;
                        pc += 4 * N - 4 * c - 6;
                    }
                    for (; c <= r; ++c) {
                        cx<R> u = {0, 0};
                        u = cmadcb(u, vr, sw[c]);
                        u = cmadcb(u, wr, sv[c]);
                        A[pc + r - c] = A[pc + r - c] - u;
                        pc += N - c;
                    }
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) vq[q] = vnext[q];
    }
    // full Hermitian rho_k, column-major
    cx<double>* o = rhok + (size_t)kk * NN;
    for (int c = 0; c < N; ++c) {
        for (int r = lane; r < N; r += kWave) {
            const cx<R> x = herm_at(A, N, r, c);
            o[(size_t)c * N + r] = mkc<double>((double)x.re, (double)x.im);
        }
    }
}

}  // namespace lqm
