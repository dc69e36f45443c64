// Large-N Hermitian eigensolver (FP64) over a batch of k-points, one k-point per batch
// entry, all matrices column-major with leading dimension N:
//   k_hk_full        A = Hermitian(H(k)) (upper triangle of H(k)), full storage
//   k_trd_col        zlatrd column step j: pending panel update of column j, zlarfg
//   k_trd_hemv       partial products of y = A22 v over column chunks
//   k_trd_w          w = tau (A22 v - V (W^H v) - W (V^H v)) + alpha v (zlatrd)
//   k_zgemm_gen      batched C = alpha sum_p op(A_p) op(B_p) + beta C on the FP64 matrix
//                    cores: the trailing her2k update, compact-WY products, rho_k
//   k_sturm_bisect   eigenvalue n of the tridiagonal by Sturm-count bisection
//   k_invit          eigenvectors of the tridiagonal by inverse iteration, clusters
//                    orthogonalised within every iteration
//   k_make_y, k_wy_T, k_z_to_v, k_occ_eb   compact-WY and rho_k helpers
//
// The tridiagonalization is the one-stage blocked Householder reduction (LAPACK zhetrd,
// lower, with zlatrd panels of kBnb columns): inside a panel the trailing matrix is left
// as it was at the panel start and the pending rank-2i update enters every column step
// through V and W; after the panel A22 -= V W^H + W V^H on the matrix cores.  The
// reflectors H_j = I - tau_j v_j v_j^H (v_j(j+1) = 1) give A = Q T Q^H, Q = H_0 ... H_{N-2}.
#pragma once
#include "lqm_common.hpp"
#include "eig_kernels.hpp"

namespace lqm {

constexpr int kBnb = 32;      // tridiagonalization panel width
constexpr int kBwy = 64;      // reflectors per compact-WY block of the back-transformation
constexpr int kHvCols = 64;   // column chunk of the trailing matrix-vector product
constexpr int kBRows = 16;    // rows per thread of the one-block-per-k kernels (N <= 4096)
constexpr int kBMaxN = 256 * kBRows;
constexpr int kHkK = 8;       // k-points per block of k_hk_full

// ---------------------------------------------------------------------------------
// A(r, c) = H(r, c) for r < c, conj(H(c, r)) for r > c, Re H(r, r) on the diagonal, with
// H(k) = sum_R T_R e^{2 pi i k.R}.  Block (t, g): the 32 x 32 tile pair (I, J), I <= J
// (t = J(J+1)/2 + I), for k-points g*kHkK ... ; the mirrored tile goes through LDS so both
// writes are coalesced.  Dynamic LDS: kHkK * nR phases.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_hk_full(int N, int d, int nR, const int* __restrict__ Rk,
                                                 const cx<double>* __restrict__ T, int Kc,
                                                 const double* __restrict__ ks,
                                                 cx<double>* __restrict__ A) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<double>* sph = reinterpret_cast<cx<double>*>(smem);
    __shared__ cx<double> tile[32][33];
    const int t = blockIdx.x;
    int J = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((J + 1) * (J + 2) / 2 <= t) ++J;
    while (J * (J + 1) / 2 > t) --J;
    const int I = t - J * (J + 1) / 2;
    const int k0 = blockIdx.y * kHkK;
    const int nk = min(kHkK, Kc - k0);
    for (int e = threadIdx.x; e < nk * nR; e += 256) {
        const int kq = e / nR, r = e - kq * nR;
        const double* k = ks + (size_t)(k0 + kq) * d;
        double x = 0.0;
        for (int a = 0; a < d; ++a) x += k[a] * (double)Rk[r * d + a];
        sph[kq * nR + r] = cis2pi(x);
    }
    __syncthreads();
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const size_t NN = (size_t)N * N;
    const int r = 32 * I + tx;
    for (int kq = 0; kq < nk; ++kq) {
        cx<double>* Ak = A + (size_t)(k0 + kq) * NN;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int cl = ty + 8 * q, c = 32 * J + cl;
            cx<double> h = {0.0, 0.0};
            if (r < N && c < N && r <= c) {
                for (int R = 0; R < nR; ++R) h = cmad(h, T[(size_t)R * NN + (size_t)c * N + r], sph[kq * nR + R]);
                Ak[(size_t)c * N + r] = r == c ? cx<double>{h.re, 0.0} : h;
            }
            tile[cl][tx] = h;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int rl = ty + 8 * q;
            const int c = 32 * J + tx, rr = 32 * I + rl;  // element (c, rr), c > rr
            if (c < N && rr < N && rr < c) {
                const cx<double> h = tile[tx][rl];
                Ak[(size_t)rr * N + c] = {h.re, -h.im};
            }
        }
        __syncthreads();
    }
}

template <int W>
__device__ __forceinline__ double blk_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) s += red[w];
    return s;
}

// ---------------------------------------------------------------------------------
// Column step j = p + i of a panel (zlatrd, lower).  Column j receives the panel's
// pending update A(j:N, j) -= V(j:N, 0:i) conj(W(j, 0:i))^T + W(j:N, 0:i) conj(V(j, 0:i))^T,
// then zlarfg maps x = A(j+1:N, j) to beta e_1: beta = -sign(Re x_0) ||x||,
// tau = (beta - x_0)/beta, v = [1, x(1:)/(x_0 - beta)] (tau = 0, v = e_1 when x(1:) = 0
// and x_0 is real).  Out: d[j] = Re A(j, j), e[j] = beta, tau[j]; v into V(:, i) (rows
// j+1..N-1) and into A(j+1:N, j) with A(j+1, j) = beta.  last: column N-1, no reflector.
// One 256-thread block per k-point; thread t owns rows j + t + 256q.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_trd_col(int N, int j, int i, int last, cx<double>* __restrict__ A,
                                                 cx<double>* __restrict__ V,
                                                 const cx<double>* __restrict__ W,
                                                 double* __restrict__ dd, double* __restrict__ ee,
                                                 cx<double>* __restrict__ tau) {
    __shared__ cx<double> wj[kBnb], vj[kBnb];
    __shared__ double red[4];
    __shared__ cx<double> s_alpha;
    __shared__ double s_dj;
    const int tid = threadIdx.x;
    const size_t kk = blockIdx.x;
    const size_t NN = (size_t)N * N, NP = (size_t)N * kBnb;
    cx<double>* Ak = A + kk * NN;
    cx<double>* Vk = V + kk * NP;
    const cx<double>* Wk = W + kk * NP;
    if (tid < i) {
        wj[tid] = Wk[(size_t)tid * N + j];
        vj[tid] = Vk[(size_t)tid * N + j];
    }
    if (tid == 0) s_alpha = {0.0, 0.0};
    __syncthreads();
    cx<double> a[kBRows];
    double xn = 0.0;
#pragma unroll
    for (int q = 0; q < kBRows; ++q) {
        const int r = j + tid + 256 * q;
        cx<double> x = {0.0, 0.0};
        if (r < N) {
            x = Ak[(size_t)j * N + r];
            cx<double> u = {0.0, 0.0};
            for (int t = 0; t < i; ++t) {
                u = cmadcb(u, Vk[(size_t)t * N + r], wj[t]);
                u = cmadcb(u, Wk[(size_t)t * N + r], vj[t]);
            }
            x = x - u;
            if (r >= j + 2) xn += abs2(x);
        }
        a[q] = x;
    }
    if (tid == 0) s_dj = a[0].re;
    if (tid == 1 && j + 1 < N) s_alpha = a[0];
    const double xnorm2 = blk_sum<4>(xn, red);  // (its barriers publish s_dj, s_alpha)
    const double dj = s_dj;
    if (last) {
        if (tid == 0) {
            Ak[(size_t)j * N + j] = {dj, 0.0};
            dd[kk * N + j] = dj;
            ee[kk * N + j] = 0.0;
            tau[kk * N + j] = {0.0, 0.0};
        }
        return;
    }
    const cx<double> alpha = s_alpha;
    cx<double> tj = {0.0, 0.0}, scale = {0.0, 0.0};
    double beta = alpha.re;
    if (!(xnorm2 == 0.0 && alpha.im == 0.0)) {
        const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xnorm2);
        beta = alpha.re >= 0.0 ? -nrm : nrm;
        tj = {(beta - alpha.re) / beta, -alpha.im / beta};
        scale = crecip(cx<double>{alpha.re - beta, alpha.im});
    }
#pragma unroll
    for (int q = 0; q < kBRows; ++q) {
        const int r = j + tid + 256 * q;
        if (r < N) {
            if (r == j) {
                Ak[(size_t)j * N + j] = {dj, 0.0};
            } else if (r == j + 1) {
                Ak[(size_t)j * N + r] = {beta, 0.0};
                Vk[(size_t)i * N + r] = {1.0, 0.0};
            } else {
                const cx<double> v = a[q] * scale;
                Ak[(size_t)j * N + r] = v;
                Vk[(size_t)i * N + r] = v;
            }
        }
    }
    if (tid == 0) {
        dd[kk * N + j] = dj;
        ee[kk * N + j] = beta;
        tau[kk * N + j] = tj;
    }
}

// Partial products of y = A22 v, A22 = A(j+1:N, j+1:N) as stored (the trailing matrix of
// the panel start, both triangles): block (x, y, k) sums columns j+1 + 64y ... of rows
// j+1 + 256x + t into P[k][y][row] (P stride per k-point: nchmax * N).
__global__ __launch_bounds__(256) void k_trd_hemv(int N, int j, int i, int nchmax, const cx<double>* __restrict__ A,
                                                  const cx<double>* __restrict__ V,
                                                  cx<double>* __restrict__ P) {
    __shared__ cx<double> sv[kHvCols];
    const size_t kk = blockIdx.z;
    const int c0 = j + 1 + blockIdx.y * kHvCols;
    const int nc = min(kHvCols, N - c0);
    const cx<double>* v = V + kk * (size_t)N * kBnb + (size_t)i * N;
    if ((int)threadIdx.x < nc) sv[threadIdx.x] = v[c0 + threadIdx.x];
    __syncthreads();
    const int r = j + 1 + blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    const cx<double>* Ak = A + kk * (size_t)N * N + (size_t)c0 * N + r;
    cx<double> a0 = {0.0, 0.0}, a1 = {0.0, 0.0};
    int c = 0;
#pragma unroll 4
    for (; c + 1 < nc; c += 2) {
        a0 = cmad(a0, Ak[(size_t)c * N], sv[c]);
        a1 = cmad(a1, Ak[(size_t)(c + 1) * N], sv[c + 1]);
    }
    if (c < nc) a0 = cmad(a0, Ak[(size_t)c * N], sv[c]);
    P[kk * (size_t)nchmax * N + (size_t)blockIdx.y * N + r] = a0 + a1;
}

// w = tau (y - V (W^H v) - W (V^H v)), y = sum_c P[k][c][.], alpha = -1/2 tau (w^H v),
// w += alpha v, into W(:, i) rows j+1..N-1 (zlatrd).  One block per k-point.
__global__ __launch_bounds__(256) void k_trd_w(int N, int j, int i, int nch, int nchmax,
                                               const cx<double>* __restrict__ P,
                                               const cx<double>* __restrict__ V,
                                               cx<double>* __restrict__ W,
                                               const cx<double>* __restrict__ tau) {
    __shared__ cx<double> t1[kBnb], t2[kBnb];
    __shared__ double red[8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t kk = blockIdx.x;
    const size_t NP = (size_t)N * kBnb;
    const cx<double>* Vk = V + kk * NP;
    cx<double>* Wk = W + kk * NP;
    const cx<double>* v = Vk + (size_t)i * N;
    // t1 = W(:, 0:i)^H v (even products), t2 = V(:, 0:i)^H v (odd); wave w takes the
    // products w, w + 4, ...; lanes stride the rows
    {
        cx<double> acc[2 * kBnb / 4];
#pragma unroll
        for (int q = 0; q < 2 * kBnb / 4; ++q) acc[q] = {0.0, 0.0};
        for (int r = j + 1 + lane; r < N; r += 64) {
            const cx<double> vr = v[r];
#pragma unroll
            for (int q = 0; q < 2 * kBnb / 4; ++q) {
                const int idx = wave + 4 * q;
                if (idx < 2 * i) {
                    const cx<double>* X = (idx & 1) ? Vk : Wk;
                    acc[q] = cmadc(acc[q], X[(size_t)(idx >> 1) * N + r], vr);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 2 * kBnb / 4; ++q) {
            const int idx = wave + 4 * q;
            if (idx < 2 * i) {
                const cx<double> s = wave_sum(acc[q]);
                if (lane == 0) ((idx & 1) ? t2 : t1)[idx >> 1] = s;
            }
        }
    }
    __syncthreads();
    const cx<double> tj = tau[kk * N + j];
    const cx<double>* Pk = P + kk * (size_t)nchmax * N;
    cx<double> w[kBRows];
    double dre = 0.0, dim = 0.0;
#pragma unroll
    for (int q = 0; q < kBRows; ++q) {
        const int r = j + 1 + tid + 256 * q;
        cx<double> x = {0.0, 0.0};
        if (r < N) {
            cx<double> y = {0.0, 0.0};
            for (int c = 0; c < nch; ++c) y = y + Pk[(size_t)c * N + r];
            cx<double> u = {0.0, 0.0};
            for (int t = 0; t < i; ++t) {
                u = cmad(u, Vk[(size_t)t * N + r], t1[t]);
                u = cmad(u, Wk[(size_t)t * N + r], t2[t]);
            }
            x = tj * (y - u);
            const cx<double> vr = v[r];
            dre += x.re * vr.re + x.im * vr.im;  // conj(w) v
            dim += x.re * vr.im - x.im * vr.re;
        }
        w[q] = x;
    }
    const double sre = blk_sum<4>(dre, red);
    const double sim = blk_sum<4>(dim, red + 4);
    // alpha = -1/2 tau (w^H v)
    const cx<double> al = {-0.5 * (tj.re * sre - tj.im * sim), -0.5 * (tj.re * sim + tj.im * sre)};
#pragma unroll
    for (int q = 0; q < kBRows; ++q) {
        const int r = j + 1 + tid + 256 * q;
        if (r < N) Wk[(size_t)i * N + r] = cmad(w[q], al, v[r]);
    }
}

// ---------------------------------------------------------------------------------
// Batched complex GEMM on the FP64 matrix cores:
//   C = alpha sum_{p < npair} op(A_p) op(B_p) + beta C
// op 0: as stored (column-major, leading dimension ld), op 1: conjugate transpose.
// ascale (optional): column k of op(A_p) scaled by ascale[k] (per batch entry, stride sS).
// 64 x 64 output tile per block (4 waves of 32 x 32, 2 x 2 MFMA 16x16x4 tiles); operands
// read straight from global memory (L1/L2), rows/columns outside M, N, K read as zero.
// ---------------------------------------------------------------------------------
struct ZGemmArgs {
    int M, N, K, npair;
    const cx<double>* A[2];
    const cx<double>* B[2];
    int lda, ldb, ldc;
    long long sA, sB, sC, sS;
    cx<double>* C;
    cx<double> alpha, beta;
    const double* ascale;
};

template <int OP>
__device__ __forceinline__ cx<double> opA_at(const cx<double>* A, int ld, int m, int k) {
    if (OP == 0) return A[(size_t)k * ld + m];
    const cx<double> x = A[(size_t)m * ld + k];
    return {x.re, -x.im};
}
template <int OP>
__device__ __forceinline__ cx<double> opB_at(const cx<double>* B, int ld, int k, int n) {
    if (OP == 0) return B[(size_t)n * ld + k];
    const cx<double> x = B[(size_t)k * ld + n];
    return {x.re, -x.im};
}

template <int OPA, int OPB>
__device__ __forceinline__ void gemm_acc(const cx<double>* Ab, int lda, const cx<double>* Bb, int ldb,
                                         const double* sc, int M, int N, int K, int m0, int n0, int li,
                                         int g, lqm_d4 (&cr)[2][2], lqm_d4 (&ci)[2][2]) {
    for (int k0 = 0; k0 < K; k0 += 16) {
        const int kb = k0 + 4 * g;
        double ar[2][4], ai[2][4], br[2][4], bi[2][4];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int m = m0 + 16 * t + li, n = n0 + 16 * t + li;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int k = kb + s;
                cx<double> x = {0.0, 0.0}, y = {0.0, 0.0};
                if (k < K) {
                    if (m < M) {
                        x = opA_at<OPA>(Ab, lda, m, k);
                        if (sc) {
                            const double f = sc[k];
                            x = {f * x.re, f * x.im};
                        }
                    }
                    if (n < N) y = opB_at<OPB>(Bb, ldb, k, n);
                }
                ar[t][s] = x.re;
                ai[t][s] = x.im;
                br[t][s] = y.re;
                bi[t][s] = y.im;
            }
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a][s], br[b][s], cr[a][b], 0, 0, 0);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a][s], bi[b][s], ci[a][b], 0, 0, 0);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[a][s], bi[b][s], cr[a][b], 0, 0, 0);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a][s], br[b][s], ci[a][b], 0, 0, 0);
        }
    }
}

template <int OPA, int OPB>
__global__ __launch_bounds__(256) void k_zgemm_gen(ZGemmArgs g) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, g4 = lane >> 4;
    const int m0 = blockIdx.x * 64 + (wave & 1) * 32;
    const int n0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const size_t b = blockIdx.z;
    lqm_d4 cr[2][2], ci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            cr[a][c] = lqm_d4{0.0, 0.0, 0.0, 0.0};
            ci[a][c] = lqm_d4{0.0, 0.0, 0.0, 0.0};
        }
    const double* sc = g.ascale ? g.ascale + b * g.sS : nullptr;
    for (int p = 0; p < g.npair; ++p)
        gemm_acc<OPA, OPB>(g.A[p] + b * g.sA, g.lda, g.B[p] + b * g.sB, g.ldb, sc, g.M, g.N, g.K, m0, n0, li,
                           g4, cr, ci);
    cx<double>* Cb = g.C + b * g.sC;
    const bool zero_beta = g.beta.re == 0.0 && g.beta.im == 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = m0 + 16 * a + g4 + 4 * i, n = n0 + 16 * c + li;
                if (m < g.M && n < g.N) {
                    const size_t off = (size_t)n * g.ldc + m;
                    const cx<double> acc = {cr[a][c][i], ci[a][c][i]};
                    cx<double> out = g.alpha * acc;
                    if (!zero_beta) out = cmad(out, g.beta, Cb[off]);
                    Cb[off] = out;
                }
            }
}

// ---------------------------------------------------------------------------------
// Eigenvalue n (ascending) of the symmetric tridiagonal (d, e) of k-point blockIdx.y:
// bisection on the Sturm count (negative pivots of T - x I = L D L^T, a pivot of modulus
// below pivmin replaced by -pivmin, LAPACK dstebz), inside the Gershgorin interval, until
// the interval holds no double between its ends or is below
// max(2 eps max(|a|, |b|), 1e-3 eps ||T||).  Block x = 0 writes ||T|| (Gershgorin bound)
// to tnorm[k].
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sturm_bisect(int N, const double* __restrict__ dd,
                                                      const double* __restrict__ ee,
                                                      double* __restrict__ lam,
                                                      double* __restrict__ tnorm) {
    __shared__ double rlo[4], rhi[4], re2[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t kk = blockIdx.y;
    const double* d = dd + kk * N;
    const double* e = ee + kk * N;
    double lo = 1e308, hi = -1e308, e2 = 0.0;
    for (int i = tid; i < N; i += 256) {
        const double el = i > 0 ? fabs(e[i - 1]) : 0.0, er = i + 1 < N ? fabs(e[i]) : 0.0;
        lo = fmin(lo, d[i] - el - er);
        hi = fmax(hi, d[i] + el + er);
        if (i + 1 < N) e2 = fmax(e2, e[i] * e[i]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, o));
        hi = fmax(hi, __shfl_xor(hi, o));
        e2 = fmax(e2, __shfl_xor(e2, o));
    }
    if (lane == 0) {
        rlo[wave] = lo;
        rhi[wave] = hi;
        re2[wave] = e2;
    }
    __syncthreads();
    lo = fmin(fmin(rlo[0], rlo[1]), fmin(rlo[2], rlo[3]));
    hi = fmax(fmax(rhi[0], rhi[1]), fmax(rhi[2], rhi[3]));
    e2 = fmax(fmax(re2[0], re2[1]), fmax(re2[2], re2[3]));
    const double eps = 2.220446049250313e-16;
    const double bnorm = fmax(fabs(lo), fabs(hi));
    const double pivmin = 2.2250738585072014e-308 * fmax(1.0, e2);
    if (blockIdx.x == 0 && tid == 0) tnorm[kk] = bnorm;
    const int n = blockIdx.x * 256 + tid;
    if (n >= N) return;
    double a = lo - 2.0 * eps * bnorm - 2.0 * pivmin, b = hi + 2.0 * eps * bnorm + 2.0 * pivmin;
    const double atol = 1e-3 * eps * bnorm;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (a + b);
        if (!(mid > a && mid < b)) break;
        if (b - a <= fmax(2.0 * eps * fmax(fabs(a), fabs(b)), atol)) break;
        double q = d[0] - mid;
        if (fabs(q) <= pivmin) q = -pivmin;
        int cnt = q < 0.0;
        for (int i = 1; i < N; ++i) {
            const double ei = e[i - 1];
            q = (d[i] - mid) - ei * ei / q;
            if (fabs(q) <= pivmin) q = -pivmin;
            cnt += q < 0.0;
        }
        if (cnt > n) b = mid;
        else a = mid;
    }
    lam[kk * N + n] = 0.5 * (a + b);
}

__device__ __forceinline__ double invit_start(unsigned long long seed, int i) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

// ---------------------------------------------------------------------------------
// Eigenvectors of the symmetric tridiagonal (d, e) by inverse iteration (LAPACK dstein's
// scheme).  For eigenvalue lambda_m: three solves of (T - lambda_m I) x = b, b the last x
// scaled to unit max-norm, first b a seeded uniform(-1, 1) vector; the solve factors
// T - lambda I = P L U with partial pivoting (dlagtf) fused with the forward
// substitution, and replaces U pivots below eps ||T|| by +-eps ||T|| (dlagts).
// Eigenvalues within ctol of their predecessor form a cluster; the thread of a cluster's
// first member computes all members in order, each orthogonalised against the earlier
// ones after every solve (modified Gram-Schmidt) with its shift raised to at least
// 10 eps |lambda| above the previous member's (dstein's PERTOL).  Thread m of block
// (x, k): eigenvector m of k-point k into Zt[k][i * ldz + m] (unit 2-norm), m < ncol
// (no cluster crosses ncol).  scr: U's diagonals, 3 * N * ldz doubles per k-point,
// interleaved by m.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_invit(int N, int ncol, int ldz, const double* __restrict__ dd,
                                              const double* __restrict__ ee,
                                              const double* __restrict__ lam,
                                              const double* __restrict__ tnorm, double ctol_rel,
                                              double* __restrict__ Zt, double* __restrict__ scr,
                                              int kbase, int* __restrict__ status) {
    const int m = blockIdx.x * 64 + threadIdx.x;
    const size_t kk = blockIdx.y;
    if (m >= ncol) return;
    const double* d = dd + kk * N;
    const double* e = ee + kk * N;
    const double* l = lam + kk * N;
    const double tn = tnorm[kk];
    const double ctol = ctol_rel * tn;
    if (m > 0 && l[m] - l[m - 1] <= ctol) return;
    int mend = m + 1;
    while (mend < ncol && l[mend] - l[mend - 1] <= ctol) ++mend;
    const double eps = 2.220446049250313e-16;
    const double tol = fmax(eps * tn, 2.2250738585072014e-308);
    const size_t NZ = (size_t)N * ldz;
    double* Z = Zt + kk * NZ;
    double* U0 = scr + (kk * 3 + 0) * NZ;
    double* U1 = scr + (kk * 3 + 1) * NZ;
    double* U2 = scr + (kk * 3 + 2) * NZ;
    double xprev = 0.0;
    bool bad = false;
    for (int mm = m; mm < mend; ++mm) {
        double x = l[mm];
        if (mm > m) {
            const double pert = 10.0 * eps * fabs(x);
            if (x - xprev < pert) x = xprev + pert;
        }
        xprev = x;
        const unsigned long long seed =
            0x51A7E5EEDull ^ ((unsigned long long)(kbase + (long long)kk) << 24) ^ (unsigned long long)mm;
        for (int i = 0; i < N; ++i) Z[(size_t)i * ldz + mm] = invit_start(seed, i);
        for (int it = 0; it < 3; ++it) {
            // factor + forward substitution
            double acur = d[0] - x, bcur = N > 1 ? e[0] : 0.0;
            double ycur = Z[mm];
            for (int k = 0; k + 1 < N; ++k) {
                const double c = e[k];
                const double anext = d[k + 1] - x;
                const double bnext = k + 2 < N ? e[k + 1] : 0.0;
                const double ynext = Z[(size_t)(k + 1) * ldz + mm];
                const size_t o = (size_t)k * ldz + mm;
                if (fabs(acur) >= fabs(c)) {
                    const double mult = acur != 0.0 ? c / acur : 0.0;
                    U0[o] = acur;
                    U1[o] = bcur;
                    U2[o] = 0.0;
                    Z[o] = ycur;
                    acur = anext - mult * bcur;
                    bcur = bnext;
                    ycur = ynext - mult * ycur;
                } else {
                    const double mult = acur / c;
                    U0[o] = c;
                    U1[o] = anext;
                    U2[o] = bnext;
                    Z[o] = ynext;
                    const double na = bcur - mult * anext;
                    bcur = -mult * bnext;
                    acur = na;
                    ycur = ycur - mult * ynext;
                }
            }
            U0[(size_t)(N - 1) * ldz + mm] = acur;
            Z[(size_t)(N - 1) * ldz + mm] = ycur;
            // back substitution
            double x1 = 0.0, x2 = 0.0, amax = 0.0;
            for (int k = N - 1; k >= 0; --k) {
                const size_t o = (size_t)k * ldz + mm;
                double p = U0[o];
                if (fabs(p) < tol) p = p >= 0.0 ? tol : -tol;
                double v = Z[o];
                if (k + 1 < N) v -= U1[o] * x1;
                if (k + 2 < N) v -= U2[o] * x2;
                v /= p;
                Z[o] = v;
                x2 = x1;
                x1 = v;
                amax = fmax(amax, fabs(v));
            }
            // orthogonalise against the earlier members of the cluster
            for (int mp = m; mp < mm; ++mp) {
                double dot = 0.0;
                for (int i = 0; i < N; ++i) dot += Z[(size_t)i * ldz + mp] * Z[(size_t)i * ldz + mm];
                amax = 0.0;
                for (int i = 0; i < N; ++i) {
                    const double v = Z[(size_t)i * ldz + mm] - dot * Z[(size_t)i * ldz + mp];
                    Z[(size_t)i * ldz + mm] = v;
                    amax = fmax(amax, fabs(v));
                }
            }
            if (!(amax > 0.0) || !isfinite(amax)) {
                bad = true;
                amax = 1.0;
            }
            if (it < 2) {
                const double s = 1.0 / amax;
                for (int i = 0; i < N; ++i) Z[(size_t)i * ldz + mm] *= s;
            }
        }
        double nrm = 0.0;
        for (int i = 0; i < N; ++i) {
            const double v = Z[(size_t)i * ldz + mm];
            nrm += v * v;
        }
        nrm = sqrt(nrm);
        if (!(nrm > 0.0) || !isfinite(nrm)) {
            bad = true;
            nrm = 1.0;
        }
        const double s = 1.0 / nrm;
        for (int i = 0; i < N; ++i) Z[(size_t)i * ldz + mm] *= s;
    }
    if (bad) atomicOr(status, kStatNonFinite);
}

// A (reflectors below the subdiagonal, beta on it) -> explicit Y in place: column c < N-1
// holds v_c (zero above row c+1, 1 at row c+1); column N-1 is zero.
__global__ __launch_bounds__(256) void k_make_y(int N, cx<double>* __restrict__ A) {
    const size_t NN = (size_t)N * N;
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= NN) return;
    const int c = (int)(e / N), r = (int)(e - (size_t)c * N);
    cx<double>* Ak = A + (size_t)blockIdx.y * NN;
    if (c == N - 1 || r <= c) Ak[e] = {0.0, 0.0};
    else if (r == c + 1) Ak[e] = {1.0, 0.0};
}

// Compact-WY factor of reflector block b (reflectors jb = 64b ... jb + nbb - 1): the
// upper triangular T with H_jb ... H_{jb+nbb-1} = I - Y T Y^H (LAPACK zlarft, forward,
// columnwise): T(t, t) = tau_t, T(0:t, t) = -tau_t T(0:t, 0:t) P(0:t, t), P = Y^H Y
// (rows jb+1..N-1, on the matrix cores).  Out: Tw[k][b] (64 x 64 column-major, zero
// outside 0..nbb-1).  Block (b, k).
__global__ __launch_bounds__(256) void k_wy_T(int N, int nblk, const cx<double>* __restrict__ Y,
                                              const cx<double>* __restrict__ tau,
                                              cx<double>* __restrict__ Tw) {
    __shared__ cx<double> Ps[64][65];
    __shared__ cx<double> col[64];
    const int b = blockIdx.x;
    const size_t kk = blockIdx.y;
    const int jb = kBwy * b, nbb = min(kBwy, N - 1 - jb);
    const int r0 = jb + 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, g4 = lane >> 4;
    const int m0 = (wave & 1) * 32, n0 = (wave >> 1) * 32;
    const cx<double>* Yb = Y + kk * (size_t)N * N + (size_t)jb * N + r0;
    lqm_d4 cr[2][2], ci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            cr[a][c] = lqm_d4{0.0, 0.0, 0.0, 0.0};
            ci[a][c] = lqm_d4{0.0, 0.0, 0.0, 0.0};
        }
    gemm_acc<1, 0>(Yb, N, Yb, N, nullptr, nbb, nbb, N - r0, m0, n0, li, g4, cr, ci);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = m0 + 16 * a + g4 + 4 * i, n = n0 + 16 * c + li;
                Ps[n][m] = cx<double>{cr[a][c][i], ci[a][c][i]};  // Ps[col][row]
            }
    __syncthreads();
    const cx<double>* tk = tau + kk * N + jb;
    // T overwrites P column by column (column t of P copied out first)
    for (int t = 0; t < nbb; ++t) {
        if (threadIdx.x < t) col[threadIdx.x] = Ps[t][threadIdx.x];
        __syncthreads();
        const cx<double> tt = tk[t];
        if (threadIdx.x < 64) {
            const int s = threadIdx.x;
            cx<double> v = {0.0, 0.0};
            if (s < t) {
                cx<double> acc = {0.0, 0.0};
                for (int u = s; u < t; ++u) acc = cmad(acc, Ps[u][s], col[u]);
                v = cx<double>{0.0, 0.0} - tt * acc;
            } else if (s == t) {
                v = tt;
            }
            Ps[t][s] = v;
        }
        __syncthreads();
    }
    cx<double>* To = Tw + (kk * nblk + b) * (size_t)(kBwy * kBwy);
    for (int e = threadIdx.x; e < kBwy * kBwy; e += 256) {
        const int c = e >> 6, r = e & 63;
        To[e] = (c < nbb && r <= c) ? Ps[c][r] : cx<double>{0.0, 0.0};
    }
}

// V[k](:, n) = Z[k](:, n) as complex (Zt row-major with row stride ldz), n < ncol.
__global__ __launch_bounds__(256) void k_z_to_v(int N, int ncol, int ldz, const double* __restrict__ Zt,
                                                cx<double>* __restrict__ V) {
    __shared__ double tile[32][33];
    const size_t kk = blockIdx.z;
    const int i0 = blockIdx.x * 32, n0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const double* Zk = Zt + kk * (size_t)N * ldz;
    cx<double>* Vk = V + kk * (size_t)N * ncol;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = i0 + ty + 8 * q, n = n0 + tx;
        tile[ty + 8 * q][tx] = (i < N && n < ncol) ? Zk[(size_t)i * ldz + n] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int n = n0 + ty + 8 * q, i = i0 + tx;
        if (i < N && n < ncol) Vk[(size_t)n * N + i] = {tile[tx][ty + 8 * q], 0.0};
    }
}

// f[k][n] = occ[k][n] (given occupations of the ascending eigenvalues) or
// fermidirac(lambda - mu; T), and eb[k] = w_k sum_n lambda_n f_n.  One block per k.
__global__ __launch_bounds__(256) void k_occ_eb(int N, const double* __restrict__ lam, const double* __restrict__ kw,
                                                double wuni, double mu, double T,
                                                const double* __restrict__ occ, double* __restrict__ f,
                                                double* __restrict__ eb) {
    __shared__ double red[4];
    const size_t kk = blockIdx.x;
    double s = 0.0;
    for (int n = threadIdx.x; n < N; n += 256) {
        const double l = lam[kk * N + n];
        double fn;
        if (occ) {
            fn = occ[kk * N + n];
        } else {
            fn = fermidirac(l, mu, T);
        }
        f[kk * N + n] = fn;
        s += l * fn;
    }
    s = blk_sum<4>(s, red);
    if (threadIdx.x == 0) eb[kk] = (kw ? kw[kk] : wuni) * s;
}

}  // namespace lqm
