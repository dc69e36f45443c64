// FP64 refinement of FP32 eigenpairs (Ogita & Aishima, "Iterative refinement for
// symmetric eigenvalue decomposition", JJIAM 35 (2018), Algorithm 1, in its Hermitian
// form) and the FP64 pieces around it:
//
//   k_zgemm_batched  C = alpha op(A) op(B) + beta C per k-point (N <= 128 tiles)
//   k_oa_correct     lambda_i = s_ii / g_ii, delta = 2(||S - D|| + ||A|| ||R||),
//                    E_ij = (s_ij + lambda_j r_ij)/(lambda_j - lambda_i) or r_ij / 2
//                    with R = I - G, G = X^H X, S = X^H A X
//   k_scale_occ      Y = X diag(f(lambda)), f = fermidirac(lambda - mu; T) in FP64
//                    (or given occupations),
//                    and eb[k] = w_k sum_n lambda_n f_n
//   k_sort_evals     ascending copy of the refined eigenvalues
#pragma once
#include "lqm_common.hpp"

namespace lqm {

// Batched C = alpha op(A) op(B) + beta D, square N x N column-major FP64 complex,
// batch stride N*N (D may be C, or null when beta = 0).  op 0 = as is, 1 = conjugate
// transpose.  The z-th batch entry is k-point kidx[z] (z itself when kidx is null).
// 64 x 64 output tile per block, K-steps of 16 staged in LDS, 4 x 4 complex register
// block per thread (rows ti + 16u, columns tj + 16v).  grid = (tiles, tiles, batch).
constexpr int kGT = 64, kGK = 16;

template <int OPA, int OPB>
__global__ __launch_bounds__(256) void k_zgemm_batched(int N, const cx<double>* __restrict__ A,
                                                       const cx<double>* __restrict__ B,
                                                       cx<double>* C, const cx<double>* D,
                                                       cx<double> alpha, cx<double> beta,
                                                       const int* __restrict__ kidx) {
    __shared__ cx<double> As[kGK][kGT + 1];  // As[k][i] = op(A)(i0 + i, k0 + k)
    __shared__ cx<double> Bs[kGK][kGT + 1];  // Bs[k][j] = op(B)(k0 + k, j0 + j)
    const int tid = threadIdx.x;
    const int ti = tid & 15, tj = tid >> 4;
    const int i0 = blockIdx.x * kGT, j0 = blockIdx.y * kGT;
    const size_t NN = (size_t)N * N;
    const size_t kk = kidx ? (size_t)kidx[blockIdx.z] : (size_t)blockIdx.z;
    const cx<double>* Ak = A + kk * NN;
    const cx<double>* Bk = B + kk * NN;
    cx<double>* Ck = C + kk * NN;
    const cx<double>* Dk = D ? D + kk * NN : nullptr;
    cx<double> acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = {0.0, 0.0};
    for (int k0 = 0; k0 < N; k0 += kGK) {
        // 64 x 16 tiles: 1024 elements, 4 per thread
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + 256 * q;
            // A tile element (i, k): i fastest when op = N (coalesced column reads)
            int i, kx;
            if (OPA == 0) { i = e & 63; kx = e >> 6; } else { kx = e & 15; i = e >> 4; }
            const int gi = i0 + i, gk = k0 + kx;
            cx<double> a = {0.0, 0.0};
            if (gi < N && gk < N) {
                if (OPA == 0) a = Ak[(size_t)gk * N + gi];
                else { a = Ak[(size_t)gi * N + gk]; a.im = -a.im; }
            }
            As[kx][i] = a;
            int j, ky;
            if (OPB == 0) { ky = e & 15; j = e >> 4; } else { j = e & 63; ky = e >> 6; }
            const int gj = j0 + j, gk2 = k0 + ky;
            cx<double> b = {0.0, 0.0};
            if (gj < N && gk2 < N) {
                if (OPB == 0) b = Bk[(size_t)gj * N + gk2];
                else { b = Bk[(size_t)gk2 * N + gj]; b.im = -b.im; }
            }
            Bs[ky][j] = b;
        }
        __syncthreads();
#pragma unroll 4
        for (int kx = 0; kx < kGK; ++kx) {
            cx<double> a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = As[kx][ti + 16 * u];
#pragma unroll
            for (int v = 0; v < 4; ++v) b[v] = Bs[kx][tj + 16 * v];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] = cmad(acc[u][v], a[u], b[v]);
        }
        __syncthreads();
    }
    const bool zero_beta = (beta.re == 0.0 && beta.im == 0.0) || !Dk;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int gi = i0 + ti + 16 * u, gj = j0 + tj + 16 * v;
            if (gi < N && gj < N) {
                const size_t off = (size_t)gj * N + gi;
                cx<double> r = alpha * acc[u][v];
                if (!zero_beta) r = cmad(r, beta, Dk[off]);
                Ck[off] = r;
            }
        }
}

// The same batched C = alpha op(A) op(B) + beta D on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64).  A 256-thread block computes a 64 x 64 tile of C, each
// wave a 32 x 32 quadrant as 2 x 2 MFMA tiles of 16 x 16; a complex product is four
// real products, C_re += A_re B_re - A_im B_im, C_im += A_re B_im + A_im B_re, issued
// so that consecutive instructions never share an accumulator.
// Instruction layouts (16 x 16 x 4, 64 lanes): lane l supplies A[l%16][k] and
// B[k][l%16] and holds D[l/16 + 4i][l%16], i = 0..3.  The contraction index is
// permuted so that a lane's four consecutive instructions use four consecutive k
// (k = k0 + 4 (l/16) + s, s = 0..3): the operand walking k is then read as one
// contiguous 64-byte run straight from global memory (L1/L2-resident).
typedef double lqm_d4 __attribute__((ext_vector_type(4)));

template <int OP>
__device__ __forceinline__ cx<double> op_at(const cx<double>* __restrict__ M, int N, int r, int c) {
    // element (r, c) of op(M): M itself (column-major) or its conjugate transpose
    if (OP == 0) return M[(size_t)c * N + r];
    const cx<double> x = M[(size_t)r * N + c];
    return {x.re, -x.im};
}

template <int OPA, int OPB>
__global__ __launch_bounds__(256) void k_zgemm_mfma(int N, const cx<double>* __restrict__ A,
                                                    const cx<double>* __restrict__ B, cx<double>* C,
                                                    const cx<double>* D, cx<double> alpha,
                                                    cx<double> beta, const int* __restrict__ kidx) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int m0 = blockIdx.x * 64 + (wave & 1) * 32;
    const int n0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const size_t NN = (size_t)N * N;
    const size_t kk = kidx ? (size_t)kidx[blockIdx.z] : (size_t)blockIdx.z;
    const cx<double>* Ak = A + kk * NN;
    const cx<double>* Bk = B + kk * NN;
    lqm_d4 cr[2][2], ci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            cr[a][b] = lqm_d4{0.0, 0.0, 0.0, 0.0};
            ci[a][b] = lqm_d4{0.0, 0.0, 0.0, 0.0};
        }
    for (int k0 = 0; k0 < N; k0 += 16) {
        const int kb = k0 + 4 * g;
        double ar[2][4], ai[2][4], br[2][4], bi[2][4];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int m = m0 + 16 * t + li, n = n0 + 16 * t + li;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int k = kb + s;
                cx<double> x = {0.0, 0.0}, y = {0.0, 0.0};
                if (k < N) {
                    if (m < N) x = op_at<OPA>(Ak, N, m, k);
                    if (n < N) y = op_at<OPB>(Bk, N, k, n);
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
    cx<double>* Ck = C + kk * NN;
    const cx<double>* Dk = D ? D + kk * NN : nullptr;
    const bool zero_beta = (beta.re == 0.0 && beta.im == 0.0) || !Dk;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = m0 + 16 * a + g + 4 * i, n = n0 + 16 * b + li;
                if (m < N && n < N) {
                    const size_t off = (size_t)n * N + m;
                    cx<double> r = alpha * cx<double>{cr[a][b][i], ci[a][b][i]};
                    if (!zero_beta) r = cmad(r, beta, Dk[off]);
                    Ck[off] = r;
                }
            }
}

// In-place cyclic Jacobi diagonalization of a small Hermitian matrix K (c x c,
// row-major in registers of one thread) accumulating the unitary U (K <- U^H K U).
// Each rotation first makes the (p, q) entry real with a phase, then applies the
// real symmetric Jacobi rotation t = sign(tau)/(|tau| + sqrt(1 + tau^2)).
constexpr int kClusterMax = 8;

// Cluster slots of k_oa_correct: up to 64 with their rotations in LDS, or (large N, where a
// symmetric spectrum can hold hundreds of degenerate pairs) N / 2 + 1 with the rotations in a
// global array of (N / 2 + 1) * kClusterMax^2 complex per k-point; and its dynamic LDS size.
__host__ __device__ inline int oa_max_clusters(int N) { return N / 2 + 1 < 64 ? N / 2 + 1 : 64; }
__host__ __device__ inline int oa_max_clusters_global(int N) { return N / 2 + 1; }
// Chains of eigenvalues (consecutive gaps <= delta) wider than kClusterMax are split by their
// first-order couplings (at most kWideMax such chains per k-point, each at most kWideLen long):
// a boundary stays where every pair across it has |E_pq| <= kSplitCoupling.
constexpr int kWideMax = 64;
constexpr int kWideLen = 1024;
constexpr double kSplitCoupling = 1.0 / 64.0;
constexpr int kOaT = 512;            // threads of k_oa_correct
constexpr int kOaW = kOaT / 64;      // its waves
inline size_t oa_lds_bytes(int N, bool global_u = false) {
    const int maxcl = global_u ? oa_max_clusters_global(N) : oa_max_clusters(N);
    const size_t o_red = (32 * (size_t)N + 8 * (size_t)maxcl + 8 * (size_t)kWideMax + 15) & ~(size_t)15;
    return o_red + 4 * kOaW * 8 + (global_u ? 0 : (size_t)maxcl * kClusterMax * kClusterMax * 16);
}

__device__ void jacobi_herm(int c, cx<double>* K, cx<double>* U) {
    for (int i = 0; i < c; ++i)
        for (int j = 0; j < c; ++j) U[i * kClusterMax + j] = {i == j ? 1.0 : 0.0, 0.0};
    for (int sweep = 0; sweep < 12; ++sweep) {
        double off = 0.0, dia = 0.0;
        for (int i = 0; i < c; ++i)
            for (int j = 0; j < c; ++j) {
                const double a = abs2(K[i * kClusterMax + j]);
                if (i == j) dia += a; else off += a;
            }
        if (off <= 1e-34 * dia || off == 0.0) break;
        for (int p = 0; p < c - 1; ++p)
            for (int q = p + 1; q < c; ++q) {
                const cx<double> kpq = K[p * kClusterMax + q];
                const double b = sqrt(abs2(kpq));
                if (b == 0.0) continue;
                const cx<double> w = {kpq.re / b, kpq.im / b};  // e^{i phi}
                const double a = K[p * kClusterMax + p].re, d = K[q * kClusterMax + q].re;
                const double tau = (d - a) / (2.0 * b);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                const double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
                // J = [[cs, sn], [-sn conj(w), cs conj(w)]] on columns (p, q)
                const cx<double> wc = conj(w);
                for (int r = 0; r < c; ++r) {  // columns: K <- K J, U <- U J
                    const cx<double> kp = K[r * kClusterMax + p], kq = K[r * kClusterMax + q];
                    K[r * kClusterMax + p] = cs * kp - sn * (wc * kq);
                    K[r * kClusterMax + q] = sn * kp + cs * (wc * kq);
                    const cx<double> up = U[r * kClusterMax + p], uq = U[r * kClusterMax + q];
                    U[r * kClusterMax + p] = cs * up - sn * (wc * uq);
                    U[r * kClusterMax + q] = sn * up + cs * (wc * uq);
                }
                for (int col = 0; col < c; ++col) {  // rows: K <- J^H K
                    const cx<double> kp = K[p * kClusterMax + col], kq = K[q * kClusterMax + col];
                    K[p * kClusterMax + col] = cs * kp - sn * (w * kq);
                    K[q * kClusterMax + col] = sn * kp + cs * (w * kq);
                }
                K[p * kClusterMax + q] = {0.0, 0.0};
                K[q * kClusterMax + p] = {0.0, 0.0};
            }
    }
}

// Rayleigh-Ritz on the cluster of columns i .. i + c - 1 (c <= kClusterMax): K = B^H S_CC B with
// B = I + R_CC / 2 diagonalised by Jacobi, U <- B U (the new coefficients of the cluster columns).
// Writes lnew, cid, cpos of the members and U.
__device__ void oa_rayleigh_ritz(int N, int i, int c, int ncl, const cx<double>* Sk, const cx<double>* Gk,
                                   double* lnew, int* cid, int* cpos, cx<double>* U) {
    cx<double> K[kClusterMax * kClusterMax], B[kClusterMax * kClusterMax];
    for (int a = 0; a < c; ++a)
        for (int b = 0; b < c; ++b) {
            const cx<double> g = Gk[(size_t)(i + b) * N + (i + a)];
            const double rre = (a == b ? 1.0 : 0.0) - g.re, rim = -g.im;
            B[a * kClusterMax + b] = {(a == b ? 1.0 : 0.0) + 0.5 * rre, 0.5 * rim};
        }
    for (int a = 0; a < c; ++a)
        for (int b = 0; b < c; ++b) {
            cx<double> acc = {0.0, 0.0};
            for (int x = 0; x < c; ++x)
                for (int y = 0; y < c; ++y) {
                    const cx<double> sxy = Sk[(size_t)(i + y) * N + (i + x)];
                    acc = cmad(acc, conj(B[x * kClusterMax + a]) * sxy, B[y * kClusterMax + b]);
                }
            K[a * kClusterMax + b] = acc;
        }
    jacobi_herm(c, K, U);
    for (int a = 0; a < c; ++a) {
        lnew[i + a] = K[a * kClusterMax + a].re;
        cid[i + a] = ncl;
        cpos[i + a] = a;
    }
    cx<double> BU[kClusterMax * kClusterMax];
    for (int a = 0; a < c; ++a)
        for (int b = 0; b < c; ++b) {
            cx<double> acc = {0.0, 0.0};
            for (int x = 0; x < c; ++x) acc = cmad(acc, B[a * kClusterMax + x], U[x * kClusterMax + b]);
            BU[a * kClusterMax + b] = acc;
        }
    for (int q = 0; q < c * kClusterMax; ++q) U[q] = BU[q];
}

// Near-degenerate pairs (gap <= kNearGap max|lambda|) go through the general correction; there a
// pair whose first-order coupling exceeds kCoupRR is rotated by Rayleigh-Ritz with its cluster.
constexpr double kNearGap = 1e-5;
constexpr double kCoupRR = 1e-5;

// First-order coupling of columns p and q: max(|E_pq|, |E_qp|) with E_pq = (s_pq - l_q g_pq) /
// (l_q - l_p) (off-diagonal r = -g); infinity when the Rayleigh quotients coincide.
__device__ __forceinline__ double oa_coupling(int N, int p, int q, const double* lam, const cx<double>* Sk,
                                              const cx<double>* Gk) {
    const double gap = lam[q] - lam[p];
    if (gap == 0.0) return INFINITY;
    const cx<double> spq = Sk[(size_t)q * N + p], gpq = Gk[(size_t)q * N + p];
    const cx<double> sqp = Sk[(size_t)p * N + q], gqp = Gk[(size_t)p * N + q];
    const double a = abs2(cx<double>{spq.re - lam[q] * gpq.re, spq.im - lam[q] * gpq.im});
    const double b = abs2(cx<double>{sqp.re - lam[p] * gqp.re, sqp.im - lam[p] * gqp.im});
    return sqrt(fmax(a, b)) / fabs(gap);
}

// Ogita-Aishima correction with Rayleigh-Ritz on near-degenerate clusters.
// In: S = X^H A X, G = X^H X, read from their lower triangles (each replaced by the Hermitian
// matrix of its lower triangle).  Out: E[kk]
// (X <- X + X E), lambda[kk*N + i], and
// stat[kk*4 + {0: ||R||_F, 1: delta, 2: flags, 3: max |E_ij| outside the clusters}] where
// flags bit 0 = some cluster was rotated, bit 1 = a cluster wider than kClusterMax
// that is not degenerate to working precision (left unresolved), bit 2 = another sweep
// is needed: bit 0 or 1, or the first-order correction is so large that its square,
// the error left after this sweep, exceeds thr (thr <= 0: always).
// Block z works on k-point kidx[z] (z when kidx is null).
// Clusters are maximal runs of consecutive columns whose lambda gaps are <= delta
// (columns arrive in ascending FP32 eigenvalue order).  For a cluster C the
// orthonormalised block K = (I + R/2)^H S_CC (I + R/2) is diagonalised, K = U L U^H,
// E_CC = (I + R_CC/2) U - I, lambda_C = L, and E_iC (i outside C) = E^std_iC U.
__global__ __launch_bounds__(kOaT) void k_oa_correct(int N, cx<double>* __restrict__ S,
                                                    cx<double>* __restrict__ G,
                                                    cx<double>* __restrict__ E,
                                                    double* __restrict__ lambda,
                                                    double* __restrict__ stat, double thr,
                                                    const int* __restrict__ kidx,
                                                    cx<double>* __restrict__ Ug = nullptr) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int maxcl = Ug ? oa_max_clusters_global(N) : oa_max_clusters(N);
    const size_t o_lnew = 8 * (size_t)N, o_cid = o_lnew + 8 * (size_t)N, o_cpos = o_cid + 4 * (size_t)N;
    const size_t o_grp = o_cpos + 4 * (size_t)N, o_reach = o_grp + 4 * (size_t)N;
    const size_t o_cst = o_reach + 4 * (size_t)N, o_csz = o_cst + 4 * (size_t)maxcl;
    const size_t o_wide = o_csz + 4 * (size_t)maxcl;
    const size_t o_red = (o_wide + 8 * (size_t)kWideMax + 15) & ~(size_t)15;
    const size_t o_u = o_red + 4 * kOaW * 8;
    double* lam = reinterpret_cast<double*>(smem);               // [N]  Rayleigh quotients
    double* lnew = reinterpret_cast<double*>(smem + o_lnew);     // [N]  final lambda
    int* cid = reinterpret_cast<int*>(smem + o_cid);             // [N]  cluster ordinal or -1
    int* cpos = reinterpret_cast<int*>(smem + o_cpos);           // [N]  position in the cluster
    int* cstart = reinterpret_cast<int*>(smem + o_cst);          // [maxcl]
    int* csize = reinterpret_cast<int*>(smem + o_csz);           // [maxcl]
    int* grp = reinterpret_cast<int*>(smem + o_grp);             // [N]  sub-cluster of a split chain, or -1
    int* reach = reinterpret_cast<int*>(smem + o_reach);         // [N]  last column coupled to this one
    int* wide = reinterpret_cast<int*>(smem + o_wide);           // [2 kWideMax] wide chains (start, end)
    double* red = reinterpret_cast<double*>(smem + o_red);       // [4 kOaW]
    const int kk = kidx ? kidx[blockIdx.x] : (int)blockIdx.x;
    cx<double>* Ubuf = Ug ? Ug + (size_t)kk * maxcl * kClusterMax * kClusterMax
                          : reinterpret_cast<cx<double>*>(smem + o_u);  // [maxcl][kClusterMax^2]
    __shared__ int s_ncl, s_flags, s_nwide;
    const int tid = threadIdx.x;
    const size_t NN = (size_t)N * N;
    cx<double>* Sk = S + (size_t)kk * NN;
    cx<double>* Gk = G + (size_t)kk * NN;
    cx<double>* Ek = E + (size_t)kk * NN;
    // S and G <- the Hermitian matrices of their lower triangles, in place (the products may
    // form only the tiles on or below the diagonal): the first-order corrections then satisfy
    // E_ij + conj(E_ji) = r_ij exactly, whatever the rounding of X^H A X and X^H X; otherwise a
    // near-degenerate pair turns that rounding (about eps ||A||) divided by its gap into a loss
    // of orthogonality.  The diagonal first (imaginary parts zeroed, the Rayleigh quotients),
    // then the strict lower triangle: wave w takes column c = 8g + w of column group g, its
    // lanes the rows (coalesced reads, four row chunks in flight), and writes the conjugates
    // to row c above the diagonal -- the eight waves fill the same 128-byte lines of those rows
    // together.  The norms ||R||_F and ||S - D||_F ride along (an off-diagonal pair counts
    // twice).
    double r2 = 0.0, s2all = 0.0, amax = 0.0;
    for (int i = tid; i < N; i += kOaT) {
        const size_t d = (size_t)i * N + i;
        const double sd = Sk[d].re, gd = Gk[d].re;
        Sk[d].im = 0.0;
        Gk[d].im = 0.0;
        lam[i] = sd / gd;
        lnew[i] = lam[i];
        cid[i] = -1;
        grp[i] = -1;
        r2 += (1.0 - gd) * (1.0 - gd);
        s2all += (sd - lam[i]) * (sd - lam[i]);
        amax = fmax(amax, fabs(lam[i]));
    }
    {
        static_assert(kOaW == 8, "one column per wave of a group of eight");
        const int w = tid >> 6, ln = tid & 63;
        for (int c = w; c < N - 1; c += kOaW) {
            for (int r0 = c + 1; r0 < N; r0 += 4 * 64) {
                cx<double> vs[4], vg[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const size_t e = (size_t)c * N + min(r0 + 64 * q + ln, N - 1);
                    vs[q] = Sk[e];
                    vg[q] = Gk[e];
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = r0 + 64 * q + ln;
                    if (r < N) {
                        r2 += 2.0 * abs2(vg[q]);
                        s2all += 2.0 * abs2(vs[q]);
                        const size_t eu = (size_t)r * N + c;  // element (c, r) above the diagonal
                        Sk[eu] = conj(vs[q]);
                        Gk[eu] = conj(vg[q]);
                    }
                }
            }
        }
    }
    r2 = wave_sum(r2);
    s2all = wave_sum(s2all);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) amax = fmax(amax, __shfl_xor(amax, o));
    if ((tid & 63) == 0) {
        red[tid / 64] = r2;
        red[kOaW + tid / 64] = s2all;
        red[2 * kOaW + tid / 64] = amax;
    }
    __syncthreads();
    double rsum = 0.0, ssum = 0.0, an = 0.0;
#pragma unroll
    for (int w = 0; w < kOaW; ++w) {
        rsum += red[w];
        ssum += red[kOaW + w];
        an = fmax(an, red[2 * kOaW + w]);
    }
    const double rn = sqrt(rsum);
    const double sn_all = sqrt(ssum);
    const double delta = 2.0 * (sn_all + an * rn);
    const double tiny = 64.0 * 2.220446049250313e-16 * (an > 0.0 ? an : 1.0);
    // cluster detection and Rayleigh-Ritz (one thread: clusters are rare and small); chains
    // wider than kClusterMax are listed for splitting
    if (tid == 0) {
        int ncl = 0, flags = 0, nw = 0;
        int i = 0;
        while (i < N) {
            int j = i;
            // a near-degenerate pair (gap <= kNearGap ||A||) whose first-order coupling is
            // above kCoupRR also joins: its first-order step would leave an error of the
            // coupling's square, which the next sweep then has to undo
            while (j + 1 < N && (fabs(lam[j + 1] - lam[j]) <= delta ||
                                 (fabs(lam[j + 1] - lam[j]) <= kNearGap * an &&
                                  oa_coupling(N, j, j + 1, lam, Sk, Gk) > kCoupRR)))
                ++j;
            const int c = j - i + 1;
            if (c >= 2) {
                double spread = 0.0;
                for (int q = i + 1; q <= j; ++q) spread = fmax(spread, fabs(lam[q] - lam[i]));
                if (c > kClusterMax && c <= kWideLen && nw < kWideMax) {
                    wide[2 * nw] = i;
                    wide[2 * nw + 1] = j;
                    ++nw;
                } else if (c > kClusterMax || ncl >= maxcl) {
                    if (spread > tiny) flags |= 2;
                } else {
                    cstart[ncl] = i;  // Rayleigh-Ritz below, one thread per cluster
                    csize[ncl] = c;
                    ++ncl;
                    if (spread > tiny) flags |= 1;
                }
            }
            i = j + 1;
        }
        s_ncl = ncl;
        s_flags = flags;
        s_nwide = nw;
    }
    __syncthreads();
    if (s_nwide > 0) {
        // reach[p]: the last column q of p's chain with a first-order coupling above
        // kSplitCoupling (p itself when none), one thread per column
        for (int w = 0; w < s_nwide; ++w) {
            const int i0 = wide[2 * w], j0 = wide[2 * w + 1];
            for (int p = i0 + tid; p <= j0; p += kOaT) {
                int rq = p;
                for (int q = j0; q > p; --q)
                    if (!(oa_coupling(N, p, q, lam, Sk, Gk) <= kSplitCoupling)) {
                        rq = q;
                        break;
                    }
                reach[p] = rq;
            }
        }
        __syncthreads();
        // sub-clusters: a boundary after p where no column from the sub-cluster's start to p
        // reaches beyond p; each gets its own group (pairs across groups take the first-order
        // correction), and groups of 2 .. kClusterMax columns the Rayleigh-Ritz step
        if (tid == 0) {
            int ncl = s_ncl, flags = s_flags, gid = 0;
            for (int w = 0; w < s_nwide; ++w) {
                const int i0 = wide[2 * w], j0 = wide[2 * w + 1];
                int b0 = i0, rm = i0;
                for (int p = i0; p <= j0; ++p) {
                    rm = max(rm, reach[p]);
                    if (rm > p) continue;
                    const int c = p - b0 + 1;
                    for (int q = b0; q <= p; ++q) grp[q] = gid;
                    ++gid;
                    if (c >= 2) {
                        double spread = 0.0;
                        for (int q = b0 + 1; q <= p; ++q) spread = fmax(spread, fabs(lam[q] - lam[b0]));
                        if (c > kClusterMax || ncl >= maxcl) {
                            if (spread > tiny) flags |= 2;
                        } else {
                            cstart[ncl] = b0;
                            csize[ncl] = c;
                            ++ncl;
                            if (spread > tiny) flags |= 1;
                        }
                    }
                    b0 = p + 1;
                    rm = p + 1;
                }
            }
            s_ncl = ncl;
            s_flags = flags;
        }
        __syncthreads();
    }
    // Rayleigh-Ritz of the clusters found above, one thread each (disjoint columns and slots)
    for (int cl = tid; cl < s_ncl; cl += kOaT)
        oa_rayleigh_ritz(N, cstart[cl], csize[cl], cl, Sk, Gk, lnew, cid, cpos,
                         Ubuf + (size_t)cl * kClusterMax * kClusterMax);
    __syncthreads();
    // pair (i, j) takes the first-order correction when its Rayleigh quotients are more than
    // delta apart or it lies in different sub-clusters of a split chain
    auto separated = [&](int a, int b) {
        return fabs(lam[b] - lam[a]) > delta || (grp[a] >= 0 && grp[b] >= 0 && grp[a] != grp[b]);
    };
    // E entries and the second-order (Rayleigh-Schroedinger) correction of every eigenvalue
    // from the pencil (S, G), one wave per column j, its lanes over the rows i (coalesced
    // reads of column j of S and G, four rows in flight):
    //   lambda_j += sum_{i separated from j} |s_ij - lambda_j g_ij|^2 / (lambda_j - lambda_i)
    // (for a rotated cluster column, its Rayleigh-Ritz value and rotated couplings), so a single
    // sweep leaves an eigenvalue error of third order in the FP32 error.
    double emax = 0.0;
    {
        const int cw = tid >> 6, ln = tid & 63;
        for (int j = cw; j < N; j += kOaW) {
            const int cj = cid[j];
            double corr = 0.0;
            cx<double>* Ej = Ek + (size_t)j * N;
            if (cj >= 0) {
                // column j of a rotated cluster (x~_a = sum_l x_{c0+l} BU_la): E_ij = sum_l Estd_il
                // BU_la, and the correction with the rotated couplings t_i = sum_l (s_{i,c0+l} -
                // l~_a g_{i,c0+l}) BU_la to the columns outside the cluster separated from it; row i
                // of the cluster's columns of S and G is loaded at once (the cluster's other columns
                // find it in L2)
                const int c0 = cstart[cj], cs = csize[cj], a = cpos[j];
                const cx<double>* BU = Ubuf + (size_t)cj * kClusterMax * kClusterMax;
                const double lj = lnew[j];
                for (int i = ln; i < N; i += 64) {
                    const int ci = cid[i];
                    cx<double> out;
                    if (ci == cj) {
                        const cx<double> bu = BU[cpos[i] * kClusterMax + a];
                        out = {bu.re - (i == j ? 1.0 : 0.0), bu.im};
                    } else {
                        cx<double> gv[kClusterMax], sv[kClusterMax];
#pragma unroll
                        for (int l = 0; l < kClusterMax; ++l)
                            if (l < cs) {
                                gv[l] = Gk[(size_t)(c0 + l) * N + i];
                                sv[l] = Sk[(size_t)(c0 + l) * N + i];
                            }
                        out = {0.0, 0.0};
                        cx<double> t = {0.0, 0.0};
#pragma unroll
                        for (int l = 0; l < kClusterMax; ++l)
                            if (l < cs) {
                                const int ll = c0 + l;
                                const double rre = -gv[l].re, rim = -gv[l].im;  // i != ll
                                const double gap = lam[ll] - lam[i];
                                const cx<double> est = separated(i, ll)
                                                           ? cx<double>{(sv[l].re + lam[ll] * rre) / gap,
                                                                        (sv[l].im + lam[ll] * rim) / gap}
                                                           : cx<double>{0.5 * rre, 0.5 * rim};
                                const cx<double> bl = BU[l * kClusterMax + a];
                                out = cmad(out, est, bl);
                                t = cmad(t, cx<double>{sv[l].re - lj * gv[l].re, sv[l].im - lj * gv[l].im}, bl);
                            }
                        if (separated(i, j)) corr += abs2(t) / (lj - lam[i]);
                        emax = fmax(emax, abs2(out));
                    }
                    Ej[i] = out;
                }
                corr = wave_sum(corr);
                if (ln == 0) lnew[j] = lj + corr;
                continue;
            }
            const double lj = lam[j];
            for (int i0 = ln; i0 < N; i0 += 4 * 64) {
                cx<double> gq[4], sq[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const size_t e = (size_t)j * N + min(i0 + 64 * q, N - 1);
                    gq[q] = Gk[e];
                    sq[q] = Sk[e];
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = i0 + 64 * q;
                    if (i >= N) continue;
                    const cx<double> g = gq[q], sv = sq[q];
                    const double rre = (i == j ? 1.0 : 0.0) - g.re, rim = -g.im;
                    const double gap = lj - lam[i];
                    cx<double> out;
                    if (i != j && separated(i, j)) {
                        out = {(sv.re + lj * rre) / gap, (sv.im + lj * rim) / gap};
                        const double tre = sv.re - lj * g.re, tim = sv.im - lj * g.im;
                        corr += (tre * tre + tim * tim) / gap;
                    } else {
                        out = {0.5 * rre, 0.5 * rim};
                    }
                    Ej[i] = out;
                    if (i != j) emax = fmax(emax, abs2(out));
                }
            }
            corr = wave_sum(corr);
            if (ln == 0) lnew[j] = lj + corr;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) emax = fmax(emax, __shfl_xor(emax, o));
    if ((tid & 63) == 0) red[3 * kOaW + tid / 64] = emax;
    __syncthreads();
    for (int i = tid; i < N; i += kOaT) lambda[(size_t)kk * N + i] = lnew[i];
    if (tid == 0) {
        double em2 = 0.0;
        for (int w = 0; w < kOaW; ++w) em2 = fmax(em2, red[3 * kOaW + w]);
        const double em = sqrt(em2);
        const double lead = fmax(em, rn);
        int flags = s_flags;
        if ((flags & 3) || thr <= 0.0 || lead * lead > thr) flags |= 4;
        stat[(size_t)kk * 4 + 0] = rn;
        stat[(size_t)kk * 4 + 1] = delta;
        stat[(size_t)kk * 4 + 2] = flags;
        stat[(size_t)kk * 4 + 3] = em;
    }
    (void)s_ncl;
}

// Y = X diag(f), f_n = fermidirac(lambda_n - mu; T) -- or, when occ is given, the
// occupation occ[kk*N + r] of the rank-r eigenvalue (lambda is in X's column order,
// ties ranked by column); eb[kk] = w_k sum_n lambda_n f_n.  Dynamic LDS: N doubles.
// Fermi-Dirac occupation of an eigenvalue (T = 0: step with 1/2 at lambda = mu).
// Occupations of magnitude at or below 2^-70 are dropped from rho_k = X f X^H, in chunks of
// 16 bands: each element of rho_k then moves by at most 64 * 2^-70 < 6e-20, far below the
// rounding of its FP64 sum (occupations given by the caller may be negative; NaN is kept).
// The reference forms U Diagonal(f) U' over every band (Operators/densitymatrix.jl:139-140).
constexpr double kOccNegligible = 0x1p-70;

__device__ __forceinline__ double fermi_occ(double l, double mu, double T) {
    const double x = l - mu;
    return T > 0.0 ? 1.0 / (exp(x / T) + 1.0) : (x < 0.0 ? 1.0 : (x > 0.0 ? 0.0 : 0.5));
}

__global__ __launch_bounds__(256) void k_scale_occ(int N, const cx<double>* __restrict__ X,
                                                   const double* __restrict__ lambda,
                                                   const double* __restrict__ kw, double wuni,
                                                   double mu, double T,
                                                   const double* __restrict__ occ,
                                                   cx<double>* __restrict__ Y,
                                                   double* __restrict__ eb,
                                                   const int* __restrict__ kidx) {
    extern __shared__ __align__(16) unsigned char smem[];
    double* sf = reinterpret_cast<double*>(smem);
    __shared__ double sred[4];
    const int kk = kidx ? kidx[blockIdx.x] : (int)blockIdx.x, tid = threadIdx.x;
    const double* lk = lambda + (size_t)kk * N;
    double e = 0.0;
    for (int n = tid; n < N; n += 256) {
        const double ln = lk[n];
        double f;
        if (occ) {
            int rank = 0;
            for (int j = 0; j < N; ++j) {
                const double lj = lk[j];
                rank += (lj < ln) || (lj == ln && j < n);
            }
            f = occ[(size_t)kk * N + min(rank, N - 1)];
        } else {
            f = fermi_occ(ln, mu, T);
        }
        sf[n] = f;
        e += ln * f;
    }
    e = wave_sum(e);
    if ((tid & 63) == 0) sred[tid / 64] = e;
    __syncthreads();
    if (tid == 0) eb[kk] = (kw ? kw[kk] : wuni) * (sred[0] + sred[1] + sred[2] + sred[3]);
    const size_t NN = (size_t)N * N;
    for (int q = tid; q < N * N; q += 256) {
        const int n = q / N;
        const double f = sf[n];
        const cx<double> x = X[(size_t)kk * NN + q];
        Y[(size_t)kk * NN + q] = {f * x.re, f * x.im};
    }
}

// Ascending copy of lambda per k-point (rank counting, O(N^2) per k-point).
__global__ __launch_bounds__(256) void k_sort_evals(int N, const double* __restrict__ lambda,
                                                    double* __restrict__ out) {
    const int kk = blockIdx.x;
    const double* l = lambda + (size_t)kk * N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const double li = l[i];
        int rank = 0;
        for (int j = 0; j < N; ++j) {
            const double lj = l[j];
            rank += (lj < li) || (lj == li && j < i);
        }
        out[(size_t)kk * N + min(rank, N - 1)] = li;
    }
}

}  // namespace lqm

namespace lqm {
// dst[kidx[z]] = src[kidx[z]] for n k-points of N x N complex blocks.
__global__ __launch_bounds__(256) void k_copy_sub(int NN, int n, const int* __restrict__ kidx,
                                                  const cx<double>* __restrict__ src,
                                                  cx<double>* __restrict__ dst) {
    const size_t z = blockIdx.y;
    const size_t kk = (size_t)kidx[z];
    for (int e = blockIdx.x * 256 + threadIdx.x; e < NN; e += gridDim.x * 256)
        dst[kk * NN + e] = src[kk * NN + e];
    (void)n;
}
}  // namespace lqm

namespace lqm {

// ---------------------------------------------------------------------------------
// One Ogita-Aishima sweep of a k-point in a single kernel (N <= 64): X and W = A X
// live in LDS (two padded 64 x 65 complex tiles, 130 KB), S = X^H W and G = X^H X in
// the matrix-core accumulators, and the corrected X is written once.  A k-point whose
// Rayleigh quotients hold a cluster (consecutive gap <= delta) is left to the general
// correction: its S and G are written to global memory and flags bit 3 is set.
// Out: X[kk] <- X + X E, lambda[kk*N + i] (Rayleigh quotient + second-order
// correction), stat[kk*4 + {||R||_F, delta, flags, max |E_ij|}], flags as k_oa_correct
// (bit 2: another sweep needed) plus bit 3 and bit 4 (rho_k formed: with rq.want, a
// k-point that needs no further sweep writes rho_k[kk] and eb[kk], not X).
// ---------------------------------------------------------------------------------
constexpr int kFusedLd = 65;
inline size_t oa_fused_lds_bytes() { return sizeof(cx<double>) * 2 * 64 * kFusedLd + 8 * (64 + 2 * 64 + 32 + 64); }

// The density-matrix request of a fused sweep: a k-point that needs no further sweep
// forms rho_k = X f(lambda) X^H and eb[kk] = w_k sum_n lambda_n f_n right away (flags
// bit 4) instead of writing X back; occupations as k_scale_occ.
struct OaRho {
    int want;
    const double* kw;
    double wuni, mu, T;
    const double* occ;
    cx<double>* rhok;
    double* eb;
    int upper;  // only the elements on and above the diagonal of rho_k are read (k_oa_panel)
};

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// X (N x N column-major, FP64 or FP32 widened exactly) into the zero-padded 64 x 64 LDS
// tile Xs (leading dimension kFusedLd): all 16 loads of a thread are issued before the first
// store (clamped addresses, so every load is in bounds and unconditional).
template <typename XT>
__device__ __forceinline__ void load_x_tile(const XT* __restrict__ X, int N, cx<double>* Xs, int tid) {
    const int r = tid & 63, c0 = tid >> 6;
    XT v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = X[(size_t)min(c0 + 4 * u, N - 1) * N + min(r, N - 1)];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int c = c0 + 4 * u;
        Xs[c * kFusedLd + r] = (r < N && c < N) ? cx<double>{(double)v[u].re, (double)v[u].im} : cx<double>{0.0, 0.0};
    }
}

// A cluster among the sorted Rayleigh quotients lam[0..N): some consecutive gap <= delta, or a
// near-degenerate pair (gap <= kNearGap max|lam|) that the first-order correction would
// take.  The general correction handles both: it takes the Hermitian part of S, which the
// fused sweeps do not form (a first-order E_ij divides the rounding of S, about eps ||A||,
// by the gap: at gaps below 1e-5 ||A|| the loss of orthogonality would exceed 1e-11 and
// stall the iteration).  Wave 0 tests the N - 1 gaps at once and thread 0 records the answer.
__device__ __forceinline__ void cluster_flag(const double* lam, int N, double delta, double an, int wave,
                                             int lane, int* flag) {
    if (wave == 0) {
        const double lim = fmax(delta, kNearGap * an);
        const bool close = lane + 1 < N && fabs(lam[min(lane + 1, 63)] - lam[lane]) <= lim;
        const unsigned long long m = __ballot(close);
        if (lane == 0) *flag = m != 0ull;
    }
}

// Combines three real products into a complex result, re = P1 - P2 and im = P3 - P1 - P2,
// and keeps the result there: the empty asm statements stop the compiler from sinking the
// combination past the next product, which would keep all three accumulator sets live.
__device__ __forceinline__ void combine3(const lqm_d4 (&p1)[4], const lqm_d4 (&p2)[4], const lqm_d4 (&p3)[4],
                                         lqm_d4 (&re)[4], lqm_d4 (&im)[4]) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        re[t] = p1[t] - p2[t];
        im[t] = p3[t] - p1[t] - p2[t];
        asm volatile("" : "+v"(re[t]), "+v"(im[t]));
    }
}

// acc[a][b][i] <-> element (m0 + 16a + g + 4i, n0 + 16b + li) of a 32 x 32 wave quadrant
template <int OPA>
__device__ __forceinline__ void fused_gemm(const cx<double>* Asrc, int ldA, bool global_a, int N,
                                           const cx<double>* Bs, int m0, int n0, int li, int g,
                                           lqm_d4 (&cr)[2][2], lqm_d4 (&ci)[2][2]) {
    // C += op(A) B with op(A)(m, k) = A(m, k) (OPA 0) or conj(A(k, m)) (OPA 1);
    // A column-major with leading dimension ldA (global: N, rows/cols >= N read as 0),
    // B = Bs column-major in LDS with leading dimension kFusedLd (zero padded to 64).
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            cr[a][b] = lqm_d4{0.0, 0.0, 0.0, 0.0};
            ci[a][b] = lqm_d4{0.0, 0.0, 0.0, 0.0};
        }
    const int kend = (N + 15) & ~15;  // the inner index runs over N (zero padded)
    for (int k0 = 0; k0 < kend; k0 += 16) {
        const int kb = k0 + 4 * g;
        double ar[2][4], ai[2][4], br[2][4], bi[2][4];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int m = m0 + 16 * t + li, n = n0 + 16 * t + li;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int k = kb + s;
                cx<double> x;
                if (OPA == 0) {
                    if (global_a) x = (m < N && k < N) ? Asrc[(size_t)k * ldA + m] : cx<double>{0.0, 0.0};
                    else x = Asrc[(size_t)k * ldA + m];
                } else {
                    x = Asrc[(size_t)m * ldA + k];
                    x.im = -x.im;
                }
                const cx<double> y = Bs[(size_t)n * kFusedLd + k];
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

__global__ __launch_bounds__(256) void k_oa_fused(int N, const cx<double>* __restrict__ Ag,
                                                  cx<double>* __restrict__ Xg,
                                                  cx<double>* __restrict__ Sg,
                                                  cx<double>* __restrict__ Gg,
                                                  double* __restrict__ lambda,
                                                  double* __restrict__ stat, double thr,
                                                  const int* __restrict__ kidx, OaRho rq) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<double>* Xs = reinterpret_cast<cx<double>*>(smem);  // [64][kFusedLd] column-major
    cx<double>* Ws = Xs + 64 * kFusedLd;                     // W = A X, later E
    double* lam = reinterpret_cast<double*>(Ws + 64 * kFusedLd);  // [64]
    double* corr = lam + 64;                                       // [2][64]
    double* red = corr + 128;                                      // [32]
    double* sf = red + 32;                                         // [64]
    __shared__ int s_cluster;
    __shared__ unsigned long long s_live;  // bands whose occupation enters rho_k
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int wr = wave & 1, wc = wave >> 1;
    const int m0 = 32 * wr, n0 = 32 * wc;
    const size_t kk = kidx ? (size_t)kidx[blockIdx.x] : (size_t)blockIdx.x;
    const size_t NN = (size_t)N * N;
    const cx<double>* A = Ag + kk * NN;
    cx<double>* X = Xg + kk * NN;

    // X into LDS, zero padded to 64 x 64
    load_x_tile(X, N, Xs, tid);
    if (tid == 0) s_cluster = 0;
    __syncthreads();

    lqm_d4 ar_[2][2], ai_[2][2];
    // W = A X
    fused_gemm<0>(A, N, true, N, Xs, m0, n0, li, g, ar_, ai_);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = m0 + 16 * a + g + 4 * i, n = n0 + 16 * b + li;
                Ws[n * kFusedLd + m] = cx<double>{ar_[a][b][i], ai_[a][b][i]};
            }
    __syncthreads();
    // S = X^H W, G = X^H X
    lqm_d4 sr[2][2], si[2][2], gr[2][2], gi[2][2];
    fused_gemm<1>(Xs, kFusedLd, false, N, Ws, m0, n0, li, g, sr, si);
    fused_gemm<1>(Xs, kFusedLd, false, N, Xs, m0, n0, li, g, gr, gi);

    // Rayleigh quotients from the diagonal
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = m0 + 16 * a + g + 4 * i, n = n0 + 16 * b + li;
                if (m == n && m < N) lam[m] = sr[a][b][i] / gr[a][b][i];
            }
    __syncthreads();  // lam ready; every wave is done reading Ws
    double r2 = 0.0, s2 = 0.0, amax = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = m0 + 16 * a + g + 4 * i, n = n0 + 16 * b + li;
                if (m < N && n < N) {
                    const double rre = (m == n ? 1.0 : 0.0) - gr[a][b][i], rim = -gi[a][b][i];
                    r2 += rre * rre + rim * rim;
                    const double sre = sr[a][b][i] - (m == n ? lam[m] : 0.0), sim = si[a][b][i];
                    s2 += sre * sre + sim * sim;
                    if (m == n) amax = fmax(amax, fabs(lam[m]));
                }
            }
    r2 = wave_sum(r2);
    s2 = wave_sum(s2);
    amax = wave_max(amax);
    if (lane == 0) {
        red[wave] = r2;
        red[4 + wave] = s2;
        red[8 + wave] = amax;
    }
    __syncthreads();
    const double rn = sqrt(red[0] + red[1] + red[2] + red[3]);
    const double an = fmax(fmax(red[8], red[9]), fmax(red[10], red[11]));
    const double delta = 2.0 * (sqrt(red[4] + red[5] + red[6] + red[7]) + an * rn);
    cluster_flag(lam, N, delta, an, wave, lane, &s_cluster);
    __syncthreads();
    if (s_cluster) {
        // hand S and G to the general correction (Rayleigh-Ritz on the clusters)
        cx<double>* Sk = Sg + kk * NN;
        cx<double>* Gk = Gg + kk * NN;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int m = m0 + 16 * a + g + 4 * i, n = n0 + 16 * b + li;
                    if (m < N && n < N) {
                        Sk[(size_t)n * N + m] = cx<double>{sr[a][b][i], si[a][b][i]};
                        Gk[(size_t)n * N + m] = cx<double>{gr[a][b][i], gi[a][b][i]};
                    }
                }
        if (tid == 0) {
            stat[kk * 4 + 0] = rn;
            stat[kk * 4 + 1] = delta;
            stat[kk * 4 + 2] = 8;
            stat[kk * 4 + 3] = 0.0;
        }
        return;
    }
    // E and the second-order eigenvalue correction (column sums over m)
    double emax = 0.0;
    double cpart[2] = {0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = m0 + 16 * a + g + 4 * i, n = n0 + 16 * b + li;
                cx<double> e = {0.0, 0.0};
                if (m < N && n < N) {
                    const double rre = (m == n ? 1.0 : 0.0) - gr[a][b][i], rim = -gi[a][b][i];
                    if (m == n) {
                        e = {0.5 * rre, 0.5 * rim};
                    } else {
                        const double ln = lam[n];
                        const double gap = ln - lam[m];
                        if (fabs(gap) > delta) {
                            const double tre = sr[a][b][i] + ln * rre, tim = si[a][b][i] + ln * rim;
                            const double rg = 1.0 / gap;
                            e = {tre * rg, tim * rg};
                            cpart[b] += (tre * tre + tim * tim) * rg;
                        } else {
                            e = {0.5 * rre, 0.5 * rim};
                        }
                        emax = fmax(emax, abs2(e));
                    }
                }
                Ws[n * kFusedLd + m] = e;
            }
    // column partial sums: the four lane rows of the wave share a column, then the two
    // waves of a column quadrant meet in LDS
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        double v = cpart[b];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        if (g == 0) corr[wr * 64 + n0 + 16 * b + li] = v;
    }
    emax = wave_max(emax);
    if (lane == 0) red[12 + wave] = emax;
    __syncthreads();
    const double em = sqrt(fmax(fmax(red[12], red[13]), fmax(red[14], red[15])));
    const double lead = fmax(em, rn);
    const bool more = thr <= 0.0 || lead * lead > thr;
    const bool form_rho = rq.want && !more;
    __syncthreads();  // every wave has read lam[] and corr[]
    for (int n = tid; n < 64; n += 256) {
        const double ln = n < N ? lam[n] + (corr[n] + corr[64 + n]) : 0.0;
        if (n < N) lambda[kk * N + n] = ln;
        sf[n] = ln;  // the refined eigenvalue, read back for the occupations
    }
    if (tid == 0) {
        stat[kk * 4 + 0] = rn;
        stat[kk * 4 + 1] = delta;
        stat[kk * 4 + 2] = (more ? 4 : 0) | (form_rho ? 16 : 0);
        stat[kk * 4 + 3] = em;
    }
    // X <- X + X E
    fused_gemm<0>(Xs, kFusedLd, false, N, Ws, m0, n0, li, g, ar_, ai_);
    if (!form_rho) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int m = m0 + 16 * a + g + 4 * i, n = n0 + 16 * b + li;
                    if (m < N && n < N) {
                        const cx<double> x = Xs[n * kFusedLd + m];
                        X[(size_t)n * N + m] = cx<double>{x.re + ar_[a][b][i], x.im + ai_[a][b][i]};
                    }
                }
        return;
    }
    // rho_k = X f X^H from the refined X, staged back into LDS (rows/columns >= N stay 0)
    __syncthreads();  // every wave is done reading the old X; sf[] holds lambda
    double fn = 0.0, ebp = 0.0;
    if (tid < N) {
        const double ln = sf[tid];
        if (rq.occ) {
            int rank = 0;
            for (int j = 0; j < N; ++j) {
                const double lj = sf[j];
                rank += (lj < ln) || (lj == ln && j < tid);
            }
            fn = rq.occ[kk * N + min(rank, N - 1)];
        } else {
            fn = fermi_occ(ln, rq.mu, rq.T);
        }
        ebp = ln * fn;
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = m0 + 16 * a + g + 4 * i, n = n0 + 16 * b + li;
                const cx<double> x = Xs[n * kFusedLd + m];
                Xs[n * kFusedLd + m] = cx<double>{x.re + ar_[a][b][i], x.im + ai_[a][b][i]};
            }
    ebp = wave_sum(ebp);
    if (lane == 0) red[16 + wave] = ebp;
    __syncthreads();  // the new X is in LDS; every lambda has been read from sf[]
    if (tid < 64) {
        sf[tid] = fn;
        const unsigned long long live = __ballot(!(fabs(fn) <= kOccNegligible));
        if (tid == 0) s_live = live;
    }
    if (tid == 0) rq.eb[kk] = (rq.kw ? rq.kw[kk] : rq.wuni) * (red[16] + red[17] + red[18] + red[19]);
    __syncthreads();
    const unsigned long long live = s_live;
    lqm_d4 cr[2][2], ci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            cr[a][b] = lqm_d4{0.0, 0.0, 0.0, 0.0};
            ci[a][b] = lqm_d4{0.0, 0.0, 0.0, 0.0};
        }
    // A(m, c) = X(m, c) f_c and B(c, n) = conj X(n, c); lane (li, g) supplies c = k0 + 4g + s
    for (int k0 = 0; k0 < 64; k0 += 16) {
        if (((live >> k0) & 0xFFFFull) == 0) continue;  // 16 bands of negligible occupation
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int c = k0 + 4 * g + s;
            const double f = sf[c];
            double xr[2], xi[2], yr[2], yi[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const cx<double> x = Xs[c * kFusedLd + m0 + 16 * t + li];
                const cx<double> y = Xs[c * kFusedLd + n0 + 16 * t + li];
                xr[t] = f * x.re;
                xi[t] = f * x.im;
                yr[t] = y.re;
                yi[t] = -y.im;
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[a], yr[b], cr[a][b], 0, 0, 0);
                    ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[a], yi[b], ci[a][b], 0, 0, 0);
                }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-xi[a], yi[b], cr[a][b], 0, 0, 0);
                    ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xi[a], yr[b], ci[a][b], 0, 0, 0);
                }
        }
    }
    cx<double>* Rk = rq.rhok + kk * NN;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = m0 + 16 * a + g + 4 * i, n = n0 + 16 * b + li;
                if (m < N && n < N) Rk[(size_t)n * N + m] = cx<double>{cr[a][b][i], ci[a][b][i]};
            }
}

// ---------------------------------------------------------------------------------
// Fused Ogita-Aishima sweep, column-panel form (N <= 64): the computation of k_oa_fused
// with wave w owning columns 16w..16w+15 of every product, all held in matrix-core
// accumulators -- W = A X, G = X^H X, S = X^H W (W straight from the accumulators: the
// MFMA's K index k = 16a + g + 4s is the accumulator layout's row g + 4s of tile a), E and
// the second-order correction (a column never leaves its wave), X E (E likewise) and,
// with rq.want on a finished k-point, rho_k = X f X^H (there a panel of bands with negligible
// occupation skips its X E, and rho_k skips such bands in groups of four).  Only X lives in
// LDS (68 KB), so two blocks share a CU.  Outputs and flags as k_oa_fused.
// ---------------------------------------------------------------------------------
inline size_t oa_panel_lds_bytes() { return sizeof(cx<double>) * 64 * kFusedLd + 8 * (64 + 32 + 64 + 64); }

// XT: the element type X is read as (Xin); the refined X, and for a k-point handed to the
// general correction its unrefined X, are written to Xg in FP64 (with XT = cx<double>, Xin and
// Xg are the same array).
template <typename XT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_oa_panel(
    int N, const cx<double>* __restrict__ Ag, const XT* Xin, cx<double>* Xg, cx<double>* __restrict__ Sg,
    cx<double>* __restrict__ Gg, double* __restrict__ lambda, double* __restrict__ stat, double thr,
    const int* __restrict__ kidx, OaRho rq) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<double>* Xs = reinterpret_cast<cx<double>*>(smem);  // [64][kFusedLd] column-major
    double* lam = reinterpret_cast<double*>(Xs + 64 * kFusedLd);  // [64] Rayleigh quotients
    double* red = lam + 64;                                        // [32]
    double* sf = red + 32;                                         // [64] occupations
    double* lnew = sf + 64;                                        // [64] refined eigenvalues
    __shared__ int s_cluster;
    __shared__ unsigned long long s_live;  // bands whose occupation enters rho_k
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int n = 16 * wave + li;  // this lane's column in every product
    const size_t kk = kidx ? (size_t)kidx[blockIdx.x] : (size_t)blockIdx.x;
    const size_t NN = (size_t)N * N;
    const cx<double>* A = Ag + kk * NN;
    cx<double>* X = Xg + kk * NN;
    load_x_tile(Xin + kk * NN, N, Xs, tid);
    if (tid == 0) s_cluster = 0;
    __syncthreads();

    // Complex products as three real ones (P1 = ar br, P2 = ai bi, P3 = (ar + ai)(br + bi);
    // re = P1 - P2, im = P3 - P1 - P2): three accumulator sets live only inside a product.
    lqm_d4 wr[4], wi[4], gr[4], gi[4], sr[4], si[4];
    {
        // W = A X over k = k0 + 4g + s
        lqm_d4 p1[4], p2[4], p3[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) p1[t] = p2[t] = p3[t] = lqm_d4{0.0, 0.0, 0.0, 0.0};
        const int kend = (N + 15) & ~15;  // rows k >= N of X are zero
        for (int k0 = 0; k0 < kend; k0 += 16) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int k = k0 + 4 * g + s;
                const cx<double> b = Xs[n * kFusedLd + k];
                const double bs = b.re + b.im;
                double ar[4], ai[4], as[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int m = 16 * t + li;
                    const cx<double> a = Ag[kk * NN + (size_t)min(k, N - 1) * N + min(m, N - 1)];
                    const bool in = m < N && k < N;
                    ar[t] = in ? a.re : 0.0;
                    ai[t] = in ? a.im : 0.0;
                    as[t] = ar[t] + ai[t];
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    p1[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[t], b.re, p1[t], 0, 0, 0);
                    p2[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[t], b.im, p2[t], 0, 0, 0);
                    p3[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[t], bs, p3[t], 0, 0, 0);
                }
            }
        }
        combine3(p1, p2, p3, wr, wi);
    }
    {
        // S = X^H W: A(m, k) = conj X(k, m) from LDS, B from W's accumulators (k = 16a + g + 4s)
        lqm_d4 p1[4], p2[4], p3[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) p1[t] = p2[t] = p3[t] = lqm_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int a = 0; a < 4; ++a) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int k = 16 * a + g + 4 * s;
                const double br = wr[a][s], bi = wi[a][s], bs = br + bi;
                double xr[4], xi[4], xs[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const cx<double> x = Xs[(16 * t + li) * kFusedLd + k];
                    xr[t] = x.re;
                    xi[t] = -x.im;
                    xs[t] = x.re - x.im;
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    p1[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[t], br, p1[t], 0, 0, 0);
                    p2[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(xi[t], bi, p2[t], 0, 0, 0);
                    p3[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(xs[t], bs, p3[t], 0, 0, 0);
                }
            }
            __builtin_amdgcn_sched_barrier(0);  // bounds the LDS loads hoisted ahead (registers)
        }
        combine3(p1, p2, p3, sr, si);
    }
    {
        // G = X^H X over k = k0 + 4g + s, both operands from LDS
        lqm_d4 p1[4], p2[4], p3[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) p1[t] = p2[t] = p3[t] = lqm_d4{0.0, 0.0, 0.0, 0.0};
        const int kend = (N + 15) & ~15;
        for (int k0 = 0; k0 < kend; k0 += 16) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int k = k0 + 4 * g + s;
                const cx<double> b = Xs[n * kFusedLd + k];
                const double bs = b.re + b.im;
                double xr[4], xi[4], xs[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const cx<double> x = Xs[(16 * t + li) * kFusedLd + k];  // conj(X(k, m)) = X^H(m, k)
                    xr[t] = x.re;
                    xi[t] = -x.im;
                    xs[t] = x.re - x.im;
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    p1[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[t], b.re, p1[t], 0, 0, 0);
                    p2[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(xi[t], b.im, p2[t], 0, 0, 0);
                    p3[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(xs[t], bs, p3[t], 0, 0, 0);
                }
            }
        }
        combine3(p1, p2, p3, gr, gi);
    }
    // Rayleigh quotients from the diagonal: (n, n) sits in tile `wave`, element i = li / 4 of
    // the lanes with g = li % 4
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = 16 * t + g + 4 * i;
            if (m == n && n < N) lam[n] = sr[t][i] / gr[t][i];
        }
    __syncthreads();
    double r2 = 0.0, s2 = 0.0, amax = 0.0;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = 16 * t + g + 4 * i;
            if (m < N && n < N) {
                const double rre = (m == n ? 1.0 : 0.0) - gr[t][i], rim = -gi[t][i];
                r2 += rre * rre + rim * rim;
                const double sre = sr[t][i] - (m == n ? lam[m] : 0.0), sim = si[t][i];
                s2 += sre * sre + sim * sim;
                if (m == n) amax = fmax(amax, fabs(lam[m]));
            }
        }
    r2 = wave_sum(r2);
    s2 = wave_sum(s2);
    amax = wave_max(amax);
    if (lane == 0) {
        red[wave] = r2;
        red[4 + wave] = s2;
        red[8 + wave] = amax;
    }
    __syncthreads();
    const double rn = sqrt(red[0] + red[1] + red[2] + red[3]);
    const double an = fmax(fmax(red[8], red[9]), fmax(red[10], red[11]));
    const double delta = 2.0 * (sqrt(red[4] + red[5] + red[6] + red[7]) + an * rn);
    cluster_flag(lam, N, delta, an, wave, lane, &s_cluster);
    __syncthreads();
    if (s_cluster) {
        cx<double>* Sk = Sg + kk * NN;
        cx<double>* Gk = Gg + kk * NN;
        if constexpr (!std::is_same<XT, cx<double>>::value) {
            // the general correction reads X from Xg
            for (int e = tid; e < N * N; e += 256) {
                const int c = e / N, r = e - c * N;
                X[e] = Xs[c * kFusedLd + r];
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = 16 * t + g + 4 * i;
                if (m < N && n < N) {
                    Sk[(size_t)n * N + m] = cx<double>{sr[t][i], si[t][i]};
                    Gk[(size_t)n * N + m] = cx<double>{gr[t][i], gi[t][i]};
                }
            }
        if (tid == 0) {
            stat[kk * 4 + 0] = rn;
            stat[kk * 4 + 1] = delta;
            stat[kk * 4 + 2] = 8;
            stat[kk * 4 + 3] = 0.0;
        }
        return;
    }
    // E (into S's accumulators) and the column sums of the second-order correction
    double emax = 0.0, cpart = 0.0;
    const double ln = n < N ? lam[n] : 0.0;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = 16 * t + g + 4 * i;
            double ere = 0.0, eim = 0.0;
            if (m < N && n < N) {
                const double rre = (m == n ? 1.0 : 0.0) - gr[t][i], rim = -gi[t][i];
                if (m == n) {
                    ere = 0.5 * rre;
                    eim = 0.5 * rim;
                } else {
                    const double gap = ln - lam[m];
                    if (fabs(gap) > delta) {
                        const double tre = sr[t][i] + ln * rre, tim = si[t][i] + ln * rim;
                        const double rg = fabs(gap) >= num<double>::rot_lo ? rcp_nr(gap) : 1.0 / gap;
                        ere = tre * rg;
                        eim = tim * rg;
                        cpart += (tre * tre + tim * tim) * rg;
                    } else {
                        ere = 0.5 * rre;
                        eim = 0.5 * rim;
                    }
                    emax = fmax(emax, ere * ere + eim * eim);
                }
            }
            sr[t][i] = ere;
            si[t][i] = eim;
        }
    cpart += __shfl_xor(cpart, 16);
    cpart += __shfl_xor(cpart, 32);
    emax = wave_max(emax);
    if (lane == 0) red[12 + wave] = emax;
    if (g == 0) lnew[n] = ln + cpart;  // rows n >= N hold ln = cpart = 0
    __syncthreads();
    const double em = sqrt(fmax(fmax(red[12], red[13]), fmax(red[14], red[15])));
    const double lead = fmax(em, rn);
    const bool more = thr <= 0.0 || lead * lead > thr;
    const bool form_rho = rq.want && !more;
    if (g == 0 && n < N) lambda[kk * N + n] = lnew[n];
    if (tid == 0) {
        stat[kk * 4 + 0] = rn;
        stat[kk * 4 + 1] = delta;
        stat[kk * 4 + 2] = (more ? 4 : 0) | (form_rho ? 16 : 0);
        stat[kk * 4 + 3] = em;
    }
    // On a finished k-point with Fermi-Dirac occupations, a panel whose bands all have negligible
    // occupation does not enter rho_k, so its X E is skipped: every lane holds the refined
    // eigenvalue ln + cpart of its column, the value lnew[n] the occupations below come from.
    bool skip_xe = false;
    if (form_rho && !rq.occ) {
        const double fo = n < N ? fermi_occ(ln + cpart, rq.mu, rq.T) : 0.0;
        skip_xe = __ballot(!(fabs(fo) <= kOccNegligible)) == 0ull;
    }
    // X E: B from E's accumulators (k = 16a + g + 4s), A = X from LDS (three real products)
    lqm_d4 xr_[4], xi_[4];
    if (!skip_xe) {
        lqm_d4 p1[4], p2[4], p3[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) p1[t] = p2[t] = p3[t] = lqm_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int a = 0; a < 4; ++a) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int k = 16 * a + g + 4 * s;
                const double br = sr[a][s], bi = si[a][s], bs = br + bi;
                double xr[4], xi[4], xs[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const cx<double> x = Xs[k * kFusedLd + 16 * t + li];
                    xr[t] = x.re;
                    xi[t] = x.im;
                    xs[t] = x.re + x.im;
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    p1[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[t], br, p1[t], 0, 0, 0);
                    p2[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(xi[t], bi, p2[t], 0, 0, 0);
                    p3[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(xs[t], bs, p3[t], 0, 0, 0);
                }
            }
            __builtin_amdgcn_sched_barrier(0);  // bounds the LDS loads hoisted ahead (registers)
        }
        combine3(p1, p2, p3, xr_, xi_);
    }
    // X + X E
    if (!skip_xe) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const cx<double> x = Xs[n * kFusedLd + 16 * t + g + 4 * i];
                xr_[t][i] += x.re;
                xi_[t][i] += x.im;
            }
    }
    if (!form_rho) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = 16 * t + g + 4 * i;
                if (m < N && n < N) X[(size_t)n * N + m] = cx<double>{xr_[t][i], xi_[t][i]};
            }
        return;
    }
    // rho_k = X f X^H from the refined X, staged back into LDS (a skipped panel's columns are
    // not read)
    __syncthreads();  // every wave is done reading the old X
    if (!skip_xe) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) Xs[n * kFusedLd + 16 * t + g + 4 * i] = cx<double>{xr_[t][i], xi_[t][i]};
    }
    double fn = 0.0, ebp = 0.0;
    if (tid < N) {
        const double l = lnew[tid];
        if (rq.occ) {
            int rank = 0;
            for (int j = 0; j < N; ++j) {
                const double lj = lnew[j];
                rank += (lj < l) || (lj == l && j < tid);
            }
            fn = rq.occ[kk * N + min(rank, N - 1)];
        } else {
            fn = fermi_occ(l, rq.mu, rq.T);
        }
        ebp = l * fn;
    }
    if (tid < 64) {
        sf[tid] = fn;
        const unsigned long long live = __ballot(!(fabs(fn) <= kOccNegligible));
        if (tid == 0) s_live = live;
    }
    ebp = wave_sum(ebp);
    if (lane == 0) red[16 + wave] = ebp;
    __syncthreads();
    if (tid == 0) rq.eb[kk] = (rq.kw ? rq.kw[kk] : rq.wuni) * (red[16] + red[17] + red[18] + red[19]);
    const unsigned long long live = s_live;
    // row tiles of rho_k this wave forms: all four, or with rq.upper those on and above the
    // diagonal (t <= wave; the lower half of the diagonal tile comes along)
    const int tmax = rq.upper ? __builtin_amdgcn_readfirstlane(wave) : 3;
    lqm_d4 cr[4], ci[4];
    {
        lqm_d4 p1[4], p2[4], p3[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) p1[t] = p2[t] = p3[t] = lqm_d4{0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < 64; k0 += 4) {
            if (((live >> k0) & 0xFull) == 0) continue;  // 4 bands of negligible occupation
            {
                const int c = k0 + g;
                const double f = sf[c];
                const cx<double> y = Xs[c * kFusedLd + n];  // B(c, n) = conj X(n, c)
                const double yr = y.re, yi = -y.im, ys = y.re - y.im;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (t <= tmax) {
                        const cx<double> x = Xs[c * kFusedLd + 16 * t + li];
                        const double xr = f * x.re, xi = f * x.im, xs = xr + xi;
                        p1[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr, yr, p1[t], 0, 0, 0);
                        p2[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(xi, yi, p2[t], 0, 0, 0);
                        p3[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(xs, ys, p3[t], 0, 0, 0);
                    }
                }
            }
        }
        combine3(p1, p2, p3, cr, ci);
    }
    cx<double>* Rk = rq.rhok + kk * NN;
#pragma unroll
    for (int t = 0; t < 4; ++t)
        if (t <= tmax)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = 16 * t + g + 4 * i;
                if (m < N && n < N) Rk[(size_t)n * N + m] = cx<double>{cr[t][i], ci[t][i]};
            }
}


// ---------------------------------------------------------------------------------
// Bookkeeping between refinement sweeps, on the device.  From the status rows
// (||R||_F, delta, flags, ||E||) of the k-points the sweep refined -- the list `cur`, or every
// k-point 0..ncur-1 when it is null -- it lists, in ascending order, the k-points handed to the
// cluster correction (flag 8) in l8 and those flagged for another sweep (flag 4) in l4, and over
// all Kc k-points those whose rho_k no sweep formed (flag 16 clear) in lrest; the header takes
// the three counts, the number of refined k-points with an unresolved cluster (flag 2), whether
// any refined k-point has a non-finite ||R||_F or ||E||, and the maxima of ||R||_F and delta.
// One block; each thread owns a contiguous run of entries, so the lists keep the input order.
// ---------------------------------------------------------------------------------
struct OaScanHdr {
    int n8, n4, nrest, unresolved, nonfinite, pad[3];
    double rmax, dmax;
};
constexpr int kScanT = 1024;

__global__ __launch_bounds__(kScanT) void k_oa_scan(int Kc, const double* __restrict__ stat,
                                                    const int* __restrict__ cur, int ncur,
                                                    OaScanHdr* __restrict__ hdr, int* __restrict__ l8,
                                                    int* __restrict__ l4, int* __restrict__ lrest) {
    __shared__ int wt[3][kScanT / 64];
    __shared__ int wu[kScanT / 64], wn[kScanT / 64];
    __shared__ double wr[kScanT / 64], wd[kScanT / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int per = (ncur + kScanT - 1) / kScanT, lo = min(ncur, tid * per), hi = min(ncur, lo + per);
    const int perA = (Kc + kScanT - 1) / kScanT, loA = min(Kc, tid * perA), hiA = min(Kc, loA + perA);
    int own[3] = {0, 0, 0}, unres = 0, nonfin = 0;
    double rmax = 0.0, dmax = 0.0;
    for (int i = lo; i < hi; ++i) {
        const int k = cur ? cur[i] : i;
        const double* r = stat + 4 * (size_t)k;
        const int fl = (int)r[2];
        own[0] += (fl & 8) != 0;
        own[1] += (fl & 4) != 0;
        unres += (fl & 2) != 0;
        nonfin |= !isfinite(r[0]) || !isfinite(r[3]);
        rmax = fmax(rmax, r[0]);
        dmax = fmax(dmax, r[1]);
    }
    for (int k = loA; k < hiA; ++k) own[2] += ((int)stat[4 * (size_t)k + 2] & 16) == 0;
    // exclusive block scan of the three counts: within the wave, then over the wave totals
    int inc[3] = {own[0], own[1], own[2]};
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int t = __shfl_up(inc[j], o);
            if (lane >= o) inc[j] += t;
        }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        unres += __shfl_xor(unres, o);
        nonfin |= __shfl_xor(nonfin, o);
        rmax = fmax(rmax, __shfl_xor(rmax, o));
        dmax = fmax(dmax, __shfl_xor(dmax, o));
    }
    if (lane == 63)
#pragma unroll
        for (int j = 0; j < 3; ++j) wt[j][wave] = inc[j];
    if (lane == 0) {
        wu[wave] = unres;
        wn[wave] = nonfin;
        wr[wave] = rmax;
        wd[wave] = dmax;
    }
    __syncthreads();
    int off[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        int b = 0;
        for (int w = 0; w < wave; ++w) b += wt[j][w];
        off[j] = b + inc[j] - own[j];
    }
    for (int i = lo; i < hi; ++i) {
        const int k = cur ? cur[i] : i;
        const int fl = (int)stat[4 * (size_t)k + 2];
        if (fl & 8) l8[off[0]++] = k;
        if (fl & 4) l4[off[1]++] = k;
    }
    for (int k = loA; k < hiA; ++k)
        if (!((int)stat[4 * (size_t)k + 2] & 16)) lrest[off[2]++] = k;
    if (tid == 0) {
        OaScanHdr h{};
        for (int w = 0; w < kScanT / 64; ++w) {
            h.n8 += wt[0][w];
            h.n4 += wt[1][w];
            h.nrest += wt[2][w];
            h.unresolved += wu[w];
            h.nonfinite |= wn[w];
            h.rmax = fmax(h.rmax, wr[w]);
            h.dmax = fmax(h.dmax, wd[w]);
        }
        *hdr = h;
    }
}

}  // namespace lqm
