// Partial spectrum of sparse Bloch Hamiltonians: the num_bands eigenpairs of H(k) closest to
// sigma (or the smallest / largest), for a batch of k-points, by Chebyshev-filtered subspace
// iteration.  The model is one CSR shared by every k-point: row i holds entries
// (column j, key r, value t) and H(k)_ij = sum over the row's entries with column j of
// t e^{2 pi i k.R_r}, so a k-point costs only its phase table.
//
// Block vectors of a k-point are stored row-major, N rows of mp columns (mp = 16, 32 or 64
// >= m, the columns of a row contiguous), so the lanes of a product walk the columns of
// one row coalesced.  Read as column-major, the same storage is the mp x N matrix Z = Y^T,
// which is the operand of the dense products (Gram matrices, rotations) on the matrix cores.
#pragma once

#include "lqm_common.hpp"

namespace lqm {

// ph[s][r] = e^{2 pi i k_s.R_r} for the Kc k-points of a chunk.
__global__ __launch_bounds__(256) void k_sp_phases(int Kc, int d, int nR, const int* __restrict__ R,
                                                   const double* __restrict__ ks, cx<double>* __restrict__ ph) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= Kc * nR) return;
    const int s = t / nR, r = t % nR;
    double x = 0.0;
    for (int a = 0; a < d; ++a) x += ks[(size_t)s * d + a] * (double)R[r * d + a];
    ph[t] = cis2pi(x);
}

// y = alpha ((H(k) - shift) x) + beta1 p + beta2 q for the active k-points
// (slot = kidx[blockIdx.y]); coefficients per slot in coef[slot] = {alpha, beta1, beta2}
// (null: alpha = 1, no p, q terms); p or q may be null; q may alias y (each element is read
// before it is written by the same thread).  LR lanes per row (LR = mp), 64 / LR rows per
// wave at a time, kSpRows rows per block.  Columns c >= m are written as zero.
constexpr int kSpRows = 32;
constexpr int kSpMaxKeys = 2048;
constexpr int kSpUnroll = 4;  // gathers in flight per lane (measured: 1.9x over 1; 8 no better)
__global__ __launch_bounds__(256) void k_spmm(int N, int mp, int m, const long long* __restrict__ rowptr,
                                              const int2* __restrict__ colkey, const cx<double>* __restrict__ val,
                                              int nR, const cx<double>* __restrict__ ph, double shift,
                                              const int* __restrict__ kidx, const double* __restrict__ coef,
                                              const cx<double>* __restrict__ X, const cx<double>* p,
                                              const cx<double>* q, cx<double>* y) {
    __shared__ cx<double> sph[kSpMaxKeys];
    const int slot = kidx[blockIdx.y];
    const size_t off = (size_t)slot * N * mp;
    for (int r = threadIdx.x; r < nR; r += 256) sph[r] = ph[(size_t)slot * nR + r];
    __syncthreads();
    double alpha = 1.0, b1 = 0.0, b2 = 0.0;
    if (coef) {
        alpha = coef[3 * slot];
        b1 = coef[3 * slot + 1];
        b2 = coef[3 * slot + 2];
    }
    const int sub = 64 / mp;                       // rows per wave at a time
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = lane % mp, rs = lane / mp;
    const cx<double>* x = X + off;
    for (int r0 = blockIdx.x * kSpRows + wv * sub; r0 < min(N, (blockIdx.x + 1) * kSpRows); r0 += 4 * sub) {
        const int i = r0 + rs;
        if (i >= N || i >= (blockIdx.x + 1) * kSpRows) continue;
        const size_t o = off + (size_t)i * mp + c;
        if (c >= m) {
            y[o] = {0.0, 0.0};
            continue;
        }
        cx<double> acc = {0.0, 0.0};
        const long long e1 = rowptr[i + 1];
        long long e = rowptr[i];
        // kSpUnroll entries at a time: their column/key loads, then all their gathers in flight
        for (; e + kSpUnroll <= e1; e += kSpUnroll) {
            int2 jk[kSpUnroll];
            cx<double> t[kSpUnroll], xv[kSpUnroll];
#pragma unroll
            for (int q = 0; q < kSpUnroll; ++q) jk[q] = colkey[e + q];
#pragma unroll
            for (int q = 0; q < kSpUnroll; ++q) xv[q] = x[(size_t)jk[q].x * mp + c];
#pragma unroll
            for (int q = 0; q < kSpUnroll; ++q) t[q] = cmad(cx<double>{0.0, 0.0}, val[e + q], sph[jk[q].y]);
#pragma unroll
            for (int q = 0; q < kSpUnroll; ++q) acc = cmad(acc, t[q], xv[q]);
        }
        for (; e < e1; ++e) {
            const int2 jk = colkey[e];
            const cx<double> t = cmad(cx<double>{0.0, 0.0}, val[e], sph[jk.y]);
            acc = cmad(acc, t, x[(size_t)jk.x * mp + c]);
        }
        const cx<double> xi = x[(size_t)i * mp + c];
        acc.re = fma(-shift, xi.re, acc.re);
        acc.im = fma(-shift, xi.im, acc.im);
        cx<double> out = {alpha * acc.re, alpha * acc.im};
        if (p) {
            const cx<double> v = p[o];
            out.re = fma(b1, v.re, out.re);
            out.im = fma(b1, v.im, out.im);
        }
        if (q) {
            const cx<double> v = q[o];
            out.re = fma(b2, v.re, out.re);
            out.im = fma(b2, v.im, out.im);
        }
        y[o] = out;
    }
}

// Deterministic start vectors: X[s][i][c] uniform in [-1/2, 1/2) + i [-1/2, 1/2) from a
// counter hash of (seed, global k-point, row, column), zero in the padding columns.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__global__ __launch_bounds__(256) void k_sp_start(int N, int mp, int m, int Kc, long long k0, unsigned long long seed,
                                                  cx<double>* __restrict__ X) {
    const size_t tot = (size_t)Kc * N * mp;
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < tot; t += (size_t)gridDim.x * 256) {
        const int c = (int)(t % mp);
        const size_t ri = t / mp;
        const long long i = (long long)(ri % N), s = (long long)(ri / N);
        if (c >= m) {
            X[t] = {0.0, 0.0};
            continue;
        }
        const unsigned long long h = mix64(seed ^ mix64((unsigned long long)(k0 + s) * 0x100000001b3ull ^
                                                        mix64((unsigned long long)i * 131u + (unsigned long long)c)));
        const double a = (double)(h >> 11) * 0x1.0p-53 - 0.5;
        const double b = (double)(mix64(h) >> 11) * 0x1.0p-53 - 0.5;
        X[t] = {a, b};
    }
}

// Inverse Cholesky factor of the Gram matrix G (m x m per slot, Hermitian positive definite,
// lower triangle read): G + s I = L L^H, Linv = L^-1 (lower, upper part zero).  First
// unshifted (s = 0); a pivot that is not positive retries with s = shift_rel * tr(G) (the
// shift of shifted CholeskyQR) and, if that fails too, gives Linv = I and flag[slot] = 1.
// One block per active slot; dynamic LDS: the factor, m rows of 65 (m <= 64).
__global__ __launch_bounds__(256) void k_chol_inv(int m, const int* __restrict__ kidx, const cx<double>* __restrict__ G,
                                                  double shift_rel, int force_shift, cx<double>* __restrict__ Linv,
                                                  int* __restrict__ flag) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<double>(*L)[65] = reinterpret_cast<cx<double>(*)[65]>(smem);  // L[i][j], m <= 64 rows
    __shared__ int bad;
    const int slot = kidx[blockIdx.x];
    const cx<double>* g = G + (size_t)slot * m * m;
    cx<double>* out = Linv + (size_t)slot * m * m;
    double tr = 0.0;
    for (int i = 0; i < m; ++i) tr += g[(size_t)i * m + i].re;
    for (int attempt = force_shift ? 1 : 0; attempt < 2; ++attempt) {
        const double s = attempt ? shift_rel * tr : 0.0;
        for (int t = threadIdx.x; t < m * m; t += 256) {
            const int i = t % m, j = t / m;
            cx<double> v = i >= j ? g[(size_t)j * m + i] : cx<double>{0.0, 0.0};
            if (i == j) v = {v.re + s, 0.0};
            L[i][j] = v;
        }
        if (threadIdx.x == 0) bad = 0;
        __syncthreads();
        for (int j = 0; j < m; ++j) {
            if (threadIdx.x == 0) {
                const double dj = L[j][j].re;
                if (!(dj > 0.0)) bad = 1;
                L[j][j] = {dj > 0.0 ? sqrt(dj) : 1.0, 0.0};
            }
            __syncthreads();
            const double rj = 1.0 / L[j][j].re;
            for (int i = j + 1 + (int)threadIdx.x; i < m; i += 256) L[i][j] = {L[i][j].re * rj, L[i][j].im * rj};
            __syncthreads();
            // trailing update of the lower triangle: L[i][l] -= L[i][j] conj(L[l][j]), j < l <= i
            const int nt = m - j - 1;
            for (int t = threadIdx.x; t < nt * nt; t += 256) {
                const int i = j + 1 + t % nt, l = j + 1 + t / nt;
                if (l > i) continue;
                const cx<double> a = L[i][j], b = L[l][j];
                L[i][l] = {L[i][l].re - (a.re * b.re + a.im * b.im), L[i][l].im - (a.im * b.re - a.re * b.im)};
            }
            __syncthreads();
        }
        if (!bad) break;
        __syncthreads();
    }
    // columns of L^-1 by forward substitution, one thread per column
    const int col = threadIdx.x;
    if (col < m) {
        for (int i = 0; i < m; ++i) {
            cx<double> acc = {i == col ? 1.0 : 0.0, 0.0};
            for (int l = col; l < i; ++l) {
                const cx<double> a = L[i][l], x = out[(size_t)col * m + l];
                acc.re -= a.re * x.re - a.im * x.im;
                acc.im -= a.re * x.im + a.im * x.re;
            }
            const double r = 1.0 / L[i][i].re;
            out[(size_t)col * m + i] = i < col ? cx<double>{0.0, 0.0} : cx<double>{acc.re * r, acc.im * r};
            if (bad) out[(size_t)col * m + i] = {i == col ? 1.0 : 0.0, 0.0};
        }
    }
    if (threadIdx.x == 0) flag[slot] = bad;
}

// Residual norms of the Ritz pairs: rn[s][a] = || WQ[:, a] - theta_a X[:, a] || over the N
// rows (X, WQ in the block layout; theta[s][a] ascending).  One block per active slot.
__global__ __launch_bounds__(256) void k_sp_resid(int N, int mp, int m, const int* __restrict__ kidx,
                                                  const cx<double>* __restrict__ X, const cx<double>* __restrict__ WQ,
                                                  const double* __restrict__ theta, double* __restrict__ rn) {
    __shared__ double part[256];
    const int slot = kidx[blockIdx.x];
    const size_t off = (size_t)slot * N * mp;
    const int c = threadIdx.x % mp, rr = threadIdx.x / mp, nr = 256 / mp;
    const double th = c < m ? theta[(size_t)slot * m + c] : 0.0;
    double a = 0.0;
    if (c < m)
        for (int i = rr; i < N; i += nr) {
            const cx<double> w = WQ[off + (size_t)i * mp + c], x = X[off + (size_t)i * mp + c];
            const double dr = w.re - th * x.re, di = w.im - th * x.im;
            a += dr * dr + di * di;
        }
    part[threadIdx.x] = a;
    __syncthreads();
    if ((int)threadIdx.x < m) {
        double t = 0.0;
        for (int q = 0; q < nr; ++q) t += part[q * mp + threadIdx.x];
        rn[(size_t)slot * m + threadIdx.x] = sqrt(t);
    }
}

// Output for the listed slots (s = slots[blockIdx.y]): eigenvalues theta[s][lo[s] + b] + shift
// and eigenvectors X[s][:, lo[s] + b] as column b of an N x nb column-major block (b < nb).
__global__ __launch_bounds__(256) void k_sp_output(int N, int mp, int m, int nb, const int* __restrict__ slots,
                                                   const int* __restrict__ lo, const double* __restrict__ theta,
                                                   double shift, const cx<double>* __restrict__ X,
                                                   double* __restrict__ evals, cx<double>* __restrict__ evecs) {
    const int s = slots[blockIdx.y];
    const int l = lo[s];
    if (blockIdx.x == 0 && (int)threadIdx.x < nb) evals[(size_t)s * nb + threadIdx.x] = theta[(size_t)s * m + l + threadIdx.x] + shift;
    if (!evecs) return;
    const size_t off = (size_t)s * N * mp;
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < (size_t)N * nb; t += (size_t)gridDim.x * 256) {
        const int i = (int)(t % N), b = (int)(t / N);
        evecs[(size_t)s * N * nb + t] = X[off + (size_t)i * mp + l + b];
    }
}

}  // namespace lqm
