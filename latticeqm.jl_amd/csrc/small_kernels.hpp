// Bloch Hamiltonians of at most four orbitals (the spinful two-site models of the SCF
// configurations): one thread per k-point holds the whole matrix in registers.
//   k_small_eig<N>  H(k) = sum_R T_R e^{2 pi i k.R}, Hermitian from its upper triangle (as
//                   Hermitian(M)), cyclic complex Jacobi to convergence, eigenvalues ascending
//                   and (optionally) the eigenvectors
//   k_small_rho<N>  rho_k = U diag(f) U^H and w_k sum_n f_n e_n from those eigenpairs, with f
//                   the given occupations or fermidirac(e - mu; T), mu from the host or the
//                   device
// A Jacobi rotation on (p, q) with a_pq = |z| e^{i phi}: U = [[c, s e^{i phi}],
// [-s e^{-i phi}, c]], t = tan(theta) the smaller root of t^2 + 2 tau t - 1 = 0,
// tau = (a_qq - a_pp) / (2 |z|); then a_pp -= t |z|, a_qq += t |z|, a_pq = 0.
#pragma once

#include "lqm_common.hpp"

namespace lqm {

template <int N>
__device__ __forceinline__ void small_rotate(cx<double> (&a)[N][N], cx<double> (&v)[N][N], int p, int q,
                                             bool vectors) {
    const cx<double> z = a[p][q];
    const double az = sqrt(z.re * z.re + z.im * z.im);
    if (!(az > 0.0)) return;
    const double app = a[p][p].re, aqq = a[q][q].re;
    const double tau = (aqq - app) / (2.0 * az);
    const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
    const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
    const cx<double> e = {z.re / az, z.im / az};  // e^{i phi}
#pragma unroll
    for (int r = 0; r < N; ++r) {
        if (r == p || r == q) continue;
        const cx<double> arp = a[r][p], arq = a[r][q];
        // a_rp' = c a_rp - s e^{-i phi} a_rq, a_rq' = s e^{i phi} a_rp + c a_rq
        const cx<double> np = {c * arp.re - s * (e.re * arq.re + e.im * arq.im),
                               c * arp.im - s * (e.re * arq.im - e.im * arq.re)};
        const cx<double> nq = {s * (e.re * arp.re - e.im * arp.im) + c * arq.re,
                               s * (e.re * arp.im + e.im * arp.re) + c * arq.im};
        a[r][p] = np;
        a[r][q] = nq;
        a[p][r] = {np.re, -np.im};
        a[q][r] = {nq.re, -nq.im};
    }
    a[p][p] = {app - t * az, 0.0};
    a[q][q] = {aqq + t * az, 0.0};
    a[p][q] = {0.0, 0.0};
    a[q][p] = {0.0, 0.0};
    if (vectors) {
#pragma unroll
        for (int r = 0; r < N; ++r) {
            const cx<double> vp = v[r][p], vq = v[r][q];
            v[r][p] = {c * vp.re - s * (e.re * vq.re + e.im * vq.im), c * vp.im - s * (e.re * vq.im - e.im * vq.re)};
            v[r][q] = {s * (e.re * vp.re - e.im * vp.im) + c * vq.re, s * (e.re * vp.im + e.im * vp.re) + c * vq.im};
        }
    }
}

// One thread per k-point: eigenvalues ev[kk][n] ascending and, with vectors, the eigenvectors
// V[kk] column-major (column n = eigenvector n).  Each thread forms its own nR phases.
template <int N>
__global__ __launch_bounds__(256) void k_small_eig(int Kc, int d, int nR, const int* __restrict__ Rk,
                                                   const cx<double>* __restrict__ T,
                                                   const double* __restrict__ ks, double* __restrict__ ev,
                                                   cx<double>* __restrict__ V, int* __restrict__ status) {
    const int kk = blockIdx.x * 256 + threadIdx.x;
    if (kk >= Kc) return;
    constexpr int NN = N * N;
    // H(k) into the upper triangle of a (column-major element (i, j) of T_R at j N + i), then
    // mirrored: Hermitian(H(k))
    cx<double> a[N][N], v[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
            a[i][j] = {0.0, 0.0};
            v[i][j] = {i == j ? 1.0 : 0.0, 0.0};
        }
    const double* k = ks + (size_t)kk * d;
    for (int r = 0; r < nR; ++r) {
        double x = 0.0;
        for (int q = 0; q < d; ++q) x += k[q] * (double)Rk[r * d + q];
        const cx<double> ph = cis2pi(x);
        const cx<double>* t = T + (size_t)r * NN;
#pragma unroll
        for (int j = 0; j < N; ++j)
#pragma unroll
            for (int i = 0; i <= j; ++i) a[i][j] = cmad(a[i][j], t[j * N + i], ph);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        a[i][i] = {a[i][i].re, 0.0};
#pragma unroll
        for (int j = 0; j < i; ++j) a[i][j] = {a[j][i].re, -a[j][i].im};
    }
    const bool vectors = V != nullptr;
    bool finite = true;
    // cyclic sweeps until the off-diagonal part is below the rounding of the diagonal
    for (int sweep = 0; sweep < 16; ++sweep) {
        double off = 0.0, dia = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            dia += a[i][i].re * a[i][i].re;
#pragma unroll
            for (int j = i + 1; j < N; ++j) off += a[i][j].re * a[i][j].re + a[i][j].im * a[i][j].im;
        }
        if (!isfinite(off) || !isfinite(dia)) {
            finite = false;  // NaN / Inf in H(k): reported like the other eigensolvers
            break;
        }
        if (!(off > 1e-34 * dia) || off == 0.0) break;
#pragma unroll
        for (int p = 0; p < N; ++p)
#pragma unroll
            for (int q = p + 1; q < N; ++q) small_rotate<N>(a, v, p, q, vectors);
    }
    // ascending order: compare-exchange on the diagonal, the columns of the vectors swapped with
    // it by selects (every register index static: no stack array)
    double e[N];
#pragma unroll
    for (int n = 0; n < N; ++n) e[n] = a[n][n].re;
#pragma unroll
    for (int p = 0; p < N; ++p)
#pragma unroll
        for (int q = p + 1; q < N; ++q) {
            const bool sw = e[q] < e[p];
            const double ep = e[p], eq = e[q];
            e[p] = sw ? eq : ep;
            e[q] = sw ? ep : eq;
            if (vectors) {
#pragma unroll
                for (int r = 0; r < N; ++r) {
                    const cx<double> vp = v[r][p], vq = v[r][q];
                    v[r][p] = sw ? vq : vp;
                    v[r][q] = sw ? vp : vq;
                }
            }
        }
#pragma unroll
    for (int n = 0; n < N; ++n) ev[(size_t)kk * N + n] = e[n];
    if (!finite) atomicOr(status, kStatNonFinite);
    if (vectors) {
        cx<double>* out = V + (size_t)kk * NN;
#pragma unroll
        for (int n = 0; n < N; ++n)
#pragma unroll
            for (int m = 0; m < N; ++m) out[(size_t)n * N + m] = v[m][n];
    }
}

// rho_k = sum_n f_n u_n u_n^H (column-major, every element) and eb[kk] = w_k sum_n f_n e_n, with
// f_n = occ[kk N + n] when occ is given, else fermidirac(e_n - mu; T) with mu = *mu_dev when
// mu_dev is given, else mu.  w_k = kw[kk] or wuni.
template <int N>
__global__ __launch_bounds__(256) void k_small_rho(int Kc, const double* __restrict__ ev, const cx<double>* __restrict__ V,
                                                   const double* __restrict__ kw, double wuni, double mu,
                                                   const double* __restrict__ mu_dev, double T,
                                                   const double* __restrict__ occ, cx<double>* __restrict__ rhok,
                                                   double* __restrict__ eb) {
    const int kk = blockIdx.x * 256 + threadIdx.x;
    if (kk >= Kc) return;
    constexpr int NN = N * N;
    const double m = mu_dev ? mu_dev[0] : mu;
    double f[N], ebk = 0.0;
#pragma unroll
    for (int n = 0; n < N; ++n) {
        const double e = ev[(size_t)kk * N + n];
        f[n] = occ ? occ[(size_t)kk * N + n] : fermidirac(e, m, T);
        ebk += f[n] * e;
    }
    eb[kk] = (kw ? kw[kk] : wuni) * ebk;
    cx<double> u[N][N];
    const cx<double>* vk = V + (size_t)kk * NN;
#pragma unroll
    for (int n = 0; n < N; ++n)
#pragma unroll
        for (int i = 0; i < N; ++i) u[i][n] = vk[(size_t)n * N + i];
    cx<double>* out = rhok + (size_t)kk * NN;
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) {
            cx<double> acc = {0.0, 0.0};
#pragma unroll
            for (int n = 0; n < N; ++n) {
                const cx<double> x = u[i][n], y = u[j][n];
                // f_n x conj(y)
                acc.re = fma(f[n], x.re * y.re + x.im * y.im, acc.re);
                acc.im = fma(f[n], x.im * y.re - x.re * y.im, acc.im);
            }
            out[(size_t)j * N + i] = acc;
        }
}

}  // namespace lqm
