// Consumers of the batched eigenvectors: site-resolved spectral weights (ldos, density) and
// the Kubo current-current correlator.  Every kernel reads eigenvalues [N x Kc] and
// eigenvectors [Kc][N x N] (column n = state n) as the band pass leaves them in HBM.
#pragma once

#include "lqm_common.hpp"

namespace lqm {

// occ[i] = fermidirac(e_i - mu; T) over L eigenvalues (Utils/fermidirac.jl:1-5).
__global__ __launch_bounds__(256) void k_fd_occ(int L, const double* __restrict__ ev, double mu, double T,
                                                double* __restrict__ occ) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < L) occ[i] = fermidirac(ev[i], mu, T);
}

// Per-site spectral weights, accumulated into part[b][i] (b = blockIdx.x) over the k-points
// kk = b, b + G, ... of a chunk:
//   part[b][i] += w_k sum_n |U_k[i, n]|^2 g(e_nk)
// kind 0, ldos (Spectrum/dos.jl:57-65): g(e) = (1/Nw) sum_o G / ((w_o - e)^2 + G^2), the
//         negative imaginary part of 1/(w + iG - e) averaged over the frequency window;
// kind 1, density (dos.jl:20-28): g(e) = 1 for e <= mu, else 0.
// w_k = kw[kk] (or 1 without weights).  `first` starts the sums (the first chunk).  Dynamic
// LDS: N doubles.
__global__ __launch_bounds__(256) void k_site_weights(int N, int Kc, const double* __restrict__ bands,
                                                      const cx<double>* __restrict__ U,
                                                      const double* __restrict__ kw, int kind, int Nw,
                                                      const double* __restrict__ om, double gamma,
                                                      double mu, int first, double* __restrict__ part) {
    extern __shared__ __align__(16) unsigned char smem[];
    double* g = reinterpret_cast<double*>(smem);
    double* out = part + (size_t)blockIdx.x * N;
    const size_t NN = (size_t)N * N;
    const double g2 = gamma * gamma;
    bool fresh = first != 0;
    for (int kk = blockIdx.x; kk < Kc; kk += gridDim.x) {
        const double wk = kw ? kw[kk] : 1.0;
        __syncthreads();
        for (int n = threadIdx.x; n < N; n += 256) {
            const double e = bands[(size_t)kk * N + n];
            double v;
            if (kind == 0) {
                v = 0.0;
                for (int o = 0; o < Nw; ++o) {
                    const double x = om[o] - e;
                    v += gamma / (x * x + g2);
                }
                v /= (double)Nw;
            } else {
                v = e <= mu ? 1.0 : 0.0;
            }
            g[n] = wk * v;
        }
        __syncthreads();
        const cx<double>* u = U + (size_t)kk * NN;
        for (int i = threadIdx.x; i < N; i += 256) {
            double a = 0.0;
            for (int n = 0; n < N; ++n) {
                const cx<double> z = u[(size_t)n * N + i];
                a += (z.re * z.re + z.im * z.im) * g[n];
            }
            out[i] = fresh ? a : out[i] + a;
        }
        fresh = false;
    }
    // blocks with no k-point in this chunk still start their row on the first chunk
    if (fresh)
        for (int i = threadIdx.x; i < N; i += 256) out[i] = 0.0;
}

// Kubo pair sums (LinearResponse/kubo.jl:11-31) over the k-points of a chunk, for the 64 F
// frequencies o = 64 F blockIdx.y + 64 f + lane (f < F):
//   acc[o] += sum_k w_k sum_{i != j} (f_i - f_j) A_ij conj(C_ij) / ((e_i - e_j) - (w_o + i G))
// with A = U^H J1 U and C = U^H J2 U per k-point (A_ij conj(C_ij) = (U^H J1 U)_ij
// (U^H J2^H U)_ji, the reference's M).  Each block takes the k-points b, b + G, ...; per
// k-point its 256 threads evaluate the pair weights 256 (i, j) at a time, compact the nonzero
// ones (f_i = f_j, in particular i = j, gives exactly zero) in order into an LDS list, and
// every wave then sums a quarter of the list for its lane's F frequencies (one LDS read of an
// entry serves F terms).  A term w / (a - iG), a = D - w_o, is w (a + iG) / (a^2 + G^2) with
// the reciprocal from the hardware approximation and one Newton step.  The four wave sums
// meet in LDS at the end: part[b][o] (complex) is written or, after the first chunk, added to.
constexpr int kKuboList = 1024;
template <int F>
__global__ __launch_bounds__(256) void k_kubo(int N, int Kc, const double* __restrict__ bands,
                                              const double* __restrict__ occ, const cx<double>* __restrict__ A,
                                              const cx<double>* __restrict__ C, const double* __restrict__ kw,
                                              int Nw, const double* __restrict__ om, double gamma, int first,
                                              cx<double>* __restrict__ part) {
    __shared__ double sD[kKuboList];
    __shared__ cx<double> sW[kKuboList];
    __shared__ int swc[4];
    __shared__ double sre[4][64], sim[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int o0 = blockIdx.y * 64 * F + lane;
    double wo[F], re[F], im[F];
#pragma unroll
    for (int f = 0; f < F; ++f) {
        wo[f] = o0 + 64 * f < Nw ? om[o0 + 64 * f] : 0.0;
        re[f] = 0.0;
        im[f] = 0.0;
    }
    const int NN = N * N;
    const double g2 = gamma * gamma;
    int cnt = 0;
    // sum the listed pairs: wave wv takes entries [wv q, min((wv + 1) q, cnt))
    auto flush = [&]() {
        __syncthreads();
        const int q = (cnt + 3) / 4;
        const int e0 = wv * q, e1 = min(cnt, e0 + q);
        for (int e = e0; e < e1; ++e) {
            const double D = sD[e];
            const cx<double> w = sW[e];
            const double wgi = -w.im * gamma, wgr = w.re * gamma;
#pragma unroll
            for (int f = 0; f < F; ++f) {
                const double a = D - wo[f];
                const double r = rcp_nr(fma(a, a, g2));
                re[f] = fma(fma(w.re, a, wgi), r, re[f]);
                im[f] = fma(fma(w.im, a, wgr), r, im[f]);
            }
        }
        __syncthreads();
        cnt = 0;
    };
    for (int kk = blockIdx.x; kk < Kc; kk += gridDim.x) {
        const double wk = kw ? kw[kk] : 1.0;
        const double* e = bands + (size_t)kk * N;
        const double* fo = occ + (size_t)kk * N;
        const cx<double>* a = A + (size_t)kk * NN;
        const cx<double>* c = C + (size_t)kk * NN;
        for (int p0 = 0; p0 < NN; p0 += 256) {
            const int p = p0 + threadIdx.x;
            cx<double> w = {0.0, 0.0};
            double D = 0.0;
            if (p < NN) {
                const int i = p % N, j = p / N;
                const double df = fo[i] - fo[j];
                if (df != 0.0) {
                    const cx<double> x = a[p], y = c[p];
                    // x conj(y), times (f_i - f_j) w_k
                    const double mr = x.re * y.re + x.im * y.im, mi = x.im * y.re - x.re * y.im;
                    w = {df * mr * wk, df * mi * wk};
                    D = e[i] - e[j];
                }
            }
            const bool nz = w.re != 0.0 || w.im != 0.0;
            const unsigned long long m = __ballot(nz);
            if (lane == 0) swc[wv] = __popcll(m);
            __syncthreads();
            int base = cnt;
            for (int v = 0; v < wv; ++v) base += swc[v];
            const int tot = swc[0] + swc[1] + swc[2] + swc[3];
            if (nz) {
                const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
                sD[pos] = D;
                sW[pos] = w;
            }
            cnt += tot;
            if (cnt > kKuboList - 256) flush();
            else __syncthreads();  // swc is rewritten by the next tile
        }
    }
    if (cnt > 0) flush();
#pragma unroll
    for (int f = 0; f < F; ++f) {
        sre[wv][lane] = re[f];
        sim[wv][lane] = im[f];
        __syncthreads();
        const int o = o0 + 64 * f;
        if (wv == 0 && o < Nw) {
            const double tr = ((sre[0][lane] + sre[1][lane]) + sre[2][lane]) + sre[3][lane];
            const double ti = ((sim[0][lane] + sim[1][lane]) + sim[2][lane]) + sim[3][lane];
            cx<double>* out = part + (size_t)blockIdx.x * Nw + o;
            if (first) *out = {tr, ti};
            else *out = {out->re + tr, out->im + ti};
        }
        __syncthreads();
    }
}

// chi[o] = -i scale sum_b part[b][o] (opticalconductivity.jl:124, OC .*= -1im / N).
__global__ __launch_bounds__(256) void k_kubo_fold(int Nw, int G, const cx<double>* __restrict__ part,
                                                   double scale, cx<double>* __restrict__ chi) {
    const int o = blockIdx.x * 256 + threadIdx.x;
    if (o >= Nw) return;
    double x = 0.0, y = 0.0;
    for (int b = 0; b < G; ++b) {
        x += part[(size_t)b * Nw + o].re;
        y += part[(size_t)b * Nw + o].im;
    }
    chi[o] = {scale * y, -scale * x};
}

}  // namespace lqm
