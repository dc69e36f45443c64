// k-grid reductions of the band path on gfx950: the phase contraction that turns the
// per-k rho_k stack into rho(R), the DOS, chemical-potential reductions, and plain
// H(k) materialisation.  Every reduction is two-level (fixed split, then a fixed-order
// fold) so results are bitwise reproducible run to run.
#pragma once
#include "lqm_common.hpp"

namespace lqm {

// rho(R)[e] = sum_k w_k e^{-2 pi i k.R} rho_k[e]   (Operators/densitymatrix.jl:141-143)
// grid of the contraction: x = element tiles of 256, y = k-splits, z = key groups of RG keys;
// part[(z*S + y)*RG + r][e] receives the split's partial sum.

// Phase table of the k-grid sums: ph[k * ldp + r] = w_k e^{sgn 2 pi i k.R_r} for r < nr,
// zero for nr <= r < ldp (w_k = kw[k], or wuni where kw is null).  The sums below read it
// at wave-uniform addresses, so the phases arrive through the scalar cache into scalar
// registers and feed the FMAs directly.
__global__ __launch_bounds__(256) void k_phase_table(int Kc, int d, const double* __restrict__ ks,
                                                     const double* __restrict__ kw, double wuni, int nr,
                                                     int ldp, const int* __restrict__ Rr, double sgn,
                                                     cx<double>* __restrict__ ph) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)Kc * ldp) return;
    const int k = (int)(idx / ldp), r = (int)(idx % ldp);
    cx<double> p = {0.0, 0.0};
    if (r < nr) {
        double x = 0.0;
        for (int a = 0; a < d; ++a) x += ks[(size_t)k * d + a] * (double)Rr[r * d + a];
        const double w = kw ? kw[k] : wuni;
        p = w * cis2pi(sgn * x);
    }
    ph[idx] = p;
}

// The phase contraction with the phases from the table ph (row stride ldp; key group z
// starts at column z * RG), read at wave-uniform addresses (scalar loads, the phases feed
// the FMAs from scalar registers): thread = element e, RG accumulators, the k-points of
// split y in order, four rho_k elements in flight.
//
// Hermitian mode (Nh = N > 0, the key set closed under R -> -R): thread eu < ne = N(N+1)/2
// takes the upper-triangle element (r, c), r <= c, packed by columns (eu = c(c+1)/2 + r), and
// part has row length ne; the fold rebuilds the lower triangle from rho(R)(r, c) =
// conj(rho(-R)(c, r)), which holds because every rho_k is Hermitian and the weights are real.
// Otherwise (Nh = 0) thread e < ne = NN is element e.
__device__ __forceinline__ void upper_rc(int eu, int& r, int& c) {
    int cc = (int)((sqrt(8.0 * eu + 1.0) - 1.0) * 0.5);
    while ((cc + 1) * (cc + 2) / 2 <= eu) ++cc;
    while (cc * (cc + 1) / 2 > eu) --cc;
    c = cc;
    r = eu - cc * (cc + 1) / 2;
}
template <int RG>
__global__ __launch_bounds__(256) void k_phase_contract_s(int NN, int ne, int Nh, int Nkc,
                                                          const cx<double>* __restrict__ ph, int ldp,
                                                          const cx<double>* __restrict__ rhok, int ksplit, int nRr,
                                                          cx<double>* __restrict__ part) {
    const int eo = blockIdx.x * 256 + threadIdx.x;  // output slot
    int e = eo;
    if (Nh > 0 && eo < ne) {
        int r, c;
        upper_rc(eo, r, c);
        e = c * Nh + r;
    }
    const int s = blockIdx.y;
    const int S = gridDim.y;
    const int rg0 = blockIdx.z * RG;
    const int nr = min(RG, nRr - rg0);
    const int k0 = s * ksplit;
    const int k1 = min(k0 + ksplit, Nkc);
    const cx<double>* phg = ph + rg0;
    cx<double> acc[RG];
#pragma unroll
    for (int r = 0; r < RG; ++r) acc[r] = {0.0, 0.0};
    if (eo < ne) {
        // slot j holds k-point min(k + j, k1 - 1); the loads are unconditional (clamped
        // index) so the waits can count them
        constexpr int D = 4;
        if (k0 < k1) {
            cx<double> buf[D];
#pragma unroll
            for (int j = 0; j < D; ++j) buf[j] = rhok[(size_t)min(k0 + j, k1 - 1) * NN + e];
            int k = k0;
            for (; k + D <= k1; k += D) {
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    const cx<double> v = buf[j];
                    buf[j] = rhok[(size_t)min(k + j + D, k1 - 1) * NN + e];
                    const cx<double>* p = phg + (size_t)(k + j) * ldp;
#pragma unroll
                    for (int r = 0; r < RG; ++r) acc[r] = cmad(acc[r], p[r], v);
                }
            }
#pragma unroll
            for (int j = 0; j < D; ++j)
                if (k + j < k1) {
                    const cx<double>* p = phg + (size_t)(k + j) * ldp;
#pragma unroll
                    for (int r = 0; r < RG; ++r) acc[r] = cmad(acc[r], p[r], buf[j]);
                }
        }
#pragma unroll
        for (int r = 0; r < RG; ++r)
            if (r < nr) part[((size_t)(blockIdx.z * S + s) * RG + r) * ne + eo] = acc[r];
    }
}

// neg[r] = the index of -R_r in the key list, -1 when absent.
__global__ void k_neg_keys(int nRr, int d, const int* __restrict__ Rr, int* __restrict__ neg) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nRr) return;
    int found = -1;
    for (int q = 0; q < nRr && found < 0; ++q) {
        bool m = true;
        for (int a = 0; a < d; ++a) m = m && Rr[q * d + a] == -Rr[r * d + a];
        if (m) found = q;
    }
    neg[r] = found;
}

// A = Hermitian(H(k)), H(k) = sum_R T_R e^{2 pi i k.R} (full storage, as k_hk_full): thread
// = element e = (r, c) of the N x N matrix, which keeps the nR <= NRM hopping entries of
// its upper-triangle source (r, c) or (c, r) in registers across the block's kb k-points.
// The phases come from the table ph (row stride ldp = NRM, zero beyond nR), staged through
// LDS 32 k-points at a time and read as broadcasts.  Every term up to NRM is summed: the
// padded entries are zero and fma(0, t, h) = h, so the sums are those over the nR keys.  A
// lower element is the conjugate of the same sum its upper mirror forms, so A is
// Hermitian bitwise.
constexpr int kHkChunk = 32;
template <int NRM>
__global__ __launch_bounds__(256) void k_hk_stream(int N, int nR, const cx<double>* __restrict__ T, int Kc,
                                                   int kb, const cx<double>* __restrict__ ph,
                                                   cx<double>* __restrict__ A) {
    __shared__ cx<double> sph[kHkChunk * NRM];
    const size_t NN = (size_t)N * N;
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int k0 = blockIdx.y * kb;
    const int k1 = min(k0 + kb, Kc);
    const bool live = e < NN;
    const int c = live ? (int)(e / N) : 0;
    const int r = live ? (int)(e - (size_t)c * N) : 0;
    const bool lower = r > c;
    const size_t src = lower ? (size_t)r * N + c : e;
    cx<double> t[NRM];
#pragma unroll
    for (int q = 0; q < NRM; ++q) t[q] = (live && q < nR) ? T[(size_t)q * NN + src] : cx<double>{0.0, 0.0};
    for (int kc = k0; kc < k1; kc += kHkChunk) {
        const int nk = min(kHkChunk, k1 - kc);
        __syncthreads();
        for (int i = threadIdx.x; i < nk * NRM; i += 256) sph[i] = ph[(size_t)kc * NRM + i];
        __syncthreads();
        if (!live) continue;
        for (int kq = 0; kq < nk; ++kq) {
            const cx<double>* p = sph + kq * NRM;
            cx<double> h = {0.0, 0.0};
#pragma unroll
            for (int q = 0; q < NRM; ++q) h = cmad(h, t[q], p[q]);
            A[(size_t)(kc + kq) * NN + e] = r == c ? cx<double>{h.re, 0.0} : (lower ? cx<double>{h.re, -h.im} : h);
        }
    }
}

// rho[(rg0 + r)][e] += sum_s part[..s..][r][e], fixed order over s.
template <int RG>
__global__ void k_phase_fold(int NN, int nRr, int S, const cx<double>* __restrict__ part,
                             cx<double>* __restrict__ rho) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)nRr * NN;
    if (idx >= total) return;
    const int key = (int)(idx / NN);
    const int e = (int)(idx % NN);
    const int g = key / RG, r = key % RG;
    cx<double> acc = rho[idx];
    const cx<double>* pp = part + ((size_t)g * S * RG + r) * NN + e;  // split s at pp[s * RG * NN]
    const size_t step = (size_t)RG * NN;
    // eight partials in flight, added in split order (the sum is the same as one at a time)
    int s = 0;
    for (; s + 8 <= S; s += 8) {
        cx<double> v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = pp[(size_t)(s + q) * step];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            acc.re += v[q].re;
            acc.im += v[q].im;
        }
    }
    for (; s < S; ++s) {
        const cx<double> v = pp[(size_t)s * step];
        acc.re += v.re;
        acc.im += v.im;
    }
    rho[idx] = acc;
}

// The fold of the Hermitian mode: thread (key, eu) sums the partials of upper element
// eu = (r, c), r <= c, of key in split order (row length nu = N(N+1)/2, coalesced reads) and
// adds the sum to rho[key](r, c) and, for r < c, its conjugate to rho[neg[key]](c, r); each
// element of every key is written by exactly one thread.
template <int RG>
__global__ void k_phase_fold_h(int N, int nRr, int S, const int* __restrict__ neg,
                               const cx<double>* __restrict__ part, cx<double>* __restrict__ rho) {
    const size_t NN = (size_t)N * N;
    const size_t nu = (size_t)N * (N + 1) / 2;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)nRr * nu) return;
    const int key = (int)(idx / nu);
    const int eu = (int)(idx % nu);
    int r, c;
    upper_rc(eu, r, c);
    const int g = key / RG, rr = key % RG;
    const cx<double>* pp = part + ((size_t)g * S * RG + rr) * nu + eu;  // split s at pp[s * RG * nu]
    const size_t step = (size_t)RG * nu;
    cx<double> sum = {0.0, 0.0};
    int s = 0;
    for (; s + 8 <= S; s += 8) {
        cx<double> v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = pp[(size_t)(s + q) * step];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            sum.re += v[q].re;
            sum.im += v[q].im;
        }
    }
    for (; s < S; ++s) {
        const cx<double> v = pp[(size_t)s * step];
        sum.re += v.re;
        sum.im += v.im;
    }
    cx<double>* up = rho + (size_t)key * NN + (size_t)c * N + r;
    *up = {up->re + sum.re, up->im + sum.im};
    const int kn = neg[key];
    if (r < c && kn >= 0) {  // kn < 0 cannot occur: the host takes this mode only for closed key sets
        cx<double>* lo = rho + (size_t)kn * NN + (size_t)r * N + c;
        *lo = {lo->re + sum.re, lo->im - sum.im};
    }
}

// Sum of n doubles into out[0] (+= when accumulate), one block, fixed order.
__global__ __launch_bounds__(256) void k_sum(int n, const double* __restrict__ x,
                                             double* __restrict__ out, int accumulate) {
    __shared__ double sred[4];
    double a = 0.0;
    int i = threadIdx.x;
    // eight loads in flight, added in index order (the same sum as one at a time)
    for (; i + 7 * 256 < n; i += 8 * 256) {
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = x[i + q * 256];
#pragma unroll
        for (int q = 0; q < 8; ++q) a += v[q];
    }
    for (; i < n; i += 256) a += x[i];
    a = wave_sum(a);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x / 64] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sred[0] + sred[1] + sred[2] + sred[3];
        out[0] = accumulate ? out[0] + t : t;
    }
}

// DOS(w) partials: part[s][o] = sum over eigenvalue slice s of
// w_k * Im[1/(w_o - i G - e)] = w_k G / ((w_o - e)^2 + G^2)   (Spectrum/dos.jl:185-188)
__global__ __launch_bounds__(256) void k_dos(int N, int Ne, const double* __restrict__ bands,
                                             const double* __restrict__ kw, int Nw,
                                             const double* __restrict__ om, double gamma,
                                             int esplit, double* __restrict__ part) {
    __shared__ double se[256];
    __shared__ double sw[256];
    const int o = blockIdx.x * 256 + threadIdx.x;
    const int s = blockIdx.y;
    const int e0 = s * esplit, e1 = min(e0 + esplit, Ne);
    const double w_o = o < Nw ? om[o] : 0.0;
    const double g2 = gamma * gamma;
    double acc = 0.0;
    for (int eb = e0; eb < e1; eb += 256) {
        __syncthreads();
        const int ei = eb + threadIdx.x;
        if (ei < e1) {
            se[threadIdx.x] = bands[ei];
            sw[threadIdx.x] = kw ? kw[ei / N] : 1.0;
        }
        __syncthreads();
        const int n = min(256, e1 - eb);
        for (int t = 0; t < n; ++t) {
            const double x = w_o - se[t];
            acc += sw[t] * gamma / (x * x + g2);
        }
    }
    if (o < Nw) part[(size_t)s * Nw + o] = acc;
}

__global__ void k_dos_fold(int Nw, int S, const double* __restrict__ part, double scale,
                           double* __restrict__ dos) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= Nw) return;
    double a = 0.0;
    for (int s = 0; s < S; ++s) a += part[(size_t)s * Nw + o];
    dos[o] = a * scale;
}

// getfilling partials (Spectrum/chemicalpotential.jl:74-78): per block,
// sum_k w_k sum_n f(e_nk - mu) (w_k = 1 when kw == NULL).
__global__ __launch_bounds__(256) void k_filling(int N, int Ne, const double* __restrict__ bands,
                                                 const double* __restrict__ kw, double mu,
                                                 double T, double* __restrict__ part) {
    __shared__ double sred[4];
    double a = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < Ne; i += gridDim.x * 256) {
        const double x = bands[i] - mu;
        double f;
        if (T > 0.0) f = 1.0 / (exp(x / T) + 1.0);
        else f = x < 0.0 ? 1.0 : (x > 0.0 ? 0.0 : 0.5);
        a += kw ? kw[i / N] * f : f;
    }
    a = wave_sum(a);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x / 64] = a;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sred[0] + sred[1] + sred[2] + sred[3];
}

// min / max of n doubles: part[2*b] = min, part[2*b+1] = max over block b's slice.
__global__ __launch_bounds__(256) void k_minmax(int n, const double* __restrict__ x,
                                                double* __restrict__ part) {
    __shared__ double smin[4], smax[4];
    double lo = INFINITY, hi = -INFINITY;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        lo = fmin(lo, x[i]);
        hi = fmax(hi, x[i]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, o));
        hi = fmax(hi, __shfl_xor(hi, o));
    }
    if ((threadIdx.x & 63) == 0) {
        smin[threadIdx.x / 64] = lo;
        smax[threadIdx.x / 64] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = fmin(fmin(smin[0], smin[1]), fmin(smin[2], smin[3]));
        part[2 * blockIdx.x + 1] = fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3]));
    }
}

// Order-preserving map double -> uint64 (total order on non-NaN values).
__device__ __forceinline__ unsigned long long dkey(double v) {
    unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

// One radix-select pass: histogram of byte `shift` among keys whose bits above the
// byte equal `prefix` under `mask`.
__global__ __launch_bounds__(256) void k_radix_hist(int n, const double* __restrict__ x,
                                                    unsigned long long prefix,
                                                    unsigned long long mask, int shift,
                                                    unsigned int* __restrict__ hist) {
    __shared__ unsigned int sh[256];
    sh[threadIdx.x] = 0;
    __syncthreads();
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const unsigned long long k = dkey(x[i]);
        if ((k & mask) == prefix) atomicAdd(&sh[(k >> shift) & 0xff], 1u);
    }
    __syncthreads();
    if (sh[threadIdx.x]) atomicAdd(&hist[threadIdx.x], sh[threadIdx.x]);
}

// Plain H(k) = sum_R T_R e^{2 pi i k.R} (TightBinding/types.jl:18-25), one block per k.
__global__ __launch_bounds__(256) void k_fouriersum(int N, int d, int nR, const int* __restrict__ Rk,
                                                    const cx<double>* __restrict__ T,
                                                    const double* __restrict__ ks,
                                                    cx<double>* __restrict__ Hk) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<double>* sph = reinterpret_cast<cx<double>*>(smem);
    const int kk = blockIdx.x;
    const double* k = ks + (size_t)kk * d;
    for (int r = threadIdx.x; r < nR; r += 256) {
        double x = 0.0;
        for (int a = 0; a < d; ++a) x += k[a] * (double)Rk[r * d + a];
        sph[r] = cis2pi(x);
    }
    __syncthreads();
    const size_t NN = (size_t)N * N;
    for (size_t e = threadIdx.x; e < NN; e += 256) {
        cx<double> h = {0.0, 0.0};
        for (int r = 0; r < nR; ++r) h = cmad(h, T[r * NN + e], sph[r]);
        Hk[(size_t)kk * NN + e] = h;
    }
}

// Grand-potential partials: part[b] = block b's share of sum_k w_k sum_n omega(e_nk),
// omega(e) = -T ln(1 + e^{-(e - mu)/T}) in the overflow-free form
// -T (max(-x, 0) + log1p(e^{-|x|})), x = (e - mu)/T; at T = 0 omega = min(e - mu, 0).
__global__ __launch_bounds__(256) void k_omega(int N, int Ne, const double* __restrict__ bands,
                                               const double* __restrict__ kw, double mu, double T,
                                               double* __restrict__ part) {
    __shared__ double sred[4];
    double a = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < Ne; i += gridDim.x * 256) {
        const double e = bands[i] - mu;
        double w;
        if (T > 0.0) {
            const double x = e / T;
            w = -T * (fmax(-x, 0.0) + log1p(exp(-fabs(x))));
        } else {
            w = fmin(e, 0.0);
        }
        a += kw ? kw[i / N] * w : w;
    }
    a = wave_sum(a);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x / 64] = a;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sred[0] + sred[1] + sred[2] + sred[3];
}

// out[b] = transpose(in[b]) for nb stacked N x N complex blocks.
__global__ __launch_bounds__(256) void k_block_transpose(int N, int nb, const cx<double>* __restrict__ in,
                                                         cx<double>* __restrict__ out) {
    const size_t NN = (size_t)N * N;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= NN * nb) return;
    const size_t b = idx / NN, e = idx - b * NN;
    const size_t i = e % N, j = e / N;
    out[b * NN + i * N + j] = in[idx];
}

// Band expectation values (Spectrum/bands.jl:52-58, :107-112): obs[n] of k-point kk =
// Re(psi_n^H W_n) with W = O(k) V, V the eigenvectors (column n = psi_n).  One block per
// k-point; out points at obs[0, k0, p] of the N x Nk x nP array (column stride N).
__global__ __launch_bounds__(256) void k_expval_cols(int N, const cx<double>* __restrict__ V,
                                                     const cx<double>* __restrict__ W,
                                                     double* __restrict__ out) {
    const int kk = blockIdx.x;
    const size_t NN = (size_t)N * N;
    const cx<double>* v = V + (size_t)kk * NN;
    const cx<double>* w = W + (size_t)kk * NN;
    for (int n = threadIdx.x; n < N; n += 256) {
        double a = 0.0;
        for (int i = 0; i < N; ++i) {
            const cx<double> x = v[(size_t)n * N + i], y = w[(size_t)n * N + i];
            a += x.re * y.re + x.im * y.im;
        }
        out[(size_t)kk * N + n] = a;
    }
}

// p[0 .. n) = 0 (32-bit words): the clearing step of passes that may be captured in a graph,
// where a memset node is avoided.
__global__ __launch_bounds__(256) void k_zero_words(size_t n, uint32_t* __restrict__ p) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0u;
}

// out[i] = in[i] converted between complex precisions (n elements, grid-stride).
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_convert_cx(size_t n, const cx<TI>* __restrict__ in, cx<TO>* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        out[i] = cx<TO>{(TO)in[i].re, (TO)in[i].im};
}

// Determinant phase det(O) / |det(O)| of each M x M complex matrix O[l] (column-major, stride
// M^2), by LU with partial pivoting in place (O is scratch): the link variable of a Berry
// plaquette (Spectrum/berry.jl:116-119).  One 256-thread block per matrix; out[2l..2l+1] = the
// unit phase, or (0, 0) for a singular matrix.
__global__ __launch_bounds__(256) void k_det_phase(int M, cx<double>* __restrict__ O, double* __restrict__ out) {
    __shared__ double sv[256];
    __shared__ int si[256];
    __shared__ cx<double> s_ph;
    __shared__ int s_piv, s_sing;
    const int tid = threadIdx.x;
    cx<double>* A = O + (size_t)blockIdx.x * M * M;
    if (tid == 0) {
        s_ph = {1.0, 0.0};
        s_sing = 0;
    }
    __syncthreads();
    for (int j = 0; j < M; ++j) {
        // pivot: the largest |a_rj|, r >= j (smallest r among ties)
        double best = -1.0;
        int bi = j;
        for (int r = j + tid; r < M; r += 256) {
            const double a = abs2(A[(size_t)j * M + r]);
            if (a > best) {
                best = a;
                bi = r;
            }
        }
        sv[tid] = best;
        si[tid] = bi;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (tid < o && (sv[tid + o] > sv[tid] || (sv[tid + o] == sv[tid] && si[tid + o] < si[tid]))) {
                sv[tid] = sv[tid + o];
                si[tid] = si[tid + o];
            }
            __syncthreads();
        }
        if (tid == 0) {
            s_piv = si[0];
            if (!(sv[0] > 0.0)) s_sing = 1;
        }
        __syncthreads();
        if (s_sing) break;
        const int p = s_piv;
        if (p != j)
            for (int c = tid; c < M; c += 256) {
                const cx<double> t = A[(size_t)c * M + j];
                A[(size_t)c * M + j] = A[(size_t)c * M + p];
                A[(size_t)c * M + p] = t;
            }
        __syncthreads();
        const cx<double> d = A[(size_t)j * M + j];
        if (tid == 0) {
            const double n = sqrt(abs2(d));
            cx<double> u = {d.re / n, d.im / n};
            if (p != j) u = {-u.re, -u.im};
            const cx<double> q = s_ph * u;
            const double qn = sqrt(abs2(q));
            s_ph = {q.re / qn, q.im / qn};
        }
        const cx<double> dinv = crecip(d);
        for (int r = j + 1 + tid; r < M; r += 256) A[(size_t)j * M + r] = A[(size_t)j * M + r] * dinv;
        __syncthreads();
        const int n = M - j - 1;
        for (int e = tid; e < n * n; e += 256) {
            const int r = j + 1 + e % n, c = j + 1 + e / n;
            A[(size_t)c * M + r] = A[(size_t)c * M + r] - A[(size_t)j * M + r] * A[(size_t)c * M + j];
        }
        __syncthreads();
    }
    if (tid == 0) {
        out[2 * blockIdx.x] = s_sing ? 0.0 : s_ph.re;
        out[2 * blockIdx.x + 1] = s_sing ? 0.0 : s_ph.im;
    }
}

// Nambu (BdG) chemical potential shift (Superconductivity/operators.jl:14-17 applied with -mu,
// as getdensitymatrix!(::BdGOperator) does, green.jl:38-54): the zero-key block of the 2d x 2d
// Nambu model becomes base - mu tau_z, i.e. -mu on the electron diagonal (i < d) and +mu on
// the hole diagonal, written from the unshifted base block so repeated shifts never drift.
__global__ __launch_bounds__(256) void k_bdg_shift(int N, const cx<double>* __restrict__ base, double mu,
                                                   cx<double>* __restrict__ blk) {
    const size_t NN = (size_t)N * N;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= NN) return;
    cx<double> v = base[idx];
    const size_t i = idx % N, j = idx / N;
    if (i == j) v.re += (int)i < N / 2 ? -mu : mu;
    blk[idx] = v;
}

// Electron filling of a Nambu density matrix (Superconductivity/operators.jl:107-112):
// out = Re tr(rho[1:d, 1:d]) / d of the zero-key block rho (N = 2d, column-major).  One block.
__global__ __launch_bounds__(256) void k_bdg_filling(int N, const cx<double>* __restrict__ rho,
                                                     double* __restrict__ out) {
    __shared__ double sred[4];
    const int d = N / 2;
    double a = 0.0;
    for (int i = threadIdx.x; i < d; i += 256) a += rho[(size_t)i * N + i].re;
    a = wave_sum(a);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x / 64] = a;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = (sred[0] + sred[1] + sred[2] + sred[3]) / (double)d;
}

}  // namespace lqm
