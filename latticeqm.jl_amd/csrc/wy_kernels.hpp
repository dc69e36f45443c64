// Back-transformation V = H_0 H_1 ... H_{N-2} Z on the FP64 matrix cores (N <= 64).
//
// The N - 1 Householder reflectors H_j = I - tau_j v_j v_j^H of the tridiagonalization are
// applied to the QL rotation product Z in blocks of 16 in compact WY form (LAPACK zlarft,
// forward, columnwise): Q_b = H_{16b} ... H_{16b+15} = I - Y T Y^H with T upper triangular,
// T(i,i) = tau_i, T(0:i, i) = -tau_i T(0:i, 0:i) Y(:, 0:i)^H v_i.  The last block is applied
// first: V <- V - (Y T)(Y^H V), two matrix-core products per block instead of a rank-1
// update (and two barriers) per reflector.  V stays in LDS (FP64 complex, 64 x 65) for the
// whole k-point; the reflectors and Z may be FP32 (mixed path) or FP64.
//   MODE 0: write the eigenvectors, column n = V(:, perm[n]) (ascending eigenvalues);
//   MODE 1: rho_k = V diag(f) V^H with f = fermidirac(e - mu; T) (or the given occupations)
//           and eb[kk] = w_k sum_n e_n f_n, as k_backtransform MODE 1.
#pragma once
#include "lqm_common.hpp"
#include "refine_kernels.hpp"

namespace lqm {

constexpr int kWyLd = 65;  // leading dimension of V in LDS
constexpr int kWyB = 16;   // reflectors per block

inline size_t wy_lds_bytes() {
    return sizeof(cx<double>) * (64 * kWyLd + 64 * kWyB * 2 + 64 * (kWyB + 1) + kWyB * kWyB + kWyB) +
           sizeof(double) * (64 + 8);
}

template <typename R, int MODE>
__global__ __launch_bounds__(256) void k_backtransform_wy(
    int N, const cx<R>* __restrict__ refl, const cx<R>* __restrict__ tau, const R* __restrict__ Zg,
    const R* __restrict__ evals, const int* __restrict__ perm, const double* __restrict__ kw,
    double wuni, double mu, double Temp, const double* __restrict__ occ, cx<double>* __restrict__ out,
    double* __restrict__ eb) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<double>* Vs = reinterpret_cast<cx<double>*>(smem);  // V(r, c) at c * kWyLd + r
    cx<double>* Ys = Vs + 64 * kWyLd;                        // Y(r, i) at r * 16 + i
    cx<double>* Us = Ys + 64 * kWyB;                         // U = Y T, same layout
    cx<double>* Ws = Us + 64 * kWyB;                         // W = Y^H V: W(i, c) at c * 17 + i
    cx<double>* Ps = Ws + 64 * (kWyB + 1);                   // P = Y^H Y: P(p, q) at q * 16 + p
    cx<double>* Ts = Ps + kWyB * kWyB;                       // tau of the block (16)
    double* sf = reinterpret_cast<double*>(Ts + kWyB);       // occupations per Z column
    double* sred = sf + 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int kk = blockIdx.x;
    const size_t NN = (size_t)N * N;
    const cx<R>* rf = refl + (size_t)kk * NN;

    // V = Z (real), zero padded to 64 x 64
    const R* Z = Zg + (size_t)kk * NN;
    for (int e = tid; e < 64 * 64; e += 256) {
        const int r = e & 63, c = e >> 6;
        Vs[c * kWyLd + r] = cx<double>{(r < N && c < N) ? (double)Z[(size_t)c * N + r] : 0.0, 0.0};
    }
    const int nrefl = N - 1;
    const int nb = (nrefl + kWyB - 1) / kWyB;
    for (int b = nb - 1; b >= 0; --b) {
        const int j0 = b * kWyB;
        __syncthreads();  // V complete (load or previous block), Ys free
        // (1) reflectors of the block: Y(r, i) = v_{j0+i}(r), zero above row j+1
        for (int e = tid; e < 64 * kWyB; e += 256) {
            const int r = e >> 4, i = e & 15;
            const int j = j0 + i;
            cx<double> y = {0.0, 0.0};
            if (j < nrefl && r > j && r < N) {
                const cx<R> v = rf[(size_t)j * N + r];
                y = {(double)v.re, (double)v.im};
            }
            Ys[r * kWyB + i] = y;
        }
        if (tid < kWyB) {
            const int j = j0 + tid;
            cx<double> t = {0.0, 0.0};
            if (j < nrefl) {
                const cx<R> x = tau[(size_t)kk * N + j];
                t = {(double)x.re, (double)x.im};
            }
            Ts[tid] = t;
        }
        __syncthreads();
        // (2) P = Y^H Y (VALU, one entry per thread) and W = Y^H V (matrix cores, wave w
        //     owns columns 16w .. 16w + 15)
        {
            const int p = tid & 15, q = tid >> 4;
            cx<double> acc = {0.0, 0.0};
            for (int r = 0; r < 64; ++r) acc = cmadc(acc, Ys[r * kWyB + p], Ys[r * kWyB + q]);
            Ps[q * kWyB + p] = acc;
        }
        {
            lqm_d4 wr = {0.0, 0.0, 0.0, 0.0}, wi = {0.0, 0.0, 0.0, 0.0};
            const int c = 16 * wave + li;
            for (int k0 = 0; k0 < 64; k0 += 16) {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int r = k0 + 4 * g + s;
                    const cx<double> y = Ys[r * kWyB + li];  // A(m = li, k = r) = conj(Y(r, li))
                    const cx<double> v = Vs[c * kWyLd + r];  // B(k = r, n = c)
                    wr = __builtin_amdgcn_mfma_f64_16x16x4f64(y.re, v.re, wr, 0, 0, 0);
                    wi = __builtin_amdgcn_mfma_f64_16x16x4f64(y.re, v.im, wi, 0, 0, 0);
                    wr = __builtin_amdgcn_mfma_f64_16x16x4f64(y.im, v.im, wr, 0, 0, 0);
                    wi = __builtin_amdgcn_mfma_f64_16x16x4f64(-y.im, v.re, wi, 0, 0, 0);
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) Ws[c * (kWyB + 1) + g + 4 * e] = cx<double>{wr[e], wi[e]};
        }
        __syncthreads();
        // (3) T, row t held by lane t of wave 0: T(t, i) = -tau_i sum_{l=t}^{i-1} T(t, l) P(l, i)
        if (wave == 0 && lane < kWyB) {
            cx<double> trow[kWyB];
#pragma unroll
            for (int i = 0; i < kWyB; ++i) {
                cx<double> acc = {0.0, 0.0};
#pragma unroll
                for (int l = 0; l < i; ++l)
                    if (l >= lane) acc = cmad(acc, trow[l], Ps[i * kWyB + l]);
                const cx<double> ti = Ts[i];
                trow[i] = lane < i ? cx<double>{0.0, 0.0} - ti * acc : (lane == i ? ti : cx<double>{0.0, 0.0});
            }
            // U = Y T is formed from these rows below; park T in Ps's transpose slot (P is done)
#pragma unroll
            for (int i = 0; i < kWyB; ++i) Ps[lane * kWyB + i] = trow[i];  // T(lane, i)
        }
        __syncthreads();
        // (4) U = Y T (64 x 16), four entries per thread
        for (int e = tid; e < 64 * kWyB; e += 256) {
            const int r = e >> 4, i = e & 15;
            cx<double> acc = {0.0, 0.0};
            for (int l = 0; l <= i; ++l) acc = cmad(acc, Ys[r * kWyB + l], Ps[l * kWyB + i]);
            Us[r * kWyB + i] = acc;
        }
        __syncthreads();
        // (5) V <- V - U W on the wave's 32 x 32 quadrant
        {
            const int m0 = 32 * (wave & 1), n0 = 32 * (wave >> 1);
            lqm_d4 vr[2][2], vi[2][2];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const cx<double> x = Vs[(n0 + 16 * bb + li) * kWyLd + m0 + 16 * a + g + 4 * e];
                        vr[a][bb][e] = x.re;
                        vi[a][bb][e] = x.im;
                    }
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int i = 4 * g + s;
                double ur[2], ui[2], wr[2], wi[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const cx<double> u = Us[(m0 + 16 * t + li) * kWyB + i];  // A(m, k = i) = -U
                    const cx<double> w = Ws[(n0 + 16 * t + li) * (kWyB + 1) + i];
                    ur[t] = -u.re;
                    ui[t] = -u.im;
                    wr[t] = w.re;
                    wi[t] = w.im;
                }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int bb = 0; bb < 2; ++bb) {
                        vr[a][bb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ur[a], wr[bb], vr[a][bb], 0, 0, 0);
                        vi[a][bb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ur[a], wi[bb], vi[a][bb], 0, 0, 0);
                    }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int bb = 0; bb < 2; ++bb) {
                        vr[a][bb] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ui[a], wi[bb], vr[a][bb], 0, 0, 0);
                        vi[a][bb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ui[a], wr[bb], vi[a][bb], 0, 0, 0);
                    }
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        Vs[(n0 + 16 * bb + li) * kWyLd + m0 + 16 * a + g + 4 * e] =
                            cx<double>{vr[a][bb][e], vi[a][bb][e]};
        }
    }
    __syncthreads();

    const R* ev = evals + (size_t)kk * N;
    const int* pm = perm + (size_t)kk * N;
    if constexpr (MODE == 0) {
        cx<double>* o = out + (size_t)kk * NN;
        for (int e = tid; e < N * N; e += 256) {
            const int i = e % N, n = e / N;
            const int c = min(max(pm[n], 0), N - 1);  // perm is only garbage after a flagged NaN
            o[e] = Vs[c * kWyLd + i];
        }
    } else {
        // occupations per Z column and the band-energy term
        for (int c = tid; c < 64; c += 256) sf[c] = 0.0;
        __syncthreads();
        double ebp = 0.0;
        for (int n = tid; n < N; n += 256) {
            const double en = (double)ev[n];
            const double f = occ ? occ[(size_t)kk * N + n] : fermidirac(en, mu, Temp);
            sf[min(max(pm[n], 0), N - 1)] = f;
            ebp += en * f;
        }
        ebp = wave_sum(ebp);
        if (lane == 0) sred[wave] = ebp;
        __syncthreads();
        if (tid == 0) eb[kk] = (kw ? kw[kk] : wuni) * (sred[0] + sred[1] + sred[2] + sred[3]);
        // rho_k = (V f) V^H on the wave's quadrant
        const int m0 = 32 * (wave & 1), n0 = 32 * (wave >> 1);
        lqm_d4 cr[2][2], ci[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int bb = 0; bb < 2; ++bb) {
                cr[a][bb] = lqm_d4{0.0, 0.0, 0.0, 0.0};
                ci[a][bb] = lqm_d4{0.0, 0.0, 0.0, 0.0};
            }
        for (int k0 = 0; k0 < 64; k0 += 16) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int c = k0 + 4 * g + s;
                const double f = sf[c];
                double ar[2], ai[2], br[2], bi[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const cx<double> x = Vs[c * kWyLd + m0 + 16 * t + li];  // A(m, k = c) = V f
                    const cx<double> y = Vs[c * kWyLd + n0 + 16 * t + li];  // B(k = c, n) = conj V(n, c)
                    ar[t] = f * x.re;
                    ai[t] = f * x.im;
                    br[t] = y.re;
                    bi[t] = -y.im;
                }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int bb = 0; bb < 2; ++bb) {
                        cr[a][bb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[bb], cr[a][bb], 0, 0, 0);
                        ci[a][bb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[bb], ci[a][bb], 0, 0, 0);
                    }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int bb = 0; bb < 2; ++bb) {
                        cr[a][bb] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[a], bi[bb], cr[a][bb], 0, 0, 0);
                        ci[a][bb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[bb], ci[a][bb], 0, 0, 0);
                    }
            }
        }
        cx<double>* o = out + (size_t)kk * NN;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int m = m0 + 16 * a + g + 4 * e, n = n0 + 16 * bb + li;
                    if (m < N && n < N) o[(size_t)n * N + m] = cx<double>{cr[a][bb][e], ci[a][bb][e]};
                }
    }
}

}  // namespace lqm
