// Back-transformation V = H_0 H_1 ... H_{N-2} Z on the FP64 matrix cores (N <= 64).
//
// The N - 1 Householder reflectors H_j = I - tau_j v_j v_j^H of the tridiagonalization are
// applied to the QL rotation product Z in blocks of 16 in compact WY form (LAPACK zlarft,
// forward, columnwise): Q_b = H_{16b} ... H_{16b+15} = I - Y T Y^H with T upper triangular,
// T(i,i) = tau_i, T(0:i, i) = -tau_i T(0:i, 0:i) P(0:i, i), P = Y^H Y.  The last block is
// applied first: V <- V - (Y T)(Y^H V).
//
// V never leaves the registers: wave (wr, wc) of the 256-thread block holds the 32 x 32
// quadrant V(32 wr + ., 32 wc + .) as matrix-core accumulators (lane l, element e of tile
// (a, b): row 32 wr + 16 a + l/16 + 4 e, column 32 wc + 16 b + l%16).  Y^H V is formed from
// those registers directly (contraction row r = 16 a + l/16 + 4 s in instruction s, which is
// the row the lane already holds), the two row-halves meeting in LDS; U = Y T and the update
// V -= U W are matrix-core products too.  The T factors come from k_wy_tfac (they depend on
// the reflectors only), so a block of reflectors costs three barriers and no serial section.
// LDS holds the block's Y, U, W and T (72 KB), so two k-points share a CU.  The reflectors
// and Z may be FP32 (mixed path) or FP64; all arithmetic is FP64.
//   MODE 0: write the eigenvectors, column n = V(:, perm[n]) (ascending eigenvalues);
//   MODE 1: rho_k = V diag(f) V^H with f = fermidirac(e - mu; T) (or the given occupations)
//           and eb[kk] = w_k sum_n e_n f_n, as k_backtransform MODE 1.
#pragma once
#include "lqm_common.hpp"
#include "refine_kernels.hpp"

namespace lqm {

constexpr int kWyB = 16;   // reflectors per block
constexpr int kWyLd = 65;  // leading dimension of V when it is staged for rho_k (MODE 1)

struct WySmem {
    cx<double> Y[64 * kWyB];           // Y(r, i) at r * 16 + i
    cx<double> U[64 * kWyB];           // U = Y T, same layout
    cx<double> W[2][64 * (kWyB + 1)];  // row-half partials of W = Y^H V: W(i, c) at c * 17 + i
    cx<double> T[kWyB * kWyB];         // T(row, col) at row * 16 + col
    int invp[64];                      // output column of each Z column (MODE 0)
    double sf[64];                     // occupation of each Z column (MODE 1)
    double red[8];
};

// T factors of every block of 16 reflectors, ahead of the back-transformation (T depends
// on the reflectors only): one wave per (block, k-point) forms P = Y^H Y on the matrix
// cores (K = the 64 rows, lane (li, g) supplying row 4 step + g of column li as both
// operands) and runs the zlarft recurrence T(t, i) = -tau_i sum_{l=t}^{i-1} T(t, l) P(l, i)
// with lane t holding row t in registers.  Out: Tg[(k * nb + b) * 256 + row * 16 + col] in
// FP64.
template <typename R>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_wy_tfac(int N, const cx<R>* __restrict__ refl,
                                                const cx<R>* __restrict__ tau, cx<double>* __restrict__ Tg) {
    __shared__ cx<double> P[kWyB * kWyB];  // P(p, q) at q * 16 + p
    const int lane = threadIdx.x, li = lane & 15, g = lane >> 4;
    const int b = blockIdx.x, nb = gridDim.x;
    const size_t kk = blockIdx.y;
    const int j0 = b * kWyB, nrefl = N - 1;
    const size_t NN = (size_t)N * N;
    const cx<R>* rf = refl + kk * NN;
    const int j = j0 + li;
    // the lane's 16 entries of Y, all loads in flight together (clamped addresses: every
    // load is in bounds; entries outside the reflector are zeroed below)
    cx<R> yv[16];
#pragma unroll
    for (int st = 0; st < 16; ++st) yv[st] = rf[(size_t)min(j, N - 1) * N + min(4 * st + g, N - 1)];
    // P = Y^H Y as three real products (P1 = ar br, P2 = ai bi, P3 = (ar + ai)(br + bi) with
    // a = conj y, b = y; re = P1 - P2, im = P3 - P1 - P2); rows below j0 + 1 are zero in every
    // column of the block, so the K-steps st < j0 / 4 are skipped (block-uniform)
    lqm_d4 p1 = {0.0, 0.0, 0.0, 0.0}, p2 = p1, p3 = p1;
    const int st0 = j0 >> 2;
#pragma unroll
    for (int st = 0; st < 16; ++st) {
        if (st < st0) continue;
        const int r = 4 * st + g;
        cx<double> y = {0.0, 0.0};
        if (j < nrefl && r > j && r < N) y = {(double)yv[st].re, (double)yv[st].im};
        // A(m = li, k = r) = conj(Y(r, li)), B(k = r, n = li) = Y(r, li)
        p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(y.re, y.re, p1, 0, 0, 0);
        p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(-y.im, y.im, p2, 0, 0, 0);
        p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(y.re - y.im, y.re + y.im, p3, 0, 0, 0);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) P[li * kWyB + g + 4 * e] = cx<double>{p1[e] - p2[e], p3[e] - p1[e] - p2[e]};
    __syncthreads();
    if (lane < kWyB) {
        // row `lane` of T in registers; T(t, l) = 0 for l < t, so every sum can start at l = 0
        cx<double> trow[kWyB];
#pragma unroll
        for (int i = 0; i < kWyB; ++i) {
            cx<double> a0 = {0.0, 0.0}, a1 = {0.0, 0.0};
#pragma unroll
            for (int l = 0; l < i; ++l) {
                if (l & 1) a1 = cmad(a1, trow[l], P[i * kWyB + l]);
                else a0 = cmad(a0, trow[l], P[i * kWyB + l]);
            }
            const cx<double> acc = a0 + a1;
            cx<double> ti = {0.0, 0.0};
            if (j0 + i < nrefl) {
                const cx<R> x = tau[kk * N + j0 + i];
                ti = {(double)x.re, (double)x.im};
            }
            trow[i] = lane < i ? cx<double>{0.0, 0.0} - ti * acc : (lane == i ? ti : cx<double>{0.0, 0.0});
        }
        cx<double>* To = Tg + (kk * nb + b) * (size_t)(kWyB * kWyB) + lane * kWyB;
#pragma unroll
        for (int c = 0; c < kWyB; ++c) To[c] = trow[c];
    }
}
inline size_t wy_lds_bytes() {
    const size_t v = sizeof(cx<double>) * 64 * kWyLd;  // staged V of MODE 1 overlays Y, U, W
    return sizeof(WySmem) > v ? sizeof(WySmem) : v;
}

template <typename R, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_backtransform_wy(
    int N, const cx<R>* __restrict__ refl, const cx<R>* __restrict__ tau, const R* __restrict__ Zg,
    const R* __restrict__ evals, const int* __restrict__ perm, const double* __restrict__ kw,
    double wuni, double mu, double Temp, const double* __restrict__ occ, cx<double>* __restrict__ out,
    double* __restrict__ eb, const cx<double>* __restrict__ Tg, const int* __restrict__ zfmt, int upper) {
    extern __shared__ __align__(16) unsigned char smem[];
    WySmem& S = *reinterpret_cast<WySmem*>(smem);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int wr = wave & 1, wc = wave >> 1;
    const int m0 = 32 * wr, n0 = 32 * wc;
    const int kk = blockIdx.x;
    const size_t NN = (size_t)N * N;
    const cx<R>* rf = refl + (size_t)kk * NN;

    // V = Z (real) in the accumulator layout, zero padded to 64 x 64; Z column-major, or
    // row-major where zfmt[kk] = 1
    lqm_d4 vr[2][2], vi[2][2];
    {
        const R* Z = Zg + (size_t)kk * NN;
        const bool zrow = zfmt && zfmt[kk];
        // all 16 loads first (clamped addresses: every load is in bounds), then the padding zeroed
        R zv[2][2][4];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int rc = min(m0 + 16 * a + g + 4 * e, N - 1), cc = min(n0 + 16 * b + li, N - 1);
                    zv[a][b][e] = Z[zrow ? (size_t)rc * N + cc : (size_t)cc * N + rc];
                }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int r = m0 + 16 * a + g + 4 * e, c = n0 + 16 * b + li;
                    vr[a][b][e] = (r < N && c < N) ? (double)zv[a][b][e] : 0.0;
                    vi[a][b][e] = 0.0;
                }
    }
    const int nrefl = N - 1;
    const int nb = (nrefl + kWyB - 1) / kWyB;
    // the reflectors Y(r, i) = v_{j0+i}(r) (zero above row j + 1) and the T factor of block b,
    // held in registers one block ahead: thread tid holds Y entries e = tid + 256 u
    cx<double> ypre[4], tpre;
    auto fetch = [&](int b) {
        const int j0 = b * kWyB;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            const int r = e >> 4, j = j0 + (e & 15);
            const cx<R> v = rf[(size_t)min(j, N - 1) * N + min(r, N - 1)];
            ypre[u] = (j < nrefl && r > j && r < N) ? cx<double>{(double)v.re, (double)v.im} : cx<double>{0.0, 0.0};
        }
        tpre = Tg[((size_t)kk * nb + b) * (kWyB * kWyB) + tid];  // 256 threads, 256 entries
    };
    if (nb > 0) fetch(nb - 1);
    for (int b0 = nb - 1; b0 >= 0; --b0) {
        // (1) the block's reflectors and T factor into LDS; the next block's are fetched
#pragma unroll
        for (int u = 0; u < 4; ++u) S.Y[tid + 256 * u] = ypre[u];
        S.T[tid] = tpre;
        __syncthreads();
        if (b0 > 0) fetch(b0 - 1);
        // (2) row-half partials W_wr = Y(R_wr, :)^H V(R_wr, C_wc) (B operand = the lane's own
        //     accumulators), and U = Y T (wave w: rows 16w .. 16w + 15, one 16 x 16 tile, K = 16);
        //     complex products as three real ones (P1 = ar br, P2 = ai bi,
        //     P3 = (ar + ai)(br + bi); re = P1 - P2, im = P3 - P1 - P2)
        {
            lqm_d4 p1[2], p2[2], p3[2];
#pragma unroll
            for (int b = 0; b < 2; ++b) p1[b] = p2[b] = p3[b] = lqm_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int r = m0 + 16 * a + g + 4 * s;
                    const cx<double> y = S.Y[r * kWyB + li];  // A(m = li, k = r) = conj(Y(r, li))
                    const double ar = y.re, ai = -y.im, as = y.re - y.im;
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        const double br = vr[a][b][s], bi = vi[a][b][s];
                        p1[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, p1[b], 0, 0, 0);
                        p2[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, p2[b], 0, 0, 0);
                        p3[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(as, br + bi, p3[b], 0, 0, 0);
                    }
                }
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    S.W[wr][(n0 + 16 * b + li) * (kWyB + 1) + g + 4 * e] =
                        cx<double>{p1[b][e] - p2[b][e], p3[b][e] - p1[b][e] - p2[b][e]};
            lqm_d4 q1 = {0.0, 0.0, 0.0, 0.0}, q2 = q1, q3 = q1;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int l = 4 * g + s;
                const cx<double> y = S.Y[(16 * wave + li) * kWyB + l];  // A(m = li, k = l)
                const cx<double> t = S.T[l * kWyB + li];                // B(k = l, n = li) = T(l, li)
                q1 = __builtin_amdgcn_mfma_f64_16x16x4f64(y.re, t.re, q1, 0, 0, 0);
                q2 = __builtin_amdgcn_mfma_f64_16x16x4f64(y.im, t.im, q2, 0, 0, 0);
                q3 = __builtin_amdgcn_mfma_f64_16x16x4f64(y.re + y.im, t.re + t.im, q3, 0, 0, 0);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e)
                S.U[(16 * wave + g + 4 * e) * kWyB + li] = cx<double>{q1[e] - q2[e], q3[e] - q1[e] - q2[e]};
        }
        __syncthreads();
        // (3) V <- V - U W, W = W_0 + W_1, on the wave's quadrant
        {
            lqm_d4 p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) p1[a][b] = p2[a][b] = p3[a][b] = lqm_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int i = 4 * g + s;
                double ur[2], ui[2], us[2], wre[2], wim[2], ws[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const cx<double> u = S.U[(m0 + 16 * t + li) * kWyB + i];  // A(m, k = i) = -U
                    const cx<double> w = S.W[0][(n0 + 16 * t + li) * (kWyB + 1) + i] +
                                          S.W[1][(n0 + 16 * t + li) * (kWyB + 1) + i];
                    ur[t] = -u.re;
                    ui[t] = -u.im;
                    us[t] = ur[t] + ui[t];
                    wre[t] = w.re;
                    wim[t] = w.im;
                    ws[t] = w.re + w.im;
                }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        p1[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ur[a], wre[b], p1[a][b], 0, 0, 0);
                        p2[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ui[a], wim[b], p2[a][b], 0, 0, 0);
                        p3[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(us[a], ws[b], p3[a][b], 0, 0, 0);
                    }
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    vr[a][b] += p1[a][b] - p2[a][b];
                    vi[a][b] += p3[a][b] - p1[a][b] - p2[a][b];
                    // keeps the update here rather than sunk into the next block, where the
                    // three products would stay live beside V
                    asm volatile("" : "+v"(vr[a][b]), "+v"(vi[a][b]));
                }
        }
        __syncthreads();  // Y, U, W are rewritten by the next block
    }

    const R* ev = evals + (size_t)kk * N;
    const int* pm = perm + (size_t)kk * N;
    if constexpr (MODE == 0) {
        for (int n = tid; n < 64; n += 256) S.invp[n] = -1;
        __syncthreads();
        for (int n = tid; n < N; n += 256) S.invp[min(max(pm[n], 0), N - 1)] = n;  // perm is a permutation
        __syncthreads();
        cx<double>* o = out + (size_t)kk * NN;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int r = m0 + 16 * a + g + 4 * e, c = n0 + 16 * b + li;
                    if (r < N && c < N) {
                        const int n = S.invp[c];
                        if (n >= 0) o[(size_t)n * N + r] = cx<double>{vr[a][b][e], vi[a][b][e]};
                    }
                }
    } else {
        // stage V for the product V f V^H (overlays Y, U, W), occupations per Z column
        cx<double>* Vs = reinterpret_cast<cx<double>*>(smem);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    Vs[(n0 + 16 * b + li) * kWyLd + m0 + 16 * a + g + 4 * e] = cx<double>{vr[a][b][e], vi[a][b][e]};
        for (int c = tid; c < 64; c += 256) S.sf[c] = 0.0;
        __syncthreads();
        double ebp = 0.0;
        for (int n = tid; n < N; n += 256) {
            const double en = (double)ev[n];
            const double f = occ ? occ[(size_t)kk * N + n] : fermidirac(en, mu, Temp);
            S.sf[min(max(pm[n], 0), N - 1)] = f;
            ebp += en * f;
        }
        ebp = wave_sum(ebp);
        if (lane == 0) S.red[wave] = ebp;
        __syncthreads();
        if (tid == 0) eb[kk] = (kw ? kw[kk] : wuni) * (S.red[0] + S.red[1] + S.red[2] + S.red[3]);
        // columns whose occupation enters the product (the rest are below kOccNegligible)
        const unsigned long long live = __ballot(!(fabs(S.sf[lane]) <= kOccNegligible));
        // 16 x 16 tiles this wave forms: all four, or with `upper` (the caller reads only the
        // elements on and above the diagonal) those whose first row is not below their first
        // column -- every other tile lies wholly below the diagonal
        bool on[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                on[a][b] = !upper || __builtin_amdgcn_readfirstlane(m0) + 16 * a <= __builtin_amdgcn_readfirstlane(n0) + 16 * b;
        lqm_d4 p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) p1[a][b] = p2[a][b] = p3[a][b] = lqm_d4{0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < 64; k0 += 4) {
            if (((live >> k0) & 0xFull) == 0) continue;  // 4 columns of negligible occupation
            {
                const int c = k0 + g;
                const double f = S.sf[c];
                double ar[2], ai[2], as[2], br[2], bi[2], bs[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const cx<double> x = Vs[c * kWyLd + m0 + 16 * t + li];  // A(m, k = c) = V f
                    const cx<double> y = Vs[c * kWyLd + n0 + 16 * t + li];  // B(k = c, n) = conj V(n, c)
                    ar[t] = f * x.re;
                    ai[t] = f * x.im;
                    as[t] = ar[t] + ai[t];
                    br[t] = y.re;
                    bi[t] = -y.im;
                    bs[t] = y.re - y.im;
                }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
                        if (on[a][b]) {
                            p1[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], p1[a][b], 0, 0, 0);
                            p2[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], bi[b], p2[a][b], 0, 0, 0);
                            p3[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[a], bs[b], p3[a][b], 0, 0, 0);
                        }
            }
        }
        lqm_d4 cr[2][2], ci[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                cr[a][b] = p1[a][b] - p2[a][b];
                ci[a][b] = p3[a][b] - p1[a][b] - p2[a][b];
            }
        cx<double>* o = out + (size_t)kk * NN;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int m = m0 + 16 * a + g + 4 * e, n = n0 + 16 * b + li;
                    if (on[a][b] && m < N && n < N) o[(size_t)n * N + m] = cx<double>{cr[a][b][e], ci[a][b][e]};
                }
    }
}


// ---------------------------------------------------------------------------------
// FP32 form of the back-transformation for the mixed path (MODE 0: eigenvectors only):
// the same blocks of 16 reflectors with V, Y, U, W and T in FP32 on v_mfma_f32_16x16x4f32
// (twice the FP64 matrix-core rate; the Ogita-Aishima refinement that follows makes the
// vectors FP64-accurate).  Lane layout of that instruction: A(m, k) from lane
// (m, k-slot) = (l % 16, l / 16), B(k, n) from lane (n, k-slot), and element e of the
// accumulator holds row 4 (l / 16) + e, column l % 16 -- so in K-step s lane (li, g)
// supplies contraction index 4 g + s, the row its accumulator element s already holds,
// and V feeds Y^H V straight from the registers.  T (computed in FP64 by k_wy_tfac) is
// rounded to FP32.  The result is staged through LDS 32 columns at a time and written as
// FP64 with whole columns contiguous.
// ---------------------------------------------------------------------------------
typedef float lqm_f4 __attribute__((ext_vector_type(4)));
constexpr int kWy32Ld = 65;  // leading dimension of a staged output half (FP64 complex)
__host__ __device__ constexpr size_t wy32_region() {
    // Y, U (64 x 16 each), W[2] (64 x 17), T (16 x 16) in FP32 complex; the staged half
    // of X (32 x 65 FP64 complex) overlays them
    return (sizeof(cx<float>) * (2 * 64 * kWyB + 2 * 64 * (kWyB + 1) + kWyB * kWyB)) >
                   sizeof(cx<double>) * 32 * kWy32Ld
               ? sizeof(cx<float>) * (2 * 64 * kWyB + 2 * 64 * (kWyB + 1) + kWyB * kWyB)
               : sizeof(cx<double>) * 32 * kWy32Ld;
}
inline size_t wy32_lds_bytes() { return wy32_region() + 64 * sizeof(int); }

// OT: the element type of the eigenvectors written (cx<double>, or cx<float> for a reader
// that widens them itself -- the values are FP32 either way)
template <typename OT>
__global__ __launch_bounds__(256) void k_backtransform_wy32(int N, const cx<float>* __restrict__ refl,
                                                            const float* __restrict__ Zg,
                                                            const int* __restrict__ perm,
                                                            OT* __restrict__ out,
                                                            const cx<double>* __restrict__ Tg,
                                                            const int* __restrict__ zfmt) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<float>* Ys = reinterpret_cast<cx<float>*>(smem);  // Y(r, i) at r * 16 + i
    cx<float>* Us = Ys + 64 * kWyB;                       // U = Y T, same layout
    cx<float>* Ws = Us + 64 * kWyB;                       // [2][c * 17 + i]: row-half partials of W
    cx<float>* Ts = Ws + 2 * 64 * (kWyB + 1);             // T(row, col) at row * 16 + col
    int* invp = reinterpret_cast<int*>(smem + wy32_region());
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int wr = wave & 1, wc = wave >> 1;
    const int m0 = 32 * wr, n0 = 32 * wc;
    const int kk = blockIdx.x;
    const size_t NN = (size_t)N * N;
    const cx<float>* rf = refl + (size_t)kk * NN;

    // V = Z in the accumulator layout (row m0 + 16 a + 4 g + e, column n0 + 16 b + li)
    lqm_f4 vr[2][2], vi[2][2];
    {
        const float* Z = Zg + (size_t)kk * NN;
        const bool zrow = zfmt && zfmt[kk];
        // all 16 loads first (clamped addresses: every load is in bounds), then the padding zeroed
        float zv[2][2][4];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int rc = min(m0 + 16 * a + 4 * g + e, N - 1), cc = min(n0 + 16 * b + li, N - 1);
                    zv[a][b][e] = Z[zrow ? (size_t)rc * N + cc : (size_t)cc * N + rc];
                }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int r = m0 + 16 * a + 4 * g + e, c = n0 + 16 * b + li;
                    vr[a][b][e] = (r < N && c < N) ? zv[a][b][e] : 0.f;
                    vi[a][b][e] = 0.f;
                }
    }
    const int nrefl = N - 1;
    const int nb = (nrefl + kWyB - 1) / kWyB;
    // the reflectors (zero above row j + 1) and the T factor of block b, held in registers one
    // block ahead: thread tid holds Y entries e = tid + 256 u (row e / 16, column e % 16)
    cx<float> ypre[4];
    cx<double> tpre;
    auto fetch = [&](int b) {
        const int j0 = b * kWyB;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            const int r = e >> 4, j = j0 + (e & 15);
            const cx<float> y = rf[(size_t)min(j, N - 1) * N + min(r, N - 1)];
            ypre[u] = (j < nrefl && r > j && r < N) ? y : cx<float>{0.f, 0.f};
        }
        tpre = Tg[((size_t)kk * nb + b) * (kWyB * kWyB) + tid];
    };
    if (nb > 0) fetch(nb - 1);
    for (int b0 = nb - 1; b0 >= 0; --b0) {
        // (1) the block's reflectors and its T factor into LDS; the next block's are fetched
#pragma unroll
        for (int u = 0; u < 4; ++u) Ys[tid + 256 * u] = ypre[u];
        Ts[tid] = cx<float>{(float)tpre.re, (float)tpre.im};
        __syncthreads();
        if (b0 > 0) fetch(b0 - 1);
        // (2) W_wr = Y(R_wr, :)^H V(R_wr, C_wc) and U = Y T (wave w: rows 16 w .. 16 w + 15);
        //     complex products as three real ones (P1 = ar br, P2 = ai bi,
        //     P3 = (ar + ai)(br + bi); re = P1 - P2, im = P3 - P1 - P2)
        {
            lqm_f4 p1[2], p2[2], p3[2];
#pragma unroll
            for (int b = 0; b < 2; ++b) p1[b] = p2[b] = p3[b] = lqm_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int r = m0 + 16 * a + 4 * g + s;
                    const cx<float> y = Ys[r * kWyB + li];  // A(m = li, k) = conj(Y(r, li))
                    const float ar = y.re, ai = -y.im, as = y.re - y.im;
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        const float br = vr[a][b][s], bi = vi[a][b][s];
                        p1[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, p1[b], 0, 0, 0);
                        p2[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, bi, p2[b], 0, 0, 0);
                        p3[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(as, br + bi, p3[b], 0, 0, 0);
                    }
                }
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    Ws[wr * 64 * (kWyB + 1) + (n0 + 16 * b + li) * (kWyB + 1) + 4 * g + e] =
                        cx<float>{p1[b][e] - p2[b][e], p3[b][e] - p1[b][e] - p2[b][e]};
            lqm_f4 q1 = {0.f, 0.f, 0.f, 0.f}, q2 = q1, q3 = q1;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int l = 4 * g + s;
                const cx<float> y = Ys[(16 * wave + li) * kWyB + l];  // A(m = li, k = l)
                const cx<float> t = Ts[l * kWyB + li];                // B(k = l, n = li)
                q1 = __builtin_amdgcn_mfma_f32_16x16x4f32(y.re, t.re, q1, 0, 0, 0);
                q2 = __builtin_amdgcn_mfma_f32_16x16x4f32(y.im, t.im, q2, 0, 0, 0);
                q3 = __builtin_amdgcn_mfma_f32_16x16x4f32(y.re + y.im, t.re + t.im, q3, 0, 0, 0);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e)
                Us[(16 * wave + 4 * g + e) * kWyB + li] = cx<float>{q1[e] - q2[e], q3[e] - q1[e] - q2[e]};
        }
        __syncthreads();
        // (3) V <- V - U W, W = W_0 + W_1
        {
            lqm_f4 p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) p1[a][b] = p2[a][b] = p3[a][b] = lqm_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int i = 4 * g + s;
                float ur[2], ui[2], us[2], wre[2], wim[2], ws[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const cx<float> u = Us[(m0 + 16 * t + li) * kWyB + i];  // A(m, k = i) = -U
                    const cx<float> w0 = Ws[(n0 + 16 * t + li) * (kWyB + 1) + i];
                    const cx<float> w1 = Ws[64 * (kWyB + 1) + (n0 + 16 * t + li) * (kWyB + 1) + i];
                    ur[t] = -u.re;
                    ui[t] = -u.im;
                    us[t] = ur[t] + ui[t];
                    wre[t] = w0.re + w1.re;
                    wim[t] = w0.im + w1.im;
                    ws[t] = wre[t] + wim[t];
                }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        p1[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ur[a], wre[b], p1[a][b], 0, 0, 0);
                        p2[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ui[a], wim[b], p2[a][b], 0, 0, 0);
                        p3[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(us[a], ws[b], p3[a][b], 0, 0, 0);
                    }
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    vr[a][b] += p1[a][b] - p2[a][b];
                    vi[a][b] += p3[a][b] - p1[a][b] - p2[a][b];
                }
        }
        __syncthreads();  // Y, U, W, T are rewritten by the next block
    }
    // eigenvector n = V(:, perm[n]): output column of each V column
    if (tid < 64) invp[tid] = -1;
    __syncthreads();
    for (int n = tid; n < N; n += 256) invp[min(max(perm[(size_t)kk * N + n], 0), N - 1)] = n;
    __syncthreads();
    OT* Xs = reinterpret_cast<OT*>(smem);  // overlays Y, U, W, T
    OT* o = out + (size_t)kk * NN;
    for (int h = 0; h < 2; ++h) {
        if (wc == h) {
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        Xs[(16 * b + li) * kWy32Ld + m0 + 16 * a + 4 * g + e] = OT{vr[a][b][e], vi[a][b][e]};
        }
        __syncthreads();
        for (int idx = tid; idx < 32 * 64; idx += 256) {
            const int cl = idx >> 6, r = idx & 63, c = 32 * h + cl;
            if (c < N && r < N) {
                const int n = invp[c];
                if (n >= 0) o[(size_t)n * N + r] = Xs[cl * kWy32Ld + r];
            }
        }
        __syncthreads();
    }
}

}  // namespace lqm
