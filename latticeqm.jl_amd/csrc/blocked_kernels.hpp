// Large-N Hermitian eigensolver (FP64) over a batch of k-points, one k-point per batch
// entry, all matrices column-major with leading dimension N:
//   k_hk_full        A = Hermitian(H(k)) (upper triangle of H(k)), full storage
//   k_trd_col        zlatrd column step j: pending panel update of column j, zlarfg
//   k_trd_hemv       y = A22 v from the lower triangle, W^H v, V^H v, v^H A22 v partials
//   k_trd_w          w = tau (A22 v - V (W^H v) - W (V^H v)) + alpha v (zlatrd)
//   k_zgemm_gen      batched C = alpha sum_p op(A_p) op(B_p) + beta C on the FP64 matrix
//                    cores: the trailing her2k update, compact-WY products, rho_k
//   k_sturm_bisect   eigenvalue n of the tridiagonal by Sturm-count bisection
//   k_invit          eigenvectors of the tridiagonal by inverse iteration, one thread each
//   k_invit_orth     orthonormalisation of the vectors of each eigenvalue cluster
//   k_make_y, k_wy_T, k_z_to_v, k_occ_eb   compact-WY and rho_k helpers
//
// The tridiagonalization is the one-stage blocked Householder reduction (LAPACK zhetrd,
// lower, with zlatrd panels of kBnb columns): inside a panel the trailing matrix is left
// as it was at the panel start and the pending rank-2i update enters every column step
// through V and W; after the panel A22 -= V W^H + W V^H on the matrix cores (tiles on or
// below the diagonal: only the lower triangle of A is kept up to date and read).  The
// reflectors H_j = I - tau_j v_j v_j^H (v_j(j+1) = 1) give A = Q T Q^H, Q = H_0 ... H_{N-2}.
#pragma once
#include "lqm_common.hpp"
#include "eig_kernels.hpp"

namespace lqm {

constexpr int kBnb = 32;      // tridiagonalization panel width
constexpr int kBwy = 64;      // reflectors per compact-WY block of the back-transformation
constexpr int kHvCols = 64;   // column chunk of the trailing matrix-vector product
constexpr int kBColT = 256;   // threads of the column-step kernel (one block per k-point)
constexpr int kBRows = 16;    // rows per thread of the column step (N <= 4096)
constexpr int kBMaxN = kBColT * kBRows;
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

template <int W>
__device__ __forceinline__ double blk_max(double v, double* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) s = fmax(s, red[w]);
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
template <typename R>
__global__ __launch_bounds__(kBColT) void k_trd_col(int N, int j, int i, int last, cx<R>* __restrict__ A,
                                                 cx<R>* __restrict__ V,
                                                 const cx<R>* __restrict__ W,
                                                 double* __restrict__ dd, double* __restrict__ ee,
                                                 cx<R>* __restrict__ tau) {
    __shared__ cx<R> wj[kBnb], vj[kBnb];
    __shared__ double red[kBColT / 64];
    __shared__ cx<R> s_alpha;
    __shared__ R s_dj;
    const int tid = threadIdx.x;
    const size_t kk = blockIdx.x;
    const size_t NN = (size_t)N * N, NP = (size_t)N * kBnb;
    cx<R>* Ak = A + kk * NN;
    cx<R>* Vk = V + kk * NP;
    const cx<R>* Wk = W + kk * NP;
    const cx<R> czero = {R(0), R(0)};
    if (tid < i) {
        wj[tid] = Wk[(size_t)tid * N + j];
        vj[tid] = Vk[(size_t)tid * N + j];
    }
    if (tid == 0) s_alpha = czero;
    __syncthreads();
    cx<R> a[kBRows];
    double xn = 0.0;
#pragma unroll
    for (int q = 0; q < kBRows; ++q) {
        const int r = j + tid + kBColT * q;
        cx<R> x = czero;
        if (r < N) {
            x = Ak[(size_t)j * N + r];
            cx<R> u = czero;
            for (int t = 0; t < i; ++t) {
                u = cmadcb(u, Vk[(size_t)t * N + r], wj[t]);
                u = cmadcb(u, Wk[(size_t)t * N + r], vj[t]);
            }
            x = x - u;
            if (r >= j + 2) xn += (double)abs2(x);
        }
        a[q] = x;
    }
    if (tid == 0) s_dj = a[0].re;
    if (tid == 1 && j + 1 < N) s_alpha = a[0];
    const double xnorm2 = blk_sum<kBColT / 64>(xn, red);  // (its barriers publish s_dj, s_alpha)
    const R dj = s_dj;
    if (last) {
        if (tid == 0) {
            Ak[(size_t)j * N + j] = {dj, R(0)};
            dd[kk * N + j] = (double)dj;
            ee[kk * N + j] = 0.0;
            tau[kk * N + j] = czero;
        }
        return;
    }
    const cx<R> alpha = s_alpha;
    cx<R> tj = czero, scale = czero;
    double beta = (double)alpha.re;
    if (!(xnorm2 == 0.0 && alpha.im == R(0))) {
        const double are = alpha.re, aim = alpha.im;
        const double nrm = sqrt(are * are + aim * aim + xnorm2);
        beta = are >= 0.0 ? -nrm : nrm;
        tj = {(R)((beta - are) / beta), (R)(-aim / beta)};
        const double dre = are - beta, den = dre * dre + aim * aim;
        if (den >= num<double>::rot_lo) {
            scale = {(R)(dre / den), (R)(-aim / den)};
        } else {
            // a column of tiny entries (block-uniform branch): the scalars of x / max|x|, then
            // beta and scale back
            double mx = 0.0;
#pragma unroll
            for (int q = 0; q < kBRows; ++q) {
                const int r = j + tid + kBColT * q;
                if (r > j && r < N) mx = fmax(mx, fmax(fabs((double)a[q].re), fabs((double)a[q].im)));
            }
            mx = blk_max<kBColT / 64>(mx, red);
            const double inv = 1.0 / mx;
            double xs = 0.0;
#pragma unroll
            for (int q = 0; q < kBRows; ++q) {
                const int r = j + tid + kBColT * q;
                if (r >= j + 2 && r < N) {
                    const double xr = inv * a[q].re, xi = inv * a[q].im;
                    xs += xr * xr + xi * xi;
                }
            }
            xs = blk_sum<kBColT / 64>(xs, red);
            cx<double> ts, ss;
            double bs;
            hh_scalars_scaled(cx<double>{inv * are, inv * aim}, xs, ts, ss, bs);
            if (inv <= num<double>::huge) {
                beta = bs * mx;
                tj = {(R)ts.re, (R)ts.im};
                scale = {(R)(inv * ss.re), (R)(inv * ss.im)};
            } else {
                // entries below 1 / huge (subnormal): left in place, H_j = I
                beta = are;
                tj = czero;
            }
        }
    }
    const R rbeta = (R)beta;
#pragma unroll
    for (int q = 0; q < kBRows; ++q) {
        const int r = j + tid + kBColT * q;
        if (r < N) {
            if (r == j) {
                Ak[(size_t)j * N + j] = {dj, R(0)};
            } else if (r == j + 1) {
                Ak[(size_t)j * N + r] = {rbeta, R(0)};
                Vk[(size_t)i * N + r] = {R(1), R(0)};
            } else {
                const cx<R> v = a[q] * scale;
                Ak[(size_t)j * N + r] = v;
                Vk[(size_t)i * N + r] = v;
            }
        }
    }
    if (tid == 0) {
        dd[kk * N + j] = (double)dj;
        ee[kk * N + j] = (double)rbeta;
        tau[kk * N + j] = tj;
    }
}

// Products of the column step j over the lower triangle of A22 = A(j+1:N, j+1:N) (only
// elements on or below the diagonal are read; 64 x 64 tiles from row/column j+1).  Tile
// column J is cut into parts of at most kHvPart tiles (I = J + kHvPart p ...), one
// 256-thread block per (J, p, k); lane = row of the tile, wave w = columns 16w .. 16w + 15
// of the tile column.  A wave streams its 16-column slab of each tile of the part with no
// barrier in between (16 loads per lane in flight, the slab's row sums parked in LDS), the
// four waves' row sums are added once at the end:
//   P[k][J][r]       = sum_{c in J, c <= r} A(r, c) v_c           (r in the part's tiles)
//   P2[k][p][c]      = sum_{r > c, r in the part} conj(A(r, c)) v_r   (c in J)
//   SS[k][b]         = the block's share of s = v^H A22 v (real)
//   Q[k][J][1 + t], Q[k][J][1 + kBnb + t] (part 0): W(:, t)^H v and V(:, t)^H v over the
//   rows of tile J (t < i).
constexpr int kQStride = 1 + 2 * kBnb;
constexpr int kHvPart = 8;
constexpr int kHvWaves = 4;
constexpr int kHvCpw = kHvCols / kHvWaves;  // columns per wave
__host__ __device__ __forceinline__ int hv_parts(int nt, int J) { return (nt - J + kHvPart - 1) / kHvPart; }
// number of (J, p) blocks of a column step with nt tile columns
__host__ __device__ __forceinline__ int hv_blocks(int nt) {
    int b = 0;
    for (int J = 0; J < nt; ++J) b += hv_parts(nt, J);
    return b;
}
template <typename R>
__global__ __launch_bounds__(kHvWaves * 64) __attribute__((amdgpu_waves_per_eu(3))) void k_trd_hemv(int N, int j, int i, int nchmax, int nbmax,
                                                  const cx<R>* __restrict__ A,
                                                  const cx<R>* __restrict__ V,
                                                  const cx<R>* __restrict__ W,
                                                  cx<R>* __restrict__ P, cx<R>* __restrict__ P2,
                                                  cx<R>* __restrict__ Q, double* __restrict__ SS) {
    __shared__ cx<R> sy[kHvPart][kHvWaves][64];
    __shared__ double sred[kHvWaves];
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const size_t kk = blockIdx.y;
    const int n = N - j - 1;
    const int nt = (n + kHvCols - 1) / kHvCols;
    // block -> (J, p)
    int J = 0, p = blockIdx.x;
    while (p >= hv_parts(nt, J)) {
        p -= hv_parts(nt, J);
        ++J;
    }
    const int I0 = J + kHvPart * p, I1 = min(nt, I0 + kHvPart);
    const int c0 = j + 1 + J * kHvCols;
    const int cw = c0 + kHvCpw * wave;  // this wave's first column
    const size_t NP = (size_t)N * kBnb;
    const cx<R>* Vk = V + kk * NP;
    const cx<R>* Wk = W + kk * NP;
    const cx<R>* v = Vk + (size_t)i * N;
    const cx<R>* Ak = A + kk * (size_t)N * N;
    cx<R> vc[kHvCpw];  // v over the wave's columns (wave-uniform)
#pragma unroll
    for (int u = 0; u < kHvCpw; ++u) vc[u] = cw + u < N ? v[cw + u] : cx<R>{R(0), R(0)};
    cx<R> ycol[kHvCpw];
#pragma unroll
    for (int u = 0; u < kHvCpw; ++u) ycol[u] = {R(0), R(0)};
#pragma nounroll
    for (int I = I0; I < I1; ++I) {
        const int r = j + 1 + I * kHvCols + lane;
        const bool ok = r < N;
        const int rr = ok ? r : N - 1;  // loads stay in bounds and unconditional (batched)
        cx<R> a[kHvCpw];
#pragma unroll
        for (int u = 0; u < kHvCpw; ++u) {
            const int c = min(cw + u, N - 1);
            const cx<R>* col = Ak + (size_t)c * N;  // wave-uniform base
            a[u] = col[max(rr, c)];  // above the diagonal: a row the wave reads anyway (masked)
        }
        const cx<R> vr = ok ? v[r] : cx<R>{R(0), R(0)};
        cx<R> yrow = {R(0), R(0)};
#pragma unroll
        for (int u = 0; u < kHvCpw; ++u) {
            const int c = cw + u;
            const bool in = ok && r >= c;  // r < N and c <= r: c < N
            const cx<R> x = in ? a[u] : cx<R>{R(0), R(0)};
            yrow = cmad(yrow, x, vc[u]);
            const cx<R> xt = (in && r > c) ? x : cx<R>{R(0), R(0)};
            ycol[u] = cmadc(ycol[u], xt, vr);
        }
        sy[I - I0][wave][lane] = yrow;
    }
    __syncthreads();
    double spart = 0.0;
    const int rows = (I1 - I0) * 64;
    for (int e = tid; e < rows; e += kHvWaves * 64) {
        const int t = e >> 6, l = e & 63;
        const int r = j + 1 + (I0 + t) * kHvCols + l;
        if (r < N) {
            cx<R> y = sy[t][0][l];
#pragma unroll
            for (int w = 1; w < kHvWaves; ++w) y = y + sy[t][w][l];
            P[kk * (size_t)nchmax * N + (size_t)J * N + r] = y;
            const cx<R> vr = v[r];
            spart += (double)(vr.re * y.re + vr.im * y.im);  // Re(conj(v_r) y_r)
        }
    }
    {
        static_assert(kHvCpw == 16, "column sums by wave_sum16");
        const cx<R> y = wave_sum16(ycol, lane);
        const int c = cw + wave_sum16_index(lane);
        if ((lane & 3) == 0 && c < N) {
            P2[kk * (size_t)nchmax * N + (size_t)p * N + c] = y;
            const cx<R> vcol = v[c];
            spart += (double)(vcol.re * y.re + vcol.im * y.im);
        }
    }
    if (p == 0) {
        // W^H v (even products) and V^H v (odd) over the rows of tile J
        cx<R>* Qk = Q + (kk * nchmax + J) * kQStride;
        const int r = c0 + lane;
        const cx<R> vr = r < N ? v[r] : cx<R>{R(0), R(0)};
        const int rr = r < N ? r : N - 1;
        // the 2i <= 2 kBnb - 2 products, kHvCpw per wave: every load issued before the sums
        static_assert(2 * kBnb <= kHvCpw * kHvWaves, "one round of products");
        cx<R> x[kHvCpw];
#pragma unroll
        for (int t = 0; t < kHvCpw; ++t) {
            const int idx = max(0, min(wave + kHvWaves * t, 2 * i - 1));
            x[t] = ((idx & 1) ? Vk : Wk)[(size_t)(idx >> 1) * N + rr];
        }
#pragma unroll
        for (int t = 0; t < kHvCpw; ++t) {
            const cx<R> z = {R(0), R(0)};
            x[t] = cmadc(z, r < N ? x[t] : z, vr);
        }
        const cx<R> acc = wave_sum16(x, lane);
        const int idx = wave + kHvWaves * wave_sum16_index(lane);
        if ((lane & 3) == 0 && idx < 2 * i) Qk[1 + (idx & 1) * kBnb + (idx >> 1)] = acc;
    }
    const double st = blk_sum<kHvWaves>(spart, sred);
    if (tid == 0) SS[kk * nbmax + blockIdx.x] = st;
}

// w = tau (y - V t1 - W t2) + alpha v into W(:, i), rows j+1..N-1 (zlatrd), from the
// products of k_trd_hemv: y_r = sum_{J <= tile(r)} P[k][J][r] + sum_p P2[k][p][r],
// t1 = W^H v, t2 = V^H v and s = v^H A22 v summed over the blocks, and
// alpha = -1/2 tau (w^H v) = -1/2 |tau|^2 (s - 2 Re(t1^H t2)) (w^H v = conj(tau) v^H
// (A22 - V W^H - W V^H) v).  Block (x, k): rows j + 1 + 256x + t.
template <typename R>
__global__ __launch_bounds__(256) void k_trd_w(int N, int j, int i, int nchmax, int nbmax,
                                               const cx<R>* __restrict__ P,
                                               const cx<R>* __restrict__ P2,
                                               const cx<R>* __restrict__ Q,
                                               const double* __restrict__ SS,
                                               const cx<R>* __restrict__ V,
                                               cx<R>* __restrict__ W,
                                               const cx<R>* __restrict__ tau) {
    __shared__ cx<R> t12[kQStride];
    __shared__ double s_alpha;
    const int tid = threadIdx.x;
    const size_t kk = blockIdx.y;
    const int n = N - j - 1;
    const int nt = (n + kHvCols - 1) / kHvCols;
    const int nb = hv_blocks(nt);
    const size_t NP = (size_t)N * kBnb;
    const cx<R>* Vk = V + kk * NP;
    cx<R>* Wk = W + kk * NP;
    const cx<R>* v = Vk + (size_t)i * N;
    const cx<R>* Qk = Q + kk * nchmax * kQStride;
    const cx<R> tj = tau[kk * N + j];
    if (tid < kQStride) {
        // the block partials in batches of 8 loads (indices clamped, extra terms masked)
        cx<R> acc = {0.0, 0.0};
        if (tid == 0) {
            for (int b0 = 0; b0 < nb; b0 += 8) {
                double t8[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) t8[u] = SS[kk * nbmax + min(b0 + u, nb - 1)];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (b0 + u < nb) acc.re += t8[u];
            }
        } else if ((tid - 1) % kBnb < i) {
            for (int J0 = 0; J0 < nt; J0 += 8) {
                cx<R> t8[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) t8[u] = Qk[(size_t)min(J0 + u, nt - 1) * kQStride + tid];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (J0 + u < nt) acc = acc + t8[u];
            }
        }
        t12[tid] = acc;
    }
    __syncthreads();
    if (tid == 0) {
        double re = 0.0;
        for (int t = 0; t < i; ++t) {
            const cx<R> a = t12[1 + t], b = t12[1 + kBnb + t];
            re += a.re * b.re + a.im * b.im;  // Re(conj(t1) t2)
        }
        s_alpha = -0.5 * (tj.re * tj.re + tj.im * tj.im) * (t12[0].re - 2.0 * re);
    }
    __syncthreads();
    const int r = j + 1 + blockIdx.x * 256 + tid;
    if (r >= N) return;
    const cx<R>* Pk = P + kk * (size_t)nchmax * N;
    const cx<R>* P2k = P2 + kk * (size_t)nchmax * N;
    const int I = (r - j - 1) / kHvCols;
    // the partial sums and the panel columns in batches of 8 loads (indices clamped into range,
    // the extra terms masked), so a batch is in flight at once
    cx<R> y = {0.0, 0.0};
    for (int J0 = 0; J0 <= I; J0 += 8) {
        cx<R> t8[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t8[u] = Pk[(size_t)min(J0 + u, I) * N + r];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (J0 + u <= I) y = y + t8[u];
    }
    const int np = hv_parts(nt, I);
    for (int q0 = 0; q0 < np; q0 += 8) {
        cx<R> t8[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t8[u] = P2k[(size_t)min(q0 + u, np - 1) * N + r];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (q0 + u < np) y = y + t8[u];
    }
    cx<R> u = {0.0, 0.0};
    for (int t0 = 0; t0 < i; t0 += 8) {
        cx<R> v8[8], w8[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int t = min(t0 + q, i - 1);
            v8[q] = Vk[(size_t)t * N + r];
            w8[q] = Wk[(size_t)t * N + r];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (t0 + q < i) {
                const int t = t0 + q;
                u = cmad(u, v8[q], t12[1 + t]);
                u = cmad(u, w8[q], t12[1 + kBnb + t]);
            }
    }
    const cx<R> w = tj * (y - u);
    const cx<R> vr = v[r];
    const R sa = (R)s_alpha;
    Wk[(size_t)i * N + r] = {w.re + sa * vr.re, w.im + sa * vr.im};
}

// ---------------------------------------------------------------------------------
// Batched complex GEMM on the FP64 matrix cores:
//   C = alpha sum_{p < npair} op(A_p) op(B_p) + beta C
// op 0: as stored (column-major, leading dimension ld), op 1: conjugate transpose.
// ascale (optional): column k of op(A_p) scaled by ascale[k] (per batch entry, stride sS).
// 64 x 64 output tile per block (4 waves of 32 x 32, 2 x 2 MFMA 16x16x4 tiles); operands
// read straight from global memory (L1/L2), rows/columns outside M, N, K read as zero.
// ---------------------------------------------------------------------------------
template <typename R>
struct GemmArgs {
    int M, N, K, npair;
    int lower;  // M == N: only the 64 x 64 tiles on or below (1) / on or above (2) the diagonal
                // (grid.x = nt(nt+1)/2)
    const cx<R>* A[2];
    const cx<R>* B[2];
    int lda, ldb, ldc;
    long long sA, sB, sC, sS;
    cx<R>* C;
    cx<double> alpha, beta;
    const double* ascale;
    const int* kidx;   // batch entry of grid z (null: z itself)
    const int* pairs;  // (a, b) per grid z: A from entry a, B from entry b, C at z (null: none)
};
using ZGemmArgs = GemmArgs<double>;

// The 16 x 16 x 4 matrix-core instruction of each precision.  Operands: A(m, k) from lane
// (m, k-slot) = (l % 16, l / 16), B(k, n) from lane (n, k-slot); accumulator element i holds
// column l % 16 and row g + 4 i (FP64) or 4 g + i (FP32), g = l / 16.
typedef float lqm_cf4 __attribute__((ext_vector_type(4)));
template <typename R>
struct Mfma16;
template <>
struct Mfma16<double> {
    typedef lqm_d4 acc;
    static __device__ __forceinline__ acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int i) { return g + 4 * i; }
};
template <>
struct Mfma16<float> {
    typedef lqm_cf4 acc;
    static __device__ __forceinline__ acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int i) { return 4 * g + i; }
};

template <int OP, typename R>
__device__ __forceinline__ cx<R> opA_at(const cx<R>* A, int ld, int m, int k) {
    if (OP == 0) return A[(size_t)k * ld + m];
    const cx<R> x = A[(size_t)m * ld + k];
    return {x.re, -x.im};
}
template <int OP, typename R>
__device__ __forceinline__ cx<R> opB_at(const cx<R>* B, int ld, int k, int n) {
    if (OP == 0) return B[(size_t)n * ld + k];
    const cx<R> x = B[(size_t)k * ld + n];
    return {x.re, -x.im};
}

template <int OPA, int OPB, typename R>
__device__ __forceinline__ void gemm_acc(const cx<R>* Ab, int lda, const cx<R>* Bb, int ldb, const double* sc,
                                         int M, int N, int K, int m0, int n0, int li, int g,
                                         typename Mfma16<R>::acc (&cr)[2][2], typename Mfma16<R>::acc (&ci)[2][2]) {
    typedef Mfma16<R> MM;
    for (int k0 = 0; k0 < K; k0 += 16) {
        const int kb = k0 + 4 * g;
        R ar[2][4], ai[2][4], br[2][4], bi[2][4];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int m = m0 + 16 * t + li, n = n0 + 16 * t + li;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                // clamped indices: every load is in bounds and unconditional, so the
                // compiler can issue them together; out-of-range values are zeroed after
                const int k = kb + s, kc = min(k, K - 1);
                cx<R> x = opA_at<OPA>(Ab, lda, min(m, M - 1), kc);
                cx<R> y = opB_at<OPB>(Bb, ldb, kc, min(n, N - 1));
                const bool ina = k < K && m < M, inb = k < K && n < N;
                const R f = sc ? (R)sc[kc] : R(1);
                ar[t][s] = ina ? f * x.re : R(0);
                ai[t][s] = ina ? f * x.im : R(0);
                br[t][s] = inb ? y.re : R(0);
                bi[t][s] = inb ? y.im : R(0);
            }
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) cr[a][b] = MM::mma(ar[a][s], br[b][s], cr[a][b]);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) ci[a][b] = MM::mma(ar[a][s], bi[b][s], ci[a][b]);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) cr[a][b] = MM::mma(-ai[a][s], bi[b][s], cr[a][b]);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) ci[a][b] = MM::mma(ai[a][s], br[b][s], ci[a][b]);
        }
    }
}

// Batched complex GEMM on the matrix cores of precision R (FP64: v_mfma_f64_16x16x4, FP32:
// v_mfma_f32_16x16x4 with FP32 accumulation):
//   C = alpha sum_{p < npair} op(A_p) op(B_p) + beta C
// op 0: as stored (column-major, leading dimension ld), op 1: conjugate transpose.
// ascale (optional): column k of op(A_p) scaled by ascale[k] (per batch entry, stride sS).
// 64 x 64 output tile per block (4 waves of 32 x 32, 2 x 2 MFMA 16x16x4 tiles); operands
// read straight from global memory (L1/L2), rows/columns outside M, N, K read as zero.
template <int OPA, int OPB, typename R = double>
__global__ __launch_bounds__(256) void k_zgemm_gen(GemmArgs<R> g) {
    typedef Mfma16<R> MM;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, g4 = lane >> 4;
    int ti = blockIdx.x, tj = blockIdx.y;
    if (g.lower) {
        const int t = blockIdx.x;
        ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
        while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
        while (ti * (ti + 1) / 2 > t) --ti;
        tj = t - ti * (ti + 1) / 2;
        if (g.lower == 2) {
            const int x = ti;
            ti = tj;
            tj = x;
        }
    }
    const int m0 = ti * 64 + (wave & 1) * 32;
    const int n0 = tj * 64 + (wave >> 1) * 32;
    const size_t b = g.kidx ? (size_t)g.kidx[blockIdx.z] : (size_t)blockIdx.z;
    const size_t bA = g.pairs ? (size_t)g.pairs[2 * blockIdx.z] : b;
    const size_t bB = g.pairs ? (size_t)g.pairs[2 * blockIdx.z + 1] : b;
    typename MM::acc cr[2][2], ci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            cr[a][c] = typename MM::acc{R(0), R(0), R(0), R(0)};
            ci[a][c] = typename MM::acc{R(0), R(0), R(0), R(0)};
        }
    const double* sc = g.ascale ? g.ascale + b * g.sS : nullptr;
    for (int p = 0; p < g.npair; ++p)
        gemm_acc<OPA, OPB, R>(g.A[p] + bA * g.sA, g.lda, g.B[p] + bB * g.sB, g.ldb, sc, g.M, g.N, g.K, m0, n0, li,
                              g4, cr, ci);
    cx<R>* Cb = g.C + b * g.sC;
    const bool zero_beta = g.beta.re == 0.0 && g.beta.im == 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = m0 + 16 * a + MM::row(g4, i), n = n0 + 16 * c + li;
                if (m < g.M && n < g.N) {
                    const size_t off = (size_t)n * g.ldc + m;
                    const cx<double> acc = {(double)cr[a][c][i], (double)ci[a][c][i]};
                    cx<double> out = g.alpha * acc;
                    if (!zero_beta) {
                        const cx<R> cv = Cb[off];
                        out = cmad(out, g.beta, cx<double>{(double)cv.re, (double)cv.im});
                    }
                    Cb[off] = cx<R>{(R)out.re, (R)out.im};
                }
            }
}

// ---------------------------------------------------------------------------------
// Eigenvalue n of the tridiagonal (d, e) of k-point blockIdx.y by the Sturm count (LAPACK
// dstebz pivots: a pivot of modulus below pivmin = tiny max(1, max e^2) becomes -pivmin), one
// thread per eigenvalue, 256 per block.  d and e^2 are staged in LDS (dynamic, 16 N bytes);
// the pivots' reciprocals come from the hardware reciprocal and a Newton step (no division in
// the recurrence).  A first Sturm pass at 256 points of the Gershgorin interval (one per
// thread, counts in LDS) gives each eigenvalue a bracket 257 times narrower; then bisection,
// with a Newton step on det(T - x I) -- d/dx log det = sum_i q_i'/q_i from the same recurrence
// -- taken whenever the bracket holds eigenvalue n alone and the step lands inside it.
// Converged when the Newton step is at most 4 eps ||T|| and points into the bracket, or the
// bracket is at most max(2 eps max(|a|, |b|), eps ||T||) (||T|| = the Gershgorin bound).
// Out: lam[k * N + n] ascending, tnorm[k] = ||T||.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sturm_bisect(int N, const double* __restrict__ dd,
                                                      const double* __restrict__ ee,
                                                      double* __restrict__ lam,
                                                      double* __restrict__ tnorm) {
    extern __shared__ __align__(16) unsigned char smem[];
    double* sd = reinterpret_cast<double*>(smem);  // [N] d
    double* se2 = sd + N;                          // [N] e^2 (e_{N-1} = 0)
    __shared__ double rlo[4], rhi[4], re2[4];
    __shared__ double sx[256];
    __shared__ int scnt[256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t kk = blockIdx.y;
    const double* d = dd + kk * N;
    const double* e = ee + kk * N;
    double lo = 1e308, hi = -1e308, e2 = 0.0;
    for (int i = tid; i < N; i += 256) {
        const double el = i > 0 ? fabs(e[i - 1]) : 0.0, er = i + 1 < N ? fabs(e[i]) : 0.0;
        sd[i] = d[i];
        se2[i] = er * er;
        lo = fmin(lo, d[i] - el - er);
        hi = fmax(hi, d[i] + el + er);
        e2 = fmax(e2, er * er);
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
    const double atol = eps * bnorm, ntol = 4.0 * eps * bnorm;
    if (blockIdx.x == 0 && tid == 0) tnorm[kk] = bnorm;
    double a = lo - 2.0 * eps * bnorm - 2.0 * pivmin, b = hi + 2.0 * eps * bnorm + 2.0 * pivmin;
    {
        // the grid pass: point t = a + (t + 1) h, h = (b - a) / 257
        const double h = (b - a) / 257.0;
        const double xg = a + (double)(tid + 1) * h;
        double q = sd[0] - xg;
        if (fabs(q) <= pivmin) q = -pivmin;
        int cg = q < 0.0;
        double r = rcp_nr(q);
        for (int i = 1; i < N; ++i) {
            double qn = (sd[i] - xg) - se2[i - 1] * r;
            if (fabs(qn) <= pivmin) qn = -pivmin;
            r = rcp_nr(qn);
            cg += qn < 0.0;
        }
        sx[tid] = xg;
        scnt[tid] = cg;
    }
    __syncthreads();
    const int n = blockIdx.x * 256 + tid;
    if (n >= N) return;  // no barrier below
    int ca = 0, cb = N;
    {
        // the first grid point whose count exceeds n (counts are nondecreasing; 256 = none)
        int l0 = 0, l1 = 256;
        while (l0 < l1) {
            const int mid = (l0 + l1) >> 1;
            if (scnt[mid] > n) l1 = mid;
            else l0 = mid + 1;
        }
        if (l0 > 0) {
            a = sx[l0 - 1];
            ca = scnt[l0 - 1];
        }
        if (l0 < 256) {
            b = sx[l0];
            cb = scnt[l0];
        }
    }
    double x = 0.5 * (a + b);
    for (int it = 0; it < 256; ++it) {
        double q = sd[0] - x;
        bool bad = false;
        if (fabs(q) <= pivmin) {
            q = -pivmin;
            bad = true;
        }
        int cnt = q < 0.0;
        double r = rcp_nr(q);
        double qp = -1.0, S = -r;
        for (int i = 1; i < N; ++i) {
            const double t = se2[i - 1] * r;
            double qn = (sd[i] - x) - t;
            if (fabs(qn) <= pivmin) {
                qn = -pivmin;
                bad = true;
            }
            qp = fma(t * r, qp, -1.0);
            r = rcp_nr(qn);
            S = fma(qp, r, S);
            cnt += qn < 0.0;
        }
        const bool above = cnt > n;  // eigenvalue n lies below x
        if (above) {
            b = x;
            cb = cnt;
        } else {
            a = x;
            ca = cnt;
        }
        const double mid = 0.5 * (a + b);
        const bool iso = cb - ca == 1;
        const double step = -1.0 / S;
        const double xn = x + step;
        const bool fin = !bad && xn - xn == 0.0;
        // a small step counts only when it points into the bracket: next to an end, a step
        // toward the neighbouring eigenvalue outside is small as well
        if (iso && fin && fabs(step) <= ntol && (above ? step <= 0.0 : step >= 0.0)) {
            x = fmin(fmax(xn, a), b);
            break;
        }
        if (!(mid > a && mid < b) || b - a <= fmax(2.0 * eps * fmax(fabs(a), fabs(b)), atol)) {
            x = mid;
            break;
        }
        x = (iso && fin && xn > a && xn < b) ? xn : mid;
    }
    lam[kk * N + n] = x;
}
inline size_t sturm_lds_bytes(int N) { return 16 * (size_t)N; }

__device__ __forceinline__ double invit_start(unsigned long long seed, int i) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

// ---------------------------------------------------------------------------------
// Eigenvectors of the symmetric tridiagonal (d, e) by inverse iteration (LAPACK dstein's
// solve).  Thread m of block (x, k): three solves of (T - lambda_m I) x = b, b the last x
// scaled to unit max-norm, first b a seeded uniform(-1, 1) vector; each solve factors
// T - lambda I = P L U with partial pivoting (dlagtf) fused with the forward
// substitution and replaces U pivots below eps ||T|| by +-eps ||T|| (dlagts).  Every
// eigenvector iterates on its own; the members of a cluster (eigenvalues within ctol of
// their predecessor) are orthonormalised afterwards by k_invit_orth.  Out:
// Zt[k][i * ldz + m] = x_m[i] (unit 2-norm), m < ncol.  scr: U's three diagonals,
// 3 * N * ldz doubles per k-point, interleaved by m so a wave's accesses are coalesced.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_invit(int N, int ncol, int ldz, const double* __restrict__ dd,
                                              const double* __restrict__ ee,
                                              const double* __restrict__ lam,
                                              const double* __restrict__ tnorm,
                                              double* __restrict__ Zt, double* __restrict__ scr,
                                              int kbase, int* __restrict__ status) {
    const int m = blockIdx.x * 64 + threadIdx.x;
    const size_t kk = blockIdx.y;
    if (m >= ncol) return;
    const double* d = dd + kk * N;
    const double* e = ee + kk * N;
    const double x = lam[kk * N + m];
    const double eps = 2.220446049250313e-16;
    const double tol = fmax(eps * tnorm[kk], 2.2250738585072014e-308);
    const size_t NZ = (size_t)N * ldz;
    double* Z = Zt + kk * NZ + m;
    double* U0 = scr + (kk * 3 + 0) * NZ + m;
    double* U1 = scr + (kk * 3 + 1) * NZ + m;
    double* U2 = scr + (kk * 3 + 2) * NZ + m;
    const unsigned long long seed =
        0x51A7E5EEDull ^ ((unsigned long long)(kbase + (long long)kk) << 24) ^ (unsigned long long)m;
    bool bad = false;
    for (int i = 0; i < N; ++i) Z[(size_t)i * ldz] = invit_start(seed, i);
    // the previous iterate's 1 / max |z| is applied as its entries are read (no scaling pass);
    // the last back substitution also forms ||z||^2 (no norm pass)
    double sc = 1.0, nrm2 = 0.0;
    for (int it = 0; it < 3; ++it) {
        // factor + forward substitution (the next right-hand side entry is read one step ahead)
        double acur = d[0] - x, bcur = N > 1 ? e[0] : 0.0;
        double ycur = Z[0] * sc;
        double ynext = N > 1 ? Z[ldz] * sc : 0.0;
        for (int k = 0; k + 1 < N; ++k) {
            const double c = e[k];
            const double anext = d[k + 1] - x;
            const double bnext = k + 2 < N ? e[k + 1] : 0.0;
            const double yfar = k + 2 < N ? Z[(size_t)(k + 2) * ldz] * sc : 0.0;
            const size_t o = (size_t)k * ldz;
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
            ynext = yfar;
        }
        U0[(size_t)(N - 1) * ldz] = acur;
        Z[(size_t)(N - 1) * ldz] = ycur;
        // back substitution (row k - 1 read one step ahead)
        double x1 = 0.0, x2 = 0.0, amax = 0.0;
        size_t o = (size_t)(N - 1) * ldz;
        double p0 = U0[o], p1 = U1[o], p2 = U2[o], zz = Z[o];
        for (int k = N - 1; k >= 0; --k) {
            double q0 = 0.0, q1 = 0.0, q2 = 0.0, qz = 0.0;
            if (k > 0) {
                const size_t on = o - ldz;
                q0 = U0[on];
                q1 = U1[on];
                q2 = U2[on];
                qz = Z[on];
            }
            double p = p0;
            if (fabs(p) < tol) p = p >= 0.0 ? tol : -tol;
            double v = zz;
            if (k + 1 < N) v -= p1 * x1;
            if (k + 2 < N) v -= p2 * x2;
            v /= p;
            Z[o] = v;
            x2 = x1;
            x1 = v;
            amax = fmax(amax, fabs(v));
            nrm2 = fma(v, v, nrm2);
            o -= ldz;
            p0 = q0;
            p1 = q1;
            p2 = q2;
            zz = qz;
        }
        if (!(amax > 0.0) || !isfinite(amax)) {
            bad = true;
            amax = 1.0;
        }
        sc = 1.0 / amax;
        if (it < 2) nrm2 = 0.0;
    }
    double nrm = sqrt(nrm2);
    if (!(nrm > 0.0) || !isfinite(nrm)) {
        bad = true;
        nrm = 1.0;
    }
    const double scn = 1.0 / nrm;
    for (int i = 0; i < N; ++i) Z[(size_t)i * ldz] *= scn;
    if (bad) atomicOr(status, kStatNonFinite);
}

// Clusters of consecutive eigenvalues within ctol = ctol_rel ||T|| of their predecessor:
// the members' independently iterated vectors are orthonormalised in eigenvalue order by
// modified Gram-Schmidt, applied twice.  One wave per candidate first member (wave w of
// block x: m = 4x + w), lanes over the rows; waves whose m does not start a cluster of two
// or more leave at once.
__global__ __launch_bounds__(256) void k_invit_orth(int N, int ncol, int ldz, const double* __restrict__ lam,
                                                    const double* __restrict__ tnorm, double ctol_rel,
                                                    double* __restrict__ Zt) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = blockIdx.x * 4 + wave;
    const size_t kk = blockIdx.y;
    if (m >= ncol) return;
    const double* l = lam + kk * N;
    const double ctol = ctol_rel * tnorm[kk];
    if (m > 0 && l[m] - l[m - 1] <= ctol) return;
    int mend = m + 1;
    while (mend < ncol && l[mend] - l[mend - 1] <= ctol) ++mend;
    if (mend - m < 2) return;
    double* Z = Zt + kk * (size_t)N * ldz;
    for (int q = m; q < mend; ++q) {
        for (int pass = 0; pass < 2; ++pass)
            for (int p = m; p < q; ++p) {
                double dot = 0.0;
                for (int i = lane; i < N; i += 64) dot += Z[(size_t)i * ldz + p] * Z[(size_t)i * ldz + q];
                dot = wave_sum(dot);
                for (int i = lane; i < N; i += 64) Z[(size_t)i * ldz + q] -= dot * Z[(size_t)i * ldz + p];
            }
        double nrm = 0.0;
        for (int i = lane; i < N; i += 64) {
            const double v = Z[(size_t)i * ldz + q];
            nrm += v * v;
        }
        nrm = wave_sum(nrm);
        const double sc = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
        for (int i = lane; i < N; i += 64) Z[(size_t)i * ldz + q] *= sc;
    }
}

// A (reflectors below the subdiagonal, beta on it) -> explicit Y in place: column c < N-1
// holds v_c (zero above row c+1, 1 at row c+1); column N-1 is zero.
template <typename R>
__global__ __launch_bounds__(256) void k_make_y(int N, cx<R>* __restrict__ A) {
    const size_t NN = (size_t)N * N;
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= NN) return;
    const int c = (int)(e / N), r = (int)(e - (size_t)c * N);
    cx<R>* Ak = A + (size_t)blockIdx.y * NN;
    if (c == N - 1 || r <= c) Ak[e] = {R(0), R(0)};
    else if (r == c + 1) Ak[e] = {R(1), R(0)};
}

// Compact-WY factor of reflector block b (reflectors jb = 64b ... jb + nbb - 1): the
// upper triangular T with H_jb ... H_{jb+nbb-1} = I - Y T Y^H (LAPACK zlarft, forward,
// columnwise): T(t, t) = tau_t, T(0:t, t) = -tau_t T(0:t, 0:t) P(0:t, t), P = Y^H Y
// (rows jb+1..N-1, on the matrix cores).  Out: Tw[k][b] (64 x 64 column-major, zero
// outside 0..nbb-1).  Block (b, k).
template <typename R>
__global__ __launch_bounds__(256) void k_wy_T(int N, int nblk, const cx<R>* __restrict__ Y,
                                              const cx<R>* __restrict__ tau,
                                              cx<R>* __restrict__ Tw) {
    typedef Mfma16<R> MM;
    __shared__ cx<R> Ps[64][65];
    __shared__ cx<R> col[64];
    const int b = blockIdx.x;
    const size_t kk = blockIdx.y;
    const int jb = kBwy * b, nbb = min(kBwy, N - 1 - jb);
    const int r0 = jb + 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, g4 = lane >> 4;
    const int m0 = (wave & 1) * 32, n0 = (wave >> 1) * 32;
    const cx<R>* Yb = Y + kk * (size_t)N * N + (size_t)jb * N + r0;
    typename MM::acc cr[2][2], ci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            cr[a][c] = typename MM::acc{R(0), R(0), R(0), R(0)};
            ci[a][c] = typename MM::acc{R(0), R(0), R(0), R(0)};
        }
    gemm_acc<1, 0, R>(Yb, N, Yb, N, nullptr, nbb, nbb, N - r0, m0, n0, li, g4, cr, ci);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = m0 + 16 * a + MM::row(g4, i), n = n0 + 16 * c + li;
                Ps[n][m] = cx<R>{cr[a][c][i], ci[a][c][i]};  // Ps[col][row]
            }
    __syncthreads();
    const cx<R>* tk = tau + kk * N + jb;
    // T overwrites P column by column (column t of P copied out first): T(s, t) =
    // -tau_t sum_{s <= u < t} T(s, u) P(u, t), the sum of row s split over four consecutive
    // lanes (u = s + q, s + q + 4, ...) and added by two lane exchanges
    const int s = threadIdx.x >> 2, q = threadIdx.x & 3;
    for (int t = 0; t < nbb; ++t) {
        if (threadIdx.x < t) col[threadIdx.x] = Ps[t][threadIdx.x];
        __syncthreads();
        const cx<R> tt = tk[t];
        cx<R> acc = {R(0), R(0)};
        for (int u = s + q; u < t; u += 4) acc = cmad(acc, Ps[u][s], col[u]);
        acc.re += __shfl_xor(acc.re, 1);
        acc.im += __shfl_xor(acc.im, 1);
        acc.re += __shfl_xor(acc.re, 2);
        acc.im += __shfl_xor(acc.im, 2);
        if (q == 0) {
            cx<R> v = {R(0), R(0)};
            if (s < t) v = cx<R>{R(0), R(0)} - tt * acc;
            else if (s == t) v = tt;
            Ps[t][s] = v;
        }
        __syncthreads();
    }
    cx<R>* To = Tw + (kk * nblk + b) * (size_t)(kBwy * kBwy);
    for (int e = threadIdx.x; e < kBwy * kBwy; e += 256) {
        const int c = e >> 6, r = e & 63;
        To[e] = (c < nbb && r <= c) ? Ps[c][r] : cx<R>{R(0), R(0)};
    }
}

// V[k](:, n) = Z[k](:, n) as complex (Zt row-major with row stride ldz), n < ncol.
template <typename R>
__global__ __launch_bounds__(256) void k_z_to_v(int N, int ncol, int ldz, const double* __restrict__ Zt,
                                                cx<R>* __restrict__ V) {
    __shared__ double tile[32][33];
    const size_t kk = blockIdx.z;
    const int i0 = blockIdx.x * 32, n0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const double* Zk = Zt + kk * (size_t)N * ldz;
    cx<R>* Vk = V + kk * (size_t)N * ncol;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = i0 + ty + 8 * q, n = n0 + tx;
        tile[ty + 8 * q][tx] = (i < N && n < ncol) ? Zk[(size_t)i * ldz + n] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int n = n0 + ty + 8 * q, i = i0 + tx;
        if (i < N && n < ncol) Vk[(size_t)n * N + i] = {(R)tile[tx][ty + 8 * q], R(0)};
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
