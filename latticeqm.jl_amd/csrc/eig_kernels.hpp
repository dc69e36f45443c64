// Batched Hermitian eigensolver for whole k-grids on gfx950.
//
// The path of one k-point (Eigen.jl:14-39 replaces LAPACK zheevr with this):
//   k_hk_tridiag        H(k) = sum_R t_R e^{2 pi i k.R} assembled straight into the
//                       workgroup's matrix tile (upper triangle read, as Hermitian(M)
//                       does) and reduced to real tridiagonal form by Householder
//                       reflections H_j = I - tau_j v_j v_j^H (one workgroup per k).
//   k_tridiag_ql        implicit QL with Wilkinson shifts on (d, e), one LANE per
//                       k-point, so the serial recurrence of 64 k-points runs in one
//                       wave; optionally logs every Givens rotation.
//   k_backtransform     replays the rotation log on Z = I, applies the reflectors
//                       (V = H_0 ... H_{N-2} Z), then either writes the eigenvectors
//                       or forms rho_k = V diag(f) V^H for the density matrix.
#pragma once
#include "lqm_common.hpp"

namespace lqm {

// ---------------------------------------------------------------------------------
// K_A: Bloch assembly + Householder tridiagonalization.
// Matrix tile: column-major, leading dimension lda, interleaved complex.  In LDS
// when it fits (LDSMAT), else in a per-k global workspace.
// Outputs (per k-point kk of the chunk):
//   dd[i*Nkc + kk], ee[i*Nkc + kk]   diagonal / off-diagonal, structure-of-arrays so
//                                    the lane-per-k QL reads them coalesced;
//   refl[kk*N*N + j*N + i]           v_j at rows i = j+1..N-1 (v_j[j+1] = 1);
//   tau[kk*N + j]                    tau_j.
// ---------------------------------------------------------------------------------
template <typename R, bool LDSMAT>
__global__ __launch_bounds__(kTPB) void k_hk_tridiag(
    int N, int d, int nR, const int* __restrict__ Rk, const cx<double>* __restrict__ T,
    const double* __restrict__ ks, int Nkc, cx<R>* __restrict__ gA, R* __restrict__ dd,
    R* __restrict__ ee, cx<R>* __restrict__ refl, cx<R>* __restrict__ tau,
    const cx<double>* __restrict__ Hpre) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int kk = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int lda = LDSMAT ? N + 1 : N;

    cx<R>* A;
    unsigned char* p = smem;
    if constexpr (LDSMAT) {
        A = reinterpret_cast<cx<R>*>(p);
        p += sizeof(cx<R>) * (size_t)N * lda;
    } else {
        A = gA + (size_t)kk * N * lda;
    }
    cx<R>* sv = reinterpret_cast<cx<R>*>(p);  // [2][N]: v of step j in sv[(j & 1) * N]
    p += sizeof(cx<R>) * 2 * N;
    cx<R>* sp = reinterpret_cast<cx<R>*>(p);
    p += sizeof(cx<R>) * N;
    cx<double>* sph = reinterpret_cast<cx<double>*>(p);
    p += sizeof(cx<double>) * nR;
    cx<R>* sred = reinterpret_cast<cx<R>*>(p);  // [4] wave partials + [2] tau (double-buffered)
    auto At = [&](int i, int j) -> cx<R>& { return A[(size_t)j * lda + i]; };

    // phases e^{2 pi i k.R}
    const double* k = ks + (size_t)kk * d;
    for (int r = tid; r < nR; r += kTPB) {
        double x = 0.0;
        for (int a = 0; a < d; ++a) x += k[a] * (double)Rk[r * d + a];
        sph[r] = cis2pi(x);
    }
    __syncthreads();

    // H(k), upper triangle summed in FP64, mirrored into the lower triangle (or, with
    // Hpre, read from the FP64 Hermitian H(k) the pass has already built)
    const size_t NN = (size_t)N * N;
    if (Hpre) {
        const cx<double>* H = Hpre + (size_t)kk * NN;
        for (int e = tid; e < N * N; e += kTPB) {
            const int i = e % N, j = e / N;
            const cx<double> h = H[e];
            At(i, j) = mkc<R>((R)h.re, i == j ? (R)0 : (R)h.im);
        }
    }
    for (int e = tid; e < N * N && !Hpre; e += kTPB) {
        int i = e % N, j = e / N;
        if (i > j) continue;
        cx<double> h = {0.0, 0.0};
        const cx<double>* t = T + (size_t)j * N + i;
        for (int r = 0; r < nR; ++r) h = cmad(h, t[r * NN], sph[r]);
        if (i == j) {
            At(i, i) = mkc<R>((R)h.re, (R)0);
        } else {
            At(i, j) = mkc<R>((R)h.re, (R)h.im);
            At(j, i) = mkc<R>((R)h.re, (R)-h.im);
        }
    }
    __syncthreads();

    cx<R>* rf = refl + (size_t)kk * NN;
    // Householder vector of column j (zlarfg semantics: H^H x = beta e_1 with beta real),
    // formed by wave 0 into sv[(j & 1) N] from A as it stands.  Column j+1 is relative
    // column 0 of step j's trailing matrix, which wave 0 updates in full (columns go to
    // waves by residue mod 4, all rows of a column to one wave): so wave 0 forms v_{j+1}
    // right after its own part of the step-j update, without a barrier, while the other
    // waves finish theirs; v and tau are double-buffered by step parity.
    auto householder = [&](int j) {
        const int m = N - 1 - j;
        cx<R>* svb = sv + (j & 1) * N;
        R xn2 = 0;
        for (int i = j + 2 + lane; i < N; i += kWave) xn2 += abs2(At(i, j));
        xn2 = wave_sum(xn2);
        cx<R> alpha = At(j + 1, j);
        cx<R> tj, scale;
        R beta;
        if (!hh_scalars(alpha, xn2, tj, scale, beta)) {
            // a column of tiny entries: the scalars of x / max|x|, then beta and scale back
            R mx = 0;
            for (int i = j + 1 + lane; i < N; i += kWave) mx = fmax(mx, fmax(fabs(At(i, j).re), fabs(At(i, j).im)));
#pragma unroll
            for (int o = kWave / 2; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
            const R inv = (R)1 / mx;
            R xs2 = 0;
            for (int i = j + 2 + lane; i < N; i += kWave) xs2 += abs2(inv * At(i, j));
            xs2 = wave_sum(xs2);
            hh_scalars_scaled(inv * alpha, xs2, tj, scale, beta);
            beta *= mx;
            scale = inv * scale;
            if (!(inv <= num<R>::huge)) {
                tj = mkc<R>(0, 0);
                scale = mkc<R>(0, 0);
                beta = alpha.re;
            }
        }
        const bool triv = (tj.re == (R)0 && tj.im == (R)0);
        for (int ip = lane; ip < m; ip += kWave) {
            cx<R> v = (ip == 0) ? mkc<R>(1, 0) : (triv ? mkc<R>(0, 0) : At(j + 1 + ip, j) * scale);
            svb[ip] = v;
            rf[(size_t)j * N + j + 1 + ip] = v;
        }
        if (lane == 0) {
            sred[4 + (j & 1)] = tj;
            tau[(size_t)kk * N + j] = tj;
            dd[(size_t)j * Nkc + kk] = At(j, j).re;
            ee[(size_t)j * Nkc + kk] = beta;
        }
    };
    if (N > 1 && wave == 0) householder(0);
    for (int j = 0; j + 1 < N; ++j) {
        const int m = N - 1 - j;  // length of the reflected column x = A[j+1:N, j]
        __syncthreads();          // v_j, tau_j formed; the step j-1 update is complete
        const cx<R> tj = sred[4 + (j & 1)];
        const cx<R>* svj = sv + (j & 1) * N;
        if (tj.re == (R)0 && tj.im == (R)0) {  // H_j = I (uniform branch)
            if (wave == 0 && j + 2 < N) householder(j + 1);
            continue;
        }

        // (b) p = tau * A22 v  (4 lanes per row, shuffle-reduced) and partial p^H v
        cx<R> dotp = {0, 0};
        {
            const int q = tid & 3;
            for (int r = tid >> 2; r < m; r += kTPB / 4) {
                cx<R> acc = {0, 0};
                const cx<R>* arow = &At(j + 1 + r, j + 1);
                for (int c = q; c < m; c += 4) acc = cmad(acc, arow[(size_t)c * lda], svj[c]);
                acc.re += __shfl_xor(acc.re, 1);
                acc.im += __shfl_xor(acc.im, 1);
                acc.re += __shfl_xor(acc.re, 2);
                acc.im += __shfl_xor(acc.im, 2);
                if (q == 0) {
                    cx<R> pr = tj * acc;
                    sp[r] = pr;
                    dotp = cmadc(dotp, pr, svj[r]);
                }
            }
        }
        dotp = wave_sum(dotp);
        if (lane == 0) sred[wave] = dotp;
        __syncthreads();
        cx<R> dot = sred[0] + sred[1] + sred[2] + sred[3];
        const cx<R> alpha2 = (R)-0.5 * (tj * dot);

        // (c) A22 -= v w^H + w v^H with w = p + alpha2 v
        {
            const int rr = tid & (kWave - 1);
            const int cq = tid / kWave;
            for (int r = rr; r < m; r += kWave) {
                const cx<R> vr = svj[r];
                const cx<R> wr = cmad(sp[r], alpha2, vr);
                for (int c = cq; c < m; c += kTPB / kWave) {
                    const cx<R> vc = svj[c];
                    const cx<R> wc = cmad(sp[c], alpha2, vc);
                    cx<R> a = At(j + 1 + r, j + 1 + c);
                    cx<R> u = {0, 0};
                    u = cmadcb(u, vr, wc);
                    u = cmadcb(u, wr, vc);
                    At(j + 1 + r, j + 1 + c) = a - u;
                }
            }
        }
        if (wave == 0 && j + 2 < N) householder(j + 1);
    }
    if (tid == 0) {
        dd[(size_t)(N - 1) * Nkc + kk] = At(N - 1, N - 1).re;
        ee[(size_t)(N - 1) * Nkc + kk] = (R)0;
    }
}

// ---------------------------------------------------------------------------------
// K_B: implicit QL on the real symmetric tridiagonal (d, e), one lane per k-point.
// Working copies live in LDS as [i][lane] so that lanes at the same index hit
// consecutive banks.  Each QL sweep is logged as a header (first index << 16 | count)
// followed by its (c, s) pairs: the rotation at index i maps columns (i, i+1) of Z to
//   z_{i+1} <- s z_i + c z_{i+1},   z_i <- c z_i - s z_{i+1}.
// Eigenvalues are written ascending (evals[kk*N + n]) with perm[kk*N + n] = the Z
// column that belongs to them.
// ---------------------------------------------------------------------------------
template <typename R> struct RotPair;
template <> struct RotPair<float> { using type = float2; };
template <> struct RotPair<double> { using type = double2; };

template <typename R, bool LOG>
__global__ void k_tridiag_ql(int N, int Nkc, const R* __restrict__ dd, const R* __restrict__ ee,
                             R* __restrict__ evals, int* __restrict__ perm, R* __restrict__ rot,
                             int* __restrict__ sweeps, int* __restrict__ nsweeps, int cap_rot,
                             int cap_sw, int* __restrict__ status, const int* __restrict__ klist) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int L = blockDim.x;
    const int lane = threadIdx.x;
    int kk = blockIdx.x * L + lane;
    R* sd = reinterpret_cast<R*>(smem);
    R* se = sd + (size_t)N * L;
    // klist (optional): klist[0] k-points, listed at klist[1 ...]
    if (klist) {
        if (kk >= klist[0]) return;
        kk = klist[1 + kk];
    }
    if (kk >= Nkc) return;  // no barriers below: early exit is safe
    auto D = [&](int i) -> R& { return sd[(size_t)i * L + lane]; };
    auto E = [&](int i) -> R& { return se[(size_t)i * L + lane]; };
    for (int i = 0; i < N; ++i) {
        D(i) = dd[(size_t)i * Nkc + kk];
        E(i) = ee[(size_t)i * Nkc + kk];
    }
    int flag = 0;
    R* lrot = LOG ? rot + (size_t)kk * cap_rot * 2 : nullptr;
    int* lsw = LOG ? sweeps + (size_t)kk * cap_sw : nullptr;
    int pos = 0, nsw = 0;
    const R heps = num<R>::half_eps;
    // E(j) is negligible when |E(j)| <= heps (|D(j)| + |D(j+1)|); E(N-1) is the sentinel.
    // m is always the first negligible index at or after l.  A full sweep over [l, m]
    // produces the final D, E of every index it touches, so it records the first negligible
    // index in (l, m] (f1) on the way and the next m needs no scan; a scan runs only after
    // an interrupted sweep or when l moves past what the last sweep saw.  Each lane runs its
    // own event loop (advance l, or one sweep), so lanes at different l do not wait for
    // each other at every l.  The rotations and their order are those of the scanning form.
    auto negl = [&](int j) -> bool { return fabs(E(j)) <= heps * (fabs(D(j)) + fabs(D(j + 1))); };
    auto scan = [&](int from) -> int {
        int mm = from;
        while (mm < N - 1 && !negl(mm)) ++mm;
        return mm;
    };
    int l = 0, iter = 0, f1 = -1;
    int m = scan(0);
    while (l < N) {
        if (m == l) {
            ++l;
            iter = 0;
            if (l < N) m = f1 > l ? f1 : scan(l);
            continue;
        }
        if (++iter > 30) {
            flag |= kStatNoConv;
            break;
        }
        // Wilkinson shift from the leading 2x2 block
        R g = (D(l + 1) - D(l)) / (R(2) * E(l));
        R r = hypot(g, R(1));
        g = D(m) - D(l) + E(l) / (g + copysign(r, g));
        R s = 1, c = 1, pacc = 0;
        const bool logit = LOG && (pos + (m - l) <= cap_rot) && (nsw < cap_sw);
        if (LOG && !logit) flag |= kStatLogFull;
        int cnt = 0;
        bool broke = false;
        f1 = m;
        R dnext = D(m);  // D(i + 2) of the iteration below (final for this sweep)
        // E(i) and D(i) do not depend on the recurrence: they are read one iteration ahead,
        // and D(i + 1) is the previous iteration's D(i) (unchanged since), so no LDS read
        // sits on the critical path of a rotation
        R ei = E(m - 1), di = D(m - 1), dip1 = D(m);
        for (int i = m - 1; i >= l; --i) {
            R en = R(0), dnx = R(0);
            if (i > l) {
                en = E(i - 1);
                dnx = D(i - 1);
            }
            R f = s * ei;
            R b = c * ei;
            // r = |(f, g)|, s = f / r, c = g / r from one reciprocal square root (FP32)
            // or one square root and one reciprocal (FP64); hypot and divisions only
            // where f^2 + g^2 leaves the safe normal range
            const R q = f * f + g * g;
            R inv;
            if (q > num<R>::rot_lo && q < num<R>::rot_hi) {
                if (sizeof(R) == 4) {
                    inv = (R)__frsqrt_rn((float)q);
                    r = q * inv;
                } else {
                    r = sqrt(q);
                    inv = R(1) / r;
                }
            } else {
                r = hypot(f, g);
                inv = r == R(0) ? R(0) : R(1) / r;
            }
            E(i + 1) = r;
            if (r == R(0)) {
                D(i + 1) = dip1 - pacc;
                E(m) = 0;
                broke = true;
                break;
            }
            const R er = r;
            s = f * inv;
            c = g * inv;
            g = dip1 - pacc;
            r = (di - g) * s + R(2) * c * b;
            pacc = s * r;
            const R dn = g + pacc;
            D(i + 1) = dn;
            g = c * r - b;
            // E(i+1), D(i+1), D(i+2) are final for this sweep (E(m) is zeroed below)
            if (i + 1 < m && fabs(er) <= heps * (fabs(dn) + fabs(dnext))) f1 = i + 1;
            dnext = dn;
            if (logit) {
                // (c, s) as one vector store (the pair is 2-element aligned)
                using pair_t = typename RotPair<R>::type;
                reinterpret_cast<pair_t*>(lrot)[pos + cnt] = pair_t{c, s};
            }
            ++cnt;
            dip1 = di;
            ei = en;
            di = dnx;
        }
        if (logit && cnt > 0) {
            lsw[nsw++] = ((m - 1) << 16) | cnt;
            pos += cnt;
        }
        if (broke) {
            f1 = -1;
            m = scan(l);
            continue;
        }
        D(l) -= pacc;
        E(l) = g;
        E(m) = 0;
        m = negl(l) ? l : f1;
    }
    // ascending order by rank counting (ties broken by index)
    for (int i = 0; i < N; ++i) {
        R di = D(i);
        if (!(di == di) || isinf(di)) flag |= kStatNonFinite;
        int rank = 0;
        for (int jx = 0; jx < N; ++jx) {
            R dj = D(jx);
            rank += (dj < di) || (dj == di && jx < i);
        }
        rank = rank < N ? rank : N - 1;
        evals[(size_t)kk * N + rank] = di;
        if (perm) perm[(size_t)kk * N + rank] = i;
    }
    if (LOG) nsweeps[kk] = nsw;
    if (flag) atomicOr(status, flag);
}

// Fermi-Dirac occupation (Utils/fermidirac.jl:1-5): T > 0 logistic, T = 0 step
// with 1/2 exactly at zero.
__device__ __forceinline__ double fermidirac(double e, double mu, double T) {
    double x = e - mu;
    if (T > 0.0) return 1.0 / (exp(x / T) + 1.0);
    return x < 0.0 ? 1.0 : (x > 0.0 ? 0.0 : 0.5);
}

// ---------------------------------------------------------------------------------
// K_C: eigenvectors (and rho_k).  V is kept planar (real and imaginary planes,
// leading dimension ldv) so lanes walking consecutive rows hit consecutive banks.
//   1. Z = I; replay the QL sweeps, one row of Z per thread (carry trick: one LDS
//      read and one write per rotation);
//   2. V = H_0 H_1 ... H_{N-2} Z, reflectors applied from the last one;
//   3. MODE 0: write evecs (ascending order); MODE 1: rho_k = V diag(f) V^H with
//      f = fermidirac(e - mu; T) (or f_n = occ[kk*N + n] for the n-th smallest
//      eigenvalue when occ is given), written as FP64, and eb[kk] = w_k sum_n e_n f_n.
// ---------------------------------------------------------------------------------
constexpr int kRotChunk = 512;  // default rotation-staging capacity (>= longest sweep)
constexpr int kHdrChunk = 256;  // sweep headers staged per LDS load

// Apply one logged sweep to row rr of Z (column-major, leading dimension ldv):
// rotations at i = i0, i0-1, ..., i0-cnt+1 with (c, s) pairs at cs[0..2cnt).
// Four rotations per group; the next group's Z loads are issued before the
// current group's stores so the carry chain is the only serial dependency.
template <typename R>
__device__ __forceinline__ void replay_sweep(R* __restrict__ zc, int ldv, int i0, int cnt,
                                             const R* __restrict__ cs) {
    R carry = zc[(size_t)(i0 + 1) * ldv];
    int t = 0;
    R z0 = 0, z1 = 0, z2 = 0, z3 = 0;
    if (cnt >= 4) {
        z0 = zc[(size_t)(i0)*ldv];
        z1 = zc[(size_t)(i0 - 1) * ldv];
        z2 = zc[(size_t)(i0 - 2) * ldv];
        z3 = zc[(size_t)(i0 - 3) * ldv];
    }
    for (; t + 4 <= cnt; t += 4) {
        R y0 = 0, y1 = 0, y2 = 0, y3 = 0;
        if (t + 8 <= cnt) {
            y0 = zc[(size_t)(i0 - t - 4) * ldv];
            y1 = zc[(size_t)(i0 - t - 5) * ldv];
            y2 = zc[(size_t)(i0 - t - 6) * ldv];
            y3 = zc[(size_t)(i0 - t - 7) * ldv];
        }
        const R c0 = cs[2 * t], s0 = cs[2 * t + 1];
        const R c1 = cs[2 * t + 2], s1 = cs[2 * t + 3];
        const R c2 = cs[2 * t + 4], s2 = cs[2 * t + 5];
        const R c3 = cs[2 * t + 6], s3 = cs[2 * t + 7];
        const R o0 = s0 * z0 + c0 * carry;
        carry = c0 * z0 - s0 * carry;
        const R o1 = s1 * z1 + c1 * carry;
        carry = c1 * z1 - s1 * carry;
        const R o2 = s2 * z2 + c2 * carry;
        carry = c2 * z2 - s2 * carry;
        const R o3 = s3 * z3 + c3 * carry;
        carry = c3 * z3 - s3 * carry;
        zc[(size_t)(i0 - t + 1) * ldv] = o0;
        zc[(size_t)(i0 - t) * ldv] = o1;
        zc[(size_t)(i0 - t - 1) * ldv] = o2;
        zc[(size_t)(i0 - t - 2) * ldv] = o3;
        z0 = y0;
        z1 = y1;
        z2 = y2;
        z3 = y3;
    }
    for (; t < cnt; ++t) {
        const R zi = zc[(size_t)(i0 - t) * ldv];
        const R c = cs[2 * t], s = cs[2 * t + 1];
        zc[(size_t)(i0 - t + 1) * ldv] = s * zi + c * carry;
        carry = c * zi - s * carry;
    }
    zc[(size_t)(i0 - cnt + 1) * ldv] = carry;
}

// 64-bit broadcast from a wave-uniform lane (v_readlane into scalar registers).
template <typename R>
__device__ __forceinline__ R bcast_lane(R v, int lane);
template <>
__device__ __forceinline__ double bcast_lane<double>(double v, int lane) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)x, lane);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <>
__device__ __forceinline__ float bcast_lane<float>(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// ---------------------------------------------------------------------------------
// K_R: QL rotation replay for N <= NM <= 64, one wave per k-point.  Lane r keeps row r
// of Z in registers (fully unrolled over the NM-1 rotation positions; positions
// outside a sweep are skipped by a wave-uniform branch).  Each sweep (<= 63
// rotations) arrives with one coalesced load -- lane t holds pair t -- issued one
// sweep ahead, and pairs are broadcast with readlane, so no LDS is used and many
// waves share a CU.  Output: Z[kk][c*N + r] (column-major, real).
// ---------------------------------------------------------------------------------
template <typename R, int NM>
__global__ __launch_bounds__(kWave) void k_replay(int N, const R* __restrict__ rot,
                                                   const int* __restrict__ sweeps,
                                                   const int* __restrict__ nsweeps, int cap_rot,
                                                   int cap_sw, R* __restrict__ Z,
                                                   const int* __restrict__ zfmt) {
    const int kk = blockIdx.x;
    const int lane = threadIdx.x;
    if (zfmt && zfmt[kk]) return;  // eigenvectors already formed by k_tri_twist
    R z[NM];
#pragma unroll
    for (int c = 0; c < NM; ++c) z[c] = (c == lane) ? R(1) : R(0);
    const R* lrot = rot + (size_t)kk * cap_rot * 2;
    const int* lsw = sweeps + (size_t)kk * cap_sw;
    const int nsw = nsweeps[kk];
    int pos = 0;
    // pairs of the current sweep: lane t holds (c_t, s_t)
    R cc = 0, ss = 0;
    if (nsw > 0 && lane < min(kWave, cap_rot)) {
        cc = lrot[2 * lane];
        ss = lrot[2 * lane + 1];
    }
    for (int sidx = 0; sidx < nsw; ++sidx) {
        const int h = lsw[sidx];
        const int i0 = h >> 16, cnt = h & 0xffff;
        // prefetch the next sweep's pairs
        const int pn = pos + cnt;
        R nc = 0, ns = 0;
        if (sidx + 1 < nsw && pn + lane < cap_rot) {
            nc = lrot[2 * (size_t)(pn + lane)];
            ns = lrot[2 * (size_t)(pn + lane) + 1];
        }
#pragma unroll
        for (int i = NM - 2; i >= 0; --i) {
            if (i <= i0 && i > i0 - cnt) {
                const int t = i0 - i;
                const R c = bcast_lane(cc, t), s = bcast_lane(ss, t);
                const R zi = z[i], zi1 = z[i + 1];
                z[i + 1] = s * zi + c * zi1;
                z[i] = c * zi - s * zi1;
            }
        }
        cc = nc;
        ss = ns;
        pos = pn;
    }
    if (lane < N) {
        R* zo = Z + (size_t)kk * N * N + lane;
#pragma unroll
        for (int c = 0; c < NM; ++c)
            if (c < N) zo[(size_t)c * N] = z[c];
    }
}

template <typename R, bool LDSMAT, int MODE, bool ZIN>
__global__ __launch_bounds__(kTPB) void k_backtransform(
    int N, int Nkc, const cx<R>* __restrict__ refl, const cx<R>* __restrict__ tau,
    const R* __restrict__ rot, const int* __restrict__ sweeps, const int* __restrict__ nsweeps,
    int cap_rot, int cap_sw, int rch, const R* __restrict__ evals, const int* __restrict__ perm,
    const double* __restrict__ kw, double wuni, double mu, double Temp,
    const double* __restrict__ occ, R* __restrict__ gV, const R* __restrict__ zin,
    cx<double>* __restrict__ out, double* __restrict__ eb) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int kk = blockIdx.x;
    const int tid = threadIdx.x;
    const int ldv = LDSMAT ? N + 1 : N;
    const size_t plane = (size_t)N * ldv;
    unsigned char* p = smem;
    R* Vr;
    if constexpr (LDSMAT) {
        Vr = reinterpret_cast<R*>(p);
        p += sizeof(R) * 2 * plane;
    } else {
        Vr = gV + (size_t)kk * 2 * plane;
    }
    R* Vi = Vr + plane;
    R* srot = reinterpret_cast<R*>(p);          // [rch][2]
    p += sizeof(R) * 2 * rch;
    cx<R>* stau = reinterpret_cast<cx<R>*>(p);  // [N]
    p += sizeof(cx<R>) * N;
    cx<R>* sv = reinterpret_cast<cx<R>*>(p);    // [2][N] double-buffered reflector
    p += sizeof(cx<R>) * 2 * N;
    cx<R>* sy = reinterpret_cast<cx<R>*>(p);    // [N]
    p += sizeof(cx<R>) * N;
    double* sf = reinterpret_cast<double*>(p);  // [N]
    p += sizeof(double) * N;
    double* sred = reinterpret_cast<double*>(p);  // [4]
    p += sizeof(double) * 4;
    int* shdr = reinterpret_cast<int*>(p);        // [kHdrChunk]

    // 1. Z: replayed already (ZIN, from k_replay) or Z = I and the logged QL rotations
    const cx<R>* rf = refl + (size_t)kk * N * N;
    for (int j = tid; j < N; j += kTPB) stau[j] = tau[(size_t)kk * N + j];
    if constexpr (ZIN) {
        const R* zk = zin + (size_t)kk * N * N;
        for (int e = tid; e < N * N; e += kTPB) {
            int i = e % N, c = e / N;
            Vr[(size_t)c * ldv + i] = zk[e];
            Vi[(size_t)c * ldv + i] = R(0);
        }
    } else {
    for (int e = tid; e < N * N; e += kTPB) {
        int i = e % N, c = e / N;
        Vr[(size_t)c * ldv + i] = (i == c) ? R(1) : R(0);
        Vi[(size_t)c * ldv + i] = R(0);
    }
    const R* lrot = rot + (size_t)kk * cap_rot * 2;
    const int* lsw = sweeps + (size_t)kk * cap_sw;
    const int nsw = nsweeps[kk];
    int pos = 0, cbase = 0, cend = 0;  // srot holds log positions [cbase, cend)
    int hbase = 0, hend = 0;           // shdr holds sweep headers [hbase, hend)
    __syncthreads();
    for (int sidx = 0; sidx < nsw; ++sidx) {
        if (sidx >= hend) {  // uniform: every thread walks the same sweeps
            __syncthreads();
            const int nh = min(kHdrChunk, cap_sw - sidx);
            for (int q = tid; q < nh; q += kTPB) shdr[q] = lsw[sidx + q];
            hbase = sidx;
            hend = sidx + nh;
            __syncthreads();
        }
        const int h = shdr[sidx - hbase];
        const int i0 = h >> 16, cnt = h & 0xffff;
        if (pos + cnt > cend) {
            __syncthreads();
            const int nload = min(rch, cap_rot - pos);
            for (int q = tid; q < 2 * nload; q += kTPB) srot[q] = lrot[2 * (size_t)pos + q];
            cbase = pos;
            cend = pos + nload;
            __syncthreads();
        }
        const R* cs = srot + 2 * (pos - cbase);
        for (int rr = tid; rr < N; rr += kTPB) replay_sweep(Vr + rr, ldv, i0, cnt, cs);
        pos += cnt;
    }
    }

    // 2. V = H_0 ... H_{N-2} Z.  Reflector j-1 is fetched into the other half of sv
    //    while reflector j is applied, so no global latency sits between barriers.
    {
        const int j0 = N - 2;
        if (j0 >= 0)
            for (int i = tid; i < N - 1 - j0; i += kTPB) sv[(j0 & 1) * N + i] = rf[(size_t)j0 * N + j0 + 1 + i];
    }
    __syncthreads();
    for (int j = N - 2; j >= 0; --j) {
        const int m = N - 1 - j;
        const cx<R> tj = stau[j];
        const cx<R>* svj = sv + (j & 1) * N;
        const bool skip = (tj.re == R(0) && tj.im == R(0));
        // y_c = tau * v^H V[j+1:, c]   (4 lanes per column)
        if (!skip) {
            const int q = tid & 3;
            for (int c = tid >> 2; c < N; c += kTPB / 4) {
                cx<R> acc = {0, 0};
                const R* vr = Vr + (size_t)c * ldv + j + 1;
                const R* vi = Vi + (size_t)c * ldv + j + 1;
                for (int i = q; i < m; i += 4) acc = cmadc(acc, svj[i], mkc<R>(vr[i], vi[i]));
                acc.re += __shfl_xor(acc.re, 1);
                acc.im += __shfl_xor(acc.im, 1);
                acc.re += __shfl_xor(acc.re, 2);
                acc.im += __shfl_xor(acc.im, 2);
                if (q == 0) sy[c] = tj * acc;
            }
        }
        __syncthreads();
        // fetch reflector j-1 into registers now, store it to the other buffer (last
        // read in step j+1) after the update, so the global latency overlaps the update
        const int jn = j - 1, mn = N - 1 - jn;
        R pre_re = 0, pre_im = 0;  // scalars: a struct here ends up in scratch
        if (j >= 1 && tid < mn) {
            const R* src = reinterpret_cast<const R*>(rf + (size_t)jn * N + jn + 1 + tid);
            pre_re = src[0];
            pre_im = src[1];
        }
        // V[j+1:, :] -= v y
        if (!skip) {
            const int rr = tid & (kWave - 1);
            const int cq = tid / kWave;
            for (int r = rr; r < m; r += kWave) {
                const cx<R> vr = svj[r];
                for (int c = cq; c < N; c += kTPB / kWave) {
                    const size_t o = (size_t)c * ldv + j + 1 + r;
                    cx<R> u = vr * sy[c];
                    Vr[o] -= u.re;
                    Vi[o] -= u.im;
                }
            }
        }
        if (j >= 1) {
            cx<R>* dst = sv + (jn & 1) * N;
            if (tid < mn) dst[tid] = mkc<R>(pre_re, pre_im);
            for (int i = tid + kTPB; i < mn; i += kTPB) dst[i] = rf[(size_t)jn * N + jn + 1 + i];
        }
        __syncthreads();
    }

    const R* ev = evals + (size_t)kk * N;
    const int* pm = perm + (size_t)kk * N;
    const size_t NN = (size_t)N * N;
    if constexpr (MODE == 0) {
        cx<double>* o = out + (size_t)kk * NN;
        for (int e = tid; e < N * N; e += kTPB) {
            int i = e % N, n = e / N;
            int c = min(max(pm[n], 0), N - 1);  // perm is only garbage after a flagged NaN
            o[e] = mkc<double>((double)Vr[(size_t)c * ldv + i], (double)Vi[(size_t)c * ldv + i]);
        }
    } else {
        // occupations per Z column and the band-energy term
        double ebp = 0.0;
        for (int n = tid; n < N; n += kTPB) {
            double en = (double)ev[n];
            double f = occ ? occ[(size_t)kk * N + n] : fermidirac(en, mu, Temp);
            sf[min(max(pm[n], 0), N - 1)] = f;
            ebp += en * f;
        }
        ebp = wave_sum(ebp);
        if ((tid & (kWave - 1)) == 0) sred[tid / kWave] = ebp;
        __syncthreads();
        if (tid == 0) {
            double w = kw ? kw[kk] : wuni;
            eb[kk] = w * (sred[0] + sred[1] + sred[2] + sred[3]);
        }
        // rho_k(i, j) = sum_c f_c V(i, c) conj(V(j, c)), 4x4 register blocks
        const int nb = (N + 3) / 4;
        cx<double>* o = out + (size_t)kk * NN;
        for (int b = tid; b < nb * nb; b += kTPB) {
            const int bi = b % nb, bj = b / nb;
            cx<R> acc[4][4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] = mkc<R>(0, 0);
            for (int c = 0; c < N; ++c) {
                const R f = (R)sf[c];
                if (f == R(0)) continue;
                cx<R> a[4], bb[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    int i = 4 * bi + u;
                    int jj = 4 * bj + u;
                    a[u] = (i < N) ? mkc<R>(Vr[(size_t)c * ldv + i], Vi[(size_t)c * ldv + i]) : mkc<R>(0, 0);
                    bb[u] = (jj < N) ? f * mkc<R>(Vr[(size_t)c * ldv + jj], Vi[(size_t)c * ldv + jj]) : mkc<R>(0, 0);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[u][v] = cmadcb(acc[u][v], a[u], bb[v]);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    int i = 4 * bi + u, jj = 4 * bj + v;
                    if (i < N && jj < N)
                        o[(size_t)jj * N + i] = mkc<double>((double)acc[u][v].re, (double)acc[u][v].im);
                }
        }
    }
}

}  // namespace lqm
