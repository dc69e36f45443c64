// Eigenpairs of the real symmetric tridiagonal (d, e) of every k-point for N <= 64, in
// O(N^2) work per k-point and parallel over the eigenvalues: one wave per k-point, lane n
// owns eigenvalue n.
//   k_tri_bisect   eigenvalue n by bisection on the Sturm count, switching to Newton steps
//                  on det(T - x I) once the bracket holds eigenvalue n alone;
//   k_tri_twist    its eigenvector from a twisted factorisation of T - lambda_n I; the
//                  vectors of eigenvalues closer than ctol ||T|| to their predecessor are
//                  orthonormalised together (their error, ~ eps ||T|| / gap, lies mostly in
//                  the span of the cluster).
// A k-point where a cluster's vectors come out nearly parallel (eigenvalues equal to
// working precision) is listed for the QL iteration and its rotation replay instead, whose
// vectors are orthogonal however close the eigenvalues are.
// Layouts: d, e structure-of-arrays dd[i * Nkc + k] as written by the tridiagonalization;
// eigenvalues evals[k * N + n] ascending; eigenvectors row-major Zt[k][i * N + n] (lane n
// writes consecutive addresses), flagged by zfmt[k] = 1 (0: column-major from the replay).
#pragma once
#include "lqm_common.hpp"

namespace lqm {

template <typename R>
__device__ __forceinline__ bool tri_finite(R x) {
    return x - x == R(0);
}

// One wave per k-point (4 per block), lane n < N: eigenvalue n (ascending).  The Sturm
// count is the number of negative pivots q_i of T - x I = L D L^T, a pivot of modulus below
// pivmin = tiny max(1, max e^2) replaced by -pivmin (LAPACK dstebz).  The same recurrence
// gives d/dx log det(T - x I) = sum_i q_i'/q_i, so once the bracket [a, b] holds
// eigenvalue n alone a Newton step x - 1/sum replaces the bisection whenever it lands
// inside the bracket.  Converged when the Newton step is at most 4 eps ||T|| and points
// into the bracket, or the bracket is at most max(2 eps max(|a|, |b|), eps ||T||) (||T|| =
// the Gershgorin bound).
// Out: evals, perm[k * N + n] = n, tnorm[k] = ||T||; a non-finite eigenvalue sets
// kStatNonFinite in *status.
template <typename R>
__global__ __launch_bounds__(256) void k_tri_bisect(int N, int Nkc, const R* __restrict__ dd,
                                                    const R* __restrict__ ee, R* __restrict__ evals,
                                                    int* __restrict__ perm, R* __restrict__ tnorm,
                                                    int* __restrict__ status) {
    __shared__ R sd[4][kWave], se2[4][kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6;
    const int kk = blockIdx.x * 4 + w;
    const bool act = kk < Nkc;
    R ev = R(0);
    if (act && lane < N) {
        sd[w][lane] = dd[(size_t)lane * Nkc + kk];
        ev = lane + 1 < N ? ee[(size_t)lane * Nkc + kk] : R(0);
        se2[w][lane] = ev * ev;
    }
    __syncthreads();
    if (!act) return;  // no barrier below
    // a non-finite entry: fmin/fmax below would pass over it, so it is caught here
    if (__ballot(lane < N && !(tri_finite(sd[w][lane]) && tri_finite(ev))) != 0) {
        if (lane < N) {
            evals[(size_t)kk * N + lane] = sd[w][lane] - sd[w][lane] + ev - ev;
            perm[(size_t)kk * N + lane] = lane;
        }
        if (lane == 0) {
            tnorm[kk] = R(0);
            atomicOr(status, kStatNonFinite);
        }
        return;
    }
    // Gershgorin interval and the largest e^2
    const R eabs = fabs(ev);
    const R eleft = __shfl(eabs, lane > 0 ? lane - 1 : 0);
    R lo = R(3e38), hi = R(-3e38), e2m = R(0);
    if (lane < N) {
        const R el = lane > 0 ? eleft : R(0);
        lo = sd[w][lane] - el - eabs;
        hi = sd[w][lane] + el + eabs;
        e2m = se2[w][lane];
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, o));
        hi = fmax(hi, __shfl_xor(hi, o));
        e2m = fmax(e2m, __shfl_xor(e2m, o));
    }
    const R eps = R(2) * num<R>::half_eps;
    const R bnorm = fmax(fabs(lo), fabs(hi));
    const R pivmin = num<R>::tiny * fmax(R(1), e2m);
    const R atol = eps * bnorm, ntol = R(4) * eps * bnorm;
    if (lane == 0) tnorm[kk] = bnorm;
    R a = lo - R(2) * eps * bnorm - R(2) * pivmin, b = hi + R(2) * eps * bnorm + R(2) * pivmin;
    int ca = 0, cb = N;
    {
        // one Sturm pass at 64 points x_l = a + (l + 1) h, h = (b - a) / 65, one per lane (every
        // lane of the wave, including lanes >= N): each eigenvalue's bracket narrows 65-fold
        // at the cost of one bisection step
        const R h = (b - a) / R(65);
        const R xg = a + R(lane + 1) * h;
        R q = sd[w][0] - xg;
        if (fabs(q) <= pivmin) q = -pivmin;
        int cg = q < R(0);
        R r = rcp_nr(q);
        for (int i = 1; i < N; ++i) {
            R qn = (sd[w][i] - xg) - se2[w][i - 1] * r;
            if (fabs(qn) <= pivmin) qn = -pivmin;
            r = rcp_nr(qn);
            cg += qn < R(0);
        }
        // lane n: the first grid point whose count exceeds n (lower-bound search; 64 = none)
        int l0 = 0, l1 = 64;
#pragma unroll
        for (int s = 0; s < 7; ++s) {
            const int mid = (l0 + l1) >> 1;
            const int cm = __shfl(cg, min(mid, 63));
            if (l0 < l1) {
                if (cm > lane) l1 = mid;
                else l0 = mid + 1;
            }
        }
        const int c1 = __shfl(cg, max(l0 - 1, 0)), c2 = __shfl(cg, min(l0, 63));
        const int na = l0 == 0 ? 0 : c1, nb = l0 == 64 ? N : c2;
        // kept only if it is a consistent bracket (ca <= n < cb) of eigenvalue n; its ends are
        // the grid points, recomputed with the same expression as their lanes used
        if (na <= lane && lane < nb) {
            const R a0 = a;
            if (l0 > 0) {
                ca = na;
                a = a0 + R(l0) * h;
            }
            if (l0 < 64) {
                cb = nb;
                b = a0 + R(l0 + 1) * h;
            }
        }
    }
    if (lane >= N) return;
    const int n = lane;
    R x = R(0.5) * (a + b);
    for (int it = 0; it < 256; ++it) {
        R q = sd[w][0] - x;
        bool bad = false;
        if (fabs(q) <= pivmin) {
            q = -pivmin;
            bad = true;
        }
        int cnt = q < R(0);
        R r = rcp_nr(q);
        R qp = R(-1), S = -r;
        for (int i = 1; i < N; ++i) {
            const R t = se2[w][i - 1] * r;
            R qn = (sd[w][i] - x) - t;
            if (fabs(qn) <= pivmin) {
                qn = -pivmin;
                bad = true;
            }
            qp = fma(t * r, qp, R(-1));
            r = rcp_nr(qn);
            S = fma(qp, r, S);
            cnt += qn < R(0);
        }
        const bool above = cnt > n;  // eigenvalue n lies below x
        if (above) {
            b = x;
            cb = cnt;
        } else {
            a = x;
            ca = cnt;
        }
        const R mid = R(0.5) * (a + b);
        const bool iso = cb - ca == 1;
        const R step = R(-1) / S;
        const R xn = x + step;
        const bool fin = !bad && tri_finite(xn);
        // a small step counts only when it points into the bracket: next to an end, a step
        // toward the neighbouring eigenvalue outside is small as well
        if (iso && fin && fabs(step) <= ntol && (above ? step <= R(0) : step >= R(0))) {
            x = fmin(fmax(xn, a), b);
            break;
        }
        if (!(mid > a && mid < b) || b - a <= fmax(R(2) * eps * fmax(fabs(a), fabs(b)), atol)) {
            x = mid;
            break;
        }
        x = (iso && fin && xn > a && xn < b) ? xn : mid;
    }
    evals[(size_t)kk * N + n] = x;
    perm[(size_t)kk * N + n] = n;
    if (!tri_finite(x)) atomicOr(status, kStatNonFinite);
}

// Eigenvector of lambda_n (lane n): twisted factorisation
// T - lambda I = N_r diag(gamma) N_r^T, r = argmin |gamma_r| with
// gamma_r = (d_r - lambda) - e_{r-1}^2 / D+_{r-1} - e_r^2 / D-_{r+1} (D+ forward, D-
// backward pivots, tiny pivots replaced by -pivmin), then z_r = 1,
// z_i = -e_i z_{i+1} / D+_i (i < r), z_i = -e_{i-1} z_{i-1} / D-_i (i > r), normalised.
// One register array of NM entries per lane holds 1/D+, then 1/D- above r, then z.
// Eigenvalues within ctol_rel ||T|| of their predecessor form a cluster, whose vectors are
// orthonormalised in eigenvalue order by modified Gram-Schmidt applied twice (the other
// members' entries arrive by cross-lane shuffles).  A vector that keeps less than 1/10 of
// its norm against the earlier members is numerically dependent on them: the k-point then
// gets zfmt[k] = 0 and is appended to list[1 + ...] (list[0] counts, zeroed by the caller)
// for the QL path; otherwise zfmt[k] = 1 and the vectors go to Zt.
template <typename R, int NM>
__global__ __launch_bounds__(256) void k_tri_twist(int N, int Nkc, const R* __restrict__ dd,
                                                   const R* __restrict__ ee, const R* __restrict__ evals,
                                                   const R* __restrict__ tnorm, R ctol_rel, R* __restrict__ Zt,
                                                   int* __restrict__ zfmt, int* __restrict__ list) {
    __shared__ R sd[4][NM], se[4][NM], se2[4][NM];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6;
    const int kk = blockIdx.x * 4 + w;
    const bool act = kk < Nkc;
    R e2m = R(0);
    if (act && lane < N) {
        sd[w][lane] = dd[(size_t)lane * Nkc + kk];
        const R e = lane + 1 < N ? ee[(size_t)lane * Nkc + kk] : R(0);
        se[w][lane] = e;
        se2[w][lane] = e * e;
        e2m = e * e;
    }
    __syncthreads();
    if (!act) return;  // no barrier below; every lane stays for the shuffles
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) e2m = fmax(e2m, __shfl_xor(e2m, o));
    const bool mine = lane < N;
    const R pivmin = num<R>::tiny * fmax(R(1), e2m);
    auto guard = [&](R q) { return fabs(q) <= pivmin ? -pivmin : q; };
    const R lam = mine ? evals[(size_t)kk * N + lane] : R(0);
    R z[NM];
    // forward pivots: z[i] = 1 / D+_i
    R rp = rcp_nr(guard(sd[w][0] - lam));
    z[0] = rp;
#pragma unroll
    for (int i = 1; i < NM; ++i) {
        z[i] = R(0);
        if (i < N) {
            rp = rcp_nr(guard((sd[w][i] - lam) - se2[w][i - 1] * rp));
            z[i] = rp;
        }
    }
    // backward pivots and the twist index
    int rr = N - 1;
    R gbest = R(3e38), rm = R(0);
#pragma unroll
    for (int i = NM - 1; i >= 0; --i)
        if (i < N) {
            const R dl = sd[w][i] - lam;
            R gam = dl;
            if (i > 0) gam -= se2[w][i - 1] * z[i - 1];
            if (i + 1 < N) gam -= se2[w][i] * rm;
            if (fabs(gam) < gbest) {
                gbest = fabs(gam);
                rr = i;
            }
            rm = rcp_nr(guard(i + 1 < N ? dl - se2[w][i] * rm : dl));
        }
    // 1 / D-_i above the twist
    rm = R(0);
#pragma unroll
    for (int i = NM - 1; i >= 1; --i)
        if (i < N) {
            const R dl = sd[w][i] - lam;
            rm = rcp_nr(guard(i + 1 < N ? dl - se2[w][i] * rm : dl));
            if (i > rr) z[i] = rm;
        }
    // the vector, outward from the twist
    R t = R(0);
#pragma unroll
    for (int i = NM - 1; i >= 0; --i)
        if (i < N) {
            t = i == rr ? R(1) : (i < rr ? -se[w][i] * z[i] * t : R(0));
            if (i <= rr) z[i] = t;
        }
    t = R(0);
    R s2 = R(0);
#pragma unroll
    for (int i = 0; i < NM; ++i)
        if (i < N) {
            t = i == rr ? R(1) : (i > rr ? -se[w][i - 1] * z[i] * t : R(0));
            if (i >= rr) z[i] = t;
            s2 = fma(z[i], z[i], s2);
        }
    bool bad = mine && !tri_finite(s2);
    {
        const R sc = mine && s2 > R(0) ? R(1) / sqrt(s2) : R(0);
#pragma unroll
        for (int i = 0; i < NM; ++i) z[i] = mine ? z[i] * sc : R(0);
    }
    // clusters: position of each eigenvalue after the last gap above ctol
    const R ctol = ctol_rel * tnorm[kk];
    const R lprev = __shfl(lam, lane > 0 ? lane - 1 : 0);
    const bool brk = !mine || lane == 0 || lam - lprev > ctol;
    const unsigned long long bm = __ballot(brk);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int start = 63 - __clzll(bm & upto);
    const int pos = lane - start;
    int pmax = pos;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) pmax = max(pmax, __shfl_xor(pmax, o));
    for (int tp = 1; tp <= pmax; ++tp) {
        const bool me = pos == tp;
        R keep = R(1);
        for (int pass = 0; pass < 2; ++pass) {
            for (int sft = tp; sft >= 1; --sft) {
                const int src = me ? lane - sft : lane;
                R dot = R(0);
#pragma unroll
                for (int i = 0; i < NM; ++i)
                    if (i < N) dot = fma(__shfl(z[i], src), z[i], dot);
                if (!me) dot = R(0);
#pragma unroll
                for (int i = 0; i < NM; ++i)
                    if (i < N) z[i] = fma(-dot, __shfl(z[i], src), z[i]);
            }
            R nn = R(0);
#pragma unroll
            for (int i = 0; i < NM; ++i) nn = fma(z[i], z[i], nn);
            if (pass == 0) keep = nn;
            const R sc = me && nn > R(0) ? R(1) / sqrt(nn) : R(1);
#pragma unroll
            for (int i = 0; i < NM; ++i) z[i] *= sc;
        }
        if (me && !(keep >= R(0.01))) bad = true;
    }
    const bool flag = __ballot(bad) != 0;
    if (lane == 0) {
        zfmt[kk] = flag ? 0 : 1;
        if (flag) {
            const int at = atomicAdd(list, 1);
            list[1 + at] = kk;
        }
    }
    if (flag || !mine) return;
    R* zo = Zt + (size_t)kk * N * N + lane;
#pragma unroll
    for (int i = 0; i < NM; ++i)
        if (i < N) zo[(size_t)i * N] = z[i];
}

}  // namespace lqm
