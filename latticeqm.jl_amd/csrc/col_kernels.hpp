// Column-per-lane Householder tridiagonalization for N <= NW * RW <= 64 (NW waves per
// k-point).
//
// Lane c of every wave owns column c of the Hermitian matrix; wave w holds rows
// RW w .. RW w + RW - 1 in registers, a[i] = A[RW w + i][c].  With whole columns in lanes,
// the matrix-vector product of a Householder step needs no reduction across lanes: A
// Hermitian gives (A v)_c = sum_r conj(A[r][c]) v_r, a dot product of the lane's own
// column with the broadcast vector.  The Householder vector is row j of the matrix,
// conjugated: A[r][j] = conj(A[j][r]), and A[j][c] sits in lane c (register xr, picked out
// of the trailing update of the step before), so v is formed one entry per lane, its norm
// is one wave reduction, and it reaches every lane through 64 LDS entries read at
// wave-uniform addresses (broadcast, no bank conflicts).
//
// Rows at or above j are finished.  The steps run in phases: phase P holds the steps whose
// first live row j + 1 lies in row chunk P (CH rows), so inside a phase the live chunks are
// a compile-time range and the register updates carry no branches (branches around
// register-array updates make the compiler copy the array).  Each wave runs its own
// instantiation (its row range is a compile-time constant).
//
// One wave (NW = 1) runs a whole step without a barrier.  With NW = 2 the waves exchange
// their partial products through LDS (a barrier after the Householder vector and one after
// the partial products) until the trailing matrix lies in the last wave's rows; from then
// on that wave runs alone.
//
// Step j (zhetd2, lower; the same outputs as k_tridiag_r):
//   x = A[j+1:N, j]; beta = -sign(Re x_0) ||x||; tau = (beta - x_0)/beta;
//   v = [1, x[1:]/(x_0 - beta)]                       (zlarfg: H^H x = beta e_1)
//   y = A22 v;  p = tau y; s = v^H y (real); alpha2 = -1/2 |tau|^2 s; w = p + alpha2 v
//   A22 <- A22 - v w^H - w v^H
// (tau = 0 leaves w = 0, so the update changes nothing and needs no branch.)
// H(k) comes in materialised (FP64, full Hermitian with a real diagonal, lower triangle
// the bitwise conjugate of the upper) and is rounded to the working precision R on load;
// lane c reads row c (coalesced) and conjugates it.
#pragma once
#include "lqm_common.hpp"
#include "reg_kernels.hpp"

namespace lqm {

template <typename R, int NW>
struct ColTriSmem {
    alignas(16) cx<R> sv[2][64];      // Householder vector of the step (by step parity)
    alignas(16) cx<R> sy[2][NW][64];  // per-wave partial products y (NW > 1)
    alignas(16) cx<R> sw[NW][64];     // w, one copy per wave (written and read by that wave only)
    cx<R> stau[2];
};

// Sum over the 64 lanes, the same bits in every lane (row sums by DPP rotations, then the
// four row sums in fixed order).
template <typename R>
__device__ __forceinline__ R wave_sum64(R x) {
    const R r = row16_sum(x);
    return ((rl(r, 0) + rl(r, 16)) + rl(r, 32)) + rl(r, 48);
}

// Maximum over the 64 lanes (the same in every lane).
template <typename R>
__device__ __forceinline__ R wave_max64(R x) {
    x = fmax(x, dpp_r<0x121>(x));
    x = fmax(x, dpp_r<0x122>(x));
    x = fmax(x, dpp_r<0x124>(x));
    x = fmax(x, dpp_r<0x128>(x));
    return fmax(fmax(rl(x, 0), rl(x, 16)), fmax(rl(x, 32), rl(x, 48)));
}

// Row broadcast by DPP (row_newbcast:n, gfx90a and later; for 64-bit operands too): every
// lane of a 16-lane row reads lane n of that row.  With a vector laid out so that lane l of
// every row holds entry 16 k + (l & 15) in register k, the broadcast of register k hands
// entry 16 k + n to all 64 lanes.  The multiply-adds take the broadcast operand directly
// (v_fmac_f32_dpp / v_fmac_f64_dpp), so the row values of the matrix-vector product and of
// the rank-2 update cost no instruction and no LDS read of their own.  The compiler does
// not form these from separate DPP moves, hence the asm; each block starts with two wait
// states (a VALU write of a VGPR read through DPP right after needs them).
#define LQM_FMAC_BC(OP, D, S0, S1) OP " " D ", " S0 ", " S1 " row_newbcast:%[n] row_mask:0xf bank_mask:0xf\n\t"

// y += conj(a) v_r, v_r = entry n of the broadcast vector (vr, vi): yre = Re y, and
// Im y = yp - yq, so that every multiply-add adds (no negated operands)
template <int n>
__device__ __forceinline__ void mv_dpp(float& yre, float& yp, float& yq, float vr, float vi, float are, float aim) {
    asm("s_nop 1\n\t"
        LQM_FMAC_BC("v_fmac_f32_dpp", "%[yre]", "%[vr]", "%[are]")
        LQM_FMAC_BC("v_fmac_f32_dpp", "%[yre]", "%[vi]", "%[aim]")
        LQM_FMAC_BC("v_fmac_f32_dpp", "%[yp]", "%[vi]", "%[are]")
        LQM_FMAC_BC("v_fmac_f32_dpp", "%[yq]", "%[vr]", "%[aim]")
        : [yre] "+v"(yre), [yp] "+v"(yp), [yq] "+v"(yq)
        : [vr] "v"(vr), [vi] "v"(vi), [are] "v"(are), [aim] "v"(aim), [n] "n"(n));
}
template <int n>
__device__ __forceinline__ void mv_dpp(double& yre, double& yp, double& yq, double vr, double vi, double are,
                                       double aim) {
    asm("s_nop 1\n\t"
        LQM_FMAC_BC("v_fmac_f64_dpp", "%[yre]", "%[vr]", "%[are]")
        LQM_FMAC_BC("v_fmac_f64_dpp", "%[yre]", "%[vi]", "%[aim]")
        LQM_FMAC_BC("v_fmac_f64_dpp", "%[yp]", "%[vi]", "%[are]")
        LQM_FMAC_BC("v_fmac_f64_dpp", "%[yq]", "%[vr]", "%[aim]")
        : [yre] "+v"(yre), [yp] "+v"(yp), [yq] "+v"(yq)
        : [vr] "v"(vr), [vi] "v"(vi), [are] "v"(are), [aim] "v"(aim), [n] "n"(n));
}

// a -= v_r conj(w_c) + w_r conj(v_c), v_r and w_r entry n of the broadcast vectors (vr, vi),
// (wr, wi); the column values come as nw = -w_c, nv = -v_c, wim = Im w_c, vim = Im v_c
template <int n>
__device__ __forceinline__ void up_dpp(float& are, float& aim, float vr, float vi, float wr, float wi, float nwre,
                                       float nwim, float nvre, float nvim, float wim, float vim) {
    asm("s_nop 1\n\t"
        LQM_FMAC_BC("v_fmac_f32_dpp", "%[are]", "%[vr]", "%[nwre]")
        LQM_FMAC_BC("v_fmac_f32_dpp", "%[are]", "%[vi]", "%[nwim]")
        LQM_FMAC_BC("v_fmac_f32_dpp", "%[are]", "%[wr]", "%[nvre]")
        LQM_FMAC_BC("v_fmac_f32_dpp", "%[are]", "%[wi]", "%[nvim]")
        LQM_FMAC_BC("v_fmac_f32_dpp", "%[aim]", "%[vi]", "%[nwre]")
        LQM_FMAC_BC("v_fmac_f32_dpp", "%[aim]", "%[vr]", "%[wim]")
        LQM_FMAC_BC("v_fmac_f32_dpp", "%[aim]", "%[wi]", "%[nvre]")
        LQM_FMAC_BC("v_fmac_f32_dpp", "%[aim]", "%[wr]", "%[vim]")
        : [are] "+v"(are), [aim] "+v"(aim)
        : [vr] "v"(vr), [vi] "v"(vi), [wr] "v"(wr), [wi] "v"(wi), [nwre] "v"(nwre), [nwim] "v"(nwim),
          [nvre] "v"(nvre), [nvim] "v"(nvim), [wim] "v"(wim), [vim] "v"(vim), [n] "n"(n));
}
template <int n>
__device__ __forceinline__ void up_dpp(double& are, double& aim, double vr, double vi, double wr, double wi,
                                       double nwre, double nwim, double nvre, double nvim, double wim, double vim) {
    asm("s_nop 1\n\t"
        LQM_FMAC_BC("v_fmac_f64_dpp", "%[are]", "%[vr]", "%[nwre]")
        LQM_FMAC_BC("v_fmac_f64_dpp", "%[are]", "%[vi]", "%[nwim]")
        LQM_FMAC_BC("v_fmac_f64_dpp", "%[are]", "%[wr]", "%[nvre]")
        LQM_FMAC_BC("v_fmac_f64_dpp", "%[are]", "%[wi]", "%[nvim]")
        LQM_FMAC_BC("v_fmac_f64_dpp", "%[aim]", "%[vi]", "%[nwre]")
        LQM_FMAC_BC("v_fmac_f64_dpp", "%[aim]", "%[vr]", "%[wim]")
        LQM_FMAC_BC("v_fmac_f64_dpp", "%[aim]", "%[wi]", "%[nvre]")
        LQM_FMAC_BC("v_fmac_f64_dpp", "%[aim]", "%[wr]", "%[vim]")
        : [are] "+v"(are), [aim] "+v"(aim)
        : [vr] "v"(vr), [vi] "v"(vi), [wr] "v"(wr), [wi] "v"(wi), [nwre] "v"(nwre), [nwim] "v"(nwim),
          [nvre] "v"(nvre), [nvim] "v"(nvim), [wim] "v"(wim), [vim] "v"(vim), [n] "n"(n));
}
#undef LQM_FMAC_BC

// The same with the lane index known only after unrolling (the switch folds away).
#define LQM_BC_SWITCH(CALL)      \
    switch (n & 15) {            \
        case 0: CALL(0); break;   \
        case 1: CALL(1); break;   \
        case 2: CALL(2); break;   \
        case 3: CALL(3); break;   \
        case 4: CALL(4); break;   \
        case 5: CALL(5); break;   \
        case 6: CALL(6); break;   \
        case 7: CALL(7); break;   \
        case 8: CALL(8); break;   \
        case 9: CALL(9); break;   \
        case 10: CALL(10); break; \
        case 11: CALL(11); break; \
        case 12: CALL(12); break; \
        case 13: CALL(13); break; \
        case 14: CALL(14); break; \
        default: CALL(15); break; \
    }
template <typename R>
__device__ __forceinline__ void mv_bc(int n, cx<R> vb, cx<R> a, R& yre, R& yp, R& yq) {
#define LQM_CALL(K) mv_dpp<K>(yre, yp, yq, vb.re, vb.im, a.re, a.im)
    LQM_BC_SWITCH(LQM_CALL)
#undef LQM_CALL
}
template <typename R>
__device__ __forceinline__ cx<R> up_bc(int n, cx<R> vb, cx<R> wb, cx<R> nw, cx<R> nv, R wim, R vim, cx<R> a) {
#define LQM_CALL(K) up_dpp<K>(a.re, a.im, vb.re, vb.im, wb.re, wb.im, nw.re, nw.im, nv.re, nv.im, wim, vim)
    LQM_BC_SWITCH(LQM_CALL)
#undef LQM_CALL
    return a;
}
#undef LQM_BC_SWITCH

// Step j on the rows of wave W; LQ = first live local row chunk (LQ = NQ: none live), XR =
// row j + 1 lies in local chunk LQ (it becomes the next xr).
template <typename R, int NW, int RW, int CH, int BC, int W, int LQ, bool XR>
__device__ __forceinline__ void coltri_step(cx<R> (&a)[RW], cx<R>& xr, ColTriSmem<R, NW>& S, int N, int Nkc,
                                            int kk, int c, int j, cx<R>* __restrict__ rf, R* __restrict__ dd,
                                            R* __restrict__ ee, cx<R>* __restrict__ tau) {
    constexpr int NQ = RW / CH;
    constexpr int r0 = RW * W;
    const int buf = j & 1;
    const bool solo = j >= RW * (NW - 1);  // rows j and after all in the last wave
    if (j / RW == W) {
        // (a) Householder vector of column j, one entry per lane: x_c = conj(A[j][c])
        const cx<R> xc = c > j ? conj(xr) : mkc<R>(0, 0);
        const R xn2 = wave_sum64(c > j + 1 ? abs2(xc) : (R)0);
        const cx<R> alpha = mkc<R>(rl(xc.re, j + 1), rl(xc.im, j + 1));
        cx<R> tj_, scale;
        R beta;
        if (!hh_scalars(alpha, xn2, tj_, scale, beta)) {
            // a column of tiny entries: the scalars of x / max|x|, then beta and scale back
            const R mx = wave_max64(c > j ? fmax(fabs(xc.re), fabs(xc.im)) : (R)0);
            const R inv = (R)1 / mx;
            const cx<R> xs = inv * xc;
            const R xn2s = wave_sum64(c > j + 1 ? abs2(xs) : (R)0);
            hh_scalars_scaled(mkc<R>(rl(xs.re, j + 1), rl(xs.im, j + 1)), xn2s, tj_, scale, beta);
            beta *= mx;
            scale = inv * scale;
            if (!(inv <= num<R>::huge)) {
                // entries below 1 / huge (subnormal): left in place, H_j = I
                tj_ = mkc<R>(0, 0);
                scale = mkc<R>(0, 0);
                beta = alpha.re;
            }
        }
        const bool triv = (tj_.re == (R)0 && tj_.im == (R)0);
        cx<R> vc = mkc<R>(0, 0);
        if (c == j + 1) vc = mkc<R>(1, 0);
        else if (c > j + 1 && !triv) vc = xc * scale;
        S.sv[buf][c] = vc;
        if (c > j && c < N) rf[(size_t)j * N + c] = vc;
        if (c == j) dd[(size_t)j * Nkc + kk] = xr.re;  // A[j][j]
        if (c == 0) {
            S.stau[buf] = tj_;
            tau[(size_t)kk * N + j] = tj_;
            ee[(size_t)j * Nkc + kk] = beta;
        }
    }
    if (NW > 1 && !solo) __syncthreads();
    else __builtin_amdgcn_wave_barrier();
    const cx<R> t = S.stau[buf];

    // v in the broadcast layout: lane l of every row holds v[r0 + 16 k + (l & 15)] in vb[k]
    constexpr int NB = RW / 16;
    cx<R> vb[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) vb[k] = S.sv[buf][r0 + 16 * k + (c & 15)];

    // (b) y_c = sum over this wave's live rows of conj(A[r][c]) v_r (v_r = 0 for r <= j), two
    //     partial sums (even and odd rows) for independent multiply-add chains; v_r by DPP
    //     broadcast inside the multiply-adds (BC bit 0) or read from LDS at a wave-uniform
    //     address
    cx<R> y;
    if constexpr (BC & 1) {
        R yre0 = 0, yp0 = 0, yq0 = 0, yre1 = 0, yp1 = 0, yq1 = 0;
#pragma unroll
        for (int q = LQ; q < NQ; ++q) {
#pragma unroll
            for (int u = 0; u < CH; u += 2) {
                const int i0 = CH * q + u;
                mv_bc(i0, vb[i0 >> 4], a[i0], yre0, yp0, yq0);
                mv_bc(i0 + 1, vb[(i0 + 1) >> 4], a[i0 + 1], yre1, yp1, yq1);
            }
        }
        y = mkc<R>(yre0 + yre1, (yp0 + yp1) - (yq0 + yq1));
    } else {
        cx<R> y0 = mkc<R>(0, 0), y1 = mkc<R>(0, 0);
#pragma unroll
        for (int q = LQ; q < NQ; ++q) {
            cx<R> vr[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) vr[u] = S.sv[buf][r0 + CH * q + u];
#pragma unroll
            for (int u = 0; u < CH; u += 2) {
                y0 = cmadc(y0, a[CH * q + u], vr[u]);
                y1 = cmadc(y1, a[CH * q + u + 1], vr[u + 1]);
            }
        }
        y = y0 + y1;
    }
    if constexpr (NW > 1) {
        if (!solo) {
            S.sy[buf][W][c] = y;
            __syncthreads();
            y = S.sy[buf][0][c];
#pragma unroll
            for (int ww = 1; ww < NW; ++ww) y = y + S.sy[buf][ww][c];
        }
    }
    if (c <= j) y = mkc<R>(0, 0);  // finished columns (stale entries below their diagonal)
    // (c) s = v^H y, w = tau y + alpha2 v
    const cx<R> vc = S.sv[buf][c];
    const R s = wave_sum64(y.re * vc.re + y.im * vc.im);
    const R alpha2 = (R)-0.5 * (t.re * t.re + t.im * t.im) * s;
    const cx<R> wc = cmad(t * y, mkc<R>(alpha2, 0), vc);
    S.sw[W][c] = wc;
    __builtin_amdgcn_wave_barrier();
    // (d) live rows: A[r][c] -= v_r conj(w_c) + w_r conj(v_c), v_r and w_r by DPP broadcast
    //     (BC bit 1) or read from LDS
    if constexpr (BC & 2) {
        cx<R> wb[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) wb[k] = S.sw[W][r0 + 16 * k + (c & 15)];
        const cx<R> nw = mkc<R>(-wc.re, -wc.im), nv = mkc<R>(-vc.re, -vc.im);
#pragma unroll
        for (int q = LQ; q < NQ; ++q) {
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int i = CH * q + u;
                a[i] = up_bc(i, vb[i >> 4], wb[i >> 4], nw, nv, wc.im, vc.im, a[i]);
            }
        }
    } else {
#pragma unroll
        for (int q = LQ; q < NQ; ++q) {
            cx<R> vr[CH], wr[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                vr[u] = S.sv[buf][r0 + CH * q + u];
                wr[u] = S.sw[W][r0 + CH * q + u];
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                a[CH * q + u] = cmadcb(a[CH * q + u], mkc<R>(-vr[u].re, -vr[u].im), wc);
                a[CH * q + u] = cmadcb(a[CH * q + u], mkc<R>(-wr[u].re, -wr[u].im), vc);
            }
        }
    }
    if constexpr (XR && LQ < NQ) {
        // row j + 1 (in chunk LQ) is the next step's xr
        const int jj = j + 1 - r0 - CH * LQ;
        cx<R> x = a[CH * LQ];
#pragma unroll
        for (int u = 1; u < CH; ++u) {
            x.re = jj == u ? a[CH * LQ + u].re : x.re;
            x.im = jj == u ? a[CH * LQ + u].im : x.im;
        }
        xr = x;
    }
}

// Phase P: the steps j with j + 1 in global row chunk P.  Returns false once this wave has
// no further part (its rows all finished and its last row's Householder vector formed).
template <typename R, int NW, int RW, int CH, int BC, int W, int P>
__device__ __forceinline__ bool coltri_phase(cx<R> (&a)[RW], cx<R>& xr, ColTriSmem<R, NW>& S, int N, int Nkc,
                                             int kk, int c, cx<R>* __restrict__ rf, R* __restrict__ dd,
                                             R* __restrict__ ee, cx<R>* __restrict__ tau) {
    constexpr int NQ = RW / CH;
    constexpr int G0 = RW * W / CH;  // first global chunk of this wave
    const int jlo = CH * P - 1 > 0 ? CH * P - 1 : 0;
    const int jhi = min(CH * P + CH - 2, N - 2);
    if constexpr (P >= G0 + NQ) {
        // this wave's rows are all finished: its last row's step still needs its Householder
        // vector (the first step of the phase), then it leaves
        if (jlo <= jhi) coltri_step<R, NW, RW, CH, BC, W, NQ, false>(a, xr, S, N, Nkc, kk, c, jlo, rf, dd, ee, tau);
        return false;
    } else {
        constexpr int LQ = P - G0 > 0 ? P - G0 : 0;
        constexpr bool XR = P >= G0;
        for (int j = jlo; j <= jhi; ++j)
            coltri_step<R, NW, RW, CH, BC, W, LQ, XR>(a, xr, S, N, Nkc, kk, c, j, rf, dd, ee, tau);
        return true;
    }
}

template <typename R, int NW, int RW, int CH, int BC, int W, int P>
__device__ __forceinline__ void coltri_phases(cx<R> (&a)[RW], cx<R>& xr, ColTriSmem<R, NW>& S, int N, int Nkc,
                                              int kk, int c, cx<R>* __restrict__ rf, R* __restrict__ dd,
                                              R* __restrict__ ee, cx<R>* __restrict__ tau) {
    if constexpr (P * CH < NW * RW) {
        if (CH * P - 1 > N - 2) return;  // no steps left
        if (!coltri_phase<R, NW, RW, CH, BC, W, P>(a, xr, S, N, Nkc, kk, c, rf, dd, ee, tau)) return;
        coltri_phases<R, NW, RW, CH, BC, W, P + 1>(a, xr, S, N, Nkc, kk, c, rf, dd, ee, tau);
    }
}

template <typename R, int NW, int RW, int CH, int BC, int W>
__device__ __forceinline__ void coltri_wave(ColTriSmem<R, NW>& S, const cx<double>* __restrict__ H, int N, int Nkc,
                                            int kk, int c, cx<R>* __restrict__ rf, R* __restrict__ dd,
                                            R* __restrict__ ee, cx<R>* __restrict__ tau) {
    constexpr int r0 = RW * W;
    // a[i] = A[r][c] = conj(A[c][r]) (element (c, r) at H[r N + c]: lanes read consecutive
    // entries); eight loads in flight at a time, addresses clamped into the matrix, padding
    // zeroed after
    const int cc = min(c, N - 1);
    cx<R> a[RW];
#pragma unroll
    for (int i0 = 0; i0 < RW; i0 += 8) {
        cx<double> h[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) h[i] = H[(size_t)min(r0 + i0 + i, N - 1) * N + cc];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int r = r0 + i0 + i;
            a[i0 + i] = (r < N && c < N) ? mkc<R>((R)h[i].re, -(R)h[i].im) : mkc<R>(0, 0);
        }
    }
    cx<R> xr = a[0];  // row 0 (wave 0)
    coltri_phases<R, NW, RW, CH, BC, W, 0>(a, xr, S, N, Nkc, kk, c, rf, dd, ee, tau);
    // last diagonal element (ee[N-1] = 0), by the wave holding row N - 1: after the last
    // step xr is row N - 1 (for N = 1, row 0 as loaded)
    if ((N - 1) / RW == W && c == N - 1) {
        dd[(size_t)(N - 1) * Nkc + kk] = xr.re;
        ee[(size_t)(N - 1) * Nkc + kk] = (R)0;
    }
}

// NW waves of 64 lanes per k-point, RW rows per wave (N <= NW RW <= 64), row chunks of CH,
// row vectors by DPP broadcast (BC bit 0: matrix-vector product, bit 1: update) or LDS, at
// least WPE waves per SIMD (the register budget the compiler keeps to).
template <typename R, int NW, int RW, int CH, int BC, int WPE = 2>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(WPE))) void k_tridiag_c(
    int N, int Nkc, const cx<double>* __restrict__ Hk, R* __restrict__ dd, R* __restrict__ ee,
    cx<R>* __restrict__ refl, cx<R>* __restrict__ tau) {
    static_assert(NW * RW <= 64 && RW % CH == 0 && CH % 2 == 0 && RW % 8 == 0, "tile shape");
    __shared__ ColTriSmem<R, NW> S;
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), c = tid & 63;  // w: wave-uniform
    const int kk = blockIdx.x;
    const size_t NN = (size_t)N * N;
    const cx<double>* H = Hk + (size_t)kk * NN;
    cx<R>* rf = refl + (size_t)kk * NN;
    if (w == 0) coltri_wave<R, NW, RW, CH, BC, 0>(S, H, N, Nkc, kk, c, rf, dd, ee, tau);
    if constexpr (NW > 1) {
        if (w == 1) coltri_wave<R, NW, RW, CH, BC, 1>(S, H, N, Nkc, kk, c, rf, dd, ee, tau);
    }
}

}  // namespace lqm

namespace lqm {

// ---------------------------------------------------------------------------------
// Column-per-lane tridiagonalization for the large-N path at 64 < N <= 128 (it replaces the
// zlatrd panel steps there): CG = 2 groups of 64 columns (lane l of group g holds column
// 64 g + l) times NRG = 4 groups of RW = 32 rows, one wave per (column group, row group),
// 512 threads per k-point.  In: A, N x N full Hermitian (column-major, working precision R)
// per k-point.  Out (the layout of k_trd_col): d[kk N + j] = Re A(j, j), e[kk N + j] = beta,
// tau[kk N + j], and the reflector entries v_j(r), r >= j + 2, written over A(r, j).
// Step j, three barriers:
//   A: the waves holding row j publish x_c = conj(A(j, c)) (c >= j + 2; alpha = x_{j+1}
//      apart) and their partial ||x||^2; every wave then forms the zlarfg scalars itself.
//      v_r = scale x_r for r >= j + 2 and v_{j+1} = 1 = scale (alpha - beta), so the product
//      and the update use the raw x from LDS (row j + 1 taking alpha - beta) with the scalar
//      applied once per lane.
//   C: partial products y over each wave's rows and partial s = v^H y, summed in fixed order.
//   E: w = tau y + alpha2 v by the last row group, read back for the update's rows.
// The wave holding row j + 1 picks that row out of the update (xr, the next step's x).
// ---------------------------------------------------------------------------------
template <typename R>
struct ColTriSmem2 {
    alignas(16) cx<R> sx[2][128];   // x_r for r >= j + 2 (zero elsewhere), by step parity
    alignas(16) cx<R> sy[4][128];   // partial products of the four row groups
    alignas(16) cx<R> sw[128];      // w
    cx<R> salpha[2];
    R spn[2][2];                    // partial ||x(2:)||^2 of the two column groups
    R smx[2][2];                    // partial max |x| (rescaling of tiny columns)
    R ss[8];                        // partial s of the eight waves
};

// Step j on the rows of one wave (column group cg, rows r0 .. r0 + RW - 1); LQ = its first
// live row chunk (NQ: none live), XR = row j + 1 lies in chunk LQ.
template <typename R, int RW, int CH, int LQ, bool XR>
__device__ __forceinline__ void coltri2_step(cx<R> (&a)[RW], cx<R>& xr, ColTriSmem2<R>& S, int N,
                                             int kk, int cg, int rg, int l, int j, cx<R>* __restrict__ Ak,
                                             double* __restrict__ dd, double* __restrict__ ee,
                                             cx<R>* __restrict__ tau) {
    constexpr int NQ = RW / CH;
    const int r0 = RW * rg;
    const int c = 64 * cg + l;
    const int buf = j & 1;
    const bool own = (j / RW) == rg;  // this wave holds row j
    if (own) {
        const cx<R> x = c > j ? conj(xr) : mkc<R>(0, 0);
        const R pn = wave_sum64(c > j + 1 ? abs2(x) : (R)0);
        const R pm = wave_max64(c > j ? fmax(fabs(x.re), fabs(x.im)) : (R)0);
        S.sx[buf][c] = c > j + 1 ? x : mkc<R>(0, 0);
        if (c == j + 1) S.salpha[buf] = x;
        if (l == 0) {
            S.spn[buf][cg] = pn;
            S.smx[buf][cg] = pm;
        }
        if (c == j) dd[(size_t)kk * N + j] = (double)xr.re;
    }
    __syncthreads();  // A
    // zlarfg scalars, formed by every wave from the published partials (the same bits)
    const cx<R> alpha = S.salpha[buf];
    const R xn2 = S.spn[buf][0] + S.spn[buf][1];
    cx<R> t, scale;
    R beta;
    if (!hh_scalars(alpha, xn2, t, scale, beta)) {
        // a column of tiny entries: the scalars of x / max|x|, then beta and scale back
        const R mx = fmax(S.smx[buf][0], S.smx[buf][1]);
        const R inv = (R)1 / mx;
        const cx<R> x0 = S.sx[buf][l], x1 = S.sx[buf][64 + l];
        const R xs2 = wave_sum64(abs2(inv * x0) + abs2(inv * x1));
        hh_scalars_scaled(inv * alpha, xs2, t, scale, beta);
        beta *= mx;
        scale = inv * scale;
        if (!(inv <= num<R>::huge)) {
            t = mkc<R>(0, 0);
            scale = mkc<R>(0, 0);
            beta = alpha.re;
        }
    }
    // own column's v
    cx<R> vc = mkc<R>(0, 0);
    if (c == j + 1) vc = mkc<R>(1, 0);
    else if (c > j + 1) vc = S.sx[buf][c] * scale;
    if (own && c > j + 1 && c < N) Ak[(size_t)j * N + c] = vc;
    if (own && cg == 0 && l == 0) {
        tau[(size_t)kk * N + j] = t;
        ee[(size_t)kk * N + j] = (double)beta;
    }
    // partial y_c over this wave's rows: scale sum_{r >= j + 2} conj(A(r, c)) x_r, plus
    // conj(A(j + 1, c)) from the wave holding row j + 1
    // this wave's rows of the LDS vectors from one base address (static offsets after it)
    const cx<R>* __restrict__ xb = &S.sx[buf][r0];
    const cx<R>* __restrict__ wb = &S.sw[r0];
    // v_r = scale x_r with x_{j+1} taken as alpha - beta (scale (alpha - beta) = 1), so the
    // row j + 1 term needs no separate copy of that row
    const cx<R> zab = mkc<R>(alpha.re - beta, alpha.im);
    const int jj = j + 1 - r0 - CH * LQ;  // row j + 1 in chunk LQ (XR)
    cx<R> ys = mkc<R>(0, 0), ys1 = mkc<R>(0, 0);
#pragma unroll
    for (int q = LQ; q < NQ; ++q) {
        cx<R> xq[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            xq[u] = xb[CH * q + u];
            if (XR && q == LQ) {
                xq[u].re = jj == u ? zab.re : xq[u].re;
                xq[u].im = jj == u ? zab.im : xq[u].im;
            }
        }
#pragma unroll
        for (int u = 0; u < CH; u += 2) {
            ys = cmadc(ys, a[CH * q + u], xq[u]);
            ys1 = cmadc(ys1, a[CH * q + u + 1], xq[u + 1]);
        }
        __builtin_amdgcn_sched_barrier(0);  // one chunk's operands live at a time
    }
    cx<R> y = (ys + ys1) * scale;
    if (c <= j) y = mkc<R>(0, 0);
    const R sp = wave_sum64(y.re * vc.re + y.im * vc.im);
    S.sy[rg][c] = y;
    if (l == 0) S.ss[2 * rg + cg] = sp;
    __syncthreads();  // C
    y = ((S.sy[0][c] + S.sy[1][c]) + S.sy[2][c]) + S.sy[3][c];
    if (c <= j) y = mkc<R>(0, 0);
    R s = 0;
#pragma unroll
    for (int w = 0; w < 8; ++w) s += S.ss[w];
    const R alpha2 = (R)-0.5 * (t.re * t.re + t.im * t.im) * s;
    const cx<R> wc = cmad(t * y, mkc<R>(alpha2, 0), vc);
    if (rg == 3) S.sw[c] = wc;
    __syncthreads();  // E
    // live rows: A(r, c) -= v_r conj(w_c) + w_r conj(v_c), v_r = scale x_r
    const cx<R> g = scale * conj(wc);
    const cx<R> nv = mkc<R>(-vc.re, -vc.im);
#pragma unroll
    for (int q = LQ; q < NQ; ++q) {
        cx<R> xq[CH], wq[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            xq[u] = xb[CH * q + u];
            wq[u] = wb[CH * q + u];
            if (XR && q == LQ) {
                xq[u].re = jj == u ? zab.re : xq[u].re;
                xq[u].im = jj == u ? zab.im : xq[u].im;
            }
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            a[CH * q + u] = cmad(a[CH * q + u], mkc<R>(-xq[u].re, -xq[u].im), g);
            a[CH * q + u] = cmadcb(a[CH * q + u], wq[u], nv);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (XR && LQ < NQ) {
        // row j + 1 (chunk LQ) is the next step's xr
        cx<R> x = a[CH * LQ];
#pragma unroll
        for (int u = 1; u < CH; ++u) {
            x.re = jj == u ? a[CH * LQ + u].re : x.re;
            x.im = jj == u ? a[CH * LQ + u].im : x.im;
        }
        xr = x;
    }
}

// The steps j in [jlo, jhi] with this wave's first live chunk LQ (compile time).
template <typename R, int RW, int CH, int LQ, bool XR>
__device__ __forceinline__ void coltri2_range(cx<R> (&a)[RW], cx<R>& xr, ColTriSmem2<R>& S, int N,
                                              int kk, int cg, int rg, int l, int jlo, int jhi,
                                              cx<R>* __restrict__ Ak, double* __restrict__ dd,
                                              double* __restrict__ ee, cx<R>* __restrict__ tau) {
#pragma nounroll
    for (int j = jlo; j <= jhi; ++j)
        coltri2_step<R, RW, CH, LQ, XR>(a, xr, S, N, kk, cg, rg, l, j, Ak, dd, ee, tau);
}
// Local phases of the wave's own rows: q = the chunk holding row j + 1.
template <typename R, int RW, int CH, int Q>
__device__ __forceinline__ void coltri2_own(cx<R> (&a)[RW], cx<R>& xr, ColTriSmem2<R>& S, int N,
                                            int kk, int cg, int rg, int l, cx<R>* __restrict__ Ak,
                                            double* __restrict__ dd, double* __restrict__ ee,
                                            cx<R>* __restrict__ tau) {
    if constexpr (Q < RW / CH) {
        const int r0 = RW * rg;
        const int jlo = max(r0 + CH * Q - 1, 0), jhi = min(r0 + CH * Q + CH - 2, N - 2);
        coltri2_range<R, RW, CH, Q, true>(a, xr, S, N, kk, cg, rg, l, jlo, jhi, Ak, dd, ee, tau);
        coltri2_own<R, RW, CH, Q + 1>(a, xr, S, N, kk, cg, rg, l, Ak, dd, ee, tau);
    }
}

template <typename R, int RW, int CH>
__device__ __forceinline__ void coltri2_wave(ColTriSmem2<R>& S, cx<R>* __restrict__ Ak, int N, int kk, int cg,
                                             int rg, int l, double* __restrict__ dd, double* __restrict__ ee,
                                             cx<R>* __restrict__ tau) {
    constexpr int NQ = RW / CH;
    const int r0 = RW * rg;
    const int c = 64 * cg + l;
    const int cc = min(c, N - 1);
    // a[i] = A(r, c) = conj(A(c, r)) (element (c, r) at A[r N + c]: lanes read consecutive
    // entries); eight loads in flight at a time, addresses clamped, padding zeroed.  A column j
    // is written over only by the waves holding row j, which loaded it themselves.
    cx<R> a[RW];
#pragma unroll
    for (int i0 = 0; i0 < RW; i0 += 8) {
        cx<R> h[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) h[i] = Ak[(size_t)min(r0 + i0 + i, N - 1) * N + cc];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int r = r0 + i0 + i;
            a[i0 + i] = (r < N && c < N) ? mkc<R>(h[i].re, -h[i].im) : mkc<R>(0, 0);
        }
    }
    cx<R> xr = a[0];  // row 0 (row group 0)
    // steps before this wave's rows: all rows live; then the steps whose row j + 1 is its own
    // (chunk by chunk); then the rest with no live rows (still at every barrier)
    coltri2_range<R, RW, CH, 0, false>(a, xr, S, N, kk, cg, rg, l, 0, min(r0 - 2, N - 2), Ak, dd, ee, tau);
    coltri2_own<R, RW, CH, 0>(a, xr, S, N, kk, cg, rg, l, Ak, dd, ee, tau);
    coltri2_range<R, RW, CH, NQ, false>(a, xr, S, N, kk, cg, rg, l, max(r0 + RW - 1, 0), N - 2, Ak, dd, ee, tau);
    if ((N - 1) / RW == rg && c == N - 1) {
        dd[(size_t)kk * N + N - 1] = (double)xr.re;
        ee[(size_t)kk * N + N - 1] = 0.0;
        tau[(size_t)kk * N + N - 1] = mkc<R>(0, 0);
    }
}

template <typename R>
__global__ __launch_bounds__(512) void k_tridiag_cg(int N, cx<R>* __restrict__ A, double* __restrict__ dd,
                                                    double* __restrict__ ee, cx<R>* __restrict__ tau) {
    constexpr int RW = 32, CH = 4;
    __shared__ ColTriSmem2<R> S;
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63;
    const int kk = blockIdx.x;
    coltri2_wave<R, RW, CH>(S, A + (size_t)kk * N * N, N, kk, w & 1, w >> 1, l, dd, ee, tau);
}

}  // namespace lqm
