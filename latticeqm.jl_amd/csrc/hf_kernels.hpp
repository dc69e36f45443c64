// Hartree-Fock mean-field kernels of the self-consistent loop on gfx950: the
// mean-field operator H_MF[rho], its scalar energies, linear mixing with the residual
// norms, and the T > 0 chemical-potential bisection.  All blocks are N x N complex
// column-major, stacked key after key.  Every reduction runs in a fixed order, so an
// SCF trajectory is bitwise reproducible run to run.
#pragma once
#include "lqm_common.hpp"

namespace lqm {

// V0 = sum_L v_L  (Meanfield/hartreefock.jl:159), keys summed in stack order.
__global__ __launch_bounds__(256) void k_blocksum(int NN, int nb, const cx<double>* __restrict__ v,
                                                  cx<double>* __restrict__ out) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= NN) return;
    cx<double> a{0.0, 0.0};
    for (int b = 0; b < nb; ++b) a = a + v[(size_t)b * NN + e];
    out[e] = a;
}

// H_MF[q] for every key q of the mean-field key set (Meanfield/hartreefock.jl:184-289):
//   start from 0, add h (addhops!), Fock  H_MF[L] -= v_L .* rho_L   (no conjugate),
//   Hartree H_MF[0][i,i] += (V0 * diag(rho[0]))[i].
// map[3q..3q+2] = index of key q in the h, v and rho stacks (-1 when absent); qz / rz are
// the zero key's index in the mean-field and rho key sets.  flags bit0 Hartree, bit1 Fock.
__global__ __launch_bounds__(256) void k_hf_build(int N, int nK, const int* __restrict__ map, int qz,
                                                  int rz, int flags, const cx<double>* __restrict__ h,
                                                  const cx<double>* __restrict__ v,
                                                  const cx<double>* __restrict__ V0,
                                                  const cx<double>* __restrict__ rho,
                                                  cx<double>* __restrict__ hmf) {
    const size_t NN = (size_t)N * N;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= NN * nK) return;
    const int q = (int)(idx / NN);
    const size_t e = idx - (size_t)q * NN;
    const int hi = map[3 * q], vi = map[3 * q + 1], ri = map[3 * q + 2];
    cx<double> val{0.0, 0.0};
    if (hi >= 0) val = val + h[(size_t)hi * NN + e];
    if ((flags & 2) && vi >= 0 && ri >= 0) val = val - v[(size_t)vi * NN + e] * rho[(size_t)ri * NN + e];
    if ((flags & 1) && q == qz) {
        const int i = (int)(e % N), j = (int)(e / N);
        if (i == j) {
            const cx<double>* r0 = rho + (size_t)rz * NN;
            cx<double> dsum{0.0, 0.0};
            for (int m = 0; m < N; ++m) dsum = dsum + V0[i + (size_t)m * N] * r0[m + (size_t)m * N];
            val = val + dsum;
        }
    }
    hmf[idx] = val;
}

// Pullback of the mean field onto its parameters at fixed rho.  G[q] = dOmega/dH_MF[q] for
// every mean-field key q (element-wise derivative, [G_q]_ij = rho_MF(-q)_ji); since
//   H_MF[q] = h[q] - v[q] .* rho[q] (Fock) + delta_{q0} diag(V0 n) (Hartree), V0 = sum_L v_L,
// the derivatives are gh[r] = G[hq[r]] for the keys of h and
//   gv[l]_ij = -G[vq[l]]_ij rho[vr[l]]_ij  (Fock, keys of v that rho has)
//            + G[qz]_ii n_j                (Hartree, every key of v; n = diag(rho[rz])).
// gh and gv may be null.  One thread per element of the nRh + nRv output blocks.
__global__ __launch_bounds__(256) void k_hf_adjoint(int N, int nRh, int nRv, const int* __restrict__ hq,
                                                    const int* __restrict__ vq, const int* __restrict__ vr,
                                                    int qz, int rz, int flags, const cx<double>* __restrict__ G,
                                                    const cx<double>* __restrict__ rho,
                                                    cx<double>* __restrict__ gh, cx<double>* __restrict__ gv) {
    const size_t NN = (size_t)N * N;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= NN * (size_t)(nRh + nRv)) return;
    const int b = (int)(idx / NN);
    const size_t e = idx - (size_t)b * NN;
    if (b < nRh) {
        if (gh) gh[idx] = G[(size_t)hq[b] * NN + e];
        return;
    }
    if (!gv) return;
    const int l = b - nRh;
    cx<double> g{0.0, 0.0};
    if ((flags & 2) && vr[l] >= 0) g = g - G[(size_t)vq[l] * NN + e] * rho[(size_t)vr[l] * NN + e];
    if ((flags & 1) && rz >= 0 && qz >= 0) {
        const size_t i = e % N, j = e / N;
        g = g + G[(size_t)qz * NN + i * (N + 1)] * rho[(size_t)rz * NN + j * (N + 1)];
    }
    gv[(size_t)l * NN + e] = g;
}

// Block-wide sum (blockDim = 256 * W), result returned to every thread.
template <int W>
__device__ __forceinline__ double block_sum(double a, double* sred) {
    a = wave_sum(a);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x / 64] = a;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < 4 * W; ++w) t += sred[w];
    return t;
}

// Energies of the HF functional at rho (Meanfield/hartreefock.jl:392-410), one block:
//   out[0] = eps_H = 1/2 n^T V0 n, n = diag(rho[0])        (real part)
//   out[1] = eps_F = -1/2 sum_{L in keys(v) and keys(rho)} sum_ij rho_L conj(rho_L) v_L
//   out[2], out[3] = their imaginary parts (zero for a Hermitian rho).
// vr[l] = index in the rho stack of v's key l (-1 when rho lacks it).
__global__ __launch_bounds__(256) void k_hf_energy(int N, int nRv, const int* __restrict__ vr, int rz,
                                                   int flags, const cx<double>* __restrict__ v,
                                                   const cx<double>* __restrict__ V0,
                                                   const cx<double>* __restrict__ rho,
                                                   double* __restrict__ out) {
    __shared__ double sred[4];
    const size_t NN = (size_t)N * N;
    double hre = 0.0, him = 0.0, fre = 0.0, fim = 0.0;
    if ((flags & 1) && rz >= 0) {
        const cx<double>* r0 = rho + (size_t)rz * NN;
        for (int i = threadIdx.x; i < N; i += 256) {
            cx<double> s{0.0, 0.0};
            for (int j = 0; j < N; ++j) s = s + V0[i + (size_t)j * N] * r0[j + (size_t)j * N];
            const cx<double> t = r0[i + (size_t)i * N] * s;
            hre += t.re;
            him += t.im;
        }
    }
    if (flags & 2) {
        for (int l = 0; l < nRv; ++l) {
            const int ri = vr[l];
            if (ri < 0) continue;
            const cx<double>* rl = rho + (size_t)ri * NN;
            const cx<double>* vl = v + (size_t)l * NN;
            for (size_t e = threadIdx.x; e < NN; e += 256) {
                const double a2 = abs2(rl[e]);
                fre += a2 * vl[e].re;
                fim += a2 * vl[e].im;
            }
        }
    }
    hre = block_sum<1>(hre, sred);
    him = block_sum<1>(him, sred);
    fre = block_sum<1>(fre, sred);
    fim = block_sum<1>(fim, sred);
    if (threadIdx.x == 0) {
        out[0] = 0.5 * hre;
        out[1] = -0.5 * fre;
        out[2] = 0.5 * him;
        out[3] = -0.5 * fim;
    }
}

// Linear mixing (Meanfield/fixedpoint.jl:146-161): x1 <- beta x1 + (1 - beta) x0,
// then x0 <- x1.  part[2b], part[2b+1] = block b's share of ||x1 - x0||^2 and ||x1||^2
// (x0 before the copy).
// ctl (may be null): the SCF control block of a device-driven loop; once ctl[0] (done) is set,
// the mixing leaves x0 and the partials untouched, so later iterations recompute the same state.
__global__ __launch_bounds__(256) void k_mix(size_t n, double beta, cx<double>* __restrict__ x1,
                                             cx<double>* __restrict__ x0, double* __restrict__ part,
                                             const double* __restrict__ ctl = nullptr) {
    __shared__ double sred[4];
    if (ctl && ctl[0] != 0.0) return;
    double dd = 0.0, nn = 0.0;
    const double omb = 1.0 - beta;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const cx<double> a = x1[i], b = x0[i];
        const cx<double> m = beta * a + omb * b;
        dd += abs2(m - b);
        nn += abs2(m);
        x1[i] = m;
        x0[i] = m;
    }
    dd = block_sum<1>(dd, sred);
    nn = block_sum<1>(nn, sred);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = dd;
        part[2 * blockIdx.x + 1] = nn;
    }
}

// out[0] += sum_b part[2b], out[1] += sum_b part[2b+1], folded in block order.
__global__ __launch_bounds__(256) void k_fold2(int G, const double* __restrict__ part,
                                               double* __restrict__ out) {
    __shared__ double sred[4];
    double a = 0.0, b = 0.0;
    for (int g = threadIdx.x; g < G; g += 256) {
        a += part[2 * g];
        b += part[2 * g + 1];
    }
    a = block_sum<1>(a, sred);
    b = block_sum<1>(b, sred);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
    }
}

// One step of the device-driven SCF loop (one thread), after the iteration's mean field, pass and
// mixing: scal[0..3] eps_H, eps_F and their imaginary parts, scal[4] eps_band, scal[5..6] the
// squared norms |x1 - x0|^2, |x1|^2, scal[8..9] mu and its status.  ctl: [0] done, [1] iterations,
// [2] residual, [3] error (0 ok, 1 non-finite bands, 2 filling not bracketed, 3 non-finite
// energy, 4 complex energy, 5 non-finite residual), [4] converged, [5] eps_band, [6] mu.  hist[3 (it - 1) ..] = (E,
// residual, mu), the rows the host loop would write.  Nothing changes once done is set.
__global__ void k_scf_step(const double* __restrict__ scal, double* __restrict__ hist, double* __restrict__ ctl,
                           double tol, int max_iter, double imag_tol, int mu_on_device) {
    if (threadIdx.x != 0 || ctl[0] != 0.0) return;
    const int it = (int)ctl[1] + 1;
    ctl[1] = it;
    int err = 0;
    if (mu_on_device && scal[9] == 2.0) err = 1;
    else if (mu_on_device && scal[9] == 1.0) err = 2;
    else if (!isfinite(scal[0]) || !isfinite(scal[1])) err = 3;
    else if (fabs(scal[2]) > imag_tol || fabs(scal[3]) > imag_tol) err = 4;
    const double E = scal[4] - (scal[0] + scal[1]);
    const double dn = sqrt(scal[5]), xn = sqrt(scal[6]);
    const double res = xn > 2.220446049250313e-16 ? dn / xn : dn;
    if (!err && hist) {
        hist[3 * (it - 1)] = E;
        hist[3 * (it - 1) + 1] = res;
        hist[3 * (it - 1) + 2] = scal[8];
    }
    if (!err && !isfinite(res)) err = 5;
    ctl[2] = res;
    ctl[3] = err;
    ctl[4] = (!err && res < tol) ? 1.0 : 0.0;
    ctl[5] = scal[4];  // eps_band and mu of this iteration (later frozen ones recompute others)
    ctl[6] = scal[8];
    if (err || res < tol || it >= max_iter) ctl[0] = 1.0;
}

// occ[i] = fermidirac(e_i - mu; T) with mu read from the device (the chemical potential the
// bisection below has just written), so no host round trip sits between mu and rho.
__global__ __launch_bounds__(256) void k_fermi_occ(int L, const double* __restrict__ ev,
                                                   const double* __restrict__ mu, double T,
                                                   double* __restrict__ occ) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < L) occ[i] = fermidirac(ev[i], mu[0], T);
}

// 1/d for d >= 1 (a Fermi denominator): hardware reciprocal with one Newton step, 0 beyond
// 1e300 (where 1/d is below 1e-300, and the Newton step would meet inf * 0).
__device__ __forceinline__ double recip_ge1(double d) { return d < 1e300 ? rcp_nr(d) : 0.0; }

// Roots.jl-style bracketing midpoint: 0 across a sign change, else the midpoint of the
// bit patterns of |a| and |b|, so a bisection reaches adjacent doubles in <= 64 steps.
__device__ __forceinline__ double bit_middle_dev(double a, double b) {
    if ((a < 0) != (b < 0) && a != 0 && b != 0) return 0.0;
    const unsigned long long ia = (unsigned long long)__double_as_longlong(fabs(a));
    const unsigned long long ib = (unsigned long long)__double_as_longlong(fabs(b));
    const unsigned long long im = (ia >> 1) + (ib >> 1) + (ia & ib & 1ull);
    const double mid = __longlong_as_double((long long)im);
    return (a + b) < 0 ? -mid : mid;
}

// T > 0 chemical potential (Spectrum/chemicalpotential.jl:26-40, :74-78): solve
// getfilling(mu) = filling on [min(bands), max(bands)] by bisection, entirely on the
// device in one block of 1024 threads (no host round trip per step).
// getfilling = mean of fermidirac over all L = N*Nk values, or sum_k w_k sum_n f / N.
// out[0] = mu, out[1] = 0 ok / 1 filling not bracketed / 2 non-finite bands.
// ework (L doubles, or null): scratch for exp(e_i / T - c) when the values do not fit in
// registers.
__global__ __launch_bounds__(1024) void k_mu_bisect(int N, int L, const double* __restrict__ bands,
                                                    const double* __restrict__ kw, double filling,
                                                    double T, double* __restrict__ out,
                                                    double* __restrict__ ework) {
    __shared__ double sred[16];
    __shared__ double smin[16], smax[16];
    double lo = INFINITY, hi = -INFINITY;
    for (int i = threadIdx.x; i < L; i += 1024) {
        lo = fmin(lo, bands[i]);
        hi = fmax(hi, bands[i]);
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
    lo = smin[0];
    hi = smax[0];
    for (int w = 1; w < 16; ++w) {
        lo = fmin(lo, smin[w]);
        hi = fmax(hi, smax[w]);
    }
    if (!isfinite(lo) || !isfinite(hi)) {
        if (threadIdx.x == 0) {
            out[0] = NAN;
            out[1] = 2.0;
        }
        return;
    }
    const double norm = kw ? (double)N : (double)L;
    const double invT = 1.0 / T;
    // the k-point of value i = threadIdx.x + 1024 q advances by 1024 / N per step: tracked
    // incrementally instead of an integer division per value
    const int kq = 1024 / N, rq = 1024 - kq * N;
    // Up to kMuReg values per thread stay in registers across the bisection (with their
    // weights), else they are read from memory.  Once the bracket [a, b] is within 1000 T,
    // the values are turned into e_i = exp(e_i / T - c), c = the bracket's centre / T
    // (exponents clamped to +-700; for mu in the bracket exp(c - mu / T) stays within
    // e^{+-500}), in the registers or in ework, and a step costs one exp in all:
    // f_i(mu) = 1 / (e_i exp(c - mu / T) + 1).  A clamped e_i gives f below e^-200 or above
    // 1 - e^-200, as the direct form does to double precision.
    constexpr int kMuReg = 16;
    const bool inreg = L <= kMuReg * 1024;
    double cexp = 0.0;
    bool factored = false;
    double xv[kMuReg], wv[kMuReg];
    if (inreg) {
        int k = threadIdx.x / N, n = threadIdx.x - k * N;
#pragma unroll
        for (int q = 0; q < kMuReg; ++q) {
            const int i = threadIdx.x + 1024 * q;
            xv[q] = i < L ? bands[i] : 0.0;
            wv[q] = i < L && kw ? kw[k] : 0.0;
            k += kq;
            n += rq;
            if (n >= N) {
                n -= N;
                ++k;
            }
        }
    }
    auto try_factor = [&](double ba, double bb) {
        if (factored || (bb - ba) * invT > 1000.0 || (!inreg && !ework)) return;
        cexp = 0.5 * (ba + bb) * invT;
        if (inreg) {
#pragma unroll
            for (int q = 0; q < kMuReg; ++q) xv[q] = exp(fmin(fmax(xv[q] * invT - cexp, -700.0), 700.0));
        } else {
            for (int i = threadIdx.x; i < L; i += 1024)
                ework[i] = exp(fmin(fmax(bands[i] * invT - cexp, -700.0), 700.0));
        }
        factored = true;  // each thread reads back only the values it wrote: no barrier
    };
    auto F = [&](double mu) -> double {
        double a = 0.0;
        if (factored && inreg) {
            const double em = exp(cexp - mu * invT);
#pragma unroll
            for (int q = 0; q < kMuReg; ++q)
                if (threadIdx.x + 1024 * q < L) {
                    const double f = recip_ge1(xv[q] * em + 1.0);
                    a += kw ? wv[q] * f : f;
                }
        } else if (inreg) {
#pragma unroll
            for (int q = 0; q < kMuReg; ++q)
                if (threadIdx.x + 1024 * q < L) {
                    const double f = recip_ge1(exp((xv[q] - mu) * invT) + 1.0);
                    a += kw ? wv[q] * f : f;
                }
        } else if (factored) {
            const double em = exp(cexp - mu * invT);
            int k = threadIdx.x / N, n = threadIdx.x - k * N;
            for (int i = threadIdx.x; i < L; i += 1024) {
                const double f = recip_ge1(ework[i] * em + 1.0);
                a += kw ? kw[k] * f : f;
                k += kq;
                n += rq;
                if (n >= N) {
                    n -= N;
                    ++k;
                }
            }
        } else {
            int k = threadIdx.x / N, n = threadIdx.x - k * N;
            for (int i = threadIdx.x; i < L; i += 1024) {
                const double x = bands[i] - mu;
                const double f = recip_ge1(exp(x * invT) + 1.0);
                a += kw ? kw[k] * f : f;
                k += kq;
                n += rq;
                if (n >= N) {
                    n -= N;
                    ++k;
                }
            }
        }
        return block_sum<4>(a, sred) / norm - filling;
    };
    double a = lo, b = hi;
    try_factor(a, b);
    double fa = F(a), fb = F(b);
    double mu = 0.0;
    int status = 0;
    if (fa == 0.0) {
        mu = a;
    } else if (fb == 0.0) {
        mu = b;
    } else if ((fa < 0) == (fb < 0)) {
        mu = NAN;
        status = 1;
    } else {
        bool done = false;
        // Illinois (modified regula falsi) shrinks the bracket from both sides until it spans a
        // few dozen doubles (at most 40 steps); the bit-midpoint bisection below then ends it
        // at adjacent doubles as Roots.jl's Bisection does.  Inside a gap, where the filling is
        // flat to rounding, the root it settles on may differ from a pure bisection's by up to
        // the gap, with occupations equal to exp(-gap / 2T).
        {
            double wa = fa, wb = fb;
            int side = 0;
            for (int it = 0; it < 40; ++it) {
                if (b - a <= 32.0 * 2.220446049250313e-16 * fmax(fmax(fabs(a), fabs(b)), T)) break;
                double c = (a * wb - b * wa) / (wb - wa);
                if (!(c > a && c < b)) c = 0.5 * (a + b);
                if (!(c > a && c < b)) break;
                try_factor(a, b);
                const double fc = F(c);
                if (fc == 0.0) {
                    mu = c;
                    done = true;
                    break;
                }
                if ((fc < 0) == (fa < 0)) {
                    a = c;
                    fa = wa = fc;
                    if (side == -1) wb *= 0.5;
                    side = -1;
                } else {
                    b = c;
                    fb = wb = fc;
                    if (side == 1) wa *= 0.5;
                    side = 1;
                }
            }
        }
        for (int it = 0; it < 200 && !done; ++it) {
            const double mid = bit_middle_dev(a, b);
            if (!(mid > a && mid < b)) break;
            try_factor(a, b);
            const double fm = F(mid);
            if (fm == 0.0) {
                mu = mid;
                done = true;
                break;
            }
            if ((fm < 0) == (fa < 0)) {
                a = mid;
                fa = fm;
            } else {
                b = mid;
                fb = fm;
            }
        }
        if (!done) mu = fabs(fa) <= fabs(fb) ? a : b;
    }
    if (threadIdx.x == 0) {
        out[0] = mu;
        out[1] = (double)status;
    }
}

}  // namespace lqm
