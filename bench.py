#!/usr/bin/env python3
"""Benchmark of the k-space band path on one MI355X.

A step is one full density-matrix pass -- H(k) assembly + Hermitian eigensolve
with vectors + rho(R) accumulation (getdensitymatrix!, Operators/densitymatrix.jl)
-- over the whole k-grid of the workload, with every input already resident in
HBM.  The default workload is the dense random tight-binding model with 64
orbitals, 25 hopping keys (|R_x|, |R_y| <= 2), a 128 x 128 Gamma-centred grid,
mu = 0, kT = 0.02 and rho on all 25 keys.

Printed: one JSON line with the throughput (k-points/s), the dominant kernel's
roofline figures measured with HIP events inside the timed region, and the FP64
CPU oracle timed on this machine's host cores over a bounded sample of the same
workload (cpu_baseline).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import __graft_entry__  # noqa: E402

pkg = __graft_entry__.load_package()
from latticeqm_amd import _lib, api  # noqa: E402

METRIC = "k-points/s (H(k)+eigh+ρ) mixed-prec vs FP64 GPU at ≤1e-10 eigval err"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    # name: (N, rmax, nk_side, mu, T)
    "random_tb_N64_R25_k128x128": (64, 2, 128, 0.0, 0.02),
    "random_tb_N128_R9_k64x64": (128, 1, 64, 0.0, 0.05),
    "random_tb_N32_R25_k128x128": (32, 2, 128, 0.0, 0.02),
}
DEFAULT = "random_tb_N64_R25_k128x128"


def stage_flops(stage: str, N: int, nR: int, nRr: int, nk: int) -> float:
    """Algorithmic real flops of one launch over nk k-points (complex MAC = 8)."""
    if stage == "hk_tridiag":
        return nk * (8.0 * nR * N * N + 16.0 / 3.0 * N ** 3)
    if stage == "backtransform":
        return nk * (8.0 * N ** 3 + 8.0 * N ** 3)
    if stage == "phase_contract":
        return nk * 8.0 * nRr * N * N
    return 0.0


def host_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        n = min(n, int(omp))
    return max(1, n)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-sample", type=int, default=16384, help="k-points of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the k-grid would shard "
              "over GPUs with one rho all-reduce per pass, which this build does not run.",
              file=sys.stderr)
        return 2

    import torch

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 1
    N, rmax, side, mu, T = WORKLOADS[args.workload]
    h = pkg.models.random_tb(N, rmax=rmax, seed=2)
    ks = pkg.regulargrid(nk=side * side)
    Nk = ks.shape[1]
    R, Tst = h.pack()
    nR = nRr = R.shape[1]
    Rr = np.asfortranarray(R.astype(np.int32))

    ex = api.to_executor("gpu")
    model = api.GPUModel(h, ex)
    dev = torch.device("cuda:0")
    ks_d = torch.from_numpy(np.ascontiguousarray(ks.T)).to(dev)
    rho_d = torch.zeros((nRr, N, N, 2), dtype=torch.float64, device=dev)
    eb_d = torch.zeros(1, dtype=torch.float64, device=dev)
    bands_d = torch.zeros((Nk, N), dtype=torch.float64, device=dev)
    L = _lib.lib()
    vp = ctypes.c_void_p

    def step(bands=None):
        st = L.lqm_densitymatrix_dev(model.handle, Nk, vp(ks_d.data_ptr()), None, mu, T, nRr,
                                     Rr.ctypes.data_as(vp), vp(rho_d.data_ptr()), vp(eb_d.data_ptr()),
                                     vp(bands.data_ptr()) if bands is not None else None)
        _lib.check(st, "lqm_densitymatrix_dev")

    chosen = api.autotune(model, ks_d.data_ptr(), min(Nk, 2048))
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = Nk * args.steps / (t1 - t0)

    # per-stage device time of the same passes (HIP events on the library's stream)
    _lib.call("lqm_context_set_timing", ex.handle, 1)
    nt = max(1, min(args.steps, 5))
    for _ in range(nt):
        step(bands_d)
    ms = (ctypes.c_double * _lib.NSTAGES)()
    cnt = (ctypes.c_int64 * _lib.NSTAGES)()
    _lib.call("lqm_context_stage_times", ex.handle, ms, cnt, _lib.NSTAGES)
    _lib.call("lqm_context_set_timing", ex.handle, 0)
    stages = {}
    for i in range(_lib.NSTAGES):
        if cnt[i]:
            name = L.lqm_stage_name(i).decode()
            stages[name] = {"ms_per_step": ms[i] / nt, "launches_per_step": cnt[i] / nt,
                            "avg_launch_ms": ms[i] / cnt[i]}
    dom = max(stages, key=lambda s: stages[s]["ms_per_step"])
    launches = stages[dom]["launches_per_step"]
    fl = stage_flops(dom, N, nR, nRr, Nk) / launches
    achieved = fl / (stages[dom]["avg_launch_ms"] * 1e-3) / 1e12
    path_tflops = stage_flops("hk_tridiag", N, nR, nRr, Nk) + stage_flops("backtransform", N, nR, nRr, Nk) \
        + stage_flops("phase_contract", N, nR, nRr, Nk)
    roofline = {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                "flops_per_launch": fl, "avg_launch_ms": stages[dom]["avg_launch_ms"]}

    # correctness on a sample (GPU vs oracle on the same sub-grid)
    out = {}
    cpu = None
    if not args.no_cpu:
        from oracle import oracle

        S = min(args.cpu_sample, Nk)
        sub = np.asfortranarray(ks[:, :: max(1, Nk // S)][:, :S])
        nth = host_threads()
        oracle.densitymatrix(R, Tst, sub[:, :64], Rr, mu, T, nthreads=nth)  # warm-up
        times = []
        for _ in range(3):
            c0 = time.perf_counter()
            ref, eb_ref = oracle.densitymatrix(R, Tst, sub, Rr, mu, T, nthreads=nth)
            times.append(time.perf_counter() - c0)
        cpu_t = statistics.median(times)
        cpu = {"value": S / cpu_t, "unit": "k-points/s", "cores": nth, "kind": "port",
               "sample": f"{S} of the {Nk} k-points (strided), full pass H(k)+zheevr('V')+rho on "
                         f"{nRr} keys, OpenMP static chunks, OpenBLAS 1 thread/core, median of 3"}
        rho_g = pkg.Hops({tuple(R[:, r]): np.zeros((N, N), complex) for r in range(nRr)})
        eb_g = api.getdensitymatrix(rho_g, model, sub, mu, T=T)
        err = max(np.max(np.abs(rho_g[tuple(R[:, r])] - ref[:, :, r])) for r in range(nRr))
        bands_ref = oracle.bands(R, Tst, sub, nthreads=nth)
        bands_g, _ = api.bandmatrix(model, sub)
        out["max_abs_err_rho_vs_fp64_oracle"] = float(err)
        out["abs_err_eband_vs_fp64_oracle"] = float(abs(eb_g - eb_ref))
        out["max_eigval_err_vs_fp64_oracle"] = float(np.max(np.abs(bands_g - bands_ref)))

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "k-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random tight-binding model)",
        "config": {"workload": args.workload, "N": N, "nR": nR, "nRrho": nRr, "Nk": Nk, "mu": mu,
                   "kT": T, "precision": "fp64 plain path", "parallelism": "single GPU",
                   "kernel_variants": {"tridiag": chosen[0], "rho": chosen[1]}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages": stages,
        "path_tflops_per_s": path_tflops / (ms_per_step * 1e-3) / 1e12,
        **out,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
