"""Benchmark of the H*psi hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload si64_60]
                    [--dtype f64|f32] [--no-cpu]

A step is one application of H (local + kinetic + nonlocal) to the whole band
block of the workload (si64_60: 64-atom diamond supercell, 60^3 grid,
Ecut 10 Ha, npw 6567, 256 bands, 256 projectors), inputs resident in HBM.
value = band-applications per second.  The line also carries the roofline
of the dominant kernel (the xy slab kernel, timed with HIP events on the
launch stream inside the timed region) and the FP64 CPU oracle timed on this
host's cores over a bounded sample of bands.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def slab_bytes(prob, info, nb, p):
    """Algorithmic HBM bytes of one xy-slab launch over nb bands: read the
    z-transformed sticks (nb*n3*nstick complex), write them back, read V once."""
    n3 = prob.grid[2]
    return nb * n3 * info["nstick"] * 2 * p * 2 + prob.ngrid * p


def staged_bytes(prob, p):
    """Staged per-band byte model p*(9 npw + 6 Nh + 5 N) of the dense-FFT
    reference pipeline (gather, C2R, V, R2C, scatter)."""
    n1, n2, n3 = prob.grid
    Nh = (n1 // 2 + 1) * n2 * n3
    return p * (9 * prob.npw + 6 * Nh + 5 * prob.ngrid)


def cpu_baseline(prob, nsample):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    cores = min(cores, 16)
    H = O.Hamiltonian(prob, threads=cores)
    psi = prob.psi[:nsample]
    H.apply_h(psi[: min(cores, nsample)])  # warm
    t0 = time.perf_counter()
    H.apply_h(psi)
    dt = time.perf_counter() - t0
    return {"value": nsample / dt, "unit": "band-applications/s", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle apply_h on {nsample} of {prob.nb} bands of {prob.name}, "
                      f"OpenMP over bands ({dt:.2f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="si64_60")
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32"])
    ap.add_argument("--cpu-sample", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is described in DESIGN.md, not run)"}))
        sys.exit(2)

    import numpy as np
    import torch
    from pwdft_develop_amd import workload as W
    from pwdft_develop_amd.ops import Context

    prob = W.WORKLOADS[args.workload]().build()
    ctx = Context(prob)
    info = ctx.info()
    dt = torch.complex128 if args.dtype == "f64" else torch.complex64
    p = 8 if args.dtype == "f64" else 4
    X = ctx.to_device(prob.psi, dt)
    HX = torch.empty_like(X)
    for _ in range(args.warmup):
        ctx.apply_h(X, HX)
    torch.cuda.synchronize()
    ctx.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.apply_h(X, HX)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stage_ms, ncalls = ctx.timing()
    ctx.set_timing(False)
    ms_step = (t1 - t0) * 1e3 / args.steps
    nb = prob.nb
    value = nb * args.steps / (t1 - t0)
    slab_ms = stage_ms[1] / max(ncalls, 1)
    traffic_alg = slab_bytes(prob, info, nb, p)
    achieved = traffic_alg / (slab_ms * 1e-3) / 1e9
    out = {
        "metric": "Hpsi band-applications/s (ChebFSI-RR sweep: see DESIGN.md)",
        "value": round(value, 1),
        "unit": "band-applications/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded Philox inputs, random-init wavefunctions)",
        "config": {"workload": prob.name, "grid": list(prob.grid), "npw": prob.npw, "nbands": nb,
                   "nbeta": prob.nbeta, "nstick": info["nstick"], "ecut_ha": prob.ecut},
        "stages_ms_per_step": [round(x / max(ncalls, 1), 4) for x in stage_ms],
        "roofline": {"bound": "hbm", "kernel": "k_slab_xy", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "staged_model_equiv_GBs": round(staged_bytes(prob, p) * value / 1e9, 1),
    }
    del X, HX
    ctx.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(prob, args.cpu_sample)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
