"""Benchmark of the H*psi hot path and the ChebFSI + Rayleigh-Ritz sweep on one
MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload si64_60]
                    [--no-cpu] [--copy-probe] [--no-realspace | --realspace-only]
                    [--large-cell | --large-exx] [--no-batch | --batch-only]

Workload si64_60: 64-atom diamond supercell, 60^3 grid, Ecut 10 Ha,
npw 6567, 256 bands, 256 projectors; degree-10 Chebyshev filter with bounds
from 8 Lanczos steps, then Cholesky QR + Rayleigh-Ritz.  Inputs are resident
in HBM.  A step is one full sweep (2824 band-applications of H: 8 Lanczos +
10 x 256 filter + 256 Rayleigh-Ritz).

W untimed FP64 sweeps warm the block (as in an SCF, where each sweep starts
from the previous one), then K mixed sweeps (FP32 filter, FP64 Cholesky QR and
Rayleigh-Ritz) and K FP64 sweeps are timed from that same block.

value = band-applications/s of the mixed sweep.  The line also carries the
FP64 sweep time, the mixed/FP64 speedup, the Ritz-value error of the mixed
sweep against the FP64 sweep, plain H*psi throughput (FP64 and FP32), the
roofline of the dominant kernel (the xy slab FFT kernel, timed with HIP
events on its launch stream inside the timed region) and the FP64 CPU oracle
timed on this host's cores.  A "ppcg" object times one PPCGSolveReal call
(sbSize 32, 20 iterations) from the warm block.  A "realspace" object
carries the same kind
of line for the real-space dual-grid H*psi on the SiC_large deck's shape
(62^3 / 124^3 grids, 216 atoms, 442 bands).  --large-cell times the
180 x 180 x 72, 432-atom, 1024-band cell on its own, --large-exx its
screened exchange on the 96 x 96 x 40 subgrid.  A "batch" object times the
48-cell uneven workload (cubic grids 24..72, 16..128 bands) through a
degree-8 FP32 Chebyshev filter in one persistent launch against the same
filter run cell by cell.
"""
import argparse
import json
import re
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
DEGREE, LANCZOS = 10, 8


def slab_bytes(prob, nstick, nb, p):
    """Algorithmic HBM bytes of one xy-slab launch over nb bands: read the
    z-transformed sticks (nb*n3*nstick complex), write them back, read V once."""
    return nb * prob.grid[2] * nstick * 2 * p * 2 + prob.ngrid * p


def slab_flops(prob, kxmax, nb):
    """Nominal FFT flops (5 n log2 n per complex length-n transform) of one
    xy-slab launch: per z plane the y transforms of the kx <= kxmax columns
    and the x transforms of the row pairs, both ways, and the V multiply."""
    import math
    n1, n2, n3 = prob.grid
    f = lambda n: 5.0 * n * math.log2(n)
    return nb * n3 * (2 * (kxmax + 1) * f(n2) + 2 * (n2 // 2) * f(n1) + n1 * n2)


def staged_bytes(prob, p):
    """Per-band bytes of the dense-FFT staged model p*(9 npw + 6 Nh + 5 N)."""
    n1, n2, n3 = prob.grid
    Nh = (n1 // 2 + 1) * n2 * n3
    return p * (9 * prob.npw + 6 * Nh + 5 * prob.ngrid)


def pmc_traffic(path, kernel):
    """(HBM bytes per launch, full kernel name) of the kernel whose name
    starts with `kernel`, from a committed rocprofv3 --pmc summary (a line
    'traffic <kernel name> bytes_per_launch=<B>'), or (None, None)."""
    try:
        for line in open(path):
            if line.startswith("traffic " + kernel) and "bytes_per_launch=" in line:
                name = line[len("traffic "):line.index(" bytes_per_launch=")]
                return int(float(line.split("bytes_per_launch=")[1].split()[0])), name
    except OSError:
        pass
    return None, None


def realspace_bench(args):
    """Real-space dual-grid H*psi (pw_rs_mult_spinor) on the SiC_large deck's
    shape: 62^3 coarse / 124^3 fine grid, 216 atoms x 8 projectors, 442
    bands resident in HBM.  A step is one H*psi over all bands.  The roofline
    line compares the whole apply with a streaming model (every stage reads
    and writes its arrays once)."""
    import numpy as np
    import torch
    from pwdft_develop_amd import realspace as R
    prob = R.sic216_rs().build()
    nb = prob.nb
    ctx = R.RealSpaceContext(prob)
    x = torch.from_numpy(prob.psi).to("cuda:0")
    y = torch.empty_like(x)
    for _ in range(max(args.warmup, 1)):
        ctx.mult_spinor(x, y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps = max(args.steps, 1)
    for _ in range(steps):
        ctx.mult_spinor(x, y)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    Nc, Nf = prob.ntot, prob.ntot_fine
    Hc = (prob.grid[0] // 2 + 1) * prob.grid[1] * prob.grid[2]
    Hf = (prob.grid_fine[0] // 2 + 1) * prob.grid_fine[1] * prob.grid_fine[2]
    spts = int(prob.npts.sum())
    # psi in, a3 out, two coarse and two fine transforms (each one read and
    # one write), c2f / f2c, V multiply and the projector gather / update
    model = (8 * Nc * 2 + 2 * (8 * Nc + 16 * Hc) + 2 * (8 * Nf + 16 * Hf) + (16 * Hc + 16 * Hf)
             + 48 * Hc + 16 * Nf + 8 * spts + 16 * spts)
    achieved = model * nb / t / 1e9
    res = {
        "metric": "real-space dual-grid H*psi band-applications/s",
        "value": round(nb / t, 1), "unit": "band-applications/s", "ms_per_step": round(t * 1e3, 3),
        "steps": steps, "dtype": "f64",
        "config": {"workload": prob.name, "grid": list(prob.grid), "grid_fine": list(prob.grid_fine),
                   "nbands": nb, "natom": len(prob.atoms), "nproj": int(prob.nproj.sum()),
                   "projector_points": spts, "bands_per_batch": ctx.info()["chunk"]},
        "roofline": {"bound": "hbm", "kernel": "whole apply (streaming model)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "model_bytes_per_band": model},
    }
    # one PPCGSolveReal call with the deck's settings (tests/SiC_large/
    # config.yaml: PPCG_sbSize 32, Eig_MaxIter 20, Eig_Tolerance 1e-13, so
    # all 20 iterations run; 432 occupied of 442 bands), from the same block
    X = x.clone()
    ctx.ppcg(X, 432, 2, 1, 1e-13, 32)   # warm-up: one iteration
    torch.cuda.synchronize()
    X = x.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev, rv, it = ctx.ppcg(X, 432, 1, 20, 1e-13, 32)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    res["ppcg"] = {"s_per_solve": round(tp, 4), "iterations": it, "sbSize": 32, "num_eig": 432,
                   "ms_per_iteration": round(tp / max(it, 1) * 1e3, 3),
                   "lowest_eigenvalue": float(ev[0]), "res_max_occupied": float(rv[:432].max())}
    ctx.close()
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        from concurrent.futures import ThreadPoolExecutor
        cores = host_cores()
        nbs = min(nb, 2 * cores)

        def one_band(b):  # the C call releases the GIL; no state is shared between calls
            return O.rs_mult_spinor(prob.grid, prob.grid_fine, prob.volume, prob.ekin, prob.vtot, prob.npts,
                                    prob.nproj, prob.idx, prob.val, prob.wgt, prob.psi[b:b + 1])

        t0 = time.perf_counter()
        with ThreadPoolExecutor(cores) as ex:
            list(ex.map(one_band, range(nbs)))
        tc = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(nbs / tc, 3), "unit": "band-applications/s", "cores": cores,
                               "kind": "port",
                               "sample": f"{nbs} bands of {prob.name} through the C oracle, one band per call "
                                         f"on {cores} threads: {tc:.2f} s"}
    return res


def batch_bench(args):
    """Many uneven cells in one persistent launch: the 48-cell workload
    (cubic grids 24..72, 2..32 atoms, 16..128 bands, FP32) through a
    degree-8 Chebyshev filter, once as one batched launch (graph-replayed)
    and once cell by cell through each cell's own context.  A step is one
    filter over all cells (sum of nb x 8 band-applications)."""
    import numpy as np
    import torch
    from pwdft_develop_amd import workload as W
    from pwdft_develop_amd.ops import Batch
    deg = 8
    cells = W.batch_cells()
    for c in cells:
        c.build()
    B = Batch(cells)
    tuned = B.autotune(None)
    bounds = np.array([(1.0, float(c.ekin.max()) + 6.0, -2.5) for c in cells])
    X0 = B.to_device([c.psi for c in cells])
    X = torch.empty_like(X0)
    napp = deg * sum(c.nb for c in cells)
    res = {}
    for tag, serial in (("batched", False), ("one_by_one", True)):
        for _ in range(3):
            X.copy_(X0)
            B.cheb_filter(X, deg, bounds, serial=serial)
        torch.cuda.synchronize()
        steps = max(args.steps, 1)
        t = 0.0
        for _ in range(steps):
            X.copy_(X0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            B.cheb_filter(X, deg, bounds, serial=serial)
            torch.cuda.synchronize()
            t += time.perf_counter() - t0
        res[tag] = t / steps
    # per-stage time of one batched filter (device wall-clock stamps at the
    # stage barriers; the stamps are off during the timed runs above)
    B.set_timing(True)
    X.copy_(X0)
    B.cheb_filter(X, deg, bounds)
    torch.cuda.synchronize()
    stage_ms = B.timing()
    B.set_timing(False)
    info = B.info()
    B.close()
    grids = {}
    for c in cells:
        grids[c.grid[0]] = grids.get(c.grid[0], 0) + 1
    return {
        "metric": "batched FP32 Chebyshev filter band-applications/s (48 uneven cells, one persistent launch)",
        "value": round(napp / res["batched"], 1), "unit": "band-applications/s",
        "ms_per_step": round(res["batched"] * 1e3, 3), "steps": max(args.steps, 1), "dtype": "f32",
        "one_by_one_ms": round(res["one_by_one"] * 1e3, 3),
        "stage_ms": {"zinv_proj": round(stage_ms[0], 3), "slab": round(stage_ms[1], 3),
                     "zfwd": round(stage_ms[2], 3), "apply": round(stage_ms[3], 3)},
        "speedup_vs_one_by_one": round(res["one_by_one"] / res["batched"], 3),
        "config": {"workload": "batch48", "cells": len(cells), "grids": grids,
                   "bands": sum(c.nb for c in cells), "degree": deg, "band_applications": napp,
                   "jobs_per_step": info["jobs_per_step"], "workgroups": info["workgroups"],
                   "lds_bytes": info["lds_bytes"], "autotune": tuned},
    }


def large_cell_bench(args):
    """The large single cell (180 x 180 x 72 grid, 432 atoms in three
    layers, 1728 projectors, npw 76241, 1024 bands): H*psi throughput in
    both precisions, then one mixed and one FP64 ChebFSI+RR sweep timed from
    the same block (warmed by W FP64 sweeps, as in an SCF), with the
    Ritz-value difference between them.  Screened exchange on this cell is not part of
    the line."""
    import numpy as np
    import torch
    from pwdft_develop_amd import workload as W
    from pwdft_develop_amd.ops import Context

    def note(msg):
        print(f"large cell: {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    prob = W.layered432_180().build()
    nb = prob.nb
    note(f"inputs built in {time.perf_counter() - t0:.1f} s")
    ctx = Context(prob)
    info = ctx.info()
    steps = max(1, min(args.steps, 3))
    hpsi = {}
    for dt, tag in ((torch.complex128, "f64"), (torch.complex64, "f32")):
        X = ctx.to_device(prob.psi, dt)
        HX = torch.empty_like(X)
        ctx.apply_h(X, HX)
        torch.cuda.synchronize()
        ctx.set_timing(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.apply_h(X, HX)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / steps
        ms, ncall = ctx.timing()
        ctx.set_timing(False)
        hpsi[tag] = {"band_applications_per_s": round(nb / t, 1), "ms_per_apply": round(t * 1e3, 3),
                     "stages_ms": [round(x / max(ncall, 1), 3) for x in ms]}
        note(f"H*psi {tag}: {t * 1e3:.1f} ms per {nb}-band apply")
        del X, HX
    v0 = ctx.to_device(prob.lanczos_start())
    X = ctx.to_device(prob.psi)
    prev = None
    for _ in range(max(1, args.warmup)):
        _, prev, _, _ = ctx.chebfsi_sweep(X, DEGREE, LANCZOS, v0, prev)
    torch.cuda.synchronize()
    note(f"{max(1, args.warmup)} FP64 warm-up sweeps done")
    sweeps = {}
    for mixed in (True, False):
        tag = "mixed" if mixed else "fp64"
        # an untimed call of the mode first (its lazily sized scratch), then
        # the timed one, both from the warm block
        Xs = X.clone()
        ctx.chebfsi_sweep(Xs, DEGREE, LANCZOS, v0, prev, mixed=mixed)
        Xs.copy_(X)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, e, r, _ = ctx.chebfsi_sweep(Xs, DEGREE, LANCZOS, v0, prev, mixed=mixed)
        torch.cuda.synchronize()
        sweeps[tag] = (time.perf_counter() - t0, e, r)
        note(f"{tag} sweep: {sweeps[tag][0]:.3f} s")
        del Xs
    ctx.close()
    napp = LANCZOS + DEGREE * nb + nb
    nwant = 3 * nb // 4
    err = np.abs(sweeps["mixed"][1] - sweeps["fp64"][1])
    # slab stage (HIP events on its launch stream) against HBM and VALU; the
    # projector GEMM stage (beta^T X then Y += (beta w) W: 4 m nbeta nb flops)
    # against the dense MFMA peaks
    roof = {}
    for tag, p, valu_peak, mfma_peak in (("f32", 4, 78.6, 157.3), ("f64", 8, 78.6, 78.6)):
        sl_ms = hpsi[tag]["stages_ms"][1]
        nl_ms = hpsi[tag]["stages_ms"][3]
        b = slab_bytes(prob, info["nstick"], nb, p)
        fl = slab_flops(prob, info["kxmax"], nb)
        gf = 4.0 * (2 * prob.npw) * prob.nbeta * nb
        roof[tag] = {"slab": {"launch_ms": sl_ms, "algorithmic_bytes": b,
                              "GBs": round(b / (sl_ms * 1e-3) / 1e9, 1),
                              "hbm_frac": round(b / (sl_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                              "valu_tflops": round(fl / (sl_ms * 1e-3) / 1e12, 2),
                              "valu_frac": round(fl / (sl_ms * 1e-3) / 1e12 / valu_peak, 4)},
                     "projector_gemms": {"ms": nl_ms, "flops": gf,
                                         "tflops": round(gf / (nl_ms * 1e-3) / 1e12, 1),
                                         "mfma_peak_tflops": mfma_peak,
                                         "mfma_frac": round(gf / (nl_ms * 1e-3) / 1e12 / mfma_peak, 4)}}
    return {"metric": "ChebFSI-RR sweep band-applications/s (mixed precision), Ritz err vs FP64",
            "value": round(napp / sweeps["mixed"][0], 1), "unit": "band-applications/s",
            "dtype": "f32 filter / f64 Cholesky-QR + Rayleigh-Ritz",
            "config": {"workload": prob.name, "grid": list(prob.grid), "npw": prob.npw, "nbands": nb,
                       "natom": len(prob.atoms), "nbeta": prob.nbeta, "nstick": info["nstick"],
                       "kxmax": info["kxmax"], "degree": DEGREE, "lanczos": LANCZOS,
                       "band_applications_per_sweep": napp},
            "sweep_s": {k: round(v[0], 4) for k, v in sweeps.items()},
            "speedup_mixed_vs_fp64": round(sweeps["fp64"][0] / sweeps["mixed"][0], 3),
            "ritz_err_ha": {"wanted_max": float(err[:nwant].max()), "all_max": float(err.max()),
                            "nwanted": nwant},
            "res_max_wanted": {k: float(v[2][:nwant].max()) for k, v in sweeps.items()},
            "hpsi": hpsi, "roofline": roof}


def large_exx_bench(args):
    """Screened exchange of the layered 432-atom cell on its 96 x 96 x 40
    subgrid (Ecut 2.8 Ha, npw 11340): V_x applied to the 1024-band block with
    the lowest 432 bands occupied (442368 pair densities, each through one
    forward and one inverse pair transform), FP64 and FP32, then the ACE
    build from the occupied block.  The FP32 result is compared with the
    FP64 one."""
    import numpy as np
    import torch
    from pwdft_develop_amd import workload as W
    from pwdft_develop_amd.ops import Context

    def note(msg):
        print(f"exchange: {msg}", file=sys.stderr, flush=True)

    prob = W.layered432_96(nb=1024).build()
    nb, nocc = prob.nb, 432
    ctx = Context(prob)
    mu, frac = 0.106, 0.25
    occ = np.ones(nocc)
    X = ctx.to_device(prob.psi)
    ctx.set_exx(prob.ecut, mu, frac, X[:nocc].contiguous(), occ)
    out = {}
    res = {}
    for dt, tag in ((torch.complex128, "f64"), (torch.complex64, "f32")):
        Xd = ctx.to_device(prob.psi, dt)
        Y = torch.zeros_like(Xd)
        ctx.add_exx(Xd, Y)            # accumulates: Y += V_x X
        Y.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.add_exx(Xd, Y)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        res[tag] = Y.cpu().numpy()
        out[tag] = {"s_per_apply": round(t, 4), "pairs_per_s": round(nb * nocc / t, 1),
                    "band_applications_per_s": round(nb / t, 2)}
        note(f"{tag} apply of {nb} bands x {nocc} occupied: {t:.3f} s")
        del Xd, Y
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.build_ace(X[:nocc].contiguous())
    torch.cuda.synchronize()
    t_ace = time.perf_counter() - t0
    note(f"ACE build: {t_ace:.3f} s")
    ctx.close()
    d = res["f32"] - res["f64"]
    return {"workload": prob.name, "grid": list(prob.grid), "npw": prob.npw, "nbands": nb, "nocc": nocc,
            "mu": mu, "exx_fraction": frac, "pairs": nb * nocc, "apply": out, "ace_build_s": round(t_ace, 4),
            "f32_vs_f64_rel_max": float(np.abs(d).max() / np.abs(res["f64"]).max())}


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed FP64 sweeps before the timed ones (default 3; 6 with --large-cell)")
    ap.add_argument("--workload", default="si64_60")
    ap.add_argument("--hpsi-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-realspace", action="store_true",
                    help="skip the real-space dual-grid H*psi line (sic216_rs, SiC_large deck shape)")
    ap.add_argument("--realspace-only", action="store_true",
                    help="time only the real-space dual-grid H*psi and print its line")
    ap.add_argument("--large-cell", action="store_true",
                    help="time only the 180x180x72, 432-atom, 1024-band cell (H*psi and one mixed and "
                         "one FP64 sweep) and print its line")
    ap.add_argument("--large-exx", action="store_true",
                    help="time only the screened exchange of the 432-atom cell on its 96x96x40 subgrid "
                         "(1024 bands x 432 occupied) and print its line")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the large-cell and large-exchange objects of the default line")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched 48-cell filter line")
    ap.add_argument("--batch-only", action="store_true",
                    help="time only the batched 48-cell filter and print its line")
    ap.add_argument("--copy-probe", action="store_true",
                    help="end with 3 device copies of 256 MiB (the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(the band-sharded multi-GPU layout is a design, not built)"}))
        sys.exit(2)

    import numpy as np
    import torch
    from pwdft_develop_amd import workload as W
    from pwdft_develop_amd.ops import Context

    if args.warmup is None and not args.large_cell:
        args.warmup = 3
    if args.realspace_only:
        print(json.dumps({"realspace": realspace_bench(args)}))
        return
    if args.batch_only:
        print(json.dumps({"batch": batch_bench(args)}))
        return
    if args.large_cell:
        if args.warmup is None:
            args.warmup = 6
        print(json.dumps({"large_cell": large_cell_bench(args)}))
        return
    if args.large_exx:
        print(json.dumps({"large_exx": large_exx_bench(args)}))
        return

    prob = W.WORKLOADS[args.workload]().build()
    nb = prob.nb
    ctx = Context(prob)
    info = ctx.info()
    out = {}
    # plan autotuning (slab layout, projector-product tiling) for the FP32
    # filter's and the FP64 Rayleigh-Ritz's H*psi shape, before any timing
    out["autotune"] = {"f32": ctx.autotune(nb, torch.complex64, None),
                       "f64": ctx.autotune(nb, torch.complex128, None)}

    # ---- plain H*psi throughput, FP64 and FP32 ----
    hpsi, stages = {}, {}
    for dt, p, tag in ((torch.complex128, 8, "f64"), (torch.complex64, 4, "f32")):
        X = ctx.to_device(prob.psi, dt)
        HX = torch.empty_like(X)
        for _ in range(3):
            ctx.apply_h(X, HX)
        torch.cuda.synchronize()
        ctx.set_timing(True)  # HIP events around each launch, on the launch stream
        t0 = time.perf_counter()
        for _ in range(args.hpsi_steps):
            ctx.apply_h(X, HX)
        torch.cuda.synchronize()
        hpsi[tag] = nb * args.hpsi_steps / (time.perf_counter() - t0)
        ms, ncall = ctx.timing()
        ctx.set_timing(False)
        stages[tag] = [x / max(ncall, 1) for x in ms]
        del X, HX

    # ---- ChebFSI + RR sweeps ----
    v0 = ctx.to_device(prob.lanczos_start())
    X = ctx.to_device(prob.psi)
    prev = None
    for _ in range(args.warmup):
        _, prev, _, _ = ctx.chebfsi_sweep(X, DEGREE, LANCZOS, v0, prev)
    torch.cuda.synchronize()
    Xw = X.clone()
    prev_w = prev.copy() if prev is not None else None
    napp = LANCZOS + DEGREE * nb + nb

    def timed(mixed):
        Xs = [Xw.clone() for _ in range(args.steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = None
        for i in range(args.steps):
            res = ctx.chebfsi_sweep(Xs[i], DEGREE, LANCZOS, v0, prev_w, mixed=mixed)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        return (t1 - t0) / args.steps, res

    t_mixed, (_, e_mixed, r_mixed, bounds) = timed(True)
    t_f64, (_, e_f64, r_f64, _) = timed(False)
    nwant = 3 * nb // 4
    err = np.abs(e_mixed - e_f64)

    # Roofline of the dominant H*psi kernel, the FP32 xy slab kernel (10 of
    # the 11 H*psi calls of a mixed sweep): its algorithmic bytes per launch
    # (sticks read and written once, V read once) over its average launch
    # duration from HIP events recorded on its launch stream in the FP32 H*psi
    # loop above (same shape as the filter's calls).
    slab_ms = stages["f32"][1]
    slab_b = slab_bytes(prob, info["nstick"], nb, 4)
    slab_fl = slab_flops(prob, info["kxmax"], nb)
    achieved = slab_b / (slab_ms * 1e-3) / 1e9
    # FP32 square grids of the register-FFT sizes take k_slab_xy_reg (8 bands
    # per z plane), others the LDS kernel k_slab_xy
    reg = prob.grid[0] == prob.grid[1] and prob.grid[0] in (24, 30, 32, 36, 40, 48, 60, 64, 72)
    # the autotuner's FP32 choice names the variant ("slab reg <lanes per y
    # transform> <bands per plane>"): k_slab_xy_reg<float, n, n, bands, 2
    # waves/SIMD, lanes[, 512 threads for the 16-band variant]>
    # (" p": the persistent grid, k_slab_xy_reg_p)
    mly = re.search(r"slab reg (\d) (\d+)( p)?", out["autotune"]["f32"])
    reg = reg and (mly is not None or "slab lds" not in out["autotune"]["f32"])
    ly, sb = (mly.group(1), int(mly.group(2))) if mly else ("1", 0)
    # (a name prefix: the traced name carries the thread count too)
    tail = ", 512" if sb == 16 else ""
    kbase = "k_slab_xy_reg_p" if mly and mly.group(3) else "k_slab_xy_reg"
    kname = (f"{kbase}<float, {prob.grid[0]}, {prob.grid[1]}, {sb or 8}, 2, {ly}{tail}" if reg
             else f"k_slab_xy<float, {prob.grid[0]}, {prob.grid[1]}")
    traffic, full = pmc_traffic(os.path.join(ROOT, "profiles", f"pmc_traffic_{prob.name}.md"), kname)
    kname = full or kname + ">"

    def staged(tag, p):
        local_ms = sum(stages[tag][:3])
        return staged_bytes(prob, p) * nb / (local_ms * 1e-3) / 1e9, local_ms

    st32, local_ms = staged("f32", 4)
    st64, local_ms64 = staged("f64", 8)
    out.update({
        "metric": "ChebFSI-RR sweep band-applications/s (mixed precision), Ritz err vs FP64",
        "value": round(napp / t_mixed, 1),
        "unit": "band-applications/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 filter / f64 Cholesky-QR + Rayleigh-Ritz",
        "data": "synthetic (seeded Philox inputs, random-init wavefunctions)",
        "config": {"workload": prob.name, "grid": list(prob.grid), "npw": prob.npw, "nbands": nb,
                   "nbeta": prob.nbeta, "nstick": info["nstick"], "degree": DEGREE, "lanczos": LANCZOS,
                   "band_applications_per_sweep": napp},
        "sweep_s": {"mixed": round(t_mixed, 6), "fp64": round(t_f64, 6)},
        "speedup_mixed_vs_fp64": round(t_f64 / t_mixed, 3),
        "ritz_err_ha": {"wanted_max": float(err[:nwant].max()), "all_max": float(err.max()),
                        "nwanted": nwant},
        "hpsi_band_applications_per_s": {k: round(v, 1) for k, v in hpsi.items()},
        "stages_ms_per_apply": {k: [round(x, 4) for x in v] for k, v in stages.items()},
        "roofline": {"bound": "hbm", "kernel": kname,
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": slab_b, "launch_ms": round(slab_ms, 4),
                     "model": "per launch: nb x (n3 x nstick complex sticks read + written) + V read",
                     # the kernel's own bound: FP32 VALU (packed FP32 is off in this build, so
                     # 78.6 TFLOP/s is the FMA peak it can reach)
                     "valu": {"achieved_tflops": round(slab_fl / (slab_ms * 1e-3) / 1e12, 2),
                              "peak_tflops_nonpacked_fp32": 78.6,
                              "frac": round(slab_fl / (slab_ms * 1e-3) / 1e12 / 78.6, 4),
                              "flops_per_launch": slab_fl,
                              "model": "5 n log2 n per complex transform: y on kx <= kxmax columns, "
                                       "x on row pairs, both ways, plus V multiply"}},
        # what a dense (unpruned) staged FFT pipeline would move per band over
        # the three local launches' time: an equivalent rate for comparison
        # with dense-FFT codes, not a roofline (the pruned pipeline never moves
        # those bytes)
        "dense_fft_equivalent": {"kernels": f"k_sticks_inverse_z + {kname.split('<')[0]} + k_sticks_forward_z",
                                 "bytes_per_band": "p (9 npw + 6 Nh + 5 N)",
                                 "f32": {"equivalent_GBs": round(st32, 1), "launch_ms": round(local_ms, 4)},
                                 "f64": {"equivalent_GBs": round(st64, 1), "launch_ms": round(local_ms64, 4)}},
    })

    # PPCG (the reference decks' PW_Solver) on the same warm block: one
    # PPCGSolveReal call with sbSize 32 and 20 iterations (tolerance 1e-13,
    # so all run), the setting of tests/*/config.yaml
    Xp = Xw.clone()
    ctx.ppcg(Xp, num_eig=3 * nb // 4, scf_iter=2, max_iter=1, tol=1e-13, sb_size=32)   # warm-up
    Xp = Xw.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, e_pp, r_pp, it_pp = ctx.ppcg(Xp, num_eig=3 * nb // 4, scf_iter=2, max_iter=20, tol=1e-13, sb_size=32)
    torch.cuda.synchronize()
    t_pp = time.perf_counter() - t0
    out["ppcg"] = {"s_per_solve": round(t_pp, 4), "iterations": it_pp, "sbSize": 32, "num_eig": 3 * nb // 4,
                   "ms_per_iteration": round(t_pp / max(it_pp, 1) * 1e3, 3),
                   "res_max_wanted": float(r_pp[:3 * nb // 4].max())}

    if not args.no_cpu:
        # FP64 oracle sweep on the same warm block, OpenMP over bands
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        cores = host_cores()
        H = O.Hamiltonian(prob, threads=cores)
        Xh = Xw.cpu().numpy()
        v0h = prob.lanczos_start()
        t0 = time.perf_counter()
        _, e_o, _, _ = H.chebfsi_sweep(Xh, DEGREE, LANCZOS, v0h, prev_w)
        t_cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": round(napp / t_cpu, 1), "unit": "band-applications/s", "cores": cores, "kind": "port",
            "sample": f"one full FP64 ChebFSI+RR sweep of {prob.name} ({nb} bands, {napp} band-applications) "
                      f"by the C oracle, OpenMP over bands: {t_cpu:.2f} s"}
        out["ritz_err_ha"]["fp64_gpu_vs_oracle_max"] = float(np.abs(e_f64 - e_o).max())
        out["ritz_err_ha"]["mixed_vs_oracle_wanted_max"] = float(np.abs(e_mixed - e_o)[:nwant].max())
    ctx.close()
    if not args.no_realspace:
        out["realspace"] = realspace_bench(args)
    if not args.no_batch:
        out["batch"] = batch_bench(args)
    if not args.no_large:
        # the large single cell (config 4): H*psi, one mixed and one FP64
        # sweep after 4 FP64 warm-up sweeps, and the exchange apply at its
        # stated size
        la = argparse.Namespace(**vars(args))
        la.steps, la.warmup = 1, 4
        out["large_cell"] = large_cell_bench(la)
        out["large_exx"] = large_exx_bench(la)
    if args.copy_probe:
        # plain device copies of a known size: the scale for FETCH_SIZE / WRITE_SIZE
        src = torch.empty(256 << 20, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_probe_bytes"] = 256 << 20
    print(json.dumps(out))


if __name__ == "__main__":
    main()
