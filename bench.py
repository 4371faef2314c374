#!/usr/bin/env python3
"""Benchmark of the GPU solve phase on one MI355X.

A step is one complete solve, from x0 = 0 to FP64 true relative residual
||b - A x|| / ||b|| <= 1e-10, of the plain FP64 GPU drop-in workload: the 3D
nonsymmetric block system on a 32^3 cell grid, block size 24, 7-point block
stencil (N = 786,432), RCM ordering, ILU2 (tau1 = 1e-2, tau2 = 1e-4) and
right-preconditioned BiCGStab. The matrix, its factors and the right-hand
side are resident in HBM before the timed region; host factorization time is
reported separately.

Prints one JSON line (rank 0) with the solves/sec, the HBM roofline of the
dominant kernel (HIP-event times of every SpMV / triangular-sweep launch in
the timed region against their algorithmic bytes) and a bounded CPU baseline
(the oracle's BiCGStab on the box's host, one core, scaled by the iteration
count of the GPU solve).
"""

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bcm-problems_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
EPS = 1e-10


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def build_system(B, nx, bla, seed):
    t0 = time.perf_counter()
    m, rhs = B.gen_seepage3d(nx, nx, nx, bla, 0.3, seed)
    order = m.rcm_order()
    mo = m.reorder(order)
    b = mo.reorder_vector(order, rhs)
    t1 = time.perf_counter()
    params = B.default_params()
    f = B.Factor.ilu2(mo, params)
    t2 = time.perf_counter()
    log(f"system n={mo.n} nzja={mo.info()['nzja']} generated+ordered {t1 - t0:.1f}s, ILU2 {t2 - t1:.1f}s "
        f"(fill {params.density:.3f}, second order {params.density2:.3f})")
    return m, mo, f, b, t2 - t1


def cpu_baseline(B, O, mo, f, b, iters_gpu, sample_iters):
    """Times `sample_iters` BiCGStab iterations of the oracle on one host core."""
    A = O.Block(*mo.arrays())
    nsup = A.nsup
    L = O.Block(A.sprnd, *f.arrays(0, nsup))
    U = O.Block(A.sprnd, *f.arrays(1, nsup))
    t0 = time.perf_counter()
    _, r = O.bicgstab(A, L, U, b, niter=sample_iters, eps=1e-300)
    dt = time.perf_counter() - t0
    per_iter = dt / max(r.nitperf, 1)
    solve_s = per_iter * iters_gpu
    return dict(value=1.0 / solve_s, unit="solves/s", cores=1, kind="port",
                sample=f"{r.nitperf} BiCGStab iterations of the same system by the oracle (C restatement, "
                       f"one core of {cpu_model()}) in {dt:.1f}s, scaled to the {iters_gpu} iterations of one "
                       f"GPU solve")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=32, help="cells per edge (32 is the benchmarked workload)")
    ap.add_argument("--cpu-sample-iters", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import bcmhip as B

    if B.device_count() < 1:
        raise SystemExit("no GPU visible: the solve phase runs on the GPU only")
    seed = 20261202 + 1
    m, mo, f, b, fct_s = build_system(B, args.grid, 24, seed)
    n = mo.n
    ctx = B.Context()
    db = B.DeviceArray.from_host(b)
    dx = B.DeviceArray(n)
    # check_every=1: the host tests convergence after every test kernel, so no launch is
    # queued past convergence and the per-kernel HIP-event times cover real work only
    opts = B.default_opts(kind=B.BICGSTAB, eps=EPS, ittype=3, check_every=1, niter=2000)
    log("warmup")
    for _ in range(args.warmup):
        dx.zero()
        st = ctx.solve(mo, f, db, dx, opts)
        log(f"warmup solve: {st.iterations} iterations, true relres {st.relres_true:.3e}, {st.seconds * 1e3:.1f} ms")
    ctx.sync()
    ctx.profile(True)
    ctx.profile_read(reset=True)
    iters, relres, ok = [], [], True
    t0 = time.perf_counter()
    for k in range(args.steps):
        dx.zero()
        st = ctx.solve(mo, f, db, dx, opts)
        iters.append(st.iterations)
        relres.append(st.relres_true)
        ok = ok and st.converged == 1 and st.relres_true <= EPS
    ctx.sync()
    t1 = time.perf_counter()
    prof = ctx.profile_read()
    ctx.profile(False)
    dt = t1 - t0
    log(f"timed {args.steps} solves in {dt:.3f}s: iterations {iters}, max relres {max(relres):.3e}")

    def kernel(ms, launches, nbytes):
        avg_ms = ms / max(launches, 1)
        per_launch = nbytes / max(launches, 1)
        gbs = per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
        return dict(total_ms=ms, launches=int(launches), avg_ms=avg_ms, bytes_per_launch=per_launch, gbs=gbs)

    spmv = kernel(prof.spmv_ms, prof.spmv_launches, prof.spmv_bytes)
    sweep = kernel(prof.sweep_ms, prof.sweep_launches, prof.sweep_bytes)
    dom_name, dom = ("sweep", sweep) if sweep["total_ms"] >= spmv["total_ms"] else ("spmv", spmv)
    roofline = dict(bound="hbm", achieved=dom["gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=dom["gbs"] / HBM_PEAK_GBS, traffic=None, kernel=dom_name,
                    bytes_per_launch=dom["bytes_per_launch"], avg_launch_ms=dom["avg_ms"])
    cpu = None
    if not args.no_cpu:
        import oracle as O

        log("cpu baseline (oracle, one core)")
        cpu = cpu_baseline(B, O, mo, f, b, int(np.median(iters)), args.cpu_sample_iters)
    value = args.steps / dt
    line = {
        "metric": "solves/sec at rel-res 1e-10: mixed+batched+learned vs plain FP64 GPU; HBM GB/s",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, config-1 shape)",
        "config": {
            "workload": f"plain FP64 GPU drop-in: 3D nonsymmetric {args.grid}^3 cells, block 24, 7-point, "
                        f"N={n}, RCM + ILU2(tau1=1e-2, tau2=1e-4) + BiCGStab to true rel-res 1e-10",
            "seed": seed,
        },
        "valid": bool(ok),
        "iterations": iters,
        "relres_true_max": max(relres),
        "factor_seconds_host": fct_s,
        "kernels": {"spmv": spmv, "sweep": sweep},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
