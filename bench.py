#!/usr/bin/env python3
"""Benchmark of the GPU solve phase on one MI355X.

Default workload -- learned initial guess over a stream of homogenization
cells: periodic 16^3 elasticity cells (block 24, 7-point, N = 98,304, one
spherical inclusion with radius U[0.1, 0.45], stiffness contrast
10^U[0, 3], Poisson ratio U[0.2, 0.45], centre shift U[-0.15, 0.15]), 6
macroscopic-strain right-hand sides each, IC2 factors (tau1 = 0.1,
tau2 = 0.01) computed on the host. The first 1024 cells are solved from
x0 = 0 (batched block PCG, FP32 factors, FP64 outer recurrence) and train
the initial-guess model (64 POD modes per right-hand side, MLP 4 -> 256 ->
256 -> 384 trained on the GPU with BF16 MFMA kernels).

Why IC2(0.1, 0.01) and not the reference's defaults (1e-2, 1e-4,
slvparam.h:146-147): on a 16^3 cell the defaults cost 6-9 s of host
factorization per cell instead of about 1 s (2.5x the fill, 800-870
dependency levels per sweep instead of 25), so the 1536 cells of the default
run would take about 15 minutes of host time. The run therefore also solves a sample of
cells at the reference's thresholds (`--ref-tau-cells`, batched from x0 = 0
against the plain FP64 path on the same cells) and reports it beside the
headline. Evaluation cells
are then solved in groups of `--group` cells (256 by default) x 6
right-hand sides: MLP inference,
POD reconstruction and one batched block-PCG launch sequence per group
(fallback to x0 = 0 for any pair whose ||r0|| >= ||b||), every pair driven
to FP64 true relative residual <= 1e-10.

A step is one evaluation group; the timed region covers the guess and the
solve of `--steps` groups with their matrices and factors already resident
in HBM (uploads happen between steps, outside the timed region). `value`
is (cell, right-hand side) solves per second. The plain FP64 GPU drop-in
(one right-hand side at a time, FP64 factors, x0 = 0) is timed on a sample
of the same cells for the comparison, and the reference's own serial Cg
(compiled from its sources into oracle/_ref), one independent solve per
host core on every core, gives the CPU baseline.

`--workload drop-in` runs the plain FP64 nonsymmetric drop-in instead (32^3
cells, block 24, ILU2 + BiCGStab); `--workload precision` the FP64
FGMRES(30) around FP64 / FP32 / BF16 factor sweeps on the same system;
`--workload batch` one batched solve of `--group` cells of uneven sizes
8^3..24^3; `--workload heat` the CPU-runnable 2D heat system (64 x 64
cells, block 9, IC2 + PCG) against the reference's own Cg; `--workload
complex` the complex solve phase (32^3 cells, complex blocks of 12, complex
ILU2 + block GMRES) against the reference's own complex BlockGMRES.
"""

import argparse
import json
import os
import platform
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bcm-problems_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
EPS = 1e-10
METRIC = "solves/sec at rel-res 1e-10: mixed+batched+learned vs plain FP64 GPU; HBM GB/s"
CELL_M = 16
GROUP = 256  # default cells per batched solve
TEMPLATE_SEED = 20261206
PARAM_SEED = 20261206


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


def host_threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def kernel_stats(ms, launches, nbytes):
    avg_ms = ms / max(launches, 1)
    per_launch = nbytes / max(launches, 1)
    gbs = per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    return dict(total_ms=ms, launches=int(launches), avg_ms=avg_ms, bytes_per_launch=per_launch, gbs=gbs)


def roofline_from(prof, sweep_single_launch=True):
    """Roofline of the dominant kernel. A level-scheduled sweep is one launch
    per dependency level (dozens of short launches), so its HIP-event span is
    a composite that no per-launch profiler average can confirm; then the
    roofline is that of the SpMV, one launch per product, and the sweep's
    composite rate stays in `kernels`."""
    spmv = kernel_stats(prof.spmv_ms, prof.spmv_launches, prof.spmv_bytes)
    sweep = kernel_stats(prof.sweep_ms, prof.sweep_launches, prof.sweep_bytes)
    sweep_dom = sweep["total_ms"] >= spmv["total_ms"] and sweep_single_launch
    name, dom = ("sweep", sweep) if sweep_dom else ("spmv", spmv)
    roof = dict(bound="hbm", achieved=dom["gbs"], peak=HBM_PEAK_GBS, unit="GB/s", frac=dom["gbs"] / HBM_PEAK_GBS,
                traffic=None, kernel=name, bytes_per_launch=dom["bytes_per_launch"], avg_launch_ms=dom["avg_ms"])
    return roof, {"spmv": spmv, "sweep": sweep}


# ---------------------------------------------------------------------------
# learned-guess workload


class CellFactory:
    """Generates, orders and factorizes cells on host threads, in order, at
    most `window` cells ahead of the consumer (bounded host memory)."""

    def __init__(self, B, params, threads, window):
        self.B = B
        self.params = params
        mat0, _ = B.gen_elastic_cell(CELL_M, 0.3, 10.0, 0.3, 0.0, TEMPLATE_SEED)
        self.order = mat0.rcm_order()  # every cell shares the template's pattern, hence its RCM ordering
        self.fparams = B.default_params(tau1=0.1, tau2=0.01)
        self.pool = ThreadPoolExecutor(threads)
        self.window = window
        self.futs = [None] * len(params)
        self.submitted = 0
        self.lock = threading.Lock()
        self.cv = threading.Condition()
        self.running, self.active = True, 0
        self._fill(0)

    def _fill(self, i):
        with self.lock:
            while self.submitted < min(len(self.params), i + self.window):
                self.futs[self.submitted] = self.pool.submit(self._build, self.submitted)
                self.submitted += 1

    def _build(self, i):
        with self.cv:
            self.cv.wait_for(lambda: self.running)
            self.active += 1
        try:
            r, lc, nu, sh = self.params[i]
            mat, rhs = self.B.gen_elastic_cell(CELL_M, r, 10 ** lc, nu, sh, TEMPLATE_SEED)
            mo = mat.reorder(self.order)
            f = self.B.Factor.ic2(mo, self.fparams)
            return mo, f, mo.reorder_vector(self.order, rhs)
        finally:
            with self.cv:
                self.active -= 1
                self.cv.notify_all()

    def quiet(self):
        """Holds new cells back and waits for the ones in progress: the host
        cores are idle while a timed solve runs."""
        with self.cv:
            self.running = False
            self.cv.wait_for(lambda: self.active == 0)

    def resume(self):
        with self.cv:
            self.running = True
            self.cv.notify_all()

    def get(self, i):
        self._fill(i)
        return self.futs[i].result()

    def release(self, i):
        self.futs[i] = None

    def close(self):
        self.resume()
        self.pool.shutdown(wait=False, cancel_futures=True)


def parse_tune(text):
    """Batch schedule from 'sweep_mode,waves_per_cu,level_order,graph,check_every'."""
    keys = ("sweep_mode", "waves_per_cu", "level_order", "graph", "check_every")
    vals = [int(v) for v in text.split(",")]
    if len(vals) != len(keys):
        raise SystemExit("--tune takes five comma-separated integers")
    return dict(zip(keys, vals))


def draw_params(k, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(0.1, 0.45, k), rng.uniform(0.0, 3.0, k), rng.uniform(0.2, 0.45, k),
                     rng.uniform(-0.15, 0.15, k)], axis=1)


def make_group(B, fac, idx, prec_mask):
    cells = [fac.get(i) for i in idx]
    bt = B.Batch([(mo, f) for mo, f, _ in cells], prec_mask=prec_mask)
    b = bt.gather([c[2] for c in cells])
    return bt, B.DeviceArray.from_host(b), cells


def run_learned(args, B):
    ntrain = args.train_cells
    group = args.group or GROUP
    neval = (args.warmup + args.steps) * group
    params = draw_params(ntrain + neval, PARAM_SEED)
    threads = host_threads()
    t_setup0 = time.perf_counter()
    fac = CellFactory(B, params, threads, window=group + 2 * threads)
    ctx = B.Context()
    n = 24 * CELL_M ** 3
    fprec = B.BF16 if args.factor_precision == "bf16" else B.F32
    fmask = 1 << fprec
    opts = B.default_opts(kind=B.BLOCKPCG, eps=EPS, precision=fprec, niter=5000)
    guess = B.Guess(n, 6, 4, nmodes=args.pod_modes, nhidden=args.mlp_hidden, capacity=ntrain, seed=11)
    tune = parse_tune(args.tune) if args.tune else None
    train_pairs, train_s, train_iters = 0, 0.0, []
    # prefetch the next group's device images while the current one solves
    next_group = [None]

    def prefetch(idx, mask):
        try:
            next_group[0] = make_group(B, fac, idx, mask)
        except BaseException as e:  # re-raised by the consumer
            next_group[0] = e

    def take():
        th.join()
        g = next_group[0]
        next_group[0] = None
        if isinstance(g, BaseException) or g is None:
            raise RuntimeError(f"cell preparation failed: {g!r}")
        return g

    groups = [list(range(k, min(ntrain, k + group))) for k in range(0, ntrain, group)]
    evalg = [list(range(ntrain + k * group, ntrain + (k + 1) * group)) for k in range(args.warmup + args.steps)]
    th = threading.Thread(target=prefetch, args=(groups[0], fmask))
    th.start()
    t_train0 = time.perf_counter()
    for gi, idx in enumerate(groups):
        bt, db, cells = take()
        nxt = groups[gi + 1] if gi + 1 < len(groups) else evalg[0]
        th = threading.Thread(target=prefetch, args=(nxt, fmask))
        th.start()
        if tune is None:
            t = bt.autotune(ctx, db, opts, 6, allow_graph=not args.no_graph)
            tune = dict(sweep_mode=t.sweep_mode, waves_per_cu=t.waves_per_cu, level_order=t.level_order,
                        graph=t.graph, check_every=t.check_every)
            log(f"autotuned schedule {tune} ({t.ms_per_iter:.2f} ms/iteration)")
        bt.set_tune(**tune)
        dx = B.DeviceArray(db.shape)
        dx.zero()
        t0 = time.perf_counter()
        st, its, rr = bt.solve(ctx, db, dx, opts, 6)
        train_s += time.perf_counter() - t0
        assert st.converged == st.pairs, f"training group {gi}: {st.converged}/{st.pairs} converged"
        train_pairs += st.pairs
        train_iters.append(float(its.mean()))
        guess.add(ctx, params[idx], dx)
        for i in idx:
            fac.release(i)
        if gi % 4 == 0 or gi == len(groups) - 1:
            log(f"training group {gi + 1}/{len(groups)}: {st.seconds * 1e3:.0f} ms, mean iterations {its.mean():.1f}, "
                f"host cells ready {time.perf_counter() - t_setup0:.0f}s")
        del bt, db, dx, cells
    t_train1 = time.perf_counter()
    rep = guess.fit(ctx, steps=args.fit_steps, lr=args.fit_lr)
    t_fit1 = time.perf_counter()
    log(f"guess fit: POD energy kept {rep.pod_energy:.6f}, loss {rep.loss_first:.3e} -> {rep.loss_final:.3e}, "
        f"POD {rep.seconds_pod:.2f}s, MLP {rep.seconds_train:.2f}s")

    # evaluation groups
    timed_s, timed_pairs, its_learned, its_zero, relres_max, ok = 0.0, 0, [], [], 0.0, True
    its_learned_max, its_zero_max = [], []
    prof = None
    plain_cells = []
    guess_err = None
    t_eval0 = time.perf_counter()
    for gi, idx in enumerate(evalg):
        bt, db, cells = take()
        bt.set_tune(**tune)
        dx = B.DeviceArray(db.shape)
        fac.quiet()  # nothing else runs on the host or the device during the timed solve
        ctx.sync()
        t0 = time.perf_counter()
        guess.predict(ctx, params[idx], dx)
        st, its, rr = bt.solve(ctx, db, dx, opts, 6, True, args.galerkin)
        dt = time.perf_counter() - t0
        timed = gi >= args.warmup
        if timed:
            timed_s += dt
            timed_pairs += st.pairs
            its_learned.append(float(its.mean()))
            its_learned_max.append(int(its.max()))
            relres_max = max(relres_max, float(rr.max()))
            ok = ok and st.converged == st.pairs and float(rr.max()) <= EPS
        if gi == args.warmup:
            # same group from x0 = 0 (iteration comparison) and one profiled pass (kernel roofline)
            dz = B.DeviceArray(db.shape)
            dz.zero()
            stz, itz, _ = bt.solve(ctx, db, dz, opts, 6)
            its_zero.append(float(itz.mean()))
            its_zero_max.append(int(itz.max()))
            ctx.profile(True)
            ctx.profile_read(reset=True)
            bt.set_tune(**dict(tune, graph=0))
            guess.predict(ctx, params[idx], dz)
            ctx.sync()
            x0h = dz.host()
            bt.solve(ctx, db, dz, opts, 6, True, args.galerkin)
            ctx.sync()
            prof = ctx.profile_read()
            ctx.profile(False)
            guess_err = guess_error(bt, x0h, dz.host())
            del x0h
            plain_cells = cells[:args.plain_cells]
            log(f"group {gi}: zero guess {itz.mean():.1f} iterations ({stz.seconds * 1e3:.0f} ms), learned "
                f"{its.mean():.1f} iterations ({st.seconds * 1e3:.0f} ms)")
        fac.resume()
        if gi + 1 < len(evalg):
            th = threading.Thread(target=prefetch, args=(evalg[gi + 1], fmask))
            th.start()
        log(f"eval group {gi + 1}/{len(evalg)}{'' if timed else ' (warm-up)'}: {dt * 1e3:.0f} ms, mean iterations "
            f"{its.mean():.1f}, max relres {rr.max():.2e}")
        for i in idx[args.plain_cells:] if gi == args.warmup else idx:
            fac.release(i)
        del bt, db, dx
    t_eval1 = time.perf_counter()
    eval_pairs = len(evalg) * group * 6
    fac.close()
    ref_tau = reference_thresholds_sample(args, B, ctx, params[ntrain + len(evalg) * group - args.ref_tau_cells:
                                                             ntrain + len(evalg) * group], tune) \
        if args.ref_tau_cells > 0 else None

    # plain FP64 drop-in on a sample of the evaluation cells: one right-hand side at a time
    popts = B.default_opts(kind=B.PCG, eps=EPS, ittype=3, check_every=8, niter=5000)
    pctx = B.Context()
    for mo, f, b in plain_cells:  # device images are uploaded outside the timed loop
        d1 = B.DeviceArray.from_host(np.ascontiguousarray(b[:, 0]))
        x1 = B.DeviceArray(mo.n)
        x1.zero()
        pctx.solve(mo, f, d1, x1, popts)
    nplain, plain_its = 0, []
    t0 = time.perf_counter()
    for mo, f, b in plain_cells:
        for q in range(6):
            d1 = B.DeviceArray.from_host(np.ascontiguousarray(b[:, q]))
            x1 = B.DeviceArray(mo.n)
            x1.zero()
            stp = pctx.solve(mo, f, d1, x1, popts)
            plain_its.append(stp.iterations)
            nplain += 1
            ok = ok and stp.converged == 1
    plain_s = time.perf_counter() - t0
    plain_rate = nplain / plain_s
    # one more plain solve with per-kernel HIP events (outside the timed loop;
    # polled every iteration so that no launch is queued past convergence)
    mo, f, b = plain_cells[0]
    d1 = B.DeviceArray.from_host(np.ascontiguousarray(b[:, 0]))
    x1 = B.DeviceArray(mo.n)
    x1.zero()
    pctx.profile(True)
    pctx.profile_read(reset=True)
    pctx.solve(mo, f, d1, x1, B.default_opts(kind=B.PCG, eps=EPS, ittype=3, check_every=1, niter=5000))
    pctx.sync()
    plain_roof, plain_kernels = roofline_from(pctx.profile_read())
    pctx.profile(False)

    roof, kernels = roofline_from(prof, sweep_single_launch=tune["sweep_mode"] != 0)
    traffic_file = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                                f"spmv_union_traffic_{group}cells_16cube_b24.json")
    if roof["kernel"] == "spmv" and os.path.exists(traffic_file):
        # measured HBM bytes per byte of algorithmic traffic of this SpMV, applied to this run's launches
        t = json.load(open(traffic_file))
        ratio = t["hbm_bytes_full_launch"] / t["algorithmic_bytes_full_launch"]
        roof["traffic"] = ratio * roof["bytes_per_launch"]
        roof["traffic_source"] = f"{t['source']}: {ratio:.4f} HBM bytes per algorithmic byte"
    kernels["sweep"]["launches_are"] = "sweeps (each one launch)" if tune["sweep_mode"] != 0 else \
        "sweeps (each one launch per dependency level)"
    value = timed_pairs / timed_s
    cpu = None
    t_cpu0 = time.perf_counter()
    if not args.no_cpu:
        cpu = cpu_baseline_ref(B, plain_cells, np.mean(plain_its), budget_s=20.0, cores=host_threads())
    t_cpu1 = time.perf_counter()
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": timed_s / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64", "mixed_precision": f"{args.factor_precision} factor sweeps inside the f64 recurrence; bf16 MLP",
        "data": "synthetic (seeded homogenization-cell generator)",
        "config": {
            "workload": f"learned initial guess: periodic {CELL_M}^3 elasticity cells (block 24, N={n}), 6 strain "
                        f"load cases each, IC2(0.1, 0.01) factors in {args.factor_precision.upper()}, batched block PCG in groups of {group} "
                        f"cells to FP64 true rel-res {EPS}; {ntrain} training cells, {args.steps} timed groups of "
                        f"{group} evaluation cells",
            "train_cells": ntrain,
            "eval_cells_timed": args.steps * group,
            "solve": "(cell, load case) pairs",
        },
        "valid": bool(ok),
        "plain_fp64_gpu": {"value": plain_rate, "unit": "solves/s", "sample": f"{nplain} solves "
                           f"({len(plain_cells)} evaluation cells x 6 load cases, one at a time, FP64 IC2, x0 = 0)",
                           "mean_iterations": float(np.mean(plain_its)),
                           "roofline": plain_roof, "kernels": plain_kernels},
        "speedup_vs_plain": value / plain_rate,
        "iterations": {"learned_mean": float(np.mean(its_learned)), "zero_guess_mean": float(np.mean(its_zero)),
                       "learned_group_max": its_learned_max, "zero_guess_group_max": its_zero_max,
                       "training_mean": float(np.mean(train_iters))},
        "guess_error": guess_err,
        "end_to_end": {
            "evaluation": {"value": eval_pairs / (t_eval1 - t_eval0), "unit": "solves/s",
                           "what": f"{len(evalg)} evaluation groups ({eval_pairs} solves, warm-up included), wall "
                                   "time from the first group's hand-over to the last solve: host cell generation, "
                                   "IC2 and batch build of the later groups overlap the solves, except that host "
                                   "preparation pauses while a solve is timed"},
            "training": {"value": train_pairs / (t_train1 - t_train0), "unit": "solves/s",
                         "what": "the training groups from x0 = 0, host preparation of the next group "
                                 "overlapping each solve (autotuning of the first group included)"},
        },
        "reference_thresholds": ref_tau,
        "relres_true_max": relres_max,
        "batched_zero_guess_training": {"value": train_pairs / train_s, "unit": "solves/s"},
        "guess_fit": {"pod_energy": rep.pod_energy, "loss_first": rep.loss_first, "loss_final": rep.loss_final,
                      "seconds_pod": rep.seconds_pod, "seconds_mlp": rep.seconds_train},
        "schedule": tune,
        "schedule_from": "--tune" if args.tune else "autotuner",
        "kernels": kernels,
        "roofline": roof,
        "cpu_baseline": cpu,
        "phase_seconds": {
            "first_training_group_ready": t_train0 - t_setup0,
            "training_groups": t_train1 - t_train0,
            "guess_fit": t_fit1 - t_train1,
            "evaluation_groups": t_eval1 - t_eval0,
            "cpu_baseline": t_cpu1 - t_cpu0,
            "whole_run": time.perf_counter() - t_setup0,
        },
    }
    print(json.dumps(line), flush=True)


def guess_error(bt, x0, x):
    """||x0 - x|| / ||x|| per (cell, load case) pair of one group: the learned
    initial guess against the converged solution."""
    e0 = bt.elem0
    errs = []
    for c in range(bt.ncell):
        a, b = x0[e0[c]:e0[c + 1], :], x[e0[c]:e0[c + 1], :]
        nx = np.linalg.norm(b, axis=0)
        errs.extend((np.linalg.norm(a - b, axis=0) / np.where(nx > 0, nx, 1.0)).tolist())
    errs = np.asarray(errs)
    return {"mean": float(errs.mean()), "median": float(np.median(errs)), "max": float(errs.max()),
            "pairs": int(errs.size), "what": "||x0 - x|| / ||x|| of the learned x0 against the converged x, "
                                             "every pair of the first timed group"}


def reference_thresholds_sample(args, B, ctx, params, tune):
    """A sample of evaluation-style cells factorized at the reference's IC2
    defaults (tau1 = 1e-2, tau2 = 1e-4): one batched FP32 solve of all of them
    from x0 = 0, and the plain FP64 drop-in on two of them, one load case at a
    time. The learned guess is not trained here (its 1024 training cells at
    9 s of host IC2 each are out of a benchmark's reach)."""
    k = len(params)
    t0 = time.perf_counter()
    mat0, _ = B.gen_elastic_cell(CELL_M, 0.3, 10.0, 0.3, 0.0, TEMPLATE_SEED)
    order = mat0.rcm_order()
    fpar = B.default_params(tau1=1e-2, tau2=1e-4, memory=-1.0)

    def build(i):
        r, lc, nu, sh = params[i]
        mat, rhs = B.gen_elastic_cell(CELL_M, r, 10 ** lc, nu, sh, TEMPLATE_SEED)
        mo = mat.reorder(order)
        return mo, B.Factor.ic2(mo, fpar), mo.reorder_vector(order, rhs)

    with ThreadPoolExecutor(host_threads()) as ex:
        cells = list(ex.map(build, range(k)))
    host_s = time.perf_counter() - t0
    bt = B.Batch([(mo, f) for mo, f, _ in cells], prec_mask=1 << B.F32)
    db = B.DeviceArray.from_host(bt.gather([c[2] for c in cells]))
    opts = B.default_opts(kind=B.BLOCKPCG, eps=EPS, precision=B.F32, niter=5000)
    t = bt.autotune(ctx, db, opts, 6, allow_graph=not args.no_graph)
    dx = B.DeviceArray(db.shape)
    dx.zero()
    ctx.sync()
    t1 = time.perf_counter()
    st, its, rr = bt.solve(ctx, db, dx, opts, 6)
    bs = time.perf_counter() - t1
    popts = B.default_opts(kind=B.PCG, eps=EPS, ittype=3, check_every=8, niter=5000)
    pctx = B.Context()
    for mo, f, b in cells[:1]:  # device images uploaded before timing
        d1 = B.DeviceArray.from_host(np.ascontiguousarray(b[:, 0]))
        x1 = B.DeviceArray(mo.n)
        x1.zero()
        pctx.solve(mo, f, d1, x1, popts)
    nplain, plain_s, pits, pok = 0, 0.0, [], True
    for mo, f, b in cells[:2]:
        for q in range(6):
            d1 = B.DeviceArray.from_host(np.ascontiguousarray(b[:, q]))
            x1 = B.DeviceArray(mo.n)
            x1.zero()
            pctx.sync()
            t2 = time.perf_counter()
            stp = pctx.solve(mo, f, d1, x1, popts)
            plain_s += time.perf_counter() - t2
            nplain += 1
            pits.append(stp.iterations)
            pok = pok and stp.converged == 1
    fi = cells[0][1].info(1)
    return {"tau": [1e-2, 1e-4], "cells": k,
            "batched_f32_zero_guess": {"value": st.pairs / bs, "unit": "solves/s", "mean_iterations": float(its.mean()),
                                       "relres_true_max": float(rr.max()),
                                       "valid": bool(st.converged == st.pairs and rr.max() <= EPS),
                                       "schedule": {"sweep_mode": t.sweep_mode, "graph": t.graph}},
            "plain_fp64_gpu": {"value": nplain / plain_s, "unit": "solves/s", "mean_iterations": float(np.mean(pits)),
                               "valid": bool(pok), "sample": f"{nplain} solves (2 cells x 6 load cases)"},
            "speedup_batched_vs_plain": (st.pairs / bs) / (nplain / plain_s),
            "factor": {"blocks": fi["nzja"], "values": fi["nza"], "levels": fi["levels_fwd"]},
            "host_ic2_seconds_per_cell": host_s * min(host_threads(), k) / k}


def cpu_baseline_ref(B, cells, iters_plain, budget_s=10.0, cores=1):
    """The reference's serial Cg (Bslib sources compiled into oracle/_ref),
    one independent solve per host core (`cores` threads; the ctypes call
    releases the interpreter lock, and the reference's Cg keeps no global
    state), over the load cases of the sample cells until every thread has
    done about budget_s of CPU work. value = solves / wall seconds."""
    import oracle as O

    use_ref = O.ref_available()
    tasks = []
    for mo, f, b in cells:
        A = O.Block(*mo.arrays())
        U = O.Block(A.sprnd, *f.arrays(1, A.nsup))
        for q in range(b.shape[1]):
            tasks.append((A, U, np.ascontiguousarray(b[:, q])))
    lock = threading.Lock()
    state = {"next": 0, "nsolve": 0, "nit": 0}

    def worker(t_end):
        while time.perf_counter() < t_end:
            with lock:
                k = state["next"]
                state["next"] += 1
            A, U, bq = tasks[k % len(tasks)]
            if use_ref:
                _, nit, _ = O.ref_cg(A, U, bq, eps=EPS, ittype=0, niter=5000)
            else:
                _, r = O.cg(A, U, bq, eps=EPS, ittype=3, niter=5000)
                nit = r.updates
            with lock:
                state["nsolve"] += 1
                state["nit"] += nit

    t0 = time.perf_counter()
    threads = [threading.Thread(target=worker, args=(t0 + budget_s,)) for _ in range(cores)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    nsolve = state["nsolve"]
    who = "the reference's serial Cg (compiled from its sources)" if use_ref else "the oracle's Cg"
    return dict(value=nsolve / wall, unit="solves/s", cores=cores, kind="reference" if use_ref else "port",
                sample=f"{nsolve} load-case solves of the first evaluation cells, one independent solve per core "
                       f"on {cores} host threads, by {who}: {state['nit'] / max(nsolve, 1):.1f} iterations each to "
                       f"rel-res {EPS}, {wall:.1f}s wall on {cpu_model()}")


# ---------------------------------------------------------------------------
# plain FP64 drop-in workload (nonsymmetric, 32^3, ILU2 + BiCGStab)


def run_dropin(args, B):
    import oracle as O

    seed = 20261203
    t0 = time.perf_counter()
    m, rhs = B.gen_seepage3d(32, 32, 32, 24, 0.3, seed)
    order = m.rcm_order()
    mo = m.reorder(order)
    b = mo.reorder_vector(order, rhs)
    params = B.default_params()
    f = B.Factor.ilu2(mo, params)
    fct_s = time.perf_counter() - t0
    log(f"drop-in system n={mo.n}, ILU2 {fct_s:.1f}s")
    ctx = B.Context()
    db, dx = B.DeviceArray.from_host(b), B.DeviceArray(mo.n)
    opts = B.default_opts(kind=B.BICGSTAB, eps=EPS, ittype=3, check_every=4, niter=2000)
    for _ in range(args.warmup):
        dx.zero()
        ctx.solve(mo, f, db, dx, opts)
    ctx.sync()
    iters, relres, ok = [], [], True
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dx.zero()
        st = ctx.solve(mo, f, db, dx, opts)
        iters.append(st.iterations)
        relres.append(st.relres_true)
        ok = ok and st.converged == 1 and st.relres_true <= EPS
    ctx.sync()
    dt = time.perf_counter() - t0
    # one more solve with per-kernel HIP events (outside the timed region)
    ctx.profile(True)
    ctx.profile_read(reset=True)
    dx.zero()
    ctx.solve(mo, f, db, dx, B.default_opts(kind=B.BICGSTAB, eps=EPS, ittype=3, check_every=1, niter=2000))
    ctx.sync()
    prof = ctx.profile_read()
    ctx.profile(False)
    roof, kernels = roofline_from(prof)
    cpu = None
    if not args.no_cpu:
        A = O.Block(*mo.arrays())
        L = O.Block(A.sprnd, *f.arrays(0, A.nsup))
        U = O.Block(A.sprnd, *f.arrays(1, A.nsup))
        t1 = time.perf_counter()
        _, r = O.bicgstab(A, L, U, b, niter=8, eps=1e-300)
        per = (time.perf_counter() - t1) / max(r.nitperf, 1)
        cpu = dict(value=1.0 / (per * np.median(iters)), unit="solves/s", cores=1, kind="port",
                   sample=f"{r.nitperf} BiCGStab iterations by the oracle on one core of {cpu_model()}, scaled to "
                          f"the {int(np.median(iters))} iterations of one GPU solve")
    print(json.dumps({
        "metric": METRIC, "value": args.steps / dt, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded generator)",
        "config": {"workload": f"plain FP64 GPU drop-in: 3D nonsymmetric 32^3 cells, block 24, 7-point, N={mo.n}, "
                               f"RCM + ILU2(1e-2, 1e-4) + BiCGStab to true rel-res {EPS}", "seed": seed},
        "valid": bool(ok), "iterations": iters, "relres_true_max": max(relres), "factor_seconds_host": fct_s,
        "kernels": kernels, "roofline": roof, "cpu_baseline": cpu,
    }), flush=True)


def run_complex(args, B):
    """The complex solve phase (the complex AbstractSolver's CSVectorC::Solver,
    solvers.cpp:2588-2830) on a 3D complex system: 32^3 cells, 7-point,
    complex blocks of 12 (real part one seeded seepage matrix, imaginary part
    0.4 x another of the same pattern), N = 393,216 complex unknowns, RCM +
    complex ILU2(1e-2, 1e-4) on the host (bit-exact with the reference's),
    block GMRES (gmres.cpp:1044) to ||r|| < 1e-10 ||r0|| on the GPU: the real
    block GMRES over the real-equivalent columns [b, i b], the matrix and
    factors held on the device as complex 12 x 12 blocks. The
    CPU baseline is the reference's own complex BlockGMRES on one core with
    the reference's own factors, which are also compared bit for bit with
    this library's."""
    import oracle as O

    seed = 20261207
    t0 = time.perf_counter()
    m, _ = B.gen_seepage3d(32, 32, 32, 12, 0.3, seed)
    m2, _ = B.gen_seepage3d(32, 32, 32, 12, 0.5, seed + 4)
    order = m.rcm_order()
    sp, ia, ja, a = m.reorder(order).arrays()
    ac = a + 0.4j * m2.reorder(order).arrays()[3]
    del m, m2, a
    n = int(sp[-1])
    rng = np.random.default_rng(seed)
    bz = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    params = B.default_params()
    f = B.Factor.ilu2_complex(sp, ia, ja, ac, params)
    fct_s = time.perf_counter() - t0
    mc = B.Matrix.from_complex(sp, ia, ja, ac)
    log(f"complex system n={n} (real equivalent {2 * n}), complex ILU2 {params.timefct:.1f}s")
    bb = np.empty((2 * n, 2), order="F")
    bb[:, 0] = B.complex_to_real(sp, bz)
    bb[:, 1] = B.complex_to_real(sp, 1j * bz)
    bnorm = float(np.linalg.norm(bz))
    ctx = B.Context()
    db, dx = B.DeviceArray.from_host(bb), B.DeviceArray((2 * n, 2))
    # polled every step, as the complex drop-in does: a step costs far more than the poll
    opts = B.default_opts(kind=B.BLOCKGMRES, eps=EPS, ittype=0, niter=400, check_every=1)
    for _ in range(args.warmup):
        dx.zero()
        ctx.solve_block(mc, f, db, dx, 2, opts)
    ctx.sync()
    iters, relres, ok = [], [], True
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dx.zero()
        st, resi, _ = ctx.solve_block(mc, f, db, dx, 2, opts)
        iters.append(st.iterations)
        relres.append(float(resi[0]) / bnorm)
        ok = ok and st.converged == 1
    ctx.sync()
    dt = time.perf_counter() - t0
    x = B.real_to_complex(sp, dx.host()[:, 0])
    ok = ok and max(relres) <= 2 * EPS
    # one more solve with per-kernel HIP events (outside the timed region),
    # polled every step so that no launch is queued past convergence
    ctx.profile(True)
    ctx.profile_read(reset=True)
    dx.zero()
    ctx.solve_block(mc, f, db, dx, 2, B.default_opts(kind=B.BLOCKGMRES, eps=EPS, ittype=0, niter=400, check_every=1))
    ctx.sync()
    prof = ctx.profile_read()
    ctx.profile(False)
    roof, kernels = roofline_from(prof)
    cpu, same_factor, rel = None, None, None
    if not args.no_cpu and O.ref_available():
        L, U = O.ref_cilu2(sp, ia, ja, ac, memory=10.0)
        fr = B.Factor.from_complex(sp, L, U)
        nsup = len(sp) - 1
        same_factor = all(np.array_equal(g, w) for which in (0, 1)
                          for g, w in zip(f.arrays(which, nsup), fr.arrays(which, nsup)))
        del fr, L, U
        steps = int(np.median(iters)) + 1
        probe = min(steps, 4)
        _, sec = O.ref_cblock_gmres(sp, ia, ja, ac, bz, niter=probe, eps=0.0)
        if sec / probe * steps <= 60.0:  # the whole solve fits the CPU budget: time it and compare
            xr, sec = O.ref_cblock_gmres(sp, ia, ja, ac, bz, niter=opts.niter, eps=EPS)
            rel = float(np.linalg.norm(x - xr) / np.linalg.norm(xr))
            sample = f"one whole solve ({steps} block GMRES steps)"
            value = 1.0 / sec
        else:
            sample = f"{probe} block GMRES steps, scaled to the {steps} steps of one GPU solve"
            value = probe / (sec * steps)
        cpu = dict(value=value, unit="solves/s", cores=1, kind="reference",
                   sample=f"the reference's own complex BlockGMRES (gmres.cpp:1044, compiled from its sources) with "
                          f"its own complex ILU2 factors on one core of {cpu_model()}: {sample}")
    print(json.dumps({
        "metric": METRIC, "value": args.steps / dt, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded generator)",
        "config": {"workload": f"complex solve phase: 3D nonsymmetric 32^3 cells, complex block 12, 7-point, "
                               f"N={n} complex, RCM + complex ILU2(1e-2, 1e-4) + block GMRES over the real "
                               f"equivalent [b, i b] to ||r|| < {EPS} ||r0||", "seed": seed},
        "valid": bool(ok), "iterations": iters, "relres_true_max": max(relres), "factor_seconds_host": fct_s,
        "factor_bitwise_equal_reference": same_factor, "rel_diff_vs_reference": rel,
        "kernels": kernels, "roofline": roof, "cpu_baseline": cpu,
    }), flush=True)


def run_precision(args, B):
    """FP64 flexible GMRES(30) around FP64, FP32 and BF16 ILU2 factor sweeps on
    the 32^3 nonsymmetric system; every variant driven to FP64 true relative
    residual 1e-10 and compared with the FP64-preconditioner solution."""
    import oracle as O

    seed = 20261203
    t0 = time.perf_counter()
    m, rhs = B.gen_seepage3d(32, 32, 32, 24, 0.3, seed)
    order = m.rcm_order()
    mo = m.reorder(order)
    b = mo.reorder_vector(order, rhs)
    f = B.Factor.ilu2(mo, B.default_params())
    fct_s = time.perf_counter() - t0
    log(f"precision system n={mo.n}, ILU2 {fct_s:.1f}s")
    ctx = B.Context()
    db, dx = B.DeviceArray.from_host(b), B.DeviceArray(mo.n)
    res, xref, ok = {}, None, True
    for name, prec in (("f64", B.F64), ("f32", B.F32), ("bf16", B.BF16)):
        opts = B.default_opts(kind=B.FGMRES, eps=EPS, ittype=3, restart=30, precision=prec, check_every=4,
                              niter=3000)
        for _ in range(max(1, args.warmup)):  # also uploads this precision's factor image
            dx.zero()
            ctx.solve(mo, f, db, dx, opts)
        ctx.sync()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            dx.zero()
            st = ctx.solve(mo, f, db, dx, opts)
        ctx.sync()
        dt = (time.perf_counter() - t1) / args.steps
        x = dx.host()
        if xref is None:
            xref = x
        ok = ok and st.converged == 1 and st.relres_true <= EPS
        res[name] = dict(ms_per_solve=dt * 1e3, solves_per_s=1.0 / dt, iterations=st.iterations,
                         restarts=st.restarts, relres_true=st.relres_true,
                         rel_diff_vs_f64=float(np.linalg.norm(x - xref) / np.linalg.norm(xref)))
        log(f"{name}: {dt * 1e3:.2f} ms/solve, {st.iterations} iterations, relres {st.relres_true:.2e}, "
            f"|x - x64|/|x64| {res[name]['rel_diff_vs_f64']:.2e}")
    # kernel rates of one FP32 solve (outside the timed region)
    ctx.profile(True)
    ctx.profile_read(reset=True)
    dx.zero()
    ctx.solve(mo, f, db, dx, B.default_opts(kind=B.FGMRES, eps=EPS, ittype=3, restart=30, precision=B.F32,
                                            check_every=1, niter=3000))
    ctx.sync()
    prof = ctx.profile_read()
    ctx.profile(False)
    roof, kernels = roofline_from(prof)
    cpu = None
    if not args.no_cpu:
        A = O.Block(*mo.arrays())
        L = O.Block(A.sprnd, *f.arrays(0, A.nsup))
        U = O.Block(A.sprnd, *f.arrays(1, A.nsup))
        t1 = time.perf_counter()
        _, r = O.fgmres(A, L, U, b, m=30, niter=8, eps=1e-300)
        per = (time.perf_counter() - t1) / max(r.nitperf, 1)
        its = res["f64"]["iterations"]
        cpu = dict(value=1.0 / (per * its), unit="solves/s", cores=1, kind="port",
                   sample=f"{r.nitperf} FGMRES iterations (FP64 factors) by the oracle on one core of {cpu_model()}, "
                          f"scaled to the {its} iterations of one solve")
    print(json.dumps({
        "metric": METRIC, "value": res["f32"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["f32"]["ms_per_solve"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "mixed_precision": "f32 factor sweeps",
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"mixed precision: FP64 FGMRES(30) with ILU2 factors stored and applied in FP32 "
                               f"(value) / BF16 / FP64, 3D nonsymmetric 32^3 cells, block 24, N={mo.n}, to true "
                               f"rel-res {EPS}", "seed": seed},
        "valid": bool(ok), "precisions": res, "factor_seconds_host": fct_s,
        "kernels": kernels, "roofline": roof, "cpu_baseline": cpu,
    }), flush=True)


def run_batch(args, B):
    """64 periodic elasticity cells of uneven sizes 8^3..24^3 x 6 load cases
    in one batched block-PCG solve (FP32 factors, FP64 outer recurrence,
    autotuned schedule) from x0 = 0, against the plain FP64 drop-in one load
    case at a time."""
    rng = np.random.default_rng(20261207)
    ncell = args.group or 64
    sizes = rng.integers(8, 25, ncell)
    params = draw_params(ncell, 20261207)
    fp = B.default_params(tau1=0.1, tau2=0.01)
    orders = {}
    for msz in sorted(set(int(v) for v in sizes)):
        mat0, _ = B.gen_elastic_cell(msz, 0.3, 10.0, 0.3, 0.0, TEMPLATE_SEED)
        orders[msz] = mat0.rcm_order()

    def build(i):
        msz = int(sizes[i])
        r, lc, nu, sh = params[i]
        mat, rhs = B.gen_elastic_cell(msz, r, 10 ** lc, nu, sh, TEMPLATE_SEED)
        mo = mat.reorder(orders[msz])
        return mo, B.Factor.ic2(mo, fp), mo.reorder_vector(orders[msz], rhs)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(host_threads()) as ex:
        cells = list(ex.map(build, range(ncell)))
    setup_s = time.perf_counter() - t0
    log(f"batch of {ncell} cells ({int(sizes.min())}^3..{int(sizes.max())}^3), host setup {setup_s:.1f}s")
    bt = B.Batch([(mo, f) for mo, f, _ in cells], prec_mask=1 << B.F32)
    b = bt.gather([c[2] for c in cells])
    db = B.DeviceArray.from_host(b)
    ctx = B.Context()
    opts = B.default_opts(kind=B.BLOCKPCG, eps=EPS, precision=B.F32, niter=5000)
    if args.tune:
        tune = parse_tune(args.tune)
        bt.set_tune(**tune)
    else:
        t = bt.autotune(ctx, db, opts, 6, allow_graph=not args.no_graph)
        tune = dict(sweep_mode=t.sweep_mode, waves_per_cu=t.waves_per_cu, level_order=t.level_order, graph=t.graph,
                    check_every=t.check_every)
        log(f"autotuned schedule {tune} ({t.ms_per_iter:.2f} ms/iteration)")
    dx = B.DeviceArray(db.shape)
    for _ in range(max(1, args.warmup)):
        dx.zero()
        bt.solve(ctx, db, dx, opts, 6)
    ctx.sync()
    timed, pairs, ok, its_all, rr_max = 0.0, 0, True, [], 0.0
    for _ in range(args.steps):
        dx.zero()
        ctx.sync()
        t1 = time.perf_counter()
        st, its, rr = bt.solve(ctx, db, dx, opts, 6)
        timed += time.perf_counter() - t1
        pairs += st.pairs
        its_all.append(its)
        rr_max = max(rr_max, float(rr.max()))
        ok = ok and st.converged == st.pairs and float(rr.max()) <= EPS
    value = pairs / timed
    # kernel rates of one more solve (outside the timed region)
    bt.set_tune(**dict(tune, graph=0))
    ctx.profile(True)
    ctx.profile_read(reset=True)
    dx.zero()
    bt.solve(ctx, db, dx, opts, 6)
    ctx.sync()
    prof = ctx.profile_read()
    ctx.profile(False)
    roof, kernels = roofline_from(prof, sweep_single_launch=tune["sweep_mode"] != 0)
    # plain FP64 drop-in on a sample of the cells, one load case at a time
    popts = B.default_opts(kind=B.PCG, eps=EPS, ittype=3, check_every=8, niter=5000)
    pctx = B.Context()
    sample = cells[:args.plain_cells]
    for mo, f, bb in sample:
        d1 = B.DeviceArray.from_host(np.ascontiguousarray(bb[:, 0]))
        x1 = B.DeviceArray(mo.n)
        x1.zero()
        pctx.solve(mo, f, d1, x1, popts)
    nplain, plain_s = 0, 0.0
    for mo, f, bb in sample:
        for q in range(6):
            d1 = B.DeviceArray.from_host(np.ascontiguousarray(bb[:, q]))
            x1 = B.DeviceArray(mo.n)
            x1.zero()
            pctx.sync()
            t1 = time.perf_counter()
            stp = pctx.solve(mo, f, d1, x1, popts)
            plain_s += time.perf_counter() - t1
            nplain += 1
            ok = ok and stp.converged == 1
    # the sample's share of the batch's unknowns, to compare like with like
    frac = sum(c[0].n for c in sample) / bt.n
    cpu = None if args.no_cpu else cpu_baseline_ref(B, sample, 0.0)
    its_all = np.concatenate([i.ravel() for i in its_all])
    print(json.dumps({
        "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": timed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "mixed_precision": "f32 factor sweeps",
        "data": "synthetic (seeded homogenization-cell generator)",
        "config": {"workload": f"batching: {ncell} periodic elasticity cells of uneven sizes "
                               f"{int(sizes.min())}^3..{int(sizes.max())}^3 (block 24, N={bt.n} in total), 6 load "
                               f"cases each, IC2(0.1, 0.01) in FP32, one batched block-PCG solve from x0 = 0 to FP64 "
                               f"true rel-res {EPS}", "cells": ncell, "solve": "(cell, load case) pairs"},
        "valid": bool(ok), "iterations": {"mean": float(its_all.mean()), "max": int(its_all.max())},
        "relres_true_max": rr_max, "schedule": tune,
        "plain_fp64_gpu": {"value": nplain / plain_s, "unit": "solves/s",
                           "sample": f"{nplain} solves of the first {len(sample)} cells ({frac:.3f} of the unknowns), "
                                     f"one at a time, FP64 IC2, x0 = 0",
                           "value_size_adjusted": 6 * ncell * frac / plain_s,
                           "size_adjusted": "the sample's time per unknown and load case applied to the batch's "
                                            "unknowns: solves/s the plain path would reach on this batch"},
        "speedup_vs_plain_size_adjusted": value / (6 * ncell * frac / plain_s),
        "host_cell_generation_and_ic2_seconds": setup_s, "kernels": kernels, "roofline": roof, "cpu_baseline": cpu,
    }), flush=True)


def run_heat(args, B):
    """The CPU-runnable 2D heat system (64 x 64 cells, block 9, N = 36,864,
    IC2(1e-2, 1e-4), PCG to rel-res 1e-10 relative to ||r0||) on the GPU,
    against the reference's own serial Cg (compiled from its sources) on the
    same factor: same iteration count expected."""
    import oracle as O

    seed = 20261201
    m, rhs = B.gen_heat2d(64, 64, 9, seed)
    order = m.rcm_order()
    mo = m.reorder(order)
    b = mo.reorder_vector(order, rhs)
    f = B.Factor.ic2(mo, B.default_params(tau1=1e-2, tau2=1e-4))
    ctx = B.Context()
    db, dx = B.DeviceArray.from_host(b), B.DeviceArray(mo.n)
    opts = B.default_opts(kind=B.PCG, eps=EPS, ittype=0, check_every=8, niter=5000)
    for _ in range(max(1, args.warmup)):
        dx.zero()
        ctx.solve(mo, f, db, dx, opts)
    ctx.sync()
    iters, ok = [], True
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dx.zero()
        st = ctx.solve(mo, f, db, dx, opts)
        iters.append(st.iterations)
        ok = ok and st.converged == 1
    ctx.sync()
    dt = time.perf_counter() - t0
    x = dx.host()
    ctx.profile(True)
    ctx.profile_read(reset=True)
    dx.zero()
    ctx.solve(mo, f, db, dx, B.default_opts(kind=B.PCG, eps=EPS, ittype=0, check_every=1, niter=5000))
    ctx.sync()
    prof = ctx.profile_read()
    ctx.profile(False)
    roof, kernels = roofline_from(prof)
    cpu, ref_iters, rel = None, None, None
    if not args.no_cpu:
        A = O.Block(*mo.arrays())
        U = O.Block(A.sprnd, *f.arrays(1, A.nsup))
        use_ref = O.ref_available()
        nsolve, t_sum = 0, 0.0
        while t_sum < 10.0 and nsolve < 200:
            t1 = time.perf_counter()
            if use_ref:
                xr, ref_iters, _ = O.ref_cg(A, U, b, eps=EPS, ittype=0, niter=5000)
            else:
                xr, r = O.cg(A, U, b, eps=EPS, ittype=0, niter=5000)
                ref_iters = r.updates
            t_sum += time.perf_counter() - t1
            nsolve += 1
        rel = float(np.linalg.norm(x - xr) / np.linalg.norm(xr))
        who = "the reference's serial Cg (compiled from its sources)" if use_ref else "the oracle's Cg"
        cpu = dict(value=nsolve / t_sum, unit="solves/s", cores=1, kind="reference" if use_ref else "port",
                   sample=f"{nsolve} solves of the same system by {who}, nitperf {ref_iters} each, "
                          f"{t_sum:.1f}s on one core of {cpu_model()}")
    print(json.dumps({
        "metric": METRIC, "value": args.steps / dt, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded generator)",
        "config": {"workload": f"2D heat, 64 x 64 cells, block 9, N={mo.n}, RCM + IC2(1e-2, 1e-4) + PCG to rel-res "
                               f"{EPS} (relative to ||r0||)", "seed": seed},
        "valid": bool(ok), "iterations": iters, "nitperf": [i + 1 for i in iters],
        "reference_nitperf": ref_iters,
        "rel_diff_vs_reference": rel, "kernels": kernels, "roofline": roof, "cpu_baseline": cpu,
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed evaluation groups (learned) or solves (drop-in)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["learned", "drop-in", "precision", "batch", "heat", "complex"],
                    default="learned")
    ap.add_argument("--train-cells", type=int, default=1024)
    ap.add_argument("--fit-steps", type=int, default=16000)
    ap.add_argument("--fit-lr", type=float, default=5e-3)
    ap.add_argument("--pod-modes", type=int, default=64, help="POD modes per load case of the initial-guess model")
    ap.add_argument("--ref-tau-cells", type=int, default=16,
                    help="cells solved at the reference's IC2 thresholds (1e-2, 1e-4) beside the headline (0: none)")
    ap.add_argument("--mlp-hidden", type=int, default=256, help="width of the guess MLP's two hidden layers")
    ap.add_argument("--plain-cells", type=int, default=4)
    ap.add_argument("--factor-precision", choices=["f32", "bf16"], default="f32",
                    help="storage of the batched solves' IC2 factors (the outer recurrence stays FP64)")
    ap.add_argument("--group", type=int, default=None,
                    help=f"cells per batched solve (learned: {GROUP}, batch: 64)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--galerkin", action="store_true",
                    help="start every evaluation cell from the A-optimal combination of its six learned guesses")
    ap.add_argument("--no-graph", action="store_true", help="leave hipGraph replay out of the batch schedules")
    ap.add_argument("--tune", default=None,
                    help="fixed batch schedule 'sweep_mode,waves_per_cu,level_order,graph,check_every' instead of "
                         "the autotuner (e.g. to profile the schedule an unprofiled run chose)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import bcmhip as B

    if B.device_count() < 1:
        raise SystemExit("no GPU visible: the solve phase runs on the GPU only")
    if args.workload == "learned":
        run_learned(args, B)
    elif args.workload == "precision":
        run_precision(args, B)
    elif args.workload == "batch":
        run_batch(args, B)
    elif args.workload == "heat":
        run_heat(args, B)
    elif args.workload == "complex":
        run_complex(args, B)
    else:
        run_dropin(args, B)
    # release what the library keeps for the process before the runtime's teardown
    B.shutdown()


if __name__ == "__main__":
    main()
