"""Benchmark of the hot path: forward-Euler CSR heat stepping on a 256^3 grid.

Workload (the reference's explicit-stepping configuration (a)): padded
256x256x256 grid (N = 16,777,216 rows), 7-point central Laplacian assembled
as CSR (117,047,296 non-zeros, int32 indices, FP32 values), per-cell
diffusivity alpha ~ U[0.005, 0.02) from seed 20261203 applied as a row
scaling, dt = 0.15*dx^2/0.02, Gaussian initial field (sigma 0.1, centre 0.5),
zero Dirichlet faces.  One step = one launch of the fused SpMV + interior
mask + face-BC kernel over the whole grid.  Inputs are resident in HBM
before the timed region.

Prints ONE JSON line.  `value` is cell-steps/s (padded rows x steps / s).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-steps/s + SpMV HBM GB/s (256³ heat); surrogate vs plain speedup @1e-3 L2"
HBM_PEAK_GBS = 8000.0
DIMS = (256, 256, 256)
SEED = 20261203
COPY_PROBE_FLOATS = 64 << 20   # 256 MiB read + 256 MiB written per copy launch


def algorithmic_bytes(n_rows: int, nnz: int) -> int:
    # col + val per non-zero, row_ptr (N+1), x read once, y written once
    return 8 * nnz + 4 * (n_rows + 1) + 4 * n_rows + 4 * n_rows


def load_traffic(variant_name):
    """HBM bytes per launch of the FE kernel measured with rocprofv3 counters
    (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel calibration), for
    the kernel variant that was measured; None when no measurement of this
    variant is on file."""
    path = os.path.join(ROOT, "profiles", "fe_csr_256cubed_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    rec = d.get("variants", {}).get(variant_name)
    return None if rec is None else rec.get("hbm_bytes_per_launch")


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(x0_host, x_dev_host, n_dev_steps, budget_s: float = 12.0, one_thread_s: float = 4.0):
    """The CPU oracle (C restatement of the reference's step) on the same
    256^3 workload: (1) the parity and FP64-error check of the device's final
    field -- the FP32 oracle and an FP64 twin run the same number of steps
    from the same initial field; (2) the CPU baseline, timed on this host's
    cores on a bounded sample of steps, with every thread and with one."""
    import numpy as np
    import oracle as po

    nx, ny, nz = DIMS
    norm = po.lap_norm(1.0, nx - 2)
    A = po.lap_csr(nx, ny, nz, po.CD, norm)
    alpha = po.alpha_uniform(A.n_rows, SEED, 0.005, 0.02)
    A = po.row_scale(A, alpha)
    dt = np.float32(np.float32(0.15) * norm / np.float32(0.02))
    snap = np.zeros(A.n_rows, np.float32)
    # (1) parity of the timed run's field
    t0 = time.perf_counter()
    ref32 = po.fe_csr_run(DIMS, A, dt, n_dev_steps, x0_host, snap)
    ref64 = po.fe_csr_run_f64(DIMS, A, float(dt), n_dev_steps, x0_host, snap)
    parity = {
        "steps": n_dev_steps,
        "bit_exact_vs_fp32_oracle": bool(x_dev_host.tobytes() == ref32.tobytes()),
        "rel_l2_vs_fp32_oracle": po.rel_l2(x_dev_host, ref32),
        "fp64_rel_l2": po.rel_l2(x_dev_host, ref64),
        "fp32_oracle_fp64_rel_l2": po.rel_l2(ref32, ref64),
        "seconds": time.perf_counter() - t0,
    }
    # (2) the baseline
    x = x0_host.copy()
    y = np.empty_like(x)
    threads = int(po.lib.po_omp_threads())

    def timed(budget):
        t0 = time.perf_counter()
        po.lib.po_fe_csr_step(nx, ny, nz, A.row_ptr, A.col, A.val, dt, x, None, y)
        t1 = time.perf_counter() - t0
        steps = max(1, min(3000, int(budget / max(t1, 1e-3))))
        a, b = x, y
        t0 = time.perf_counter()
        for _ in range(steps):
            po.lib.po_fe_csr_step(nx, ny, nz, A.row_ptr, A.col, A.val, dt, a, None, b)
            a, b = b, a
        return steps, time.perf_counter() - t0

    steps, el = timed(budget_s)
    po.lib.po_set_omp_threads(1)
    steps1, el1 = timed(one_thread_s)
    po.lib.po_set_omp_threads(threads)
    base = {
        "value": A.n_rows * steps / el,
        "unit": "cell-steps/s",
        "cores": threads,
        "kind": "port",
        "label": "CPU restatement of PIRO's path",
        "sample": f"{steps} FE steps of the 256^3 CSR heat step (oracle/piro_oracle.c, OpenMP over rows, "
                  f"{threads} threads), {el:.1f} s; one thread: {steps1} steps, {el1:.1f} s",
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
        "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
        "one_thread_value": A.n_rows * steps1 / el1,
    }
    return base, parity


def surrogate_leg(ctx):
    """The metric's second half: the learned k = 16 propagator, trained on
    trajectories of the batched stepper by the library's own kernels, then a
    1600-step 128^3 run with every 10th call checked against 16 plain CSR
    steps (rejected above 1e-3 relative L2), timed against the plain
    stepper on the same field."""
    from piro_amd import learned
    from piro_amd.core import Surrogate

    sur = Surrogate(ctx, seed=5)
    log = learned.train(ctx, sur, n_traj=32000, per_batch=50, side=32, segments=4, k=16, passes=3, lr=1e-3,
                        lr_final=1e-5, blob_widths=(4.0, 20.0), noise=0.03, log_every=400)
    run = learned.HeatRun(ctx, side=128)
    steps = 1600
    # each path warmed by a run of the timed length (the first ~100 launches
    # on a fresh box run slower than the steady state)
    run.plain(steps)
    ref, t_plain = run.plain(steps)
    out = {
        "workload": "heat 128^3, smooth alpha field in [0.005,0.02], Gaussian initial condition, 1600 FE steps; "
                    "surrogate calls of k=16 steps, every 10th checked with 16 plain steps (the checked call "
                    "takes the plain result; with defect correction the plain-minus-surrogate difference of the "
                    "last check is added to the following calls)",
        "plain_seconds": t_plain,
        "check_tolerance": 1e-3,
        "max_amplification": learned.max_amplification(sur.stencil()),
        "train_trajectories": log.trajectories,
        "train_seconds": log.seconds,
    }
    for key, correct in (("corrected", True), ("uncorrected", False)):
        run.accelerated(sur, steps, correct=correct)
        acc, res = run.accelerated(sur, steps, correct=correct)
        out[key] = {
            "speedup_vs_plain": t_plain / res.seconds,
            "accelerated_seconds": res.seconds,
            "checks": res.checks,
            "rejected": res.rejected,
            "check_rel_l2_max": max(res.check_errors) if res.check_errors else None,
            "final_rel_l2_vs_plain": ctx.rel_l2(acc, ref),
        }
    out["speedup_vs_plain"] = out["corrected"]["speedup_vs_plain"]
    out["final_rel_l2_vs_plain"] = out["corrected"]["final_rel_l2_vs_plain"]
    sur.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the surrogate leg (about two minutes of training on the batched stepper)")
    ap.add_argument("--variant", type=int, default=-1, help="force an FE kernel variant (-1 = autotune)")
    ap.add_argument("--copy-reps", type=int, default=10,
                    help="after the timed region, time the library's streaming copy on a 256 MiB vector this many "
                         "times (the box's practical HBM rate, and the known-size kernel that HBM counter "
                         "readings are calibrated against); 0 skips it")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: the operator would shard into z-slabs with one halo plane "
                                   "exchanged per step, which is not built", "n_gpus": args.gpus}))
        return 2

    import numpy as np

    from piro_amd.core import CD, Context

    ctx = Context(0)
    nx, ny, nz = DIMS
    norm = ctx.lap_norm(1.0, nx - 2)
    A = ctx.laplacian_csr(nx, ny, nz, CD, norm)
    alpha = ctx.alpha_uniform(A.n_rows, SEED, 0.005, 0.02)
    ctx.row_scale(A, alpha)
    dt = np.float32(np.float32(0.15) * np.float32(norm) / np.float32(0.02))
    x = ctx.ic_gaussian(DIMS, [1.0] * 3, [0.5] * 3, [0.1] * 3, 1.0)
    x0_host = x.host()
    snap = ctx.apply_face_values(DIMS, x, [0.0] * 6)
    y = ctx.empty_like(x)
    if args.variant >= 0:
        ctx.set_variant(args.variant)
        variant = args.variant
    else:
        variant, _ = ctx.autotune(DIMS, A, dt, x, snap, y)
    vname = ctx.variants()[variant]

    def step(a, b):
        ctx.fe_step_csr(DIMS, A, dt, a, snap, b)

    for _ in range(args.warmup):
        step(x, y)
        x, y = y, x
    ctx.sync()
    # HIP events on the library's stream, around exactly `steps` launches
    t_host = time.perf_counter()
    ctx.timer_start()
    for _ in range(args.steps):
        step(x, y)
        x, y = y, x
    ms = ctx.timer_stop()
    ctx.sync()
    t_host = time.perf_counter() - t_host
    ms_step = ms / max(args.steps, 1)
    n_rows, nnz = A.n_rows, A.nnz
    cell_steps = n_rows * args.steps / (ms / 1e3)
    interior = (nx - 2) * (ny - 2) * (nz - 2)
    byts = algorithmic_bytes(n_rows, nnz)
    gbs = byts / (ms_step / 1e3) / 1e9
    finite = bool(np.isfinite(ctx.reduce("sumsq", x)))
    copy_gbs = None
    if args.copy_reps > 0:
        src = ctx.full(COPY_PROBE_FLOATS, 1.0)
        dst = ctx.empty_like(src)
        ctx.copy(src, dst)
        ctx.sync()
        ctx.timer_start()
        for _ in range(args.copy_reps):
            ctx.copy(src, dst)
        cms = ctx.timer_stop()
        copy_gbs = 8.0 * COPY_PROBE_FLOATS * args.copy_reps / (cms / 1e3) / 1e9
        src.free()
        dst.free()

    base = parity = None
    if not args.no_cpu_baseline:
        base, parity = cpu_leg(x0_host, x.host(), args.warmup + args.steps)
    for arr in (A.row_ptr, A.col, A.val, x, y, snap):
        arr.free()
    sur = None if args.no_surrogate else surrogate_leg(ctx)

    out = {
        "metric": METRIC,
        "value": cell_steps,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded alpha field and Gaussian initial condition)",
        "config": {"workload": "heat 256^3 FE CSR step (7-point cd Laplacian, alpha~U[0.005,0.02), "
                               "alpha*dt/dx^2<=0.15, zero Dirichlet faces)",
                   "grid": list(DIMS), "rows": n_rows, "nnz": nnz, "kernel_variant": vname},
        "hbm_gbs": gbs,
        "interior_cell_steps_per_s": interior * args.steps / (ms / 1e3),
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": load_traffic(vname),
                     "algorithmic_bytes_per_launch": byts, "measured_copy_gbs": copy_gbs},
        "cpu_baseline": base,
        "parity": parity,
        "fp64_rel_l2": None if parity is None else parity["fp64_rel_l2"],
        "surrogate": sur,
        "field_finite": finite,
        "host_wall_s": t_host,
    }
    print(json.dumps(out))
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
