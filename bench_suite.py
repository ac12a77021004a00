"""Benchmarks of the other workloads of the path (one JSON line each).

  python bench_suite.py lu          # 4096^2 backward-Euler system: FP32-MFMA LU
                                    #   + FP64 refinement vs FP64 LU (torch/rocSOLVER,
                                    #   timed in a separate process)
  python bench_suite.py batch       # 512 uneven 2-D instances, 200 steps, hipGraph
  python bench_suite.py advection   # 256^3 upwind advection via geam3, 1000 steps
  python bench_suite.py surrogate   # train on solver trajectories, 128^3 run vs plain
  python bench_suite.py jli         # device JLI backward Euler

bench.py remains the headline (256^3 heat) measurement.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from piro_amd.core import CD, UW, Batch, Context, DevCSR, Surrogate  # noqa: E402

HBM = 8000.0
MFMA_F32 = 157.3e12
MFMA_BF16 = 2.5e15


def timed(ctx, fn, reps=1):
    """seconds per call, HIP events on the library's stream"""
    ctx.sync()
    ctx.timer_start()
    for _ in range(reps):
        fn()
    return ctx.timer_stop() / reps / 1e3


# ----------------------------------------------------------------------- LU

def bump_rhs(n_side, nrhs, seed=2026):
    rng = np.random.default_rng(seed)
    h = 1.0 / (n_side + 1)
    xs = (np.arange(n_side) + 1) * h
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    out = np.zeros((n_side * n_side, nrhs))
    for c in range(nrhs):
        f = np.zeros_like(X)
        for _ in range(8):
            cx, cy = rng.uniform(0, 1, 2)
            w = rng.uniform(0.05, 0.2)
            f += np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * w * w))
        out[:, c] = f.ravel()
    return out


FP64_SCRIPT = r"""
import json, sys, numpy as np, torch
a = torch.from_numpy(np.load(sys.argv[1])).cuda().double()
b = torch.from_numpy(np.load(sys.argv[2])).cuda().double()
def run():
    LU, piv = torch.linalg.lu_factor(a)
    return torch.linalg.lu_solve(LU, piv, b)
run(); torch.cuda.synchronize()
e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
e0.record()
for _ in range(3):
    x = run()
e1.record(); torch.cuda.synchronize()
rel = (torch.linalg.vector_norm(b - a @ x, dim=0) / torch.linalg.vector_norm(b, dim=0)).max().item()
t64 = e0.elapsed_time(e1) / 3 / 1e3
# the vendor FP32 factorisation (rocSOLVER sgetrf) of the same matrix, for
# the factor-only comparison
a32 = a.float()
torch.linalg.lu_factor(a32); torch.cuda.synchronize()
e0.record()
for _ in range(5):
    torch.linalg.lu_factor(a32)
e1.record(); torch.cuda.synchronize()
print(json.dumps({"seconds": t64, "rel": rel, "f32_factor_seconds": e0.elapsed_time(e1) / 5 / 1e3}))
"""


def bench_lu(args):
    ctx = Context(0)
    side, alpha, r = 64, 0.05, 2.0
    h = 1.0 / (side + 1)
    inv_h2 = np.float32(1.0 / (h * h))
    L = ctx.laplacian2d_csr(side, inv_h2)
    N = side * side
    I = DevCSR(ctx.from_host(np.arange(N + 1, dtype=np.int32)), ctx.from_host(np.arange(N, dtype=np.int32)),
               ctx.full(N, 1.0))
    dt = np.float32(r * h * h / alpha)
    A = ctx.geam2(1.0, I, np.float32(-(dt * np.float32(alpha))), L)
    dense = ctx.csr_to_dense(A)
    out = []
    for nrhs in (1, 32):
        bh = bump_rhs(side, nrhs)
        b = ctx.from_host(bh.astype(np.float64))
        lu = dense.clone()
        # the copy alone first: the factorisation below must be the last
        # thing written to lu before the solves use it
        t_copy = timed(ctx, lambda: lu.copy_(dense))
        ipiv = ctx.lu_factor(lu, N)  # warm-up
        t_fac = timed(ctx, lambda: (lu.copy_(dense), ctx.lu_factor(lu, N)), reps=5)
        t_fac -= t_copy
        res = {}

        def solve():
            res["x"] = ctx.lu_solve_refine(dense, lu, N, ipiv, b, tol=1e-10, max_sweeps=10)

        solve()
        t_sol = timed(ctx, solve)
        x, sweeps, rel = res["x"]
        f64 = {"seconds": None, "rel": None, "f32_factor_seconds": None}
        if not args.no_fp64:
            with tempfile.TemporaryDirectory() as td:
                fa, fb = os.path.join(td, "a.npy"), os.path.join(td, "b.npy")
                np.save(fa, dense.host().reshape(N, N))
                np.save(fb, bh)
                r64 = subprocess.run([sys.executable, "-c", FP64_SCRIPT, fa, fb], capture_output=True, text=True)
            if r64.returncode == 0:
                f64 = json.loads(r64.stdout.strip().splitlines()[-1])
        flop = 2.0 / 3.0 * N ** 3
        tot = t_fac + t_sol
        out.append({"metric": "backward-Euler LU time-to-1e-10 (4096^2)", "nrhs": nrhs,
                    "mixed_factor_s": t_fac, "mixed_solve_refine_s": t_sol, "mixed_total_s": tot,
                    "refinement_sweeps": sweeps, "mixed_max_rel_residual": max(rel),
                    "fp64_torch_total_s": f64["seconds"], "fp64_max_rel_residual": f64["rel"],
                    "speedup_vs_fp64": (f64["seconds"] / tot) if f64["seconds"] else None,
                    "torch_f32_factor_s": f64.get("f32_factor_seconds"),
                    "factor_speedup_vs_torch_f32": (f64["f32_factor_seconds"] / t_fac)
                    if f64.get("f32_factor_seconds") else None,
                    "factor_tflops": flop / t_fac / 1e12, "factor_frac_of_fp32_mfma_peak": flop / t_fac / MFMA_F32})
    for o in out:
        print(json.dumps(o), flush=True)
    ctx.close()


# -------------------------------------------------------------------- batch

def bench_batch(args):
    ctx = Context(0)
    rng = np.random.default_rng(4)
    n_inst = 512
    sides = np.exp(rng.uniform(np.log(32), np.log(256), n_inst)).astype(int)
    alphas = np.exp(rng.uniform(np.log(1e-3), np.log(1e-1), n_inst)).astype(np.float32)
    dims = [(int(s), int(s), 1) for s in sides]
    a_parts, dts, x_parts = [], [], []
    for (nx, ny, _), a in zip(dims, alphas):
        a_parts.append(np.full(nx * ny, a, np.float32))
        hh = np.float32(1.0) / np.float32(nx - 2)
        dts.append(np.float32(np.float32(0.2) * np.float32(float(hh) ** 2) / a))
        X, Y = np.meshgrid(np.linspace(0, 1, nx), np.linspace(0, 1, ny), indexing="xy")
        f = np.zeros_like(X)
        for _ in range(int(rng.integers(4, 17))):
            c = rng.uniform(0.1, 0.9, 2)
            w = rng.uniform(0.03, 0.2)
            f += np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2) / (2 * w * w))
        x_parts.append(f.astype(np.float32).ravel())
    alpha_h = np.concatenate(a_parts)
    x0h = np.concatenate(x_parts)
    alpha = ctx.from_host(alpha_h)
    b = Batch(ctx, dims, alpha, dts)
    x = ctx.from_host(x0h)
    w = ctx.empty_like(x)
    rb, g, tune_ms = b.autotune(x.clone(), w)
    steps = args.steps or 200
    x.upload(x0h)
    b.run(x, w, steps, use_graph=True)  # capture + warm
    x.upload(x0h)
    t = timed(ctx, lambda: b.run(x, w, steps, use_graph=True))
    byts = 8 * b.nnz + 12 * b.total_rows
    got = x.host()
    ok = True
    num = den = 0.0
    for i in rng.choice(n_inst, 8, replace=False):
        s, e = b.offsets[i], b.offsets[i + 1]
        single = Batch(ctx, [dims[i]], ctx.from_host(alpha_h[s:e]), [dts[i]])
        xs = ctx.from_host(x0h[s:e])
        ws = ctx.empty_like(xs)
        single.run(xs, ws, steps, use_graph=False)
        ok = ok and (xs.host().tobytes() == got[s:e].tobytes())
        single.close()
        # the same instance in FP64 (operator entries as stored, FP64 sums)
        r64 = heat2d_f64(dims[i], alpha_h[s:e], float(dts[i]), x0h[s:e], steps)
        num += float(np.sum((got[s:e].astype(np.float64) - r64) ** 2))
        den += float(np.sum(r64 ** 2))
    print(json.dumps({"metric": "batched 2-D heat, 512 uneven instances, 200 steps under hipGraph",
                      "rows": b.total_rows, "nnz": b.nnz, "seconds": t, "us_per_step": t / steps * 1e6,
                      "cell_steps_per_s": b.total_rows * steps / t, "hbm_gbs": byts * steps / t / 1e9,
                      "roofline_frac": byts * steps / t / 1e9 / HBM, "tiling_rows_per_unit": rb,
                      "tiling_grid": g, "schedule_chunk": b.schedule()[2], "xruns_in_lds": b.xruns(),
                      "bit_exact_vs_single_instance": ok,
                      "fp64_rel_l2_8_instances": (num / den) ** 0.5}),
          flush=True)
    b.close()
    ctx.close()


def heat2d_f64(dims, alpha_rows, dt, x0, steps):
    """FP64 twin of one 2-D batch instance: the FP32 operator entries (the
    build's 5-point Laplacian, row-scaled), sums and updates in FP64"""
    import scipy.sparse as sp
    nx, ny, _ = dims
    h = np.float32(1.0) / np.float32(nx - 2)
    norm = float(np.float32(float(h) ** 2))
    vs, vn = np.float32(-4.0 / norm), np.float32(1.0 / norm)
    N = nx * ny
    i = np.arange(N)
    xx, yy = i % nx, i // nx
    rows, cols, vals = [i], [i], [np.full(N, vs, np.float32)]
    for cond, off in ((xx > 0, -1), (xx < nx - 1, 1), (yy > 0, -nx), (yy < ny - 1, nx)):
        rows.append(i[cond]); cols.append(i[cond] + off); vals.append(np.full(cond.sum(), vn, np.float32))
    r, c, v = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    v = (v * alpha_rows[r]).astype(np.float32).astype(np.float64)
    A = sp.csr_matrix((v, (r, c)), shape=(N, N))
    inter = ((xx > 0) & (xx < nx - 1) & (yy > 0) & (yy < ny - 1)).astype(np.float64)
    x = x0.astype(np.float64)
    for _ in range(steps):
        x = (x + dt * (A @ x)) * inter
    return x


# ---------------------------------------------------------------- advection

def bench_advection(args):
    ctx = Context(0)
    side = args.side
    dims = (side, side, side)
    N = side ** 3
    g = ctx.grad_norm(1.0, side - 2)
    G = [ctx.gradient_csr(side, side, side, UW, d, g) for d in range(3)]
    # velocity in the reference's sign convention (dU/dt = +sum v_d G_d U with
    # backward differences): v = -(1, 0.5, 0.25) is upwind advection
    vel = (-1.0, -0.5, -0.25)
    v3 = ctx.from_host(np.concatenate([np.full(N, v, np.float32) for v in vel]))
    t0 = time.perf_counter()
    D = ctx.geam3(G[0], G[1], G[2], v3)
    ctx.sync()
    t_geam = time.perf_counter() - t0
    dx = 1.0 / (side - 2)
    dt = np.float32(0.5 * 2 * dx / sum(abs(v) for v in vel))
    x0 = ctx.ic_tophat(dims, [1.0] * 3, [0.5] * 3, 0.15)
    snap = ctx.apply_face_values(dims, x0, [0.0] * 6)
    if args.variant >= 0:  # pinned (profiling runs: counters perturb the tuner's timings)
        ctx.set_variant(args.variant)
        variant = args.variant
    else:
        variant, _ = ctx.autotune(dims, D, dt, x0, snap, ctx.empty_like(x0))
    x = x0.clone()
    w = ctx.empty_like(x)
    steps = args.steps or 1000
    t = timed(ctx, lambda: ctx.fe_run_csr(dims, D, dt, steps, x, w, snap))
    mass0 = ctx.reduce("sum", x0)
    mass1 = ctx.reduce("sum", x)
    byts = 8 * D.nnz + 12 * N
    parity = {}
    if not args.no_oracle:
        # the timed run's field (x0 advanced `steps` steps) against the CPU
        # oracle: the FP32 restatement (bit for bit) and its FP64 twin
        import oracle as po
        R = po.CSR(*D.host())
        x0h, snaph, got = x0.host(), snap.host(), x.host()
        ref32 = po.fe_csr_run(dims, R, dt, steps, x0h, snaph)
        ref64 = po.fe_csr_run_f64(dims, R, float(dt), steps, x0h, snaph)
        parity = {"bit_exact_vs_fp32_oracle": got.tobytes() == ref32.tobytes(),
                  "fp64_rel_l2": po.rel_l2(got, ref64)}
    print(json.dumps({"metric": f"advection {side}^3 upwind via geam3 concatenation, FE", "nnz": D.nnz, **parity,
                      "geam_seconds": t_geam, "steps": steps, "kernel_variant": ctx.variants()[variant],
                      "us_per_step": t / steps * 1e6,
                      "cell_steps_per_s": N * steps / t, "hbm_gbs": byts * steps / t / 1e9,
                      "roofline_frac": byts * steps / t / 1e9 / HBM, "mass_ratio": mass1 / mass0,
                      "max_after": ctx.reduce("absmax", x)}), flush=True)
    ctx.close()


# ---------------------------------------------------------------- surrogate

# per cell of one inference call: the composed stencils' useful flops
# (3 outputs x 13^3 taps x multiply-add) and the MFMA flops issued for them
# (169 tap rows x 9 products of 16x16x32 per 256 cells: Toeplitz padding and
# the three-product BF16 split included)
SUR_USEFUL_FLOP = 2 * 3 * 13 ** 3
SUR_MFMA_FLOP = 169 * 9 * 2 * 16 * 16 * 32 // 256

def bench_surrogate(args):
    from piro_amd import learned
    ctx = Context(0)
    sur = Surrogate(ctx, seed=5)
    log = learned.train(ctx, sur, n_traj=args.traj, per_batch=50, side=args.train_side, segments=4, k=16,
                        passes=args.passes, lr=1e-3, lr_final=1e-5, blob_widths=(4.0, 20.0), noise=0.03,
                        log_every=200)
    run = learned.HeatRun(ctx, side=128)
    steps = 1600
    run.plain(160)  # warm-up
    ref, t_plain = run.plain(steps)
    run.accelerated(sur, 160, correct=False)  # warm-up
    acc0, res0 = run.accelerated(sur, steps, correct=False)
    err0 = ctx.rel_l2(acc0, ref)
    run.accelerated(sur, 160)  # warm-up
    acc, res = run.accelerated(sur, steps)
    err = ctx.rel_l2(acc, ref)
    amp = learned.max_amplification(sur.stencil())
    shape = (1, 128, 128, 128)
    y = ctx.empty_like(run.u0)
    t_call = timed(ctx, lambda: sur.forward(shape, run.u0, run.alpha, 1.0, out=y), reps=10)
    cells = 128 ** 3
    print(json.dumps({"metric": "surrogate-accelerated 128^3 heat, 1600 steps (k=16, check every 10th call)",
                      "train_trajectories": log.trajectories, "train_seconds": log.seconds,
                      "train_loss_first": log.losses[0], "train_loss_last": log.losses[-1],
                      "plain_seconds": t_plain, "accelerated_seconds": res.seconds,
                      "speedup": t_plain / res.seconds, "rel_l2_vs_plain": err,
                      "checks": res.checks, "rejected": res.rejected, "max_amplification": amp,
                      "uncorrected": {"speedup": t_plain / res0.seconds, "rel_l2_vs_plain": err0,
                                      "rejected": res0.rejected,
                                      "check_errors": [float(f"{e:.3e}") for e in res0.check_errors]},
                      "check_errors": [float(f"{e:.3e}") for e in res.check_errors],
                      "surrogate_call_ms": t_call * 1e3,
                      "surrogate_useful_tflops": SUR_USEFUL_FLOP * cells / t_call / 1e12,
                      "surrogate_mfma_tflops": SUR_MFMA_FLOP * cells / t_call / 1e12,
                      "surrogate_frac_bf16_peak": SUR_MFMA_FLOP * cells / t_call / MFMA_BF16,
                      "plain_16_steps_ms": t_plain / 100 * 1e3}), flush=True)
    sur.close()
    ctx.close()


def bench_surrogate_fwd(args):
    """Forward calls only, on the 128^3 run's field (timing does not depend on
    the weights, so the untrained network is used)."""
    from piro_amd import learned
    ctx = Context(0)
    sur = Surrogate(ctx, seed=5)
    run = learned.HeatRun(ctx, side=128)
    shape = (1, 128, 128, 128)
    y = ctx.empty_like(run.u0)
    sur.forward(shape, run.u0, run.alpha, 1.0, out=y)  # warm-up (composes the stencils)
    t_call = timed(ctx, lambda: sur.forward(shape, run.u0, run.alpha, 1.0, out=y), reps=10)
    cells = 128 ** 3
    print(json.dumps({"metric": "surrogate forward, 128^3 (one k=16 call)", "ms_per_call": t_call * 1e3,
                      "useful_tflops": SUR_USEFUL_FLOP * cells / t_call / 1e12,
                      "mfma_tflops": SUR_MFMA_FLOP * cells / t_call / 1e12,
                      "frac_bf16_peak": SUR_MFMA_FLOP * cells / t_call / MFMA_BF16}), flush=True)
    sur.close()
    ctx.close()


# ------------------------------------------------------------------ formats

def bench_formats(args):
    """Assembly, the COO/HT/DENSE all-rows steps, the divergence operator and
    the reductions: device time per call (HIP events on the library's
    stream) against each operation's algorithmic bytes."""
    from piro_amd.core import _p, check, lib
    ctx = Context(0)
    rows = []

    def row(name, seconds, nbytes, note=""):
        r = {"op": name, "us": seconds * 1e6}
        if nbytes:
            r.update({"algorithmic_bytes": nbytes, "gbs": nbytes / seconds / 1e9,
                      "roofline_frac": nbytes / seconds / 1e9 / HBM})
        if note:
            r["note"] = note
        rows.append(r)
        print(json.dumps(r), flush=True)

    side = 256
    dims = (side,) * 3
    N = side ** 3
    norm = ctx.lap_norm(1.0, side - 2)
    nnz = ctx.lap_nnz(*dims)
    A = ctx.laplacian_csr(*dims, CD, norm)
    t = timed(ctx, lambda: check(lib.piro_assemble_laplacian_csr(ctx.h, *dims, CD, float(norm), _p(A.row_ptr),
                                                                 _p(A.col), _p(A.val))), reps=5)
    row("laplacian CSR assembly 256^3 (cd)", t, 4 * (N + 1) + 8 * nnz, "writes only")
    alpha = ctx.alpha_uniform(N, 7, 0.005, 0.02)
    t = timed(ctx, lambda: ctx.row_scale(A, alpha), reps=5)
    row("alpha row scaling 256^3", t, 4 * (N + 1) + 4 * N + 8 * nnz, "row_ptr, alpha and val read, val written")
    g = ctx.grad_norm(1.0, side - 2)
    G0 = ctx.gradient_csr(*dims, UW, 0, g)
    t = timed(ctx, lambda: check(lib.piro_assemble_gradient_csr(ctx.h, *dims, UW, 0, float(g), _p(G0.row_ptr),
                                                                _p(G0.col), _p(G0.val))), reps=5)
    row("gradient CSR assembly 256^3 (uw, x)", t, 4 * (N + 1) + 8 * G0.nnz, "writes only")
    u3 = ctx.from_host(np.full(3 * N, 0.5, np.float32))
    ctx.sync()
    t0 = time.perf_counter()
    Dv = ctx.divergence_csr(*dims, CD, g, u3)
    ctx.sync()
    row("divergence operator 256^3 (3 gradients, row scaling, merge), host wall", time.perf_counter() - t0, 0,
        f"nnz {Dv.nnz}; includes allocation of the intermediates")
    for arr in (Dv.row_ptr, Dv.col, Dv.val, G0.row_ptr, G0.col, G0.val, u3):
        arr.free()

    # COO all-rows FE step (row_ptr recovered from the sorted row indices)
    rws, col, val = ctx.laplacian_coo(*dims, CD, norm)
    t = timed(ctx, lambda: check(lib.piro_coo_row_ptr(ctx.h, N, nnz, _p(rws), _p(A.row_ptr))), reps=5)
    row("COO row_ptr recovery 256^3", t, 4 * nnz + 4 * (N + 1))
    Ac = ctx.coo_to_csr(N, rws, col, val)
    x = ctx.ic_gaussian(dims, [1.0] * 3, [0.5] * 3, [0.1] * 3, 1.0)
    y = ctx.empty_like(x)
    dt = np.float32(0.1 * float(norm) / 0.02)
    ctx.fe_step_allrows(Ac, dt, x, y)
    t = timed(ctx, lambda: ctx.fe_step_allrows(Ac, dt, x, y), reps=20)
    row("COO all-rows FE step 256^3", t, 8 * nnz + 12 * N + 4, "SpMVCOO + addVectors, fused")
    for arr in (rws, col, val, Ac.row_ptr, A.row_ptr, A.col, A.val, x, y, alpha):
        arr.free()

    # HT near the largest grid the shipped table-size fit covers (N ~ 18652)
    hs = 26
    hd = (hs,) * 3
    Nh = hs ** 3
    hn = ctx.lap_norm(1.0, hs - 2)
    Ah = ctx.laplacian_csr(*hd, CD, hn)
    ts = ctx.ht_table_size(Ah.nnz, Nh)
    keys, vals = ctx.ht_build(Ah, ts)
    t = timed(ctx, lambda: check(lib.piro_ht_build(ctx.h, Nh, _p(Ah.row_ptr), _p(Ah.col), _p(Ah.val), int(ts),
                                                   _p(keys), _p(vals))), reps=5)
    row(f"HT build 26^3 (table {ts} slots)", t, 8 * ts, "clearing the table (keys -1, values 0) is most of it; the parallel insert ~70 us")
    Ahc = ctx.ht_to_csr(Nh, keys, vals)
    xh = ctx.ic_gaussian(hd, [1.0] * 3, [0.5] * 3, [0.1] * 3, 1.0)
    yh = ctx.empty_like(xh)
    t = timed(ctx, lambda: ctx.fe_step_allrows(Ahc, dt, xh, yh), reps=50)
    row("HT all-rows FE step 26^3 (unpacked to CSR)", t, 8 * Ahc.nnz + 12 * Nh + 4, "launch-bound at this size")

    # DENSE: the N x N matrix streams once per step
    ds = 16
    dd = (ds,) * 3
    Nd = ds ** 3
    dn = ctx.lap_norm(1.0, ds - 2)
    Ad = ctx.laplacian_dense(*dd, 3, dn)
    t = timed(ctx, lambda: check(lib.piro_assemble_laplacian_dense(ctx.h, *dd, 3, float(dn), _p(Ad), Nd)), reps=5)
    row("DENSE Laplacian assembly 16^3 (4096 x 4096)", t, 4 * Nd * Nd, "clear + entries, writes only")
    xd = ctx.ic_gaussian(dd, [1.0] * 3, [0.5] * 3, [0.1] * 3, 1.0)
    yd = ctx.empty(Nd)
    t = timed(ctx, lambda: ctx.fe_step_dense(Ad, Nd, dt, xd, yd), reps=20)
    row("DENSE all-rows FE step 16^3", t, 4 * Nd * Nd + 12 * Nd, "SpMVDENSE + addVectors")

    # reductions on 64 Mi floats (FP64 accumulation); each call ends in an
    # 8-byte read-back, included in its device time
    n = 1 << 26
    a = ctx.alpha_uniform(n, 1, -1, 1)
    b = ctx.alpha_uniform(n, 2, -1, 1)
    for op, nb in (("sum", 4 * n), ("sumsq", 4 * n), ("absmax", 4 * n), ("dot", 8 * n), ("diffsq", 8 * n)):
        ctx.reduce(op, a, b)
        t = sum(timed(ctx, lambda: ctx.reduce(op, a, b)) for _ in range(5)) / 5
        row(f"reduce {op} 64Mi floats", t, nb)
    a.free()
    b.free()

    # 2-D operators on a 4096^2 grid: the Laplacian and the true merge
    # I - dt*alpha*L (count, scan, fill), host wall with the outputs'
    # allocation
    s2 = 4096
    n2 = s2 * s2
    h2 = 1.0 / (s2 + 1)
    ctx.sync()
    t0 = time.perf_counter()
    L2 = ctx.laplacian2d_csr(s2, np.float32(1.0 / (h2 * h2)))
    ctx.sync()
    row("2-D Laplacian CSR 4096^2, host wall", time.perf_counter() - t0, 0,
        f"nnz {L2.nnz}; includes allocation")
    I2 = DevCSR(ctx.from_host(np.arange(n2 + 1, dtype=np.int32)), ctx.from_host(np.arange(n2, dtype=np.int32)),
                ctx.full(n2, 1.0))
    ctx.sync()
    t0 = time.perf_counter()
    M2 = ctx.geam2(1.0, I2, np.float32(-0.1), L2)
    ctx.sync()
    row("geam2 merge I - dt*alpha*L 4096^2, host wall", time.perf_counter() - t0, 0,
        f"nnz {M2.nnz}; count, scan and fill, includes allocation")
    for m in (L2, I2, M2):
        for arr in (m.row_ptr, m.col, m.val):
            arr.free()
    print(json.dumps({"metric": "formats, assembly and reductions", "rows": len(rows)}), flush=True)
    ctx.close()

    # host paths: the case driver on PIRO's own HeatEq case and the Paraview
    # writer on a 128^3 scalar field
    import ctypes as C
    from piro_amd.core import case_run
    cases = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "cases")
    with tempfile.TemporaryDirectory() as td:
        case_run(cases, "HeatEq.ini", td, -1)
        t0 = time.perf_counter()
        steps, _ = case_run(cases, "HeatEq.ini", td, -1)
        row(f"case driver, PIRO's HeatEq case ({steps} steps, 10x10x31 cells, with its Paraview exports), "
            "host wall", time.perf_counter() - t0, 0, "context, assembly, stepping and export")
        pd = 130
        field = np.random.default_rng(3).random(pd ** 3, dtype=np.float32)
        names = (C.c_char_p * 1)(b"T")
        types = (C.c_int * 1)(1)
        ptrs = (C.c_void_p * 1)(field.ctypes.data)
        dims = (C.c_int * 3)(pd, pd, pd)
        t0 = time.perf_counter()
        check(lib.piro_paraview_write(td.encode(), b"bench", 1, dims, None, None, 1, names, types, ptrs))
        tw = time.perf_counter() - t0
        nbytes = sum(os.path.getsize(os.path.join(r, f)) for r, _, fs in os.walk(os.path.join(td, "bench"))
                     for f in fs)
        row("Paraview VTH/VTI writer, 128^3 interior scalar field, host wall", tw, 0,
            f"{nbytes / 1e6:.1f} MB of text, {nbytes / tw / 1e6:.0f} MB/s")


# ---------------------------------------------------------------------- JLI

def bench_jli(args):
    ctx = Context(0)
    side = args.side
    dims = (side, side, side)
    norm = ctx.lap_norm(1.0, side - 2)
    A = ctx.laplacian_csr(side, side, side, 2, norm)
    dt = np.float32(0.05 * norm)
    b = ctx.ic_gaussian(dims, [1.0] * 3, [0.5] * 3, [0.1] * 3, 1.0)
    ctx.jli_solve(dims, A, dt, b, 1.0, 1e-12, 50)
    out = {}

    def run():
        out["r"] = ctx.jli_solve(dims, A, dt, b, 1.0, 1e-12, 200)

    t = timed(ctx, run)
    x, it, res = out["r"]
    # steady state: a tolerance no residual change meets runs exactly 64 iterations
    fixed = 64

    def run_fixed():
        out["f"] = ctx.jli_solve(dims, A, dt, b, 1.0, -1.0, fixed)

    run_fixed()
    tf = timed(ctx, run_fixed)
    assert out["f"][1] == fixed
    N = side ** 3
    byts = 8 * A.nnz + 16 * N + 4   # col+val, row_ptr, x and b read, x' written
    per_it = tf / fixed
    print(json.dumps({"metric": f"device JLI backward Euler {side}^3", "iterations_to_converge": it,
                      "residual": res, "seconds_to_converge": t, "fixed_iterations": fixed,
                      "us_per_iteration": per_it * 1e6, "hbm_gbs": byts / per_it / 1e9,
                      "roofline_frac": byts / per_it / 1e9 / HBM,
                      "algorithmic_bytes_per_iteration": byts,
                      "kernel_variant": "untuned default (pfx_rb256_c64)"}), flush=True)
    ctx.close()


def calibrate_copy(n_floats=64 << 20, reps=10):
    ctx = Context(0)
    src = ctx.full(n_floats, 1.0)
    dst = ctx.empty_like(src)
    t = timed(ctx, lambda: ctx.copy(src, dst), reps)
    print(json.dumps({"calibration": "k_copy4 (piro_vec_copy)", "bytes_read": 4 * n_floats,
                      "bytes_written": 4 * n_floats, "dispatches": reps, "gbs": 8 * n_floats / t / 1e9}), flush=True)
    src.free()
    dst.free()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["lu", "batch", "advection", "surrogate", "surrogate_fwd", "jli", "formats"])
    ap.add_argument("--side", type=int, default=256)
    ap.add_argument("--steps", type=int, default=0)
    ap.add_argument("--traj", type=int, default=32000)
    ap.add_argument("--train-side", type=int, default=32)
    ap.add_argument("--passes", type=int, default=3)
    ap.add_argument("--no-oracle", action="store_true", help="skip the CPU oracle checks of the advection run")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 LU comparison (profiling runs)")
    ap.add_argument("--variant", type=int, default=-1,
                    help="advection: force an FE kernel variant (-1 = autotune)")
    ap.add_argument("--calibrate", action="store_true",
                    help="first time the library's streaming copy of a 256 MiB vector 10 times (the known-size "
                         "kernel that HBM counter readings are calibrated against)")
    args = ap.parse_args()
    if args.calibrate:
        calibrate_copy()
    {"lu": bench_lu, "batch": bench_batch, "advection": bench_advection, "surrogate": bench_surrogate,
     "surrogate_fwd": bench_surrogate_fwd, "jli": bench_jli, "formats": bench_formats}[args.what](args)


if __name__ == "__main__":
    main()
