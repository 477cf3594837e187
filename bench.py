"""Benchmark: RHS solves/s at 128^3 to a 1e-6 residual, mixed batched form vs
the plain FP64 form, one adjoint phase gradient, the stencil's HBM roofline
and a CPU baseline.

A "step" is one batched solve of `--chunk` columns of the phase response
basis G (transducers r .. r+chunk-1 at unit amplitude) on the heterogeneous
128^3 scene (c = 343 (1 + 0.05 g), rigid sphere, PML on -x, +-y, +z,
16x16 array).  A column counts as solved when it passes the reference's
acceptance test (scaled residual <= 1e-6, backward error <= 1e-4) and its
FP64 relative residual ||b - A x||/||b|| is <= 1e-6.  Inputs (coefficients,
source columns) are resident on the device before the timed region.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


ROOFLINE_SHAPE = 16  # paired tiles, 4 z-planes per lane, 4 rows, non-temporal stores: k_slab2<MODE, 0, 4, 4, 1>
ROOFLINE_RHS = 32


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=128)
    p.add_argument("--chunk", type=int, default=0, help="RHS per step; 0 = autotuned")
    p.add_argument("--restart", type=int, default=30, help="plain FP64 form (the reference's GMRES restart)")
    p.add_argument("--mixed-restart", type=int, default=10,
                   help="restart length of the mixed form's inner complex64 GMRES")
    p.add_argument("--maxiter", type=int, default=60000)
    p.add_argument("--plain-maxiter", type=int, default=300000,
                   help="inner-step cap of the plain form (Jacobi needs far more steps than multigrid)")
    p.add_argument("--rtol", type=float, default=1e-6)
    p.add_argument("--plain-rhs", type=int, default=1, help="columns timed in the plain FP64 form")
    p.add_argument("--cpu-iters", type=int, default=60, help="scipy GMRES iterations in the CPU sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--stencil-reps", type=int, default=20)
    p.add_argument("--stencil-rhs", type=int, default=0,
                   help="columns per roofline stencil launch (0: 32, whatever chunk the solve uses)")
    p.add_argument("--stencil-shape", type=int, default=-1,
                   help="roofline-only: the tile shape (index) of the launches (default: the solve's paired shape)")
    p.add_argument("--mg-shift", type=float, default=1.2)
    p.add_argument("--mg-pre", type=int, default=1)
    p.add_argument("--mg-post", type=int, default=1, help="post-sweeps on the finest multigrid level")
    p.add_argument("--mg-post-coarse", type=int, default=5, help="post-sweeps on the coarser levels")
    p.add_argument("--mg-coarse-sweeps", type=int, default=30, help="sweeps on the coarsest level")
    p.add_argument("--mg-levels", type=int, default=0, help="multigrid levels (0: the library's default, down to the first level of <= 1024 nodes)")
    p.add_argument("--mg-weight", type=float, default=0.7)
    p.add_argument("--mg-cycle", type=int, default=1, choices=[1, 2], help="1: V-cycle, 2: W below the finest level")
    p.add_argument("--graph", type=int, default=0, help="replay each GMRES cycle from a captured hipGraph")
    p.add_argument("--arnoldi-block", type=int, default=2, choices=[1, 2],
                   help="mixed form: basis vectors per Gram-Schmidt pass")
    p.add_argument("--workload", choices=["g128", "uneven", "large", "surface", "online"], default="g128",
                   help="g128: G columns of the heterogeneous 128^3 scene (the headline); uneven: the "
                        "32-scene batch; large: the 256^3 two-layer scene; surface: homogeneous target responses; "
                        "online: PhaseObjective evaluations on a device basis of the archived 41^3 x 144 shape")
    p.add_argument("--scenes", type=int, default=32, help="uneven: scenes in the batch")
    p.add_argument("--scene-rhs", type=int, default=16, help="uneven: right-hand sides per scene")
    p.add_argument("--workers", type=int, default=4, help="uneven: concurrent scene streams")
    p.add_argument("--uneven-legs", choices=["all", "fused"], default="all",
                   help="uneven: all legs (scene by scene, streams, fused) or the fused solve alone")
    p.add_argument("--fused-graph", type=int, default=1,
                   help="uneven: capture each lockstep cycle body of the fused solve as a graph")
    p.add_argument("--precond", choices=["multigrid", "jacobi"], default="multigrid",
                   help="inner preconditioner of the mixed form (the plain form always uses Jacobi)")
    p.add_argument("--jacobi-rhs", type=int, default=16,
                   help="columns of the mixed form with the Jacobi preconditioner (GMRES(30)); 0 skips it")
    p.add_argument("--no-gradient", action="store_true", help="skip the timed adjoint phase gradient")
    p.add_argument("--roofline-only", action="store_true",
                   help="only the roofline launches: the solve's Krylov apply over 32 columns with the tile "
                        "shape pinned, no autotune and no solve (for profiler runs)")
    p.add_argument("--transfers-only", action="store_true",
                   help="only the finest-level multigrid transfers (restriction and prolongation of 64 "
                        "columns, hz_mg_transfer), timed with events, then the counter-calibration copy "
                        "(for profiler runs)")
    p.add_argument("--cpu-samples", type=int, default=3, help="timed CPU GMRES samples per thread setting")
    p.add_argument("--evaluations", type=int, default=2000, help="online: timed evaluations per loss")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards by "
              "right-hand side across GPUs, with no collectives in the solve loop).", file=sys.stderr)
        sys.exit(2)
    import torch
    from ultrasound_solver_amd import scenes
    from ultrasound_solver_amd.operator import HelmholtzOperator

    if args.workload == "uneven":
        return bench_uneven(args)
    if args.workload == "large":
        return bench_large(args)
    if args.workload == "surface":
        return bench_surface(args)
    if args.workload == "online":
        return bench_online(args, args.evaluations)
    t0 = time.perf_counter()
    scene = scenes.config_heterogeneous(n=args.n)
    op = HelmholtzOperator(scene.rho, scene.c, scene.dx, scene.omega, scene.boundaries, scene.pml,
                           device="cuda:0")
    node, col, val = scenes.source_triplets(scene)
    if np.any(op.pml_mask[:, :, 0].ravel(order="F")[node]):
        raise ValueError("PML overlaps the deposited sinc aperture")
    op.set_sources(node, col, val, scene.num_sources)
    if args.roofline_only:
        roofline_only(op, args)
        return
    if args.transfers_only:
        transfers_only(op, args)
        return
    if args.precond == "multigrid":
        op.set_multigrid(**_mg_args(args))
    m = scene.num_sources
    tuning = op.autotune(m, max_chunk=args.chunk or 64, restart=args.mixed_restart, preconditioner=args.precond)
    if not args.chunk:
        args.chunk = tuning["chunk"]
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    log(f"setup {setup_s:.1f}s n={args.n} N={op.n} M={m} ({'; '.join(scene.notes)}); autotune {tuning}")

    def rhs_block(start, k):
        w = torch.zeros((k, m), dtype=torch.complex128, device="cuda:0")
        idx = torch.arange(k, device="cuda:0")
        w[idx, (start + idx) % m] = 1.0
        return op.build_rhs(w)

    common = dict(restart=args.restart, maxiter=args.maxiter, rtol=args.rtol, backward_error_tol=1e-4,
                  relative_rtol=args.rtol)
    mixed = dict(common, restart=args.mixed_restart, precision="c64", chunk=args.chunk, preconditioner=args.precond,
                 use_graph=bool(args.graph), arnoldi_block=args.arnoldi_block)
    # G is kept as its rows on a 9 x 9 x 9 node patch between the array and
    # the obstacle (trilinear samples at the nodes), the form the
    # optimisers consume; each step samples its chunk of columns.
    prows, pw = patch_rows(scene.shape, (args.n//2, args.n//2, int(0.4*args.n)), 9)
    g_patch = torch.zeros((m, len(prows)), dtype=torch.complex128, device="cuda:0")
    for s in range(args.warmup):
        b = rhs_block(s*args.chunk, args.chunk)
        _, st = op.gmres(b, **mixed)
        log(f"warmup step {s}: iterations {[x['iterations'] for x in st]}")
    torch.cuda.synchronize()
    stats = []
    t_start = time.perf_counter()
    for s in range(args.steps):
        start = (args.warmup + s)*args.chunk
        b = rhs_block(start, args.chunk)
        x, st = op.gmres(b, **mixed)
        cols = (start + torch.arange(args.chunk, device="cuda:0")) % m
        g_patch[cols] = op.sample(prows, pw, x)
        stats += st
        log(f"step {s}: converged {sum(x['converged'] for x in st)}/{len(st)}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    solved = sum(x["converged"] for x in stats)
    mixed_rate = solved/elapsed
    log(f"mixed: {solved} solves in {elapsed:.2f}s -> {mixed_rate:.3f}/s; median iterations "
        f"{np.median([x['iterations'] for x in stats])}")

    # Plain FP64 form (reference semantics, one right-hand side per launch).
    plain = []
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for r in range(args.plain_rhs):
        _, st = op.gmres(rhs_block(r, 1), precision="c128", chunk=1, use_graph=bool(args.graph),
                         **dict(common, maxiter=args.plain_maxiter))
        plain += st
    torch.cuda.synchronize()
    plain_elapsed = time.perf_counter() - tp
    plain_solved = sum(x["converged"] for x in plain)
    plain_rate = plain_solved/plain_elapsed if plain_elapsed > 0 else 0.0
    plain_iters = [x["iterations"] for x in plain]
    log(f"plain: {plain_solved} solves in {plain_elapsed:.2f}s -> {plain_rate:.3f}/s; iterations {plain_iters}")

    # The mixed form with the reference's Jacobi preconditioner: precision and
    # batching alone, without multigrid.
    jacobi = None
    if args.jacobi_rhs > 0:
        k = args.jacobi_rhs
        torch.cuda.synchronize()
        tj = time.perf_counter()
        _, st = op.gmres(rhs_block(0, k), **dict(common, precision="c64", chunk=k, preconditioner="jacobi",
                                                 maxiter=args.plain_maxiter))
        torch.cuda.synchronize()
        tj = time.perf_counter() - tj
        ok = sum(x["converged"] for x in st)
        jacobi = {"value": ok/tj, "unit": "solves/s", "rhs": k, "restart": args.restart,
                  "iterations_median": float(np.median([x["iterations"] for x in st])),
                  "worst_relative_residual": max(x["relative_residual"] for x in st)}
        log(f"mixed Jacobi: {ok}/{k} solves in {tj:.2f}s -> {ok/tj:.3f}/s")

    gradient = None if args.no_gradient else adjoint_gradient(op, scene, args)

    # Stencil roofline: the Krylov apply of the timed solve (R A with the
    # multigrid preconditioner, R A M with Jacobi) over 32 complex64 columns
    # with the tile shape pinned, so that the figure moves neither with the
    # chunk nor with the shapes the autotuner picks; timed with events on the
    # stream it is launched on.
    op.set_tile_shapes([ROOFLINE_SHAPE] + tuning["shapes"][1:])
    roof = roofline_launches(op, args)
    roof["solve_apply_shape"] = tuning["shapes"][0]
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scene, args, plain_iters)
    g_cols = m/mixed_rate if mixed_rate > 0 else None
    plain_col = plain_elapsed/max(plain_solved, 1)
    out = {
        "metric": "RHS solves/s at 128^3 to 1e-6 rel. residual (mixed vs FP64); stencil HBM GB/s",
        "value": mixed_rate, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed/args.steps*1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 (FP64 refinement)", "data": "synthetic",
        "config": {"workload": f"heterogeneous {args.n}^3 scene, G columns in chunks of {args.chunk}, "
                               f"{args.precond}-preconditioned GMRES({args.mixed_restart}) c64 + FP64 refinement to "
                               f"{args.rtol:g} (scaled and relative residual)",
                   "grid": [args.n]*3, "rhs_per_step": args.chunk, "sources": m,
                   "autotune": tuning},
        "plain_fp64": {"value": plain_rate, "unit": "solves/s", "rhs": args.plain_rhs,
                       "iterations": plain_iters,
                       "worst_relative_residual": max(x["relative_residual"] for x in plain) if plain else None},
        "speedup_mixed_vs_plain": mixed_rate/plain_rate if plain_rate > 0 else None,
        "mixed_jacobi": jacobi,
        "iterations_median": float(np.median([x["iterations"] for x in stats])) if stats else None,
        "refinements_median": float(np.median([x["refinements"] for x in stats])) if stats else None,
        "worst_scaled_residual": max(x["scaled_residual"] for x in stats) if stats else None,
        "worst_relative_residual": max(x["relative_residual"] for x in stats) if stats else None,
        "worst_backward_error": max(x["backward_error"] for x in stats) if stats else None,
        "g_patch": {"rows": len(prows), "columns_sampled": args.steps*args.chunk,
                    "what": "trilinear samples of each solved column at a 9x9x9 node patch centred at "
                            f"({args.n//2}, {args.n//2}, {int(0.4*args.n)}), inside the timed steps"},
        "adjoint_gradient": gradient,
        "g_plus_gradient": None if g_cols is None or gradient is None else {
            "mixed_s": g_cols + gradient["seconds"],
            "plain_s_estimate": (m + 2)*plain_col,
            "speedup": (m + 2)*plain_col/(g_cols + gradient["seconds"]),
            "note": f"G = {m} columns at the mixed rate; plain = {m} + 2 solves (forward and adjoint) at the "
                    "measured plain time per column"},
        "setup_s": setup_s,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def roofline_launches(op, args, shape=ROOFLINE_SHAPE):
    """`--stencil-reps` launches of the solve's Krylov apply over 32 columns,
    timed with events on the launching stream; a device copy of the same 32
    columns beside it."""
    import torch
    k = args.stencil_rhs or ROOFLINE_RHS
    mode = stencil_mode(args)
    x = torch.randn((k, op.n), dtype=torch.complex64, device="cuda:0")
    # The GPU's clocks ramp up under sustained load and drop within a second
    # of idling: the first launches after a pause run up to 12 % slower than
    # the steady state a solve works in.  Half a second of the same stencil
    # in its unscaled mode (another kernel instance, so a profiler's
    # statistics row for the timed kernel holds only the launches below)
    # brings the clocks up first.
    t_warm = time.perf_counter()
    while time.perf_counter() - t_warm < 0.5:
        op.apply(x, mode="A", precision="c64")
    torch.cuda.synchronize()
    for _ in range(3):
        op.apply(x, mode=mode, precision="c64")
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.stencil_reps):
        op.apply(x, mode=mode, precision="c64")
    e1.record(stream)
    torch.cuda.synchronize()
    stencil_s = e0.elapsed_time(e1)/1e3/args.stencil_reps
    alg_bytes = (16*k + 8)*op.n
    achieved = alg_bytes/stencil_s/1e9
    log(f"stencil k={k}: {stencil_s*1e6:.1f} us/launch, {achieved:.0f} GB/s algorithmic")
    # Reference stream: a device-to-device copy of the same k columns (read +
    # write 8 B per node and column), the rate a plain copy reaches here.
    dst = torch.empty_like(x)
    for _ in range(3):
        dst.copy_(x)
    e0.record(stream)
    for _ in range(args.stencil_reps):
        dst.copy_(x)
    e1.record(stream)
    torch.cuda.synchronize()
    copy_s = e0.elapsed_time(e1)/1e3/args.stencil_reps
    copy_gbps = 16*k*op.n/copy_s/1e9
    log(f"copy of {k} columns: {copy_s*1e6:.1f} us, {copy_gbps:.0f} GB/s")
    zb, ty, rp, nt, flat = op.SLAB_SHAPES[shape]
    return {"kernel": f"{slab_kernel_name(mode, zb, ty, rp, nt, flat)} (complex64 {' '.join(mode)} apply, {k} columns)",
            "bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s", "frac": achieved/8000.0,
            "traffic": measured_traffic(args, mode, k, shape), "alg_bytes_per_launch": alg_bytes,
            "launch_us": stencil_s*1e6, "launches": args.stencil_reps,
            "copy_reference": {"what": f"device copy of the same {k} complex64 columns",
                               "achieved": copy_gbps, "frac": copy_gbps/8000.0}}


def roofline_only(op, args):
    """The roofline launches alone (pinned shape, no autotune, no solve), so
    that a profiler run sees only these launches of the kernel; then a copy
    of known size (2 GiB each way, from a source evicted from the 256 MB
    last-level cache by a 1 GiB fill) that scales the hardware byte counters."""
    import torch
    shape = ROOFLINE_SHAPE if args.stencil_shape < 0 else args.stencil_shape
    op.set_tile_shapes([shape])
    roof = roofline_launches(op, args, shape)
    copied = calibration_copy()
    print(json.dumps({"roofline": roof, "shape": shape, "calibration_copy_bytes_each_way": copied}), flush=True)


def calibration_copy():
    """A device copy of 2 GiB each way from a source evicted from the 256 MB
    last-level cache by a 1 GiB fill: the known byte count that scales the
    hardware byte counters of a profiler run."""
    import torch
    src = torch.ones(512*1024*1024, dtype=torch.float32, device="cuda:0")
    flush = torch.empty(256*1024*1024, dtype=torch.float32, device="cuda:0")
    flush.fill_(2.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    return src.numel()*4


def transfers_only(op, args, k=64):
    """The finest-level multigrid transfers of the 128^3 scene alone:
    `--stencil-reps` restrictions P^T r (128^3 -> 65^3) and prolongations
    x += P e of k complex64 columns through hz_mg_transfer, timed with events
    after half a second of stencil launches (loaded clocks), then the
    calibration copy.  Algorithmic bytes: restriction 8 k (N_f + N_c),
    prolongation 8 k (2 N_f + N_c)."""
    import torch
    op.set_multigrid(**_mg_args(args))
    nf, nc = (int(np.prod(lev["shape"])) for lev in op.hierarchy.levels[:2])
    g = torch.Generator(device="cuda:0")
    g.manual_seed(5)
    xf = torch.randn((k, nf), dtype=torch.complex64, device="cuda:0", generator=g)
    xc = torch.randn((k, nc), dtype=torch.complex64, device="cuda:0", generator=g)
    bc = torch.empty_like(xc)
    t_warm = time.perf_counter()
    while time.perf_counter() - t_warm < 0.5:
        op.apply(xf[:32], mode="A", precision="c64")
    stream = torch.cuda.current_stream()
    out = {}
    for name, launch, alg in (("restrict", lambda: op.mg_transfer(0, xf, out=bc), 8*k*(nf + nc)),
                              ("prolong", lambda: op.mg_transfer(0, xc, prolong=True, out=xf), 8*k*(2*nf + nc))):
        for _ in range(3):
            launch()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.stencil_reps):
            launch()
        e1.record(stream)
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1)*1e3/args.stencil_reps
        out[name] = {"launch_us": us, "alg_bytes_per_launch": alg, "achieved": alg/us/1e3, "peak": 8000.0,
                     "unit": "GB/s", "frac": alg/us/1e3/8000.0, "launches": args.stencil_reps}
        log(f"{name} k={k}: {us:.1f} us/launch, {alg/us/1e3:.0f} GB/s algorithmic")
    copied = calibration_copy()
    print(json.dumps({"transfers": out, "columns": k, "fine": nf, "coarse": nc,
                      "calibration_copy_bytes_each_way": copied}), flush=True)


def adjoint_gradient(op, scene, args):
    """One adjoint phase gradient at 128^3, timed: the forward solve of one
    seeded phase vector through PhaseField, the contrast-ratio loss (mean
    |u|^2 on a 5-cell ball around a seeded target at z ~ 0.7 Lz over mean
    |u|^2 on the rest of the upper half), and its backward (adjoint solve
    and source gather).  Mixed multigrid form to the same tolerances."""
    import torch
    from ultrasound_solver_amd import autograd as hza
    n = args.n
    m = scene.num_sources
    rng = np.random.default_rng(3)
    tgt = np.array([rng.integers(n//2 - 10, n//2 + 10), rng.integers(n//2 - 10, n//2 + 10), int(0.7*(n - 1))])
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ball_np = (ii - tgt[0])**2 + (jj - tgt[1])**2 + (kk - tgt[2])**2 <= 25
    ball = torch.as_tensor(ball_np, device="cuda:0")
    rest = torch.as_tensor((~ball_np) & (kk >= n//2), device="cuda:0")
    eng = hza.PhaseEngine(op, np.ones(m), precision="c64", preconditioner=args.precond, rtol=args.rtol,
                          restart=args.mixed_restart, maxiter=args.maxiter, chunk=1)
    eng.kw["relative_rtol"] = args.rtol
    phases = torch.tensor(np.random.default_rng(2).uniform(0, 2*np.pi, m), device="cuda:0", requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u = hza.phase_field(phases, eng)
    loss = hza.focal_contrast_ratio_loss(hza.grid_view(u, scene.shape), ball, rest)
    loss.backward()
    grad = phases.grad
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    fw, bw = eng.last_forward_stats[0], eng.last_adjoint_stats[0]
    value = float(loss.detach())
    log(f"adjoint gradient: {sec:.2f}s (forward {fw['iterations']} + adjoint {bw['iterations']} iterations), "
        f"loss {value:.4e}")
    return {"seconds": sec, "loss": value, "gradient_norm": float(torch.linalg.vector_norm(grad)),
            "forward_iterations": fw["iterations"], "adjoint_iterations": bw["iterations"],
            "forward_relative_residual": fw["relative_residual"], "adjoint_relative_residual": bw["relative_residual"],
            "what": "one PhaseField forward solve + contrast-ratio loss + backward (adjoint solve, source gather), "
                    f"{args.precond} GMRES({args.mixed_restart}) c64 + FP64 refinement, 1 phase vector"}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "stencil_128_roofline.md")
STENCIL_MODE_ID = {"RAM": 1, "RA": 4}


def slab_kernel_name(mode, zb, ty, rp, nt, flat):
    """The slab-stencil kernel a tile shape launches (flat = 4: the paired
    kernel, template <MODE, EPI, ZB, TY, NT>)."""
    if flat == 4:
        return f"hz::k_slab2<{STENCIL_MODE_ID[mode]}, 0, {zb}, {ty}, {nt}>"
    return f"hz::k_slab<{STENCIL_MODE_ID[mode]}, 0, {zb}, {ty}, {rp}, {nt}, {flat}>"


def stencil_mode(args):
    """The Krylov apply of the solve: R A under the multigrid preconditioner
    (which is applied as P R^-1), R A M with Jacobi."""
    return "RA" if args.precond == "multigrid" else "RAM"


def measured_traffic(args, mode, k, shape):
    """HBM bytes per launch of the roofline kernel (tile shape, k columns of
    the R A apply on the 128^3 scene) from the FETCH_SIZE / WRITE_SIZE
    counter runs summarised beside its kernel trace, scaled by the copy
    calibration recorded there; None for another kernel or grid."""
    if args.n != 128 or mode != "RA" or not os.path.exists(TRAFFIC_PROFILE):
        return None
    key = f"traffic_bytes_per_launch[shape={shape},k={k}]:"
    with open(TRAFFIC_PROFILE) as f:
        for line in f:
            if line.startswith(key):
                return int(line.split(":")[1])
    return None


def _unit_rhs(op, m, start, k):
    import torch
    w = torch.zeros((k, m), dtype=torch.complex128, device="cuda:0")
    idx = torch.arange(k, device="cuda:0")
    w[idx, (start + idx) % m] = 1.0
    return op.build_rhs(w)


def _mg_args(args):
    """The multigrid settings of the command line, for every workload."""
    return dict(shift=args.mg_shift, pre=args.mg_pre, post=args.mg_post, post_coarse=args.mg_post_coarse,
                weight=args.mg_weight, coarse_sweeps=args.mg_coarse_sweeps,
                levels=args.mg_levels or None, cycle=args.mg_cycle)


def _scene_operator(scene, multigrid=True, **mg):
    from ultrasound_solver_amd import scenes
    from ultrasound_solver_amd.operator import HelmholtzOperator
    op = HelmholtzOperator(scene.rho, scene.c, scene.dx, scene.omega, scene.boundaries, scene.pml, device="cuda:0")
    node, col, val = scenes.source_triplets(scene)
    op.set_sources(node, col, val, scene.num_sources)
    if multigrid:
        op.set_multigrid(**mg)
    return op


def bench_uneven(args):
    """The uneven batch: `--scenes` scenes (edge 64/96/128, 1-3 obstacles,
    impedance walls, 8x8 arrays), `--scene-rhs` columns of G each, solved
    scene after scene, as one hz_gmres_batch call on `--workers` streams and
    as one hz_gmres_fused call (all scenes in lockstep: each kernel of a
    cycle one launch for the whole batch, each cycle body one graph launch).
    `value` is the fused solve's solves/s."""
    import torch
    from ultrasound_solver_amd import scenes
    from ultrasound_solver_amd.operator import gmres_batch, gmres_fused, launch_counts
    t0 = time.perf_counter()
    sc, ops = [], []
    for i in range(args.scenes):
        sc.append(scenes.config_uneven(i))
        ops.append(_scene_operator(sc[-1], **_mg_args(args)))
        if i % 8 == 7:
            log(f"uneven: {i + 1} scenes set up")
    bs = [_unit_rhs(op, s.num_sources, 0, args.scene_rhs) for op, s in zip(ops, sc)]
    # Tile shapes: autotuned on the largest scene and used by every scene
    # (the fused solve runs all scenes with the largest scene's shapes).
    lead = max(range(len(ops)), key=lambda i: ops[i].n)
    tuned = ops[lead].autotune(args.scene_rhs, max_chunk=args.chunk or 16, restart=args.mixed_restart,
                               preconditioner="multigrid")
    for op in ops:
        op.set_tile_shapes(tuned["shapes"][:1 + 3*len(op.hierarchy.levels)])
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    edges = [s.shape[0] for s in sc]
    log(f"uneven setup {setup_s:.1f}s (shapes {tuned['shapes']}): edges {edges}")
    kw = dict(restart=args.mixed_restart, maxiter=args.maxiter, rtol=args.rtol, backward_error_tol=1e-4,
              precision="c64", chunk=args.chunk or 16, preconditioner="multigrid", use_graph=bool(args.graph))
    legs = {}
    if args.uneven_legs == "all":
        # Warm-up: one small batch (kernel loading, graph capture paths).
        gmres_batch(ops[:2], [b[:1] for b in bs[:2]], workers=2, **kw)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        seq_ok = 0
        for op, b in zip(ops, bs):
            _, st = op.gmres(b, **kw)
            seq_ok += sum(x["converged"] for x in st)
            op.release_workspace()
        torch.cuda.synchronize()
        seq_s = time.perf_counter() - ts
        log(f"uneven sequential: {seq_ok} solves in {seq_s:.2f}s")
        tb = time.perf_counter()
        out = gmres_batch(ops, bs, workers=args.workers, **kw)
        torch.cuda.synchronize()
        batch_s = time.perf_counter() - tb
        ok = sum(x["converged"] for _, st in out for x in st)
        log(f"uneven batch ({args.workers} streams): {ok} solves in {batch_s:.2f}s")
        legs = {"streams": {"value": ok/batch_s, "unit": "solves/s", "workers": args.workers},
                "sequential": {"value": seq_ok/seq_s, "unit": "solves/s"}}
        for op in ops:
            op.release_workspace()
    # Lockstep: one fused solve of every scene, each cycle body one graph
    # launch (warm-up on two scenes first: kernel loading).
    fkw = dict(kw, use_graph=bool(args.fused_graph))
    gmres_fused(ops[:2], [b[:1] for b in bs[:2]], **fkw)
    torch.cuda.synchronize()
    launch_counts(reset=True)
    tf = time.perf_counter()
    fout = gmres_fused(ops, bs, **fkw)
    torch.cuda.synchronize()
    fused_s = time.perf_counter() - tf
    counts = launch_counts(reset=True)
    fok = sum(x["converged"] for _, st in fout for x in st)
    fworst = max(x["relative_residual"] for _, st in fout for x in st)
    cycles = [x["cycles"] for _, st in fout for x in st]
    iters = [x["iterations"] for _, st in fout for x in st]
    log(f"uneven fused: {fok} solves in {fused_s:.2f}s; launches {counts}")
    print(json.dumps({
        "metric": "uneven 32-scene batch: RHS solves/s to 1e-6 rel. residual (mixed form)",
        "value": fok/fused_s, "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "c64 (FP64 refinement)", "data": "synthetic",
        "config": {"workload": f"{args.scenes} uneven scenes (edges {sorted(set(edges))}), {args.scene_rhs} G "
                               f"columns each, multigrid GMRES({args.mixed_restart}), all scenes in lockstep "
                               f"(hz_gmres_fused, cycle bodies as graphs)",
                   "edges": edges, "rhs_per_scene": args.scene_rhs},
        "fused": {"seconds": fused_s, "worst_relative_residual": fworst,
                  "graph_launches": counts["graphs"], "fused_kernel_launches": counts["fused"],
                  "single_kernel_launches": counts["single"], "max_scene_cycles": max(cycles),
                  "what": "one graph launch per lockstep cycle body for all scenes still iterating; the "
                          "cycle tops and FP64 refinement checks are the fused and single launches"},
        **legs, "iterations_median": float(np.median(iters)), "setup_s": setup_s}), flush=True)


def bench_large(args):
    """The 256^3 two-layer scene: one chunk of 16 G columns (multigrid GMRES,
    FP64 refinement), the stencil's algorithmic GB/s at 16 columns, and the
    peak device memory."""
    import torch
    from ultrasound_solver_amd import scenes
    t0 = time.perf_counter()
    scene = scenes.config_two_layer(256)
    op = _scene_operator(scene, **_mg_args(args))
    m = scene.num_sources
    tuning = op.autotune(m, max_chunk=16, restart=args.mixed_restart, preconditioner="multigrid")
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    log(f"large setup {setup_s:.1f}s N={op.n}; autotune {tuning}")
    k = 16
    b = _unit_rhs(op, m, 0, k)
    torch.cuda.reset_peak_memory_stats()
    torch.cuda.synchronize()
    ts = time.perf_counter()
    _, st = op.gmres(b, restart=args.mixed_restart, maxiter=args.maxiter, rtol=args.rtol, backward_error_tol=1e-4,
                     precision="c64", chunk=k, preconditioner="multigrid", use_graph=bool(args.graph))
    torch.cuda.synchronize()
    solve_s = time.perf_counter() - ts
    ok = sum(x["converged"] for x in st)
    peak = torch.cuda.max_memory_allocated() + op.device_bytes()
    log(f"large: {ok}/{k} solves in {solve_s:.2f}s; median iterations {np.median([x['iterations'] for x in st])}")
    x = torch.randn((k, op.n), dtype=torch.complex64, device="cuda:0")
    for _ in range(3):
        op.apply(x, mode="RA", precision="c64")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.stencil_reps):
        op.apply(x, mode="RA", precision="c64")
    e1.record()
    torch.cuda.synchronize()
    st_s = e0.elapsed_time(e1)/1e3/args.stencil_reps
    alg = (16*k + 8)*op.n
    print(json.dumps({
        "metric": "256^3 two-layer scene: RHS solves/s to 1e-6 rel. residual (mixed form); stencil HBM GB/s",
        "value": ok/solve_s, "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "c64 (FP64 refinement)", "data": "synthetic",
        "config": {"workload": "two-layer 256^3 scene with rigid ellipsoid, 16 G columns, multigrid "
                               f"GMRES({args.mixed_restart}) c64 + FP64 refinement", "grid": [256]*3, "rhs": k,
                   "autotune": tuning},
        "iterations_median": float(np.median([x["iterations"] for x in st])),
        "stencil": {"achieved_GBps": alg/st_s/1e9, "peak_GBps": 8000.0, "frac": alg/st_s/1e9/8000.0,
                    "launch_us": st_s*1e6, "alg_bytes_per_launch": alg},
        "peak_device_bytes": int(peak), "setup_s": setup_s}), flush=True)


def bench_surface(args, points=64):
    """Homogeneous 128^3 air scene with a 16 x 16 array: the exact target
    response of `points` target points to all 256 emitters through the
    surface condensation (HelmholtzSolver.target_response's path).  Timed:
    the responses (observation kernels, two-sided modal products, emitter
    sums); reported beside it: the setup (host eigendecompositions plus the
    device transfer-ratio recurrence) and the modal-product kernel against
    the FP64 peak.  CPU baseline: the oracle's numpy path on 4 points."""
    import scipy.sparse as sps
    import torch
    from ultrasound_solver_amd import physics, scenes
    from ultrasound_solver_amd.surface import HomogeneousSurfaceSolver
    sc = scenes.config_small(args.n, array_n=16)
    freq = sc.omega/(2*np.pi)
    c0, rho0 = float(sc.c.flat[0]), float(sc.rho.flat[0])
    node, col, mass = physics.source_deposition(sc.shape[:2], sc.dx, sc.centers, 2*sc.radius)
    masses = sps.csr_matrix((mass, (node, col)), shape=(sc.shape[0]*sc.shape[1], sc.num_sources))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev = HomogeneousSurfaceSolver(sc.shape, sc.dx, freq, c0, rho0, sc.boundaries, sc.pml, device="cuda:0")
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    ext = (np.asarray(sc.shape) - 1)*sc.dx
    rng = np.random.default_rng(11)
    pts = (0.3 + 0.4*rng.random((points, 3)))*ext
    for _ in range(args.warmup):
        dev.flux_target_response(masses, pts)
    # A step is well under a millisecond: time at least 50 of them, so that
    # host jitter does not move the figure.
    reps = max(args.steps, 50)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for _ in range(reps):
        resp = dev.flux_target_response(masses, pts)
    torch.cuda.synchronize()
    step_s = (time.perf_counter() - ts)/reps
    # The modal product alone, through the C ABI on device planes already in
    # its layout: 2 launches of `points` (nx x ny) products, 8 flops per complex MAC.
    import ctypes
    from ultrasound_solver_amd import _capi
    nx, ny, _ = sc.shape
    src = torch.randn((points, ny, nx), dtype=torch.complex128, device="cuda:0")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()
    sp_ = ctypes.c_void_p(stream.cuda_stream)

    def transform():
        _capi.check(_capi.lib().hz_surface_transform(dev.handle, 0, ctypes.c_void_p(src.data_ptr()),
                                                     ctypes.c_void_p(dst.data_ptr()), points, sp_),
                    "hz_surface_transform")
    for _ in range(3):
        transform()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record(stream)
    for _ in range(reps):
        transform()
    e1.record(stream)
    torch.cuda.synchronize()
    launch_s = e0.elapsed_time(e1)/1e3/reps/2
    flops = 8.0*nx*ny*ny*points  # stage 1 (inner ny); stage 2 has inner nx = ny here
    cpu = None
    if not args.no_cpu:
        from oracle import surface as osu
        ref = osu.SurfaceSolver(sc.shape, sc.dx, freq, c0, rho0, sc.boundaries, sc.pml)
        tc = time.perf_counter()
        want = ref.flux_target_response(masses, pts[:4])
        cpu_s = time.perf_counter() - tc
        err = float(np.abs(resp[:4] - want).max()/np.abs(want).max())
        cpu = {"value": 4/cpu_s, "unit": "target points/s", "cores": _threads(), "kind": "port",
               "sample": "oracle numpy flux_target_response on 4 of the points, 256 emitters, 128^3",
               "max_rel_diff_vs_device": err}
    print(json.dumps({
        "metric": "homogeneous 128^3 exact target response: target points/s (256 emitters each)",
        "value": points/step_s, "unit": "target points/s", "n_gpus": 1, "steps": reps,
        "warmup": args.warmup, "ms_per_step": step_s*1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128", "data": "synthetic",
        "config": {"workload": f"surface condensation, {args.n}^3 air, PML 8, 16x16 array, {points} points"},
        "setup_s": setup_s,
        "roofline": {"kernel": "hz::k_zgemm", "bound": "fp64", "achieved": flops/launch_s/1e12, "peak": 78.6,
                     "unit": "TFLOP/s", "frac": flops/launch_s/1e12/78.6, "launch_us": launch_s*1e6},
        "cpu_baseline": cpu}), flush=True)


def patch_rows(shape, centre, width):
    """(rows (P, 8), weights (P, 8)) sampling the width^3 nodes centred at
    `centre` (node indices): each point is a node, so one weight is 1."""
    nx, ny, nz = shape
    h = width//2
    ii, jj, kk = np.meshgrid(*[np.arange(c - h, c - h + width) for c in centre], indexing="ij")
    node = (ii + nx*(jj + ny*kk)).ravel(order="F").astype(np.int64)
    rows = np.repeat(node[:, None], 8, axis=1)
    w = np.zeros(rows.shape)
    w[:, 0] = 1.0
    return rows, w


class _DeviceBasis:
    """A phase response basis already on the device, for PhaseObjective:
    rows = nodes (Fortran order), trilinear target rows from it."""

    def __init__(self, g, cfg):
        self.basis_gpu = g
        self.cfg = cfg

    def target_response(self, points):
        from ultrasound_solver_amd import training
        d = self.cfg.domain
        rows, w = training.trilinear_rows(points, d.grid_size, d.origin_m, d.dx, d.box_size)
        import torch
        corner = self.basis_gpu.index_select(0, torch.as_tensor(rows.ravel(), device=self.basis_gpu.device))
        corner = corner.reshape(len(rows), 8, -1)
        return (torch.as_tensor(w, device=corner.device)[:, :, None]*corner).sum(dim=1).cpu().numpy()


def bench_online(args, evaluations=2000):
    """PhaseObjective.evaluate (loss and phase gradient, one fused pass over
    G) on a device basis the shape of the archived 41^3 run with a 12 x 12
    array (68,921 x 144 complex128): the inner loop of every phase optimiser
    (src/algorithm/baselines/common.py:275-343).  The archived scene itself
    has 2.9 points per wavelength, below the solver's 6, so G holds seeded
    random values; the cost of an evaluation depends only on the shape.
    Losses as the archived run configures them (focal_contrast, T = 5,
    beta = 1.3, target ball 8 mm, 4 source layers excluded), plus
    focal_pressure and field_match on the same basis."""
    import torch
    from types import SimpleNamespace
    from ultrasound_solver_amd import training
    from ultrasound_solver_amd.config import SimulationConfig
    raw = {"mode": "phase_optimization",
           "physics": {"sound_speed": 343.0, "medium_density": 1.21},
           "transducer_specs": {"frequency": 40000.0, "diameter": 0.01, "pitch": 0.01, "array_n": 12},
           "domain": {"box_size": [0.12]*3, "grid_size": [41]*3},
           "boundary_conditions": {"-x": "open", "+x": "reflecting", "-y": "open", "+y": "open",
                                   "-z": "reflecting", "+z": "open"},
           "targets": [{"point": [0.09, 0.09, 0.09]}],
           "training": {"loss_type": "focal_contrast", "target_peak_pressure": 600.0, "target_sigma": 0.006,
                        "background_weight": 1.3, "target_radius": 0.008, "source_exclusion_layers": 4,
                        "sidelobe_temperature": 5.0},
           "solver": {"backend": "gpu_iterative"}}
    n, m = 41**3, 144
    gen = torch.Generator(device="cuda:0").manual_seed(12)
    g = torch.complex(torch.randn((n, m), generator=gen, device="cuda:0", dtype=torch.float64),
                      torch.randn((n, m), generator=gen, device="cuda:0", dtype=torch.float64))*50
    rng = np.random.default_rng(13)
    out = {}
    for loss in ("focal_contrast", "focal_pressure", "field_match"):
        cfg = SimulationConfig.from_dict(dict(raw, training=dict(raw["training"], loss_type=loss)))
        solver = SimpleNamespace(cfg=cfg, transducers=SimpleNamespace(num_transducers=m))
        phases = rng.uniform(0, 2*np.pi, (evaluations, m))
        for graph in (False, True):
            obj = training.PhaseObjective(solver, basis=_DeviceBasis(g, cfg), use_graph=graph)
            for q in range(20):
                obj.evaluate(phases[q], gradient=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for q in range(evaluations):
                obj.evaluate(phases[q], gradient=True)
            torch.cuda.synchronize()
            sec = time.perf_counter() - t0
            leg = "graph" if graph else "launches"
            out.setdefault(loss, {})[leg] = {"evaluations_per_s": evaluations/sec,
                                             "ms_per_evaluation": sec/evaluations*1e3}
            log(f"online {loss} ({leg}): {evaluations/sec:.0f} evaluations/s")
    basis_bytes = n*m*16
    print(json.dumps({
        "metric": "phase-optimiser evaluations/s (loss + phase gradient) on a 41^3 x 144 device basis",
        "value": out["focal_contrast"]["graph"]["evaluations_per_s"], "unit": "evaluations/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "c128", "data": "synthetic basis of the archived run's shape",
        "config": {"workload": f"PhaseObjective.evaluate(gradient=True), {evaluations} seeded phase vectors, "
                               "archived focal_contrast settings; G read once per evaluation; value: each "
                               "evaluation one graph replay (use_graph=True), 'launches': the same step "
                               "launched op by op"},
        "losses": out, "basis_bytes": basis_bytes,
        "basis_read_GBps": basis_bytes*out["focal_contrast"]["graph"]["evaluations_per_s"]/1e9,
        "context": "the reference's archived run: 1,701 field + VJP evaluations in 12.54 s on its own GPU "
                   "(research_archive/algorithm_comparison/results/phase_oblique_reflecting_x_12x12)"}),
        flush=True)


def _threads():
    try:
        from threadpoolctl import threadpool_info
        return max((i.get("num_threads", 1) for i in threadpool_info()), default=1)
    except ImportError:
        return 1


def _host_info():
    """CPU model, os.cpu_count() and the BLAS threadpools of this host."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    pools = []
    try:
        from threadpoolctl import threadpool_info
        pools = [{k: i.get(k) for k in ("internal_api", "num_threads", "version")} for i in threadpool_info()]
    except ImportError:
        pass
    return {"cpu_model": model, "os_cpu_count": os.cpu_count(), "threadpools": pools,
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def cpu_baseline(scene, args, plain_iters):
    """scipy FP64 GMRES(30) on the oracle's CSR matrix, composed as R A M with
    the reference's row scale and Jacobi (iterative_solver.py:37-69), on
    column 0: one warm-up, then the median of `--cpu-samples` (>= 3) timed
    samples of one restart cycle (30 inner steps) each, at 1 BLAS thread and
    at all threads.  scipy's SpMV is single-threaded; the Gram-Schmidt BLAS calls
    use the threads.  The per-step time is scaled by the inner steps the
    plain FP64 form needed for the same column (the same algorithm), which
    128^3 makes infeasible to run to the end on the CPU."""
    import scipy.sparse.linalg as spla
    from oracle import helmholtz as ohz, krylov as okr
    from ultrasound_solver_amd import scenes
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:
        threadpool_limits = None
    t0 = time.perf_counter()
    a = ohz.OracleOperator(scene.rho, scene.c, scene.dx, scene.omega, scene.boundaries, scene.pml).matrix().tocsr()
    assemble_s = time.perf_counter() - t0
    node, col, val = scenes.source_triplets(scene)
    b = np.zeros(a.shape[0], dtype=np.complex128)
    sel = col == 0
    b[node[sel]] = val[sel]
    scale = okr.row_scale(a)
    minv = okr.jacobi_inverse(a, scale)
    ram = spla.LinearOperator(a.shape, matvec=lambda v: scale*(a @ (minv*v)), dtype=np.complex128)
    bh = scale*b
    restart = args.restart

    def one_cycle():
        steps = []
        t = time.perf_counter()
        spla.gmres(ram, bh, rtol=1e-30, atol=0.0, restart=restart, maxiter=1,
                   callback=lambda r: steps.append(r), callback_type="pr_norm")
        return (time.perf_counter() - t)/max(len(steps), 1)

    def timed(threads):
        ctx = threadpool_limits(threads) if (threadpool_limits and threads) else None
        if ctx:
            ctx.__enter__()
        try:
            return [one_cycle() for _ in range(max(args.cpu_samples, 3))]
        finally:
            if ctx:
                ctx.__exit__(None, None, None)

    one_cycle()  # warm-up (page-in of the matrix and the vectors)
    single = timed(1)
    multi = timed(None)
    per1, perm = float(np.median(single)), float(np.median(multi))
    threads = _threads()
    needed = plain_iters[0] if plain_iters else None
    log(f"cpu: assemble {assemble_s:.1f}s, {per1*1e3:.1f} ms/step (1 thread), {perm*1e3:.1f} ms/step "
        f"({threads} threads), needed {needed}")
    best, cores = (perm, threads) if perm < per1 else (per1, 1)
    return {"value": 1.0/(best*needed) if needed else None, "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"scipy.sparse.linalg.gmres, restart {restart}, on the oracle's {args.n}^3 complex128 CSR "
                      f"composed as R A M (row scale, Jacobi), column 0: median of {len(multi)} one-cycle samples "
                      f"({restart} inner steps each) after a warm-up, scaled by the {needed} inner steps the plain "
                      "FP64 form needed for that column",
            "seconds_per_step": {"1_thread": per1, f"{threads}_threads": perm,
                                 "samples_1_thread": single, f"samples_{threads}_threads": multi},
            "host": _host_info(), "assemble_s": assemble_s}


if __name__ == "__main__":
    main()
