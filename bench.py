"""Benchmark: RHS solves/s at 128^3 to a 1e-6 scaled residual, mixed batched
form vs the plain FP64 form, with the stencil's HBM roofline and a CPU
baseline.

A "step" is one batched solve of `--chunk` columns of the phase response
basis G (transducers r .. r+chunk-1 at unit amplitude) to the reference's
acceptance test (scaled residual <= 1e-6, backward error <= 1e-4) on the
heterogeneous 128^3 scene (c = 343 (1 + 0.05 g), rigid sphere, PML on -x,
+-y, +z, 16x16 array).  Inputs (coefficients, source columns) are resident
on the device before the timed region.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


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
                   help="stencil-only: pin the fine apply's tile shape (index) after autotuning")
    p.add_argument("--stencil-only", action="store_true",
                   help="only the stencil launches and a 512 MiB copy (for hardware counter runs)")
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
    p.add_argument("--workload", choices=["g128", "uneven", "large", "surface"], default="g128",
                   help="g128: G columns of the heterogeneous 128^3 scene (the headline); uneven: the "
                        "32-scene batch; large: the 256^3 two-layer scene")
    p.add_argument("--scenes", type=int, default=32, help="uneven: scenes in the batch")
    p.add_argument("--scene-rhs", type=int, default=16, help="uneven: right-hand sides per scene")
    p.add_argument("--workers", type=int, default=4, help="uneven: concurrent scene streams")
    p.add_argument("--precond", choices=["multigrid", "jacobi"], default="multigrid",
                   help="inner preconditioner of the mixed form (the plain form always uses Jacobi)")
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
    t0 = time.perf_counter()
    scene = scenes.config_heterogeneous(n=args.n)
    op = HelmholtzOperator(scene.rho, scene.c, scene.dx, scene.omega, scene.boundaries, scene.pml,
                           device="cuda:0")
    node, col, val = scenes.source_triplets(scene)
    if np.any(op.pml_mask[:, :, 0].ravel(order="F")[node]):
        raise ValueError("PML overlaps the deposited sinc aperture")
    op.set_sources(node, col, val, scene.num_sources)
    if args.stencil_only:
        stencil_only(op, args, scene.num_sources)
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

    common = dict(restart=args.restart, maxiter=args.maxiter, rtol=args.rtol, backward_error_tol=1e-4)
    mixed = dict(common, restart=args.mixed_restart, precision="c64", chunk=args.chunk, preconditioner=args.precond,
                 use_graph=bool(args.graph), arnoldi_block=args.arnoldi_block)
    for s in range(args.warmup):
        b = rhs_block(s*args.chunk, args.chunk)
        _, st = op.gmres(b, **mixed)
        log(f"warmup step {s}: iterations {[x['iterations'] for x in st]}")
    torch.cuda.synchronize()
    solved, iters, refines, worst = 0, [], [], 0.0
    t_start = time.perf_counter()
    for s in range(args.steps):
        b = rhs_block((args.warmup + s)*args.chunk, args.chunk)
        _, st = op.gmres(b, **mixed)
        solved += sum(x["converged"] for x in st)
        iters += [x["iterations"] for x in st]
        refines += [x["refinements"] for x in st]
        worst = max(worst, max(x["scaled_residual"] for x in st))
        log(f"step {s}: converged {sum(x['converged'] for x in st)}/{len(st)}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    mixed_rate = solved/elapsed
    log(f"mixed: {solved} solves in {elapsed:.2f}s -> {mixed_rate:.3f}/s; median iterations {np.median(iters)}")

    # Plain FP64 form (reference semantics, one right-hand side per launch).
    plain_iters, plain_solved = [], 0
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for r in range(args.plain_rhs):
        b = rhs_block(r, 1)
        _, st = op.gmres(b, precision="c128", chunk=1, use_graph=bool(args.graph),
                         **dict(common, maxiter=args.plain_maxiter))
        plain_solved += st[0]["converged"]
        plain_iters.append(st[0]["iterations"])
    torch.cuda.synchronize()
    plain_elapsed = time.perf_counter() - tp
    plain_rate = plain_solved/plain_elapsed if plain_elapsed > 0 else 0.0
    log(f"plain: {plain_solved} solves in {plain_elapsed:.2f}s -> {plain_rate:.3f}/s; iterations {plain_iters}")

    # Stencil roofline: the Krylov apply of the timed solve (R A with the
    # multigrid preconditioner, R A M with Jacobi) with the solve's tile shape,
    # over 32 complex64 columns (a fixed width, so that the figure does not
    # move with the chunk the autotuner picks), timed with events on the
    # stream it is launched on.
    k = args.stencil_rhs or 32
    mode = stencil_mode(args)
    x = torch.randn((k, op.n), dtype=torch.complex64, device="cuda:0")
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
    del dst

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scene, args, plain_iters)
    zb, ty, rp, nt, flat = HelmholtzOperator.SLAB_SHAPES[tuning["shapes"][0]]
    traffic = measured_traffic(args, mode, k, zb, ty, rp, nt, flat)

    out = {
        "metric": "RHS solves/s at 128^3 to 1e-6 rel. residual (mixed vs FP64); stencil HBM GB/s",
        "value": mixed_rate, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed/args.steps*1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 (FP64 refinement)", "data": "synthetic",
        "config": {"workload": f"heterogeneous {args.n}^3 scene, G columns in chunks of {args.chunk}, "
                               f"{args.precond}-preconditioned GMRES({args.mixed_restart}) c64 + FP64 refinement to {args.rtol:g}",
                   "grid": [args.n]*3, "rhs_per_step": args.chunk, "sources": m,
                   "autotune": tuning},
        "plain_fp64": {"value": plain_rate, "unit": "solves/s", "rhs": args.plain_rhs,
                       "iterations": plain_iters},
        "speedup_mixed_vs_plain": mixed_rate/plain_rate if plain_rate > 0 else None,
        "iterations_median": float(np.median(iters)) if iters else None,
        "refinements_median": float(np.median(refines)) if refines else None,
        "worst_scaled_residual": worst,
        "setup_s": setup_s,
        "roofline": {"kernel": f"{slab_kernel_name(mode, zb, ty, rp, nt, flat)} "
                               f"(complex64 {' '.join(mode)} apply, {k} columns)",
                     "bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved/8000.0, "traffic": traffic, "alg_bytes_per_launch": alg_bytes,
                     "launch_us": stencil_s*1e6,
                     "copy_reference": {"what": f"device copy of the same {k} complex64 columns",
                                        "achieved": copy_gbps, "frac": copy_gbps/8000.0}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "stencil_128_traffic.md")
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


def measured_traffic(args, mode, k, zb, ty, rp, nt, flat):
    """HBM bytes per stencil launch from the FETCH_SIZE/WRITE_SIZE counter runs
    of the same kernel, mode, column count and tile shape on the 128^3 scene,
    scaled by the copy calibration recorded beside them; None otherwise."""
    if args.n != 128 or not os.path.exists(TRAFFIC_PROFILE):
        return None
    key = f"traffic_bytes_per_launch[mode={mode},k={k},zb={zb},ty={ty},rp={rp},nt={nt},flat={flat}]:"
    with open(TRAFFIC_PROFILE) as f:
        for line in f:
            if line.startswith(key):
                return int(line.split(":")[1])
    return None


def stencil_only(op, args, m):
    """Launches of the solve's Krylov apply (k complex64 columns; R A with the
    multigrid preconditioner) and a copy of known size, so that hardware byte
    counters can be scaled.  The copy moves 2 GiB each
    way from a source evicted from the 256 MB last-level cache by a 1 GiB
    fill in between."""
    import torch
    k = args.stencil_rhs or 32
    mode = stencil_mode(args)
    if args.precond == "multigrid":
        op.set_multigrid(**_mg_args(args))
    tuning = op.autotune(m, max_chunk=k, restart=args.mixed_restart, preconditioner=args.precond)
    if args.stencil_shape >= 0:
        op.set_tile_shapes([args.stencil_shape] + tuning["shapes"][1:])
    print(json.dumps({"autotune": tuning, "stencil_shape": args.stencil_shape}), flush=True)
    x = torch.randn((k, op.n), dtype=torch.complex64, device="cuda:0")
    for _ in range(args.stencil_reps):
        op.apply(x, mode=mode, precision="c64")
    src = torch.ones(512*1024*1024, dtype=torch.float32, device="cuda:0")
    flush = torch.empty(256*1024*1024, dtype=torch.float32, device="cuda:0")
    flush.fill_(2.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"stencil_launches": args.stencil_reps, "mode": mode, "rhs": k, "n": op.n,
                      "alg_bytes_per_launch": (16*k + 8)*op.n, "copy_bytes_each_way": src.numel()*4}))


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
    scene after scene and then as one hz_gmres_batch call on `--workers`
    streams.  `value` is the batch's solves/s."""
    import torch
    from ultrasound_solver_amd import scenes
    from ultrasound_solver_amd.operator import gmres_batch
    t0 = time.perf_counter()
    sc, ops = [], []
    for i in range(args.scenes):
        sc.append(scenes.config_uneven(i))
        ops.append(_scene_operator(sc[-1], **_mg_args(args)))
        if i % 8 == 7:
            log(f"uneven: {i + 1} scenes set up")
    bs = [_unit_rhs(op, s.num_sources, 0, args.scene_rhs) for op, s in zip(ops, sc)]
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    edges = [s.shape[0] for s in sc]
    log(f"uneven setup {setup_s:.1f}s: edges {edges}")
    kw = dict(restart=args.mixed_restart, maxiter=args.maxiter, rtol=args.rtol, backward_error_tol=1e-4,
              precision="c64", chunk=args.chunk or 16, preconditioner="multigrid", use_graph=bool(args.graph))
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
    iters = [x["iterations"] for _, st in out for x in st]
    log(f"uneven batch ({args.workers} streams): {ok} solves in {batch_s:.2f}s")
    print(json.dumps({
        "metric": "uneven 32-scene batch: RHS solves/s to 1e-6 rel. residual (mixed form)",
        "value": ok/batch_s, "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "c64 (FP64 refinement)", "data": "synthetic",
        "config": {"workload": f"{args.scenes} uneven scenes (edges {sorted(set(edges))}), {args.scene_rhs} G "
                               f"columns each, multigrid GMRES({args.mixed_restart}), {args.workers} streams",
                   "edges": edges, "rhs_per_scene": args.scene_rhs},
        "sequential": {"value": seq_ok/seq_s, "unit": "solves/s"},
        "iterations_median": float(np.median(iters)), "setup_s": setup_s}), flush=True)


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


def _threads():
    try:
        from threadpoolctl import threadpool_info
        return max((i.get("num_threads", 1) for i in threadpool_info()), default=1)
    except ImportError:
        return 1


def cpu_baseline(scene, args, plain_iters):
    """scipy/numpy FP64 Jacobi-GMRES on the oracle CSR, timed on a bounded
    sample of iterations and scaled by the iteration count the plain form
    needed for the same column."""
    from oracle import helmholtz as ohz, krylov as okr
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:
        threadpool_limits = None
    t0 = time.perf_counter()
    a = ohz.OracleOperator(scene.rho, scene.c, scene.dx, scene.omega, scene.boundaries, scene.pml).matrix()
    assemble_s = time.perf_counter() - t0
    from ultrasound_solver_amd import scenes
    node, col, val = scenes.source_triplets(scene)
    b = np.zeros(a.shape[0], dtype=np.complex128)
    sel = col == 0
    b[node[sel]] = val[sel]
    scale = okr.row_scale(a)
    minv = okr.jacobi_inverse(a, scale)
    bh = scale*b
    ctx = threadpool_limits(1) if threadpool_limits else None
    try:
        if ctx:
            ctx.__enter__()
        t1 = time.perf_counter()
        okr.cupy_style_gmres(lambda v: scale*(a @ v), lambda v: minv*v, bh, 1e-30, 0.0, args.restart,
                             args.cpu_iters)
        per_iter = (time.perf_counter() - t1)/args.cpu_iters
    finally:
        if ctx:
            ctx.__exit__(None, None, None)
    needed = plain_iters[0] if plain_iters else None
    if needed is None:
        return {"value": None, "unit": "solves/s", "cores": 1, "kind": "port",
                "sample": "not scaled: the plain form was not run", "seconds_per_iteration": per_iter}
    value = 1.0/(per_iter*needed) if needed else None
    log(f"cpu: assemble {assemble_s:.1f}s, {per_iter*1e3:.1f} ms/iteration, needed {needed}")
    return {"value": value, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"{args.cpu_iters} Jacobi-GMRES({args.restart}) iterations (scipy CSR SpMV + numpy CGS, "
                      f"1 thread) on the oracle's {args.n}^3 complex128 matrix, column 0; scaled by the "
                      f"{needed} iterations the plain FP64 form needed",
            "seconds_per_iteration": per_iter, "assemble_s": assemble_s}


if __name__ == "__main__":
    main()
