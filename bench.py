"""Benchmark of the tt-fort local-solver hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints one JSON line.  Workloads:

  eigb    (default) -- tt_eigb on the Heisenberg XXX chain d=64, 5 lowest
          eigenpairs, rmax 128, tol 1e-10, rank-8 N(0,1) start (seed 0), in
          two legs: FP64 throughout, and fp32 matrix-core products of the
          search directions with FP64 refinement.  Each leg: W warm-up
          sweeps, then K timed sweeps of one run; one step = one sweep.
          `value` = algorithmic local-matvec flops of the mixed leg's timed
          sweeps / their wall time; both legs' sweep times, ratio and final
          FP64 eigen-residuals are in `config`.  All cores and interfaces
          stay in HBM.
  matvec  -- one block matvec y = Phi_L A_k Phi_R x of the largest local
          problem of that sweep, (r_left, n, r_right, R) = (128, 2, 128, 5),
          block of 8 vectors (the block eigensolver's width), inputs resident
          in HBM.

The `roofline` object describes the dominant kernel, the FP64 local matvec
at that local shape (one fused launch of the three contraction stages plus
the zero fill of y), timed with HIP events on the library stream over 50
calls; `traffic` is its HBM bytes per call from the committed counter
summary.  `cpu_baseline` times the
oracle's sweep (numpy + host BLAS) on a bounded sample of the same workload
on this host's cores (eigb), or its local matvec (matvec).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tt-fort_amd"))

MI355X_PEAK_F64_TFLOPS = 78.6
MI355X_PEAK_F32_TFLOPS = 157.3
MI355X_HBM_GBS = 8000.0

# Largest local problem of the d=64 Heisenberg sweep at rmax 128:
# (rx1, m, rx2, ry1, n, ry2, ra1, ra2), and the block eigensolver's width
# (5 wanted eigenpairs + 3 guard vectors).
C2_LOCAL_SHAPE = (128, 2, 128, 128, 2, 128, 5, 5)
C2_BLOCK = 8

# HBM bytes per call of the FP64 local matvec, from two rocprofv3 --pmc
# passes (FETCH_SIZE, WRITE_SIZE) scaled by the same counters of a
# known-size copy.
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "bfun3_fused_r128_block8_hbm_traffic.md")


def measured_traffic():
    """HBM bytes per call of the FP64 local matvec at the C2 local shape,
    block 8, as recorded in the committed counter summary; None when the
    summary has no figure."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes ="):
                    return int(line.split("=")[1])
    except OSError:
        pass
    return None


def bfun3_stage_flops(rx1, m, rx2, ry1, n, ry2, ra1, ra2):
    return (2.0 * rx1 * m * rx2 * ra2 * ry2,
            2.0 * ra1 * n * m * ra2 * ry2 * rx1,
            2.0 * ry1 * n * ry2 * rx1 * ra1)


def _blas_threads_options():
    n = os.cpu_count() or 1
    lim = int(os.environ.get("OMP_NUM_THREADS", n) or n)
    opts = [t for t in (1, 2, 4, 8, 16, 32) if t <= max(1, min(n, lim))]
    return opts or [1]


def cpu_baseline_matvec(shape, nvec, budget_s=12.0):
    """Time the oracle's bfun3 (the reference's GEMM chain on host BLAS).

    The host BLAS thread count is calibrated first (about 0.5 s per option)
    and the sample then runs with the fastest count, which is reported."""
    from oracle import kernels as K
    from threadpoolctl import threadpool_limits
    rng = np.random.default_rng(1)
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    phi1 = rng.standard_normal(ry1 * rx1 * ra1)
    A = rng.standard_normal(ra1 * n * m * ra2)
    phi2 = rng.standard_normal(rx2 * ra2 * ry2)
    X = rng.standard_normal((rx1 * m * rx2, nvec))

    def rate(threads, budget):
        with threadpool_limits(limits=threads, user_api="blas"):
            K.dbfun3(*shape, phi1, A, phi2, X[:, 0])   # warm
            reps, t0 = 0, time.perf_counter()
            while True:
                for v in range(nvec):
                    K.dbfun3(*shape, phi1, A, phi2, X[:, v])
                reps += 1
                el = time.perf_counter() - t0
                if el > budget:
                    break
        return reps, el

    best = None
    for t in _blas_threads_options():
        reps, el = rate(t, 0.5)
        if best is None or reps / el > best[1]:
            best = (t, reps / el)
    cores = best[0]
    reps, el = rate(cores, budget_s)
    flops = sum(bfun3_stage_flops(*shape)) * nvec * reps
    return dict(value=flops / el / 1e12, unit="TFLOP/s", cores=int(cores), kind="port",
                sample="%d block matvecs (%d vectors each) of the oracle's dbfun3 GEMM chain "
                       "in %.1f s, host BLAS on %d threads (fastest of %s)"
                       % (reps, nvec, el, cores, _blas_threads_options()))


def run_matvec(args):
    """One block matvec of the C2 local problem per step (block of 8, the
    block eigensolver's width), inputs resident in HBM."""
    import ttfort
    lib = ttfort.load()
    ttfort.check(lib.ttf_init(0), lib)
    shape = C2_LOCAL_SHAPE
    nvec = C2_BLOCK
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    N = rx1 * m * rx2
    rng = np.random.default_rng(0)
    d = dict(phi1=ttfort.DeviceArray.from_numpy(rng.standard_normal(ry1 * rx1 * ra1)),
             A=ttfort.DeviceArray.from_numpy(rng.standard_normal(ra1 * n * m * ra2)),
             phi2=ttfort.DeviceArray.from_numpy(rng.standard_normal(rx2 * ra2 * ry2)),
             x=ttfort.DeviceArray.from_numpy(rng.standard_normal(N * nvec)),
             y=ttfort.DeviceArray((N * nvec,)))

    def step():
        ttfort.check(lib.ttf_dbfun3(*shape, nvec, d["phi1"].ptr, d["A"].ptr, d["phi2"].ptr,
                                    d["x"].ptr, N, d["y"].ptr, N), lib)

    for _ in range(args.warmup):
        step()
    ttfort.sync(lib)
    t0 = time.perf_counter()
    with ttfort.Timer(lib) as tm:
        for _ in range(args.steps):
            step()
    ttfort.sync(lib)
    wall = time.perf_counter() - t0
    ms_step = tm.ms / args.steps
    flops_step = sum(bfun3_stage_flops(*shape)) * nvec
    value = flops_step / (ms_step * 1e-3) / 1e12
    out = dict(metric="local-matvec TFLOP/s (FP64, Heisenberg d=64 r=128 local problem)",
               value=round(value, 4), unit="TFLOP/s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=round(ms_step, 5), higher_is_better=True,
               scaling="weak", vs_baseline=None, dtype="f64", data="synthetic",
               config=dict(workload="eigb-local-matvec", r_left=rx1, n=m, r_right=rx2, R=ra2,
                           block=nvec, wall_s=round(wall, 3)),
               roofline=dominant_kernel_roofline(lib, shape, nvec))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_matvec(shape, nvec, budget_s=args.cpu_budget)
    return out


def dominant_kernel_roofline(lib, shape, nvec, iters=50):
    """HIP-event timing of the FP64 local matvec as the library runs it (at
    this shape the fused one-launch bfun3 kernel and its zero fill of y),
    `iters` calls back to back between two events on the library stream.
    achieved = algorithmic flops of the three stages per call / mean call
    time; traffic = HBM bytes per call from the committed counter summary
    (against the algorithmic bytes phi1 + phi2 + A + block (x + y))."""
    import ctypes
    import ttfort
    rx1, m, rx2, ry1, n, ry2, ra1, ra2 = shape
    nx, ny = rx1 * m * rx2, ry1 * n * ry2
    rng = np.random.default_rng(3)
    d = dict(phi1=ttfort.DeviceArray.from_numpy(rng.standard_normal(ry1 * rx1 * ra1)),
             A=ttfort.DeviceArray.from_numpy(rng.standard_normal(ra1 * n * m * ra2)),
             phi2=ttfort.DeviceArray.from_numpy(rng.standard_normal(rx2 * ra2 * ry2)),
             x=ttfort.DeviceArray.from_numpy(rng.standard_normal(nx * nvec)),
             y=ttfort.DeviceArray((ny * nvec,)))
    ms = ctypes.c_float()
    fused = ctypes.c_int()
    for it in (5, iters):
        ttfort.check(lib.ttf_dbfun3_chain_profile(*shape, nvec, d["phi1"].ptr, d["A"].ptr, d["phi2"].ptr,
                                                  d["x"].ptr, nx, d["y"].ptr, ny, it, ctypes.byref(ms),
                                                  ctypes.byref(fused)), lib)
    flops = sum(bfun3_stage_flops(*shape)) * nvec
    achieved = flops / (ms.value * 1e-3) / 1e12
    alg_bytes = 8 * (ry1 * rx1 * ra1 + rx2 * ra2 * ry2 + ra1 * n * m * ra2 + nvec * (nx + ny))
    traffic = measured_traffic() if (shape, nvec) == (C2_LOCAL_SHAPE, C2_BLOCK) else None
    name = ("bfun3_fused_kernel<3> + zero fill of y" if fused.value else "bfun3 GEMM chain (three stages)")
    return dict(bound="mfma", achieved=round(achieved, 3), peak=MI355X_PEAK_F64_TFLOPS,
                unit="TFLOP/s", frac=round(achieved / MI355X_PEAK_F64_TFLOPS, 4), traffic=traffic,
                kernel="%s: y = Phi1 A Phi2 x at (r,n,r,R)=(%d,%d,%d,%d), block %d, %.1f MFLOP per call"
                       % (name, rx1, m, rx2, ra2, nvec, flops / 1e6),
                us_per_call=round(ms.value * 1e3, 3), algorithmic_bytes=alg_bytes)


def _eigb_leg(args, mixed):
    """One tt_eigb run of the benchmark workload: W warm-up sweeps then K
    timed sweeps, FP64 throughout or with fp32 search-direction products
    and FP64 refinement; the final FP64 eigen-residuals of the result are
    then computed in TT format on the host (outside the timed region)."""
    import ttfort
    import mpo
    import ttops
    d, B, rmax, tol = args.eigb_d, 5, args.eigb_rmax, 1e-10
    crA, ra, n = mpo.xxz(d)
    crY, ry = mpo.random_block_tt(d, n, 8, B, seed=0)
    nsw = args.warmup + args.steps
    ttfort.set_local_precision(32 if mixed else 64)
    try:
        t0 = time.perf_counter()
        lam, res, ry_out, st = ttfort.tt_eigb(n, ra, crA, crY, ry, tol, rmax, B, nswp=nsw, verb=args.verb)
        wall = time.perf_counter() - t0
    finally:
        ttfort.set_local_precision(64)
    secs, flops = st["sweep_seconds"], st["sweep_flops"]
    timed = list(range(min(args.warmup, len(secs)), len(secs))) or list(range(len(secs)))
    t_timed = sum(secs[i] for i in timed)
    f_timed = sum(flops[i] for i in timed)
    mv = sum(st["sweep_matvecs"][i] for i in timed)
    mv32 = sum(st["sweep_matvecs32"][i] for i in timed)
    t0 = time.perf_counter()
    resid, G = ttops.block_eig_residuals(crA, ra, n, res, ry_out, lam)
    t_res = time.perf_counter() - t0
    leg = dict(ms_per_sweep=round(1e3 * t_timed / len(timed), 2), tflops=round(f_timed / t_timed / 1e12, 4),
               sweep_seconds=[round(v, 3) for v in secs], matvecs_timed=int(mv), matvecs_fp32_timed=int(mv32),
               fp64_refreshes=st["refreshes"], lambda_=[float("%.12f" % v) for v in lam],
               eigen_residuals=[float("%.3e" % v) for v in resid],
               orthonormality=float("%.2e" % np.abs(G - np.eye(B)).max()),
               sweep_ermax=[float("%.3e" % v) for v in st["sweep_ermax"]], last_ermax=st["last_ermax"],
               converged=bool(st["converged"]), reached_tol_at_sweep=next(
                   (i + 1 for i, e in enumerate(st["sweep_ermax"]) if e < tol), None),
               max_rank=st["max_rank"], solve_s=round(st["solve_seconds"], 3),
               trunc_s=round(st["trunc_seconds"], 3), phi_s=round(st["phi_seconds"], 3), wall_s=round(wall, 2),
               host_residual_s=round(t_res, 1))
    return leg, (crA, ra, n, res, ry_out, d, B, rmax, tol, len(timed), f_timed, t_timed)


def cpu_baseline_sweep(ctx, budget_s):
    """The oracle's tt_eigb (the reference's sweep restated in numpy: the
    dbfun3 GEMM chain on host BLAS, ARPACK as the local eigensolver with the
    reference's tolerance eps2/resid_damp) on the same operator, started
    from the GPU run's rank-128 result with the block index moved to the
    middle of the chain, for a bounded number of consecutive backward local
    steps at the full local size (128, 2, 128).  The host BLAS thread count
    is calibrated on the local matvec first and the fastest is used."""
    from oracle import drivers as D
    from oracle import kernels as K
    from threadpoolctl import threadpool_limits
    import scipy.sparse.linalg as sla
    crA, ra, n, res, ry_out, d, B, rmax, tol, nsteps_gpu, f_gpu, t_gpu = ctx
    shape = (rmax, 2, rmax, rmax, 2, rmax, 5, 5)
    N = rmax * 2 * rmax
    rng = np.random.default_rng(1)
    P1 = rng.standard_normal((rmax, rmax, 5))
    P2 = rng.standard_normal((rmax, 5, rmax))
    Ac = rng.standard_normal((5, 2, 2, 5))
    p1 = (P1 + P1.transpose(1, 0, 2)).reshape(-1, order="F")
    p2 = (P2 + P2.transpose(2, 1, 0)).reshape(-1, order="F")
    A = (Ac + Ac.transpose(3, 2, 1, 0)).reshape(-1, order="F")
    op = sla.LinearOperator((N, N), dtype=float,
                            matvec=lambda v: K.dbfun3(*shape, p1, A, p2, np.asarray(v, float).reshape(-1)))
    v0 = rng.standard_normal(N)
    # calibration on what the sample runs: an ARPACK restart cycle of local
    # matvecs (BLAS threads that pay off on one large product can lose
    # many-fold inside the eigensolver's loop)
    best = None
    for t in _blas_threads_options():
        with threadpool_limits(limits=t, user_api="blas"):
            t0 = time.perf_counter()
            try:
                sla.eigsh(op, k=B, which="SA", v0=v0, ncv=20, maxiter=1, tol=1e-12)
            except sla.ArpackNoConvergence:
                pass
            el = time.perf_counter() - t0
        if best is None or el < best[1]:
            best = (t, el)
    cores = best[0]
    with threadpool_limits(limits=cores, user_api="blas"):
        lam, _, _, st = D.tt_eigb(d, n, ra, crA, res, list(ry_out), tol, rmax, B, nswp=1, start_site=d // 2,
                                  time_budget=budget_s)
    steps = st["steps"]
    secs = sum(s[1] for s in steps)
    mv = sum(s[2] for s in steps)
    fl = sum(bfun3_stage_flops(*shape)) * mv
    full = [s for s in steps if s[0] == rmax * 2 * rmax]
    per_step = (sum(s[1] for s in full) / len(full)) if full else secs / max(len(steps), 1)
    gpu_per_step = t_gpu / nsteps_gpu / (2 * (d - 1))
    return dict(value=round(fl / secs / 1e12, 6), unit="TFLOP/s", cores=int(cores), kind="port",
                sample="%d consecutive backward local steps of the oracle's tt_eigb sweep (local eigensolve to "
                       "eps2/resid_damp, SVD truncation, interface update) at local size %d, from the GPU "
                       "result moved to site %d; %.1f s, %d matvecs; host BLAS on %d threads (fastest of %s)"
                       % (len(steps), rmax * 2 * rmax, d // 2 + 1, secs, mv, cores, _blas_threads_options()),
                s_per_local_step=round(per_step, 3), gpu_s_per_local_step=round(gpu_per_step, 5),
                sweep_seconds_estimate=round(per_step * 2 * (d - 1), 1),
                note="sweep estimate = mean full-size local step x %d steps (an upper estimate: the steps "
                     "near the chain ends are smaller)" % (2 * (d - 1)))


def run_eigb(args):
    """Heisenberg XXX d=64, 5 lowest eigenpairs, rmax 128, tol 1e-10, rank-8
    N(0,1) start: the FP64 leg and the mixed-precision leg (fp32 matrix-core
    search directions, FP64 refinement), each W warm-up sweeps then K timed
    sweeps of one tt_eigb run, with their final FP64 eigen-residuals."""
    import ttfort
    lib = ttfort.load()
    ttfort.check(lib.ttf_init(0), lib)
    legs = {}
    leg64, ctx = _eigb_leg(args, mixed=False)
    legs["fp64"] = leg64
    if not args.fp64_only:
        legs["mixed"], _ = _eigb_leg(args, mixed=True)
    main = legs.get("mixed", leg64)
    d, rmax = args.eigb_d, args.eigb_rmax
    roof = dominant_kernel_roofline(lib, (rmax, 2, rmax, rmax, 2, rmax, 5, 5), C2_BLOCK)
    cfg = dict(workload="tt_eigb Heisenberg XXX d=%d B=5 rmax=%d tol=1e-10 rank-8 start, FP64 leg and "
                        "mixed leg (fp32 search directions + FP64 refinement)" % (d, rmax),
               d=d, B=5, rmax=rmax, tol=1e-10, legs=legs)
    if "mixed" in legs:
        cfg["mixed_speedup_vs_fp64"] = round(leg64["ms_per_sweep"] / legs["mixed"]["ms_per_sweep"], 3)
        cfg["max_eigen_residual"] = dict(fp64=max(leg64["eigen_residuals"]),
                                         mixed=max(legs["mixed"]["eigen_residuals"]))
    out = dict(metric="sweep time + local-matvec TFLOP/s: mixed/learned vs FP64 GPU at equal residual",
               value=main["tflops"], unit="TFLOP/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=main["ms_per_sweep"], higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="f32+f64" if "mixed" in legs else "f64", data="synthetic", config=cfg, roofline=roof)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_sweep(ctx, args.cpu_budget)
    return out


def _tt_expect(d, n, ra, crA, cr, ry):
    """<y|A|y> / <y|y> for a TT y and TT-matrix A (host numpy contraction)."""
    pa, py = 0, 0
    L = np.ones((1, 1, 1), dtype=np.complex128)     # (ry, ra, ry)
    N = np.ones((1, 1), dtype=np.complex128)
    for k in range(d):
        a = crA[pa:pa + ra[k] * n[k] * n[k] * ra[k + 1]].reshape((ra[k], n[k], n[k], ra[k + 1]), order="F")
        y = cr[py:py + ry[k] * n[k] * ry[k + 1]].reshape((ry[k], n[k], ry[k + 1]), order="F")
        pa += a.size
        py += y.size
        T = np.tensordot(L, y, axes=([2], [0]))             # (x, a, j, Y)
        T = np.tensordot(T, a, axes=([1, 2], [0, 2]))       # (x, Y, i, A)
        L = np.tensordot(y.conj(), T, axes=([0, 1], [0, 2])).transpose(0, 2, 1)   # (X, A, Y)
        N = np.tensordot(y.conj(), np.tensordot(N, y, axes=([1], [0])), axes=([0, 1], [0, 1]))
    return complex(L.reshape(-1)[0] / N.reshape(-1)[0])


def run_ksl(args):
    """Transverse-field Ising chain g=1, d=40, A = -iH, all-up product state
    zero-padded to fixed TT rank 64, dt 0.05, Krylov order 30: two legs, the
    core products of the local exponentials in complex128 and in complex64
    (Krylov basis, Hessenberg exponential and S-steps complex128 in both);
    each W untimed then K timed ztt_ksl steps (one step = one typ-2
    symmetric KSL step; the driver twin takes and returns host cores, so
    each step includes their upload and download, about 5 MB each way).
    The complex64 leg's final state is compared with the complex128 one."""
    import ttfort
    import mpo
    import ttops
    lib = ttfort.load()
    ttfort.check(lib.ttf_init(0), lib)
    d, r, tau, order = 40, 64, 0.05, 30
    crA, ra, n = mpo.tfim_schroedinger(d)
    cr0, ry0 = mpo.product_state_padded(d, r)
    sz0 = int(ra[0] * n[0] * n[0] * ra[1])
    crH = np.concatenate([1j * crA[:sz0], crA[sz0:]])      # H = i A (A carries -i on core 1)
    e0 = _tt_expect(d, n, ra, crH, cr0, ry0)
    legs, finals = {}, {}
    for name, bits in (("c128", 64), ("c64", 32)):
        cr, ry = cr0, ry0
        ttfort.set_local_precision(bits)
        try:
            for _ in range(args.warmup):
                cr, ry, _ = ttfort.ztt_ksl(n, ra, crA, cr, ry, tau, order=order)
            t0 = time.perf_counter()
            mv = mv32 = 0
            fl = 0.0
            for step in range(args.steps):
                cr, ry, st = ttfort.ztt_ksl(n, ra, crA, cr, ry, tau, order=order)
                if args.verb:
                    print("ksl %s step %d: %.3f s, %d matvecs" % (name, step, st["seconds"], st["matvecs"]),
                          file=sys.stderr, flush=True)
                mv += st["matvecs"]
                mv32 += st["matvecs32"]
                fl += st["flops"]
            wall = time.perf_counter() - t0
        finally:
            ttfort.set_local_precision(64)
        e1 = _tt_expect(d, n, ra, crH, cr, ry)
        finals[name] = (cr, ry)
        legs[name] = dict(ms_per_step=round(1e3 * wall / args.steps, 2), tflops=round(fl / wall / 1e12, 4),
                          matvecs=mv, matvecs_c64=mv32, energy_start=round(e0.real, 10),
                          energy_end=round(e1.real, 10), wall_s=round(wall, 2))
    (ca, ra_), (cb, rb_) = finals["c128"], finals["c64"]
    A = ttops.split(ca, ra_, n)
    Bc = ttops.split(cb, rb_, n)
    diff = ttops.norm(ttops.sub(Bc, A)) / ttops.norm(A)
    main = legs["c64"]
    return dict(metric="sweep time + local-matvec TFLOP/s: mixed/learned vs FP64 GPU at equal residual",
                value=main["tflops"], unit="TFLOP/s", n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=main["ms_per_step"], higher_is_better=True, scaling="weak",
                vs_baseline=None, dtype="c64+c128", data="synthetic",
                config=dict(workload="ztt_ksl TFIM g=1 d=40 rank 64 dt=0.05 order 30 all-up start, complex128 "
                                     "and complex64 core products", d=d, r=r, tau=tau, order=order, legs=legs,
                            c64_speedup_vs_c128=round(legs["c128"]["ms_per_step"] / main["ms_per_step"], 3),
                            c64_state_rel_diff_vs_c128=float("%.3e" % diff),
                            host_transfer_bytes_per_step=int(2 * 16 * cr0.size)))


def run_amen(args):
    """Variable-coefficient diffusion -div(a grad u) = 1 in 6D, n=64 per mode,
    a = 1 + 0.5 tanh(sum c_k x_k) (c_k ~ U(-1,1), seed 0) as a rank-3 TT
    coefficient, 7-point stencil TT-matrix of rank 6; rank-1 all-ones rhs,
    rank-2 N(0,1) start (seed 1), tol 1e-8, max rank 64, nrestart 40,
    niters 2, kick 4.  Legs: block-Jacobi 'c' in FP64, block-Jacobi 'c' with
    fp32 Krylov matvecs and FP64 refinement, and the learned preconditioner
    'm' in FP64 and mixed (when its parameters are present; --prec picks a
    single preconditioner).  Each leg: W untimed solves, then K timed solves
    to convergence (a step is one solve); the global FP64 residual of each
    leg's result is computed in TT format on the host."""
    import ttfort
    import mpo
    import ttops
    lib = ttfort.load()
    ttfort.check(lib.ttf_init(0), lib)
    d, n, tol, rmax = 6, 64, 1e-8, 64
    crA, ra, nn, info = mpo.diffusion_7pt(d, n, seed=0, coef_rank=3)
    crY, ry = mpo.ones_tt(d, n)
    crX, rx = mpo.random_tt(d, n, 2, seed=1)
    precs = [args.prec] if args.prec_given else ["c", "m"]
    if "m" in precs:
        import learned
        if learned.load(args.model) is None:
            if args.prec_given:
                raise SystemExit("bench.py: --prec m needs trained parameters at %s "
                                 "(python tt-fort_amd/learned.py)" % args.model)
            precs = [p for p in precs if p != "m"]
    legs = {}
    for prec in precs:
        for mixed in ((args.mixed,) if args.prec_given and args.mixed else (False, True)):
            ttfort.set_local_precision(32 if mixed else 64)

            def solve():
                return ttfort.tt_amen_solve(nn, ra, crA, ry, crY, rx, crX, tol, kickrank=4, nswp=20, prec=prec,
                                            nrestart=40, niters=2, rmax=rmax, verb=args.verb)
            try:
                for _ in range(args.warmup):
                    solve()
                secs, flops, sweeps = 0.0, 0.0, 0
                t0 = time.perf_counter()
                for _ in range(args.steps):
                    res, rx_out, st = solve()
                    secs += st["seconds"]
                    flops += st["flops"]
                    sweeps += st["sweeps"]
                wall = time.perf_counter() - t0
            finally:
                ttfort.set_local_precision(64)
            resid = ttops.residual(crA, ra, nn, res, rx_out, crY, ry)
            legs["%s_%s" % (prec, "mixed" if mixed else "fp64")] = dict(
                ms_per_solve=round(1e3 * wall / args.steps, 1), tflops=round(flops / secs / 1e12, 4),
                sweeps_per_solve=sweeps / args.steps, ms_per_sweep=round(1e3 * secs / max(sweeps, 1), 1),
                max_rank=st["max_rank"], matvecs=st["total_matvecs"], global_residual=float("%.3e" % resid),
                last_local_residual=st["last_res_max"], learned_solves=st["learned_solves"],
                learned_fallbacks=st["learned_fallbacks"], solve_s=round(st["solve_seconds"], 3),
                trunc_s=round(st["trunc_seconds"], 3), phi_s=round(st["phi_seconds"], 3))
    base = legs.get("c_fp64") or next(iter(legs.values()))
    best_key = min(legs, key=lambda k: legs[k]["ms_per_solve"])
    best = legs[best_key]
    return dict(metric="sweep time + local-matvec TFLOP/s: mixed/learned vs FP64 GPU at equal residual",
                value=best["tflops"], unit="TFLOP/s", n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=best["ms_per_solve"], higher_is_better=True,
                scaling="weak", vs_baseline=None, dtype="f32+f64" if "mixed" in best_key else "f64",
                data="synthetic",
                config=dict(workload="tt_amen_solve 6D diffusion n=64 coef rank 3 (operator rank 6) tol 1e-8 rmax 64",
                            d=d, n=n, tol=tol, rmax=rmax, coef_maxerr=info["coef_maxerr"], legs=legs,
                            fastest_leg=best_key,
                            speedup_vs_c_fp64={k: round(base["ms_per_solve"] / v["ms_per_solve"], 3)
                                               for k, v in legs.items()}))


def run_batch(args):
    """Many-instance eigb: independent open XXZ chains, d ~ U{16..48},
    Delta ~ U(0.5, 1.5), J ~ U(0.8, 1.2), max rank r ~ U{32..96} from
    PCG64(seed 5), 3 lowest eigenpairs each, tol 1e-8, random rank-4 start;
    all chains solved concurrently by ttf_tt_eigb_batch.  One step = the
    whole batch."""
    if args.hw_queues:
        # hardware queues the HIP runtime gives this process (read when it
        # initialises): the worker streams share them
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    import ttfort
    import mpo
    lib = ttfort.load()
    ttfort.check(lib.ttf_init(0), lib)
    rng = np.random.default_rng(5)
    probs = []
    for c in range(args.chains):
        d = int(rng.integers(16, 49))
        delta = float(rng.uniform(0.5, 1.5))
        J = float(rng.uniform(0.8, 1.2))
        r = int(rng.integers(32, 97))
        crA, ra, n = mpo.xxz(d, J=J, delta=delta)
        crY, ry = mpo.random_block_tt(d, n, 4, 3, seed=1000 + c)
        probs.append(dict(n=n, ra=ra, crA=crA, crY0=crY, ry=ry, eps=1e-8, rmax=r))
    if args.verb:
        os.environ["TTF_PROGRESS"] = "1"   # a stderr line per finished chain
    for _ in range(args.warmup):
        ttfort.tt_eigb_batch(probs[:min(8, len(probs))], 3, nswp=20, nworkers=args.workers)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = ttfort.tt_eigb_batch(probs, 3, nswp=20, nworkers=args.workers)
    wall = (time.perf_counter() - t0) / args.steps
    flops = sum(sum(o[3]["sweep_flops"]) for o in out)
    chain_s = sum(o[3]["seconds"] for o in out)
    sweeps = [o[3]["sweeps"] for o in out]
    return dict(metric="batch time + local-matvec TFLOP/s (FP64 tt_eigb on %d independent XXZ chains)"
                       % args.chains,
                value=round(flops / wall / 1e12, 4), unit="TFLOP/s", n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=round(1e3 * wall, 1), higher_is_better=True, scaling="weak",
                vs_baseline=None, dtype="f64", data="synthetic",
                config=dict(workload="tt_eigb batch: %d XXZ chains d~U{16..48} r~U{32..96} B=3 tol 1e-8"
                                     % args.chains,
                            chains=args.chains, workers=args.workers,
                            hw_queues=int(os.environ.get("GPU_MAX_HW_QUEUES", "4")), B=3, tol=1e-8,
                            sum_chain_seconds=round(chain_s, 2), concurrency=round(chain_s / wall, 2),
                            mean_sweeps=float(np.mean(sweeps)), max_rank=max(o[3]["max_rank"] for o in out),
                            matvecs=int(sum(o[3]["total_matvecs"] for o in out)), wall_s=round(wall, 2)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="eigb", choices=["eigb", "matvec", "ksl", "amen", "batch"])
    ap.add_argument("--chains", type=int, default=256, help="batch: number of chains")
    ap.add_argument("--workers", type=int, default=16, help="batch: concurrent worker streams")
    ap.add_argument("--hw-queues", type=int, default=None,
                    help="batch: GPU_MAX_HW_QUEUES for this process (4..32; default: the runtime's)")
    ap.add_argument("--prec", default=None,
                    help="AMEn local preconditioner: n, c, l, r, m (learned); default: legs c and m")
    ap.add_argument("--mixed", action="store_true",
                    help="AMEn: fp32 Krylov matvecs with FP64 refinement in the local GMRES")
    ap.add_argument("--model", default=os.path.join(ROOT, "tt-fort_amd", "models", "mlp_precond_n64.npy"),
                    help="parameters of the learned preconditioner (--prec m)")
    ap.add_argument("--eigb-d", type=int, default=64)
    ap.add_argument("--eigb-rmax", type=int, default=128)
    ap.add_argument("--verb", type=int, default=1, help="per-step progress lines on stderr")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--fp64-only", action="store_true", help="eigb: the FP64 leg only")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    args = ap.parse_args()
    args.prec_given = args.prec is not None
    if args.prec is None:
        args.prec = "c"
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the local-solver path has no "
              "multi-GPU mode); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    default_steps = {"eigb": (2, 2), "matvec": (200, 20), "ksl": (10, 1), "amen": (1, 1),
                     "batch": (1, 0)}[args.workload]
    if args.steps is None:
        args.steps = default_steps[0]
    if args.warmup is None:
        args.warmup = default_steps[1]
    out = {"matvec": run_matvec, "eigb": run_eigb, "ksl": run_ksl, "amen": run_amen,
           "batch": run_batch}[args.workload](args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
