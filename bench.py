"""Per-Newton-step linear solve on one MI355X: assembly -> numeric factorization -> solve (+FP64
refinement to <= 1e-10 relative residual) on the 64^3 Kuhn-tet cube (811,200 free DOFs).

One step = one Newton-like re-perturbation (new seed, sigma = 1e-3 h): element Hessians (the 78
local-upper doubles per element the assembly reads, or full 12x12 with --he-layout full12) and the
RHS (= -gradient) for every seed are generated beforehand by the device element kernel and are
resident in HBM; the timed region covers Ax := 0, assembly, value injection + numeric
factorization, solve + refinement.  Symbolic analysis (ordering, supernodes, schedule) runs once
before the timed region and is reported separately.  Prints one JSON line.

The default FP64 run also
  * repeats the steps with an event pair around every launch (per-class device time; the tile
    GEMM's average launch gives `roofline`, assembly and solves give `roofline_assembly` and
    `roofline_solve` against the HBM peak);
  * times the precision configuration (per-element E log-uniform in [1e3, 1e7], nu uniform in
    [0.2, 0.45]) in FP64 and in the --stress-legs forms (FP32 and BF16x3 factor + FP64
    refinement), with refinement steps, forward error against the FP64 solution and speed-up;
  * measures HBM traffic with two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) over a child
    run, each scaled by the same counter on a 2 GiB device copy in that pass;
  * times the CPU baseline (the oracle's supernodal Cholesky on the host) for one 64^3 step and,
    as the median of five, the 16^3 configuration.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 64] [--precision fp64|fp32|bf16|bf16x3]
                  [--stress] [--stress-legs fp32,bf16x3] [--no-cpu] [--no-traffic]
                  [--workload cube|batch|learned|replay]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
PEAK_HBM_GBS = 8000.0
# kernel classes of mfs_kernel_profile
KCLASSES = ["fill_L", "inject", "unused", "extend", "potrf", "trsm", "update", "syrk", "fgather",
            "fdiag", "fupdate", "bpartial", "bdiag", "spmv", "maps", "assemble", "blkinv", "bgather",
            "convert", "refine", "fused", "panel"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n_sample, flops_target, threads, reps=1):
    """The oracle's multifrontal supernodal Cholesky (oracle/supernodal.c: relaxed supernodes,
    per-supernode POTRF/TRSM/SYRK from the OpenBLAS that ships with scipy, elimination tree level
    by level over `threads` OpenMP threads with a single-threaded BLAS, the BLAS threaded only at
    the top levels that have fewer supernodes than threads) timed on this host for one Newton
    step of the n_sample^3 cube with the GPU path's ordering: serial reference-semantics assembly,
    value injection into P A P^T, numeric factorization, forward and backward solves.  The median
    of `reps` timed steps is reported; it is scaled to the full workload by the ratio of
    factorization flops (1 when n_sample is the full size)."""
    from tests import oracle_lib as orc
    from meshfemsparse_amd import workload as wl
    from meshfemsparse_amd import _lib
    tets, Xr = orc.cube_mesh(n_sample)
    X = orc.perturb(n_sample, Xr, 0)
    He, g, _ = wl.neohookean(Xr, X, tets)
    nb = len(Xr)
    Ap, Ai = orc.pattern(tets, nb)
    pins = wl.z0_pins(n_sample)
    iperm = np.empty(nb, np.int64)
    st = np.zeros(8, np.int64)
    fl = np.zeros(2)
    _lib.call("mfs_analyze_host", _lib.ptr(Ap, _lib.c_i64), _lib.ptr(Ai, _lib.c_i64), nb, _lib.ptr(pins, _lib.c_i64),
              len(pins), _lib.ptr(iperm, _lib.c_i64), _lib.ptr(st, _lib.c_i64), _lib.ptr(fl, _lib.c_double))
    iperm = iperm[:st[0]]
    nred = len(iperm)
    pos = np.empty(nred, np.int32)
    pos[iperm] = np.arange(nred)
    red = -np.ones(nb, np.int64)
    red[(n_sample + 1) ** 2:] = np.arange(nred)
    free = np.arange(3 * (n_sample + 1) ** 2, 3 * nb)
    ip = (3 * iperm[:, None] + np.arange(3)).ravel()
    bp = -g[free][ip]
    # symbolic analysis and the first touch of the factor storage stay outside the timing (the
    # reference refactors into fixed storage)
    Cp, Ci, Cx = orc.permuted_lower(Ap, Ai, nb, orc.assemble(Ap, Ai, nb, tets, He), red, pos)
    S = orc.Supernodal(Cp, Ci, relax=32)
    S.factor(Cx, threads)
    runs = []
    for _ in range(reps):
        t0 = time.perf_counter()
        Ax = orc.assemble(Ap, Ai, nb, tets, He)
        t1 = time.perf_counter()
        Cp, Ci, Cx = orc.permuted_lower(Ap, Ai, nb, Ax, red, pos)
        S.factor(Cx, threads)
        t2 = time.perf_counter()
        x = S.solve(bp)
        t3 = time.perf_counter()
        runs.append((t3 - t0, t1 - t0, t2 - t1, t3 - t2))
    if nred * 3 <= 3000:      # residual check on small samples
        A = orc.dense_from_lower(Cp, Ci, Cx)
        assert np.linalg.norm(A @ x - bp) <= 1e-10 * np.linalg.norm(bp)
    runs.sort()
    tot, ta, tf, ts = runs[len(runs) // 2]
    scale = flops_target / fl[0]
    return dict(t_assembly=ta, t_factor=tf, t_solve=ts, sample_flops=fl[0], cpu_flops=S.info["flops"],
                factor_gflops=S.info["flops"] / tf / 1e9, scale=scale, s_sample=tot, s_per_step=tot * scale,
                supernodes=S.info["supernodes"], spread=(runs[-1][0] - runs[0][0]) / tot if reps > 1 else 0.0)


def cpu_threads():
    """Host threads for the CPU baseline: the CPUs this process may run on, capped by
    OMP_NUM_THREADS when it is set (the share of the host that belongs to this GPU)."""
    aff = len(os.sched_getaffinity(0))
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(aff, env) if env > 0 else aff), aff


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def bench_batch(args):
    """512 independent cubes (per-side n ~ U{6..20}, E log-uniform [1e3, 1e6] per element, nu 0.3,
    z = 0 pinned) factored together as one block-diagonal system: one host symbolic analysis
    (each mesh its own nested dissection), an autotuned schedule (fused small-supernode kernel
    bound x big-tile threshold), the numeric factorization captured in a hipGraph, and per-mesh
    refinement to <= 1e-10.  One step = new perturbation of every mesh."""
    import torch
    import meshfemsparse_amd as mfs
    from meshfemsparse_amd import workload as wl
    dev = torch.device("cuda:0")
    nseeds = args.steps + args.warmup
    t0 = time.perf_counter()
    B = wl.batch_meshes(args.batch, 6, 20, seed=0)
    nb = int(B["node_off"][-1])
    tets = B["tets"]
    asm = mfs.SystemAssembler(nb)
    H = asm.blockSparsityPattern(tets)
    H.setZero()
    rng = np.random.default_rng(12345)
    Ee = 10 ** rng.uniform(3, 6, len(tets))
    log(f"[bench] batch of {args.batch}: {len(tets)} tets, {nb} nodes, {H.nz} blocks; {time.perf_counter() - t0:.1f}s")
    He_d, b_d = [], []
    t0 = time.perf_counter()
    gasm = mfs.DeviceAssembler(H, tets)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    Xr_d, Ee_d = T(B["X_rest"]), T(Ee)
    for sd in range(nseeds):
        He, g12 = gasm.neohookean(Xr_d, T(wl.batch_perturb(B, sd)), E_per_elem=Ee_d, nu=0.3)
        g = torch.zeros(3 * nb, dtype=torch.float64, device=dev)
        gasm.assembleGradient(g12, g)
        He_d.append(He)
        b_d.append(-g)
        del g12
    torch.cuda.synchronize()
    log(f"[bench] generated + uploaded {nseeds} seeds in {time.perf_counter() - t0:.1f}s")
    single = args.precision != "fp64"
    f = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, single,
                                     args.precision if args.precision.startswith("bf16") else None)
    dasm = asm._assembler(H, np.ascontiguousarray(tets, np.int32))
    dasm.zero(H.Ax)
    dasm.assemble(He_d[0], H.Ax)
    t0 = time.perf_counter()
    f.factorizeSymbolic(H, B["pins"])
    t_sym = time.perf_counter() - t0
    t0 = time.perf_counter()
    tune = f.autotune(H, B["pins"], fuse=(0, 192, 384, 768), tiles=(128, 64), reps=3)
    t_tune = time.perf_counter() - t0
    f.setGraphCapture(not args.no_graph)
    log(f"[bench] symbolic {t_sym:.2f}s, autotune {t_tune:.1f}s: {tune}")
    flops = f.getFlopEstimate()

    def step(sd):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        e[0].record()
        dasm.zero(H.Ax)
        dasm.assemble(He_d[sd], H.Ax)
        e[1].record()
        f.factorizeNumeric(H)
        e[2].record()
        x, it, rr, per = f.solveRefinedSegments(H, b_d[sd], B["seg"], tol=1e-10, maxIters=30 if single else 10)
        e[3].record()
        return e, it, rr

    for sd in range(args.warmup):
        step(sd)
    torch.cuda.synchronize()
    recs = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        recs.append(step(args.warmup + k))
        log(f"[bench] step {k} done")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_asm = np.mean([r[0][0].elapsed_time(r[0][1]) for r in recs]) / 1e3
    t_fac = np.mean([r[0][1].elapsed_time(r[0][2]) for r in recs]) / 1e3
    t_sol = np.mean([r[0][2].elapsed_time(r[0][3]) for r in recs]) / 1e3
    # per-class device time from a pass with an event pair around every launch (eager launches,
    # outside the timed region)
    kms = {}
    if not args.no_kernel_timing:
        from meshfemsparse_amd import _lib
        f.setGraphCapture(False)
        ms = np.zeros(len(KCLASSES))
        cnt = np.zeros(len(KCLASSES), np.int64)
        _lib.call("mfs_kernel_profiling", f._h, 1, 1)
        for k in range(args.steps):
            step(args.warmup + k)
        torch.cuda.synchronize()
        _lib.call("mfs_kernel_profiling", f._h, 0, 0)
        _lib.call("mfs_kernel_profile", f._h, _lib.ptr(ms, _lib.c_double), _lib.ptr(cnt, _lib.c_i64))
        kms = {c: round(v / args.steps, 3) for c, v in zip(KCLASSES, ms) if v > 0}
        log("[bench] kernel-timing pass done")
    out = {
        "metric": "batched Newton linear-solve s/step (512 meshes, every mesh to <=1e-10 rel. residual); factor TFLOP/s",
        "value": wall / args.steps, "unit": "s/step", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if not single else "f32 factor/solve, f64 residual+refinement",
        "data": "synthetic (seeded perturbations of each cube, neo-Hookean He and RHS from the device element kernel)",
        "config": {"workload": f"{args.batch} unit cubes, per-side n ~ U{{6..20}} hexes x 6 Kuhn tets, E log-uniform "
                               f"[1e3,1e6], nu 0.3, z=0 pinned, one block-diagonal system",
                   "free_dofs": 3 * nb - len(B["pins"]), "tets": len(tets), "factor_nnz": f.getFactorNNZ(),
                   "factor_flops": flops, "factor_flops_unrelaxed": f.flopBreakdown()["unrelaxed"],
                   "schedule": f.scheduleStats(), "autotune": {
                       "fuse_degree": tune["fuse_degree"], "bucket_tiles": tune["bucket_tiles"],
                       "ms": {"fuse": {str(a): round(v, 3) for a, v in tune["ms"]["fuse"].items()},
                              "bucket_tiles": [{str(t): round(v, 3) for t, v in d.items()}
                                               for d in tune["ms"]["bucket_tiles"]]}},
                   "graph": not args.no_graph},
        "factor_tflops": flops / t_fac / 1e12, "t_assembly_s": t_asm, "t_factor_s": t_fac, "t_solve_s": t_sol,
        "refine_iters": [r[1] for r in recs], "relres_max": max(r[2] for r in recs), "symbolic_s": t_sym,
        "autotune_s": t_tune, "kernel_ms_per_step": kms,
    }
    print(json.dumps(out), flush=True)


def bench_learned(args):
    """Learned block-IC preconditioner: train the message-passing network on the device on
    `--train` generated Hessians of the 48^3 cube (log E a Gaussian random field, mean ln 1e5,
    std 1.5, correlation length 0.1; nu 0.3; sigma = 1e-3 h), then solve held-out 64^3 steps by
    FP32 PCG + FP64 refinement to 1e-10 with the trained factor, falling back to the supernodal
    Cholesky after 200 CG iterations; block-Jacobi PCG and the untrained network are the
    baselines.  One step = one held-out 64^3 solve."""
    import torch
    import meshfemsparse_amd as mfs
    from meshfemsparse_amd import workload as wl, learned
    dev = torch.device("cuda:0")

    def setup(N):
        tets, Xr = wl.cube_mesh(N)
        asm = mfs.SystemAssembler(len(Xr))
        H = asm.blockSparsityPattern(tets)
        H.setZero()
        da = mfs.DeviceAssembler(H, tets)
        return dict(N=N, tets=tets, Xr=Xr, H=H, da=da, pins=wl.z0_pins(N),
                    Xr_d=torch.from_numpy(Xr).to(dev), tets_d=torch.from_numpy(tets).to(dev))

    def sample(m, seed):
        """assemble the Hessian of mesh m for (perturbation, stiffness field) seed; returns -grad"""
        X = torch.from_numpy(wl.perturb(m["N"], m["Xr"], seed)).to(dev)
        Ee = wl.grf_log_E(m["Xr_d"], m["tets_d"], seed).contiguous()
        He, g12 = m["da"].neohookean(m["Xr_d"], X, E_per_elem=Ee, nu=0.3)
        m["H"].Ax.zero_()
        m["da"].assemble(He, m["H"].Ax)
        g = torch.zeros(3 * len(m["Xr"]), dtype=torch.float64, device=dev)
        m["da"].assembleGradient(g12, g)
        return -g

    tr = setup(args.learn_n)
    lp = learned.LearnedPreconditioner(tr["H"], tr["pins"], seed=1)
    rng = np.random.default_rng(0)
    t0 = time.perf_counter()
    losses = []
    for k in range(args.train):
        sample(tr, 10000 + k)
        losses.append(lp.trainStep(tr["H"], nprobe=4, lr=args.lr if k < args.train // 2 else args.lr / 4, rng=rng))
        if k % 100 == 0:
            log(f"[bench] train step {k}: loss {losses[-1]:.4f} ({time.perf_counter() - t0:.1f}s)")
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    params = lp.params()
    del lp, tr
    ev = setup(args.n)
    lpe = learned.LearnedPreconditioner(ev["H"], ev["pins"], seed=1)
    f = mfs.make_cholesky_factorizer()
    f.factorizeSymbolic(ev["H"], ev["pins"])
    untrained = lpe.params()
    res = {"trained": [], "untrained": [], "jacobi": []}
    times = []
    for k in range(args.warmup + args.steps):
        b = sample(ev, 20000 + k)
        for name, prm, pc in (("jacobi", None, "jacobi"), ("untrained", untrained, "learned"),
                              ("trained", params, "learned")):
            if prm is not None:
                lpe.setParams(prm)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            x, it, outer, rr, fb = lpe.solve(ev["H"], b, precond=pc, tol=1e-10, maxPCG=args.max_pcg, fallback=f)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t1
            if k >= args.warmup:
                res[name].append(dict(cg_iters=it, refinements=outer, relres=rr, fallback=fb, s=dt))
                if name == "trained":
                    times.append(dt)
        log(f"[bench] held-out step {k}: " + ", ".join(f"{n}: {r[-1]['cg_iters'] if r else '-'} it"
                                                      for n, r in res.items()))
    # CG iterations without the fallback limit, to report what each preconditioner needs
    b = sample(ev, 20000)
    full = {}
    for name, prm, pc in (("jacobi", None, "jacobi"), ("trained", params, "learned")):
        if prm is not None:
            lpe.setParams(prm)
        try:
            _, it, outer, rr, fb = lpe.solve(ev["H"], b, precond=pc, tol=1e-10, maxPCG=args.max_pcg_free)
            full[name] = dict(cg_iters=it, refinements=outer, relres=rr)
        except mfs.MFSError:
            full[name] = dict(cg_iters=None, note=f"not converged in {args.max_pcg_free} CG iterations")
    mean = lambda name, key: float(np.mean([r[key] for r in res[name]]))
    out = {
        "metric": "learned-preconditioner Newton linear-solve s/step (64^3 held-out, <=1e-10 rel. residual)",
        "value": float(np.mean(times)), "unit": "s/step", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": float(np.mean(times)) * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 PCG (bf16-rounded network weights), f64 refinement",
        "data": "synthetic (GRF stiffness fields, seeded perturbations; device element kernel)",
        "config": {"workload": f"train {args.train} Hessians of the {args.learn_n}^3 cube, evaluate held-out "
                               f"{args.n}^3 steps; FP32 PCG + FP64 refinement to 1e-10, Cholesky fallback after "
                               f"{args.max_pcg} CG iterations", "network": lpe.info()},
        "train_s": t_train, "train_loss_first": float(np.mean(losses[:20])), "train_loss_last": float(np.mean(losses[-20:])),
        "held_out": {n: dict(cg_iters_mean=mean(n, "cg_iters"), fallback_rate=mean(n, "fallback"), s_mean=mean(n, "s"))
                     for n in res},
        "cg_iters_without_limit": full,
    }
    print(json.dumps(out), flush=True)


def bench_replay(args):
    """benchmark_linear_systems' replay (tests/benchmark_linear_systems.cc:216-342) of a recorded
    Newton sequence: the n^3 cube's pattern and `--steps` Hessians (new perturbation each) are
    written in the reference's MatrixRecorder layout, then read back and replayed through the
    GPU factorizer (symbolic per pattern; numeric factorization timed per matrix; backward error of
    a random-solution solve).  --replay-dir replays an existing recording instead."""
    import tempfile
    import torch
    import meshfemsparse_amd as mfs
    from meshfemsparse_amd import workload as wl
    from meshfemsparse_amd.recorder import MatrixRecorder, replay
    dev = torch.device("cuda:0")
    tmp = None
    d = args.replay_dir
    if not d:
        tmp = tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp"))
        d = tmp.name
        N = args.n
        tets, Xr = wl.cube_mesh(N)
        H = mfs.SystemAssembler(len(Xr)).blockSparsityPattern(tets)
        H.setZero()
        da = mfs.DeviceAssembler(H, tets)
        rec = MatrixRecorder(d)
        rec.countSparsityUpdateCall()
        rec.recordSymbolic(H, wl.z0_pins(N))
        Xr_d = torch.from_numpy(Xr).to(dev)
        for k in range(args.steps):
            He, _ = da.neohookean(Xr_d, torch.from_numpy(wl.perturb(N, Xr, k)).to(dev), gradients=False)
            H.Ax.zero_()
            da.assemble(He, H.Ax)
            rec.recordNumeric(H)
            del He
        log(f"[bench] recorded {args.steps} matrices of the {N}^3 cube in {d}")
    f = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, args.precision != "fp64",
                                     args.precision if args.precision.startswith("bf16") else None)
    t0 = time.perf_counter()
    out = replay(d, f, repeats=1 + args.warmup, device=dev)
    wall = time.perf_counter() - t0
    nf = out["numeric_s"]
    res = {
        "metric": "recorded-sequence replay: numeric factorization s per matrix (benchmark_linear_systems)",
        "value": float(np.mean(nf)) if nf else None, "unit": "s/factorization", "n_gpus": 1,
        "steps": len(nf), "warmup": 0, "ms_per_step": float(np.mean(nf)) * 1e3 if nf else None,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "recorded synthetic Newton sequence (reference MatrixRecorder layout)",
        "config": {"workload": f"replay of {args.replay_dir or f'{args.n}^3 cube, {args.steps} recorded Hessians'}"},
        "posdef": out["posdef"], "indef": out["indef"], "symbolic": out["symbolic"],
        "backward_error_max": max(out["backward_error"]) if out["backward_error"] else None,
        "replay_wall_s": wall,
    }
    print(json.dumps(res), flush=True)
    if tmp:
        tmp.cleanup()


def cube_inputs(dev, N, nseeds, stress, layout):
    """Pattern, assembler, pins and per-seed element Hessians (layout) and RHS = -gradient of the
    N^3 cube, generated by the device element kernel and resident in HBM."""
    import torch
    import meshfemsparse_amd as mfs
    from meshfemsparse_amd import workload as wl
    tets, Xr = wl.cube_mesh(N)
    nb = len(Xr)
    asm = mfs.SystemAssembler(nb)
    H = asm.blockSparsityPattern(tets)
    H.setZero()
    Ee = nue = None
    if stress:
        rng = np.random.default_rng(12345)
        Ee = 10 ** rng.uniform(3, 7, len(tets))
        nue = rng.uniform(0.2, 0.45, len(tets))
    gasm = mfs.DeviceAssembler(H, tets)
    T = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    Xr_d, Ee_d, nue_d = T(Xr), T(Ee), T(nue)
    He_d, b_d = [], []
    for sd in range(nseeds):
        He, g12 = gasm.neohookean(Xr_d, T(wl.perturb(N, Xr, seed=sd)), E_per_elem=Ee_d, nu_per_elem=nue_d,
                                  layout=layout)
        g = torch.zeros(3 * nb, dtype=torch.float64, device=dev)
        gasm.assembleGradient(g12, g)
        He_d.append(He)
        b_d.append(-g)
        del g12
    torch.cuda.synchronize()
    return dict(tets=tets, H=H, pins=wl.z0_pins(N), asm=gasm, He=He_d, b=b_d, nb=nb)


def make_factorizer(precision):
    import meshfemsparse_amd as mfs
    return mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, precision != "fp64",
                                        precision if precision.startswith("bf16") else None)


MAX_REFINE = {"fp64": 10, "fp32": 30, "bf16x3": 30, "bf16": 80}


def run_steps(inp, f, layout, seeds, max_iters):
    """One Newton step per seed (Ax := 0, assembly, numeric factorization, solve + refinement to
    1e-10), timed by the wall clock around all of them; per-stage device times from events on
    the current stream (the library launches on it)."""
    import torch
    H, asm = inp["H"], inp["asm"]
    recs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for sd in seeds:
        e = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        e[0].record()
        asm.zero(H.Ax)
        e[1].record()
        asm.assemble(inp["He"][sd], H.Ax, layout)
        e[2].record()
        f.factorizeNumeric(H)
        e[3].record()
        x, it, rr = f.solveRefined(H, inp["b"][sd], tol=1e-10, maxIters=max_iters)
        e[4].record()
        recs.append((e, it, rr, x))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = lambda a, b: float(np.mean([r[0][a].elapsed_time(r[0][b]) for r in recs]))
    return dict(wall=wall, s_per_step=wall / len(seeds), t_zero_ms=ms(0, 1), t_assembly_ms=ms(1, 2),
                t_factor_s=ms(2, 3) / 1e3, t_solve_s=ms(3, 4) / 1e3, iters=[r[1] for r in recs],
                relres_max=max(r[2] for r in recs), x=[r[3] for r in recs])


def stress_legs(args, dev, layout):
    """The precision configuration: the 64^3 cube with per-element E log-uniform in [1e3, 1e7]
    and nu uniform in [0.2, 0.45], timed in FP64 and in each reduced-precision form (factor +
    FP64 refinement to 1e-10) on the same inputs; forward error of every refined solution against
    the FP64 factor's solution of the same step."""
    import torch
    forms = [p for p in args.stress_legs.split(",") if p]
    nseeds = args.stress_steps + 1
    inp = cube_inputs(dev, args.n, nseeds, True, layout)
    seeds = list(range(1, nseeds))
    out = {}
    x64 = None
    for form in ["fp64"] + forms:
        f = make_factorizer(form)
        f.factorizeSymbolic(inp["H"], inp["pins"])
        run_steps(inp, f, layout, [0], MAX_REFINE[form])                   # warmup
        r = run_steps(inp, f, layout, seeds, MAX_REFINE[form])
        leg = dict(s_per_step=r["s_per_step"], t_factor_s=r["t_factor_s"], t_solve_s=r["t_solve_s"],
                   refine_iters=r["iters"], relres_max=r["relres_max"],
                   factor_tflops=f.getFlopEstimate() / r["t_factor_s"] / 1e12)
        if form == "fp64":
            x64 = [x.clone() for x in r["x"]]
        else:
            leg["forward_error_vs_fp64_max"] = max(float(torch.linalg.norm(x - y) / torch.linalg.norm(y))
                                                   for x, y in zip(r["x"], x64))
            leg["speedup_vs_fp64"] = out["fp64"]["s_per_step"] / r["s_per_step"]
        out[form] = leg
        log(f"[bench] stress {form}: {r['s_per_step']:.4f} s/step, refinement {r['iters']}")
        del f, r
    del inp
    torch.cuda.synchronize()
    return out


def _counter_sums(csv_dir):
    """Per kernel name: (dispatches, counter sum in KiB, summed duration ns) from a rocprofv3
    counter-collection CSV (one counter per pass)."""
    import csv
    import glob
    files = glob.glob(os.path.join(csv_dir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError("no counter-collection CSV")
    sums = {}
    for fn in files:
        for r in csv.DictReader(open(fn)):
            name = r["Kernel_Name"]
            v = float(r["Counter_Value"])
            d = float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
            c, tot, dur, mx = sums.get(name, (0, 0.0, 0.0, 0.0))
            sums[name] = (c + 1, tot + v, dur + d, max(mx, v))
    return sums


def traffic_passes(args, precision):
    """HBM traffic of the kernels from two rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE) over
    a child run of this script (--traffic-probe: a 2 GiB device copy, then one Newton step), each
    counter scaled by what it reads for the copy (2 GiB read, 2 GiB written) in the same pass.
    Returns per kernel family bytes per dispatch, or None with the reason."""
    import shutil
    import signal
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not found"
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix=f"mfs_{counter.lower()}_", dir=os.environ.get("TMPDIR", "/tmp"))
        cmd = [prof, "--pmc", counter, "--kernel-trace", "--output-format", "csv", "-d", d, "-o", "run", "--",
               sys.executable, os.path.abspath(__file__), "--traffic-probe", "--n", str(args.n), "--precision",
               precision, "--he-layout", args.he_layout]
        log(f"[bench] traffic pass {counter} ...")
        p = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, start_new_session=True,
                             cwd=os.environ.get("TMPDIR", "/tmp"))
        try:
            _, err = p.communicate(timeout=args.traffic_timeout)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.wait()
            shutil.rmtree(d, ignore_errors=True)
            return None, f"{counter} pass exceeded {args.traffic_timeout}s"
        try:
            res[counter] = _counter_sums(d)
        except Exception as e:   # the profiler may end with a signal after writing complete results
            shutil.rmtree(d, ignore_errors=True)
            return None, f"{counter} pass (rc {p.returncode}): {e}: {err.decode(errors='replace')[-300:]}"
        shutil.rmtree(d, ignore_errors=True)
    calib = {}
    for counter in res:
        cands = [(v[3], name) for name, v in res[counter].items() if "copyBuffer" in name]
        if not cands:
            return None, "calibration copy not found"
        calib[counter] = (2 << 30) / (max(cands)[0] * 1024.0)
    fams = {"k_gemm": lambda n: n.startswith("void mfs::k_gemm<") or n.startswith("void mfs::k_gemm_dl<"),
            "k_gemm_bf": lambda n: n.startswith("void mfs::k_gemm_bf<"),
            "k_assemble": lambda n: "mfs::k_assemble" in n,
            "solve": lambda n: any(k in n for k in ("mfs::k_sf", "mfs::k_sb", "mfs::k_fgather", "mfs::k_bgather")),
            "k_fill": lambda n: "mfs::k_fill" in n, "k_extend": lambda n: "mfs::k_extend<" in n,
            "k_panel_pull": lambda n: "mfs::k_panel_pull<" in n}
    out = {"scale_fetch": calib["FETCH_SIZE"], "scale_write": calib["WRITE_SIZE"]}
    for fam, pred in fams.items():
        rb = wb = dur = 0.0
        cnt = 0
        for name, (c, tot, du, _) in res["FETCH_SIZE"].items():
            if pred(name):
                rb += tot * 1024.0 * calib["FETCH_SIZE"]
                dur += du
                cnt += c
        for name, (c, tot, _, _) in res["WRITE_SIZE"].items():
            if pred(name):
                wb += tot * 1024.0 * calib["WRITE_SIZE"]
        if cnt:
            out[fam] = dict(dispatches=cnt, read_bytes=rb, write_bytes=wb, bytes_per_dispatch=(rb + wb) / cnt,
                            implied_GBps=(rb + wb) / dur if dur > 0 else None)
    return out, None


def traffic_probe(args):
    """Child of traffic_passes: a 2 GiB device copy (the counters' calibration), then one Newton
    step of the cube (assembly, factorization, solve) with symbolic analysis beforehand."""
    import torch
    dev = torch.device("cuda:0")
    a = torch.ones(1 << 28, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    layout = 1 if args.he_layout == "upper78" else 0
    inp = cube_inputs(dev, args.n, 1, False, layout)
    f = make_factorizer(args.precision)
    f.factorizeSymbolic(inp["H"], inp["pins"])
    run_steps(inp, f, layout, [0], MAX_REFINE[args.precision])
    torch.cuda.synchronize()


def hbm_copy_gbs(dev, gib=2, reps=5):
    """Bytes moved (read + write) per second by a plain 2 GiB device-to-device copy: what this
    GPU's HBM delivers to a streaming kernel, beside the published 8 TB/s."""
    import torch
    a = torch.empty(gib << 27, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2 * (gib << 30) * reps / (e0.elapsed_time(e1) / 1e3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return gbs


def bench_cube(args):
    import torch
    from meshfemsparse_amd import _lib
    dev = torch.device("cuda:0")
    layout = _lib.MFS_HE_UPPER78 if args.he_layout == "upper78" else _lib.MFS_HE_FULL12
    N = args.n
    nseeds = args.steps + args.warmup
    t0 = time.perf_counter()
    inp = cube_inputs(dev, N, nseeds, args.stress, layout)
    H, pins, nb = inp["H"], inp["pins"], inp["nb"]
    log(f"[bench] mesh N={N}: {len(inp['tets'])} tets, {nb} nodes, {H.nz} blocks; inputs for {nseeds} seeds "
        f"in {time.perf_counter() - t0:.1f}s")
    single = args.precision != "fp64"
    f = make_factorizer(args.precision)
    t0 = time.perf_counter()
    f.factorizeSymbolic(H, pins)
    t_sym = time.perf_counter() - t0
    st = f.supernodeStats()
    flops = f.getFlopEstimate()
    nred = f.n_reduced()
    nnz = f.getFactorNNZ()
    log(f"[bench] symbolic {t_sym:.2f}s: n_red={nred} nnz(L)={nnz:.3e} flops={flops:.3e} {st}")
    f.setGraphCapture(args.graph)
    maxit = MAX_REFINE[args.precision]
    run_steps(inp, f, layout, list(range(args.warmup)), maxit)
    log("[bench] warmup done")
    r = run_steps(inp, f, layout, list(range(args.warmup, nseeds)), maxit)
    log(f"[bench] {args.steps} timed steps: {r['s_per_step']:.4f} s/step")
    wall = r["wall"]
    timed = {k: r[k] for k in ("t_zero_ms", "t_assembly_ms", "t_factor_s", "t_solve_s", "iters", "relres_max")}
    ntets = len(inp["tets"])
    # the same steps again with a HIP event pair around every kernel launch (per-class device
    # time and the dominant kernel's average launch duration); kept out of the timed region
    # because the events add a few microseconds between dependent launches
    kt = not args.no_kernel_timing
    ms = np.zeros(len(KCLASSES))
    cnt = np.zeros(len(KCLASSES), np.int64)
    work = np.zeros(len(KCLASSES))
    gbytes = np.zeros(len(KCLASSES))
    asm_ms = None
    nsolves = 0
    if kt:
        f.setGraphCapture(False)            # per-launch events need the launches themselves
        _lib.call("mfs_kernel_profiling", f._h, 1, 1)
        rk = run_steps(inp, f, layout, list(range(args.warmup, nseeds)), maxit)
        torch.cuda.synchronize()
        _lib.call("mfs_kernel_profiling", f._h, 0, 0)
        _lib.call("mfs_kernel_profile", f._h, _lib.ptr(ms, _lib.c_double), _lib.ptr(cnt, _lib.c_i64))
        _lib.call("mfs_kernel_profile_work", f._h, _lib.ptr(work, _lib.c_double))
        _lib.call("mfs_kernel_profile_bytes", f._h, _lib.ptr(gbytes, _lib.c_double))
        nsolves = sum(1 + i for i in rk["iters"])
        asm_ms = rk["t_assembly_ms"]
        log("[bench] kernel-timing pass done")
    s_per_step = wall / args.steps
    factor_tflops = flops / timed["t_factor_s"] / 1e12
    fs = 4 if single else 8
    # roofline of the dominant kernel: the tile GEMM (TRSM + in-panel updates + Schur updates);
    # for the BF16 variants the bfloat16 Schur-update kernel (in-panel + update-matrix classes)
    bf = args.precision.startswith("bf16")
    gcls = [6, 7] if bf else [5, 6, 7]
    gemm_ms = float(sum(ms[c] for c in gcls))
    gemm_launches = int(sum(cnt[c] for c in gcls))
    # algorithmic GEMM flops of a step (supernodal count minus the diagonal-tile POTRF share),
    # apportioned to the classes by their executed tile flops
    potrf_flops = f.flopBreakdown()["potrf"]
    unrelaxed_flops = f.flopBreakdown()["unrelaxed"]   # sum of c_j^2 over the columns of L
    exec_all = float(work[5] + work[6] + work[7] + work[KCLASSES.index("panel")])
    share = float(sum(work[c] for c in gcls)) / exec_all if exec_all > 0 else 1.0
    gemm_work = (flops - potrf_flops) * share * args.steps
    if kt and gemm_launches:
        avg_launch_s = gemm_ms / 1e3 / gemm_launches
        flops_per_launch = gemm_work / gemm_launches
        achieved = flops_per_launch / avg_launch_s / 1e12
        noreuse_per_launch = float(sum(gbytes[c] for c in gcls)) / gemm_launches
    else:
        avg_launch_s, achieved, noreuse_per_launch = None, factor_tflops, None
    peak = PEAK_TFLOPS["bf16" if bf else args.precision]
    kname = (f"k_gemm_bf ({'hi+lo split ' if args.precision == 'bf16x3' else ''}bfloat16 MFMA Schur updates, "
             f"FP32 accumulation; algorithmic flops)" if bf else
             (f"k_gemm_dl (FP64 MFMA tile GEMM, operands loaded straight into LDS: TRSM, in-panel update, "
              f"Schur update)" if args.precision == "fp64" else
              f"k_gemm_dl / k_gemm ({args.precision.upper()} MFMA tile GEMM, 128x128 tiles loaded straight into "
              f"LDS: TRSM, in-panel update, Schur update)"))
    roofline = {"bound": "mfma", "kernel": kname,
                "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None,
                "avg_launch_ms": None if avg_launch_s is None else avg_launch_s * 1e3,
                "launches_per_step": gemm_launches / max(args.steps, 1),
                "noreuse_bytes_per_launch": noreuse_per_launch}
    # HBM rooflines of assembly and solves (algorithmic bytes / measured time / 8 TB/s)
    nz = H.nz
    per_elem = 78 if layout == _lib.MFS_HE_UPPER78 else 144
    asm_bytes = ntets * (per_elem * 8 + 10 * 4) + 4 * (nz + 1) + nz + 2 * 9 * nz * 8
    roof_asm = None
    if kt and asm_ms:
        a_gbs = asm_bytes / (asm_ms / 1e3) / 1e9
        roof_asm = {"bound": "hbm", "kernel": "k_assemble (slot-map gather, fixed element order per 3x3 slot)",
                    "achieved": a_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": a_gbs / PEAK_HBM_GBS,
                    "algorithmic_bytes": asm_bytes, "avg_launch_ms": asm_ms,
                    "bytes_basis": f"He {per_elem} doubles + 10 slot-map entries per element, segment pointers "
                                   f"and diagonal flags per slot, Ax read + write"}
    solve_cls = [KCLASSES.index(c) for c in ("fgather", "fdiag", "fupdate", "bpartial", "bdiag", "bgather")]
    roof_solve = None
    if kt and nsolves:
        t_solve_k = float(sum(ms[c] for c in solve_cls)) / nsolves / 1e3
        sbytes = 2.0 * nnz * fs
        s_gbs = sbytes / t_solve_k / 1e9
        roof_solve = {"bound": "hbm", "kernel": "supernodal forward + backward sweeps (one right-hand side)",
                      "achieved": s_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": s_gbs / PEAK_HBM_GBS,
                      "algorithmic_bytes": sbytes, "kernel_ms_per_solve": t_solve_k * 1e3,
                      "bytes_basis": "2 nnz(L) x element size (the factor read once per sweep)"}
    kernel_ms = {k: round(float(m) / args.steps, 3) for k, m in zip(KCLASSES, ms) if cnt[KCLASSES.index(k)]}
    del r
    # the precision configuration (FP64 vs FP32 / BF16x3 + refinement on the stressed cube)
    legs = None
    if args.stress_legs and not args.stress and args.precision == "fp64":
        f.clearFactors()
        legs = stress_legs(args, dev, layout)
    del inp
    del f
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    copy_gbs = hbm_copy_gbs(dev)
    for rf in (roof_asm, roof_solve):
        if rf:
            rf["copy_gbs"] = copy_gbs
            rf["frac_of_copy"] = rf["achieved"] / copy_gbs
    traffic, why = (None, "skipped (--no-traffic)")
    if not args.no_traffic:
        traffic, why = traffic_passes(args, args.precision)
    if traffic:
        fam = "k_gemm_bf" if bf else "k_gemm"
        if fam in traffic:
            roofline["traffic"] = traffic[fam]["bytes_per_dispatch"]
            if noreuse_per_launch:
                roofline["traffic_vs_noreuse"] = traffic[fam]["bytes_per_dispatch"] / noreuse_per_launch
        if roof_asm and "k_assemble" in traffic:
            roof_asm["traffic"] = traffic["k_assemble"]["bytes_per_dispatch"]
        if roof_solve and "solve" in traffic:
            roof_solve["traffic"] = traffic["solve"]["read_bytes"] + traffic["solve"]["write_bytes"]
    cpu = None
    if not args.no_cpu:
        threads, aff = cpu_threads()
        log(f"[bench] CPU baseline: one {args.cpu_n}^3 step, {threads} threads ...")
        c = cpu_baseline(args.cpu_n, flops, threads)
        # the 16^3 configuration is small enough that fewer threads can be faster: the thread count
        # with the fastest of three trial steps, then the median of seven steps with it
        t16 = min(((cpu_baseline(16, 1.0, t, reps=3)["s_sample"], t) for t in (1, 2, 4, 8, 16) if t <= threads))[1]
        c16 = cpu_baseline(16, 1.0, t16, reps=7)
        cpu = {"value": c["s_per_step"], "unit": "s/step", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(), "affinity_cpus": aff,
               "sample": f"oracle multifrontal supernodal Cholesky (relaxed supernodes, per-supernode POTRF/TRSM/"
                         f"SYRK from scipy's OpenBLAS single-threaded inside the OpenMP tree levels and threaded at "
                         f"the top levels, {threads} threads), same ordering as the GPU path, one Newton step of the "
                         f"{args.cpu_n}^3 cube: assembly {c['t_assembly']:.2f}s (serial) + injection and factor "
                         f"{c['t_factor']:.2f}s ({c['factor_gflops']:.0f} GFLOP/s) + solve {c['t_solve']:.2f}s = "
                         f"{c['s_sample']:.2f}s" + (f", scaled x{c['scale']:.3f} by the factorization flop ratio"
                                                    if abs(c["scale"] - 1.0) > 1e-9 else " (the full workload)"),
               "s_per_step_16cube": c16["s_sample"], "threads_16cube": t16,
               "s_16cube_spread_of_7": c16["spread"]}
    out = {
        "metric": "Newton linear-solve s/step @811k DOF to <=1e-10 rel. residual; factor TFLOP/s",
        "value": s_per_step, "unit": "s/step", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": s_per_step * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if not single else ("f32 factor/solve, " + ("bf16 Schur updates (f32 accumulation), "
                                                                  if args.precision == "bf16" else
                                                                  "bf16x3 Schur updates (f32 accumulation), "
                                                                  if args.precision == "bf16x3" else "")
                                           + "f64 residual+refinement"),
        "data": "synthetic (seeded perturbations of the unit cube, neo-Hookean He and RHS from the device element kernel)",
        "config": {"workload": f"{N}^3 hex cube x 6 Kuhn tets, neo-Hookean "
                               + ("E log-uniform [1e3,1e7] nu U[0.2,0.45]" if args.stress else "E=1e5 nu=0.3")
                               + ", sigma=1e-3 h, z=0 pinned; assembly + "
                               + ("FP32 scaled supernodal Cholesky + FP64 refinement to 1e-10" if single
                                  else "FP64 supernodal Cholesky + solve/refine to 1e-10"),
                   "free_dofs": nred, "tets": ntets, "factor_nnz": nnz,
                   "factor_flops": flops, "factor_flops_unrelaxed": unrelaxed_flops,
                   "supernodes": st["num_supernodes"], "levels": st["num_levels"],
                   "he_layout": args.he_layout},
        "factor_tflops": factor_tflops, "t_assembly_s": (timed["t_zero_ms"] + timed["t_assembly_ms"]) / 1e3,
        "t_factor_s": timed["t_factor_s"], "t_solve_s": timed["t_solve_s"],
        "refine_iters": timed["iters"], "relres_max": timed["relres_max"],
        "symbolic_s": t_sym, "kernel_ms_per_step": kernel_ms, "hbm_copy_gbs": copy_gbs,
        "roofline": roofline, "roofline_assembly": roof_asm, "roofline_solve": roof_solve,
        "traffic_note": why if traffic is None else
        f"rocprofv3 FETCH_SIZE x{traffic['scale_fetch']:.3f} and WRITE_SIZE x{traffic['scale_write']:.3f} "
        f"(the same counters on a 2 GiB device copy in each pass); L2-to-memory bytes",
        "traffic_kernels": traffic,
        "precision_config": legs,
        "cpu_baseline": cpu,
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="hexes per cube side (64 = 811,200 free DOFs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-n", type=int, default=64, help="cube side of the CPU baseline step (64: the full workload)")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--precision", choices=["fp64", "fp32", "bf16", "bf16x3"], default="fp64",
                    help="bf16 / bf16x3: FP32 factor with bfloat16 Schur updates (x3: hi+lo operand split)")
    ap.add_argument("--stress", action="store_true",
                    help="per-element E log-uniform [1e3, 1e7], nu uniform [0.2, 0.45]")
    ap.add_argument("--stress-legs", default="fp32,bf16x3",
                    help="fp64 headline runs: also time the stressed cube in FP64 and these forms (empty: none)")
    ap.add_argument("--stress-steps", type=int, default=3)
    ap.add_argument("--he-layout", choices=["upper78", "full12"], default="upper78",
                    help="element Hessians as the 78 local-upper doubles the assembly reads, or full 12x12")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 counter passes")
    ap.add_argument("--traffic-timeout", type=int, default=240)
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--workload", choices=["cube", "batch", "learned", "replay"], default="cube",
                    help="cube: the 64^3 Newton step (headline); batch: 512 uneven meshes in one system; "
                         "learned: train the learned preconditioner, evaluate held-out solves; replay: "
                         "record and replay a Newton sequence like benchmark_linear_systems")
    ap.add_argument("--replay-dir", default="", help="replay: an existing recording directory")
    ap.add_argument("--train", type=int, default=2000, help="learned: training Hessians")
    ap.add_argument("--learn-n", type=int, default=48, help="learned: cube side of the training meshes")
    ap.add_argument("--lr", type=float, default=2e-3, help="learned: Adam learning rate")
    ap.add_argument("--max-pcg", type=int, default=200, help="learned: CG iterations before the Cholesky fallback")
    ap.add_argument("--max-pcg-free", type=int, default=5000, help="learned: CG limit of the unlimited comparison")
    ap.add_argument("--batch", type=int, default=512, help="number of meshes of --workload batch")
    ap.add_argument("--no-graph", action="store_true", help="batch: launch the factorization without a hipGraph")
    ap.add_argument("--graph", action="store_true", help="cube: replay the numeric factorization as a captured hipGraph")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU sharding of the factorization "
              "is a design, not an implementation); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.traffic_probe:
        traffic_probe(args)
        return
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "learned":
        bench_learned(args)
        return
    if args.workload == "replay":
        bench_replay(args)
        return
    out = bench_cube(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
