"""Per-Newton-step linear solve on one MI355X: assembly -> numeric factorization -> solve (+FP64
refinement to <= 1e-10 relative residual) on the 64^3 Kuhn-tet cube (811,200 free DOFs).

One step = one Newton-like re-perturbation (new seed, sigma = 1e-3 h): element Hessians and the
RHS (= -gradient) for every seed are generated beforehand and resident in HBM; the timed region
covers Ax := 0, assembly, value injection + numeric factorization, solve + refinement.  Symbolic
analysis (ordering, supernodes, schedule) runs once before the timed region and is reported
separately.  Prints one JSON line.

With --precision fp32 the factorization runs in FP32 on MFMA after symmetric diagonal scaling and
FP64 iterative refinement brings the residual to <= 1e-10; --stress draws per-element E
log-uniform in [1e3, 1e7] and nu uniform in [0.2, 0.45] (the conditioning stress case), and the
forward error against the FP64 factor's solution is reported per step.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 64] [--precision fp64|fp32] [--stress]
                  [--no-cpu]
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


def cpu_baseline(n_sample, flops_target, threads):
    """The oracle's multifrontal supernodal Cholesky (oracle/supernodal.c: relaxed supernodes,
    per-supernode POTRF/TRSM/SYRK from the OpenBLAS that ships with scipy, elimination tree level
    by level over `threads` OpenMP threads, the BLAS threaded at the top levels) timed on this
    host for one Newton step of the n_sample^3 cube with the GPU path's ordering: serial
    reference-semantics assembly, value injection into P A P^T, numeric factorization, forward
    and backward solves.  Scaled to the full workload by the ratio of factorization flops."""
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
    t0 = time.perf_counter()
    Ax = orc.assemble(Ap, Ai, nb, tets, He)
    t1 = time.perf_counter()
    Cp, Ci, Cx = orc.permuted_lower(Ap, Ai, nb, Ax, red, pos)
    S.factor(Cx, threads)
    t2 = time.perf_counter()
    x = S.solve(bp)
    t3 = time.perf_counter()
    A = None
    if nred * 3 <= 3000:      # residual check on small samples
        A = orc.dense_from_lower(Cp, Ci, Cx)
        assert np.linalg.norm(A @ x - bp) <= 1e-10 * np.linalg.norm(bp)
    scale = flops_target / fl[0]
    return dict(t_assembly=t1 - t0, t_factor=t2 - t1, t_solve=t3 - t2, sample_flops=fl[0], cpu_flops=S.info["flops"],
                factor_gflops=S.info["flops"] / (t2 - t1) / 1e9, scale=scale, s_sample=t3 - t0,
                s_per_step=(t3 - t0) * scale, supernodes=S.info["supernodes"])


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
    tune = f.autotune(H, B["pins"], fuse=(0, 192, 384, 768), tiles=(128, 256), reps=1)
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
                   "factor_flops": flops, "schedule": f.scheduleStats(), "autotune": {
                       "fuse_degree": tune["fuse_degree"], "big_tile": tune["big_tile"],
                       "ms": {f"{a},{b}": round(v, 3) for (a, b), v in tune["ms"].items()}},
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="hexes per cube side (64 = 811,200 free DOFs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-n", type=int, default=48, help="cube side of the CPU baseline sample")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run a device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--precision", choices=["fp64", "fp32", "bf16", "bf16x3"], default="fp64",
                    help="bf16 / bf16x3: FP32 factor with bfloat16 Schur updates (x3: hi+lo operand split)")
    ap.add_argument("--stress", action="store_true",
                    help="per-element E log-uniform [1e3, 1e7], nu uniform [0.2, 0.45]")
    ap.add_argument("--forward-error", action="store_true",
                    help="also factor in FP64 and report the forward error of every step (fp32)")
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

    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "learned":
        bench_learned(args)
        return
    if args.workload == "replay":
        bench_replay(args)
        return

    import torch
    import meshfemsparse_amd as mfs
    from meshfemsparse_amd import _lib, workload as wl

    dev = torch.device("cuda:0")
    if args.calibrate_copy_mib:
        src = torch.ones(args.calibrate_copy_mib * 131072, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.calibrate_copy_mib} MiB done")
        del src, dst
    N = args.n
    nseeds = args.steps + args.warmup
    t0 = time.perf_counter()
    tets, Xr = wl.cube_mesh(N)
    nb = len(Xr)
    asm = mfs.SystemAssembler(nb)
    H = asm.blockSparsityPattern(tets)
    H.setZero()
    pins = wl.z0_pins(N)
    log(f"[bench] mesh N={N}: {len(tets)} tets, {nb} nodes, {H.nz} blocks; pattern {time.perf_counter() - t0:.2f}s")
    # inputs for every step, resident in HBM
    He_d, b_d = [], []
    t0 = time.perf_counter()
    Ee = nue = None
    if args.stress:
        rng = np.random.default_rng(12345)
        Ee = 10 ** rng.uniform(3, 7, len(tets))
        nue = rng.uniform(0.2, 0.45, len(tets))
    # element Hessians and gradients by the device element kernel (the same formulas as the host
    # generator), gradients assembled on the device; RHS = -gradient
    gasm = mfs.DeviceAssembler(H, tets)
    T = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    Xr_d, Ee_d, nue_d = T(Xr), T(Ee), T(nue)
    for s in range(nseeds):
        He, g12 = gasm.neohookean(Xr_d, T(wl.perturb(N, Xr, seed=s)), E_per_elem=Ee_d, nu_per_elem=nue_d)
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
    t0 = time.perf_counter()
    f.factorizeSymbolic(H, pins)
    t_sym = time.perf_counter() - t0
    st = f.supernodeStats()
    flops = f.getFlopEstimate()
    nred = f.n_reduced()
    log(f"[bench] symbolic {t_sym:.2f}s: n_red={nred} nnz(L)={f.getFactorNNZ():.3e} flops={flops:.3e} {st}")
    dasm = asm._assembler(H, np.ascontiguousarray(tets, np.int32))
    f.setGraphCapture(args.graph)

    def step(s):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        e[0].record()
        dasm.zero(H.Ax)
        dasm.assemble(He_d[s], H.Ax)
        e[1].record()
        f.factorizeNumeric(H)
        e[2].record()
        x, iters, rr = f.solveRefined(H, b_d[s], tol=1e-10, maxIters=30 if single else 10)
        e[3].record()
        return e, iters, rr, x

    for s in range(args.warmup):
        step(s)
        log(f"[bench] warmup step {s} done")
    torch.cuda.synchronize()
    recs = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        recs.append(step(args.warmup + k))
        log(f"[bench] step {k} done")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the same steps again with a HIP event pair around every kernel launch (per-class device
    # time and the dominant kernel's average launch duration); kept out of the timed region
    # because the events add a few microseconds between dependent launches
    kt = not args.no_kernel_timing
    ms = np.zeros(len(KCLASSES))
    cnt = np.zeros(len(KCLASSES), np.int64)
    work = np.zeros(len(KCLASSES))
    if kt:
        f.setGraphCapture(False)            # per-launch events need the launches themselves
        _lib.call("mfs_kernel_profiling", f._h, 1, 1)
        for k in range(args.steps):
            step(args.warmup + k)
        torch.cuda.synchronize()
        _lib.call("mfs_kernel_profiling", f._h, 0, 0)
        _lib.call("mfs_kernel_profile", f._h, _lib.ptr(ms, _lib.c_double), _lib.ptr(cnt, _lib.c_i64))
        _lib.call("mfs_kernel_profile_work", f._h, _lib.ptr(work, _lib.c_double))
        log("[bench] kernel-timing pass done")
    t_asm = np.mean([r[0][0].elapsed_time(r[0][1]) for r in recs]) / 1e3
    t_fac = np.mean([r[0][1].elapsed_time(r[0][2]) for r in recs]) / 1e3
    t_sol = np.mean([r[0][2].elapsed_time(r[0][3]) for r in recs]) / 1e3
    iters = [r[1] for r in recs]
    rrs = [r[2] for r in recs]
    fwd_err = None
    if single and args.forward_error:
        # forward error of each step's refined FP32-factor solution against the FP64 factor's solve
        xs = [r[3].clone() for r in recs]
        del recs
        f64 = mfs.make_cholesky_factorizer()
        f64.factorizeSymbolic(H, pins)
        fwd_err = []
        for k in range(args.steps):
            dasm.zero(H.Ax)
            dasm.assemble(He_d[args.warmup + k], H.Ax)
            f64.factorizeNumeric(H)
            x64, _, _ = f64.solveRefined(H, b_d[args.warmup + k], tol=1e-12, maxIters=5)
            fwd_err.append(float(torch.linalg.norm(xs[k] - x64) / torch.linalg.norm(x64)))
        del f64
        log(f"[bench] forward errors vs FP64 factor: {fwd_err}")
    s_per_step = wall / args.steps
    factor_tflops = flops / t_fac / 1e12
    # roofline of the dominant kernel: the tile GEMM (TRSM + in-panel updates + Schur updates);
    # for the BF16 variants the bfloat16 Schur-update kernel (in-panel + update-matrix classes)
    bf = args.precision.startswith("bf16")
    gcls = [6, 7] if bf else [5, 6, 7]
    gemm_ms = float(sum(ms[c] for c in gcls))
    gemm_launches = int(sum(cnt[c] for c in gcls))
    # algorithmic GEMM flops of a step (supernodal count minus the diagonal-tile POTRF share),
    # apportioned to the classes by their executed tile flops
    potrf_flops = f.flopBreakdown()["potrf"]
    exec_all = float(work[5] + work[6] + work[7] + work[KCLASSES.index("panel")])
    share = float(sum(work[c] for c in gcls)) / exec_all if exec_all > 0 else 1.0
    gemm_work = (flops - potrf_flops) * share * args.steps
    if kt and gemm_launches:
        avg_launch_s = gemm_ms / 1e3 / gemm_launches
        flops_per_launch = gemm_work / gemm_launches
        achieved = flops_per_launch / avg_launch_s / 1e12
    else:
        avg_launch_s, achieved = None, factor_tflops
    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE passes on this workload
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"k_gemm_traffic_64cube_{args.precision}.json")
    if N == 64 and os.path.exists(tpath):
        traffic = json.load(open(tpath))["traffic_bytes_per_launch"]
    peak = PEAK_TFLOPS["bf16" if bf else args.precision]
    kname = (f"k_gemm_bf ({'hi+lo split ' if args.precision == 'bf16x3' else ''}bfloat16 MFMA Schur updates, "
             f"FP32 accumulation; algorithmic flops)" if bf else
             f"k_gemm ({args.precision.upper()} MFMA tile GEMM: TRSM, in-panel update, Schur update)")
    roofline = {"bound": "mfma", "kernel": kname,
                "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": traffic,
                "avg_launch_ms": None if avg_launch_s is None else avg_launch_s * 1e3,
                "launches_per_step": gemm_launches / max(args.steps, 1)}
    kernel_ms = {k: round(float(m) / args.steps, 3) for k, m in zip(KCLASSES, ms) if cnt[KCLASSES.index(k)]}
    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
        log(f"[bench] CPU baseline on a {args.cpu_n}^3 sample, {threads} threads ...")
        c = cpu_baseline(args.cpu_n, flops, threads)
        c16 = cpu_baseline(16, 1.0, threads)
        cpu = {"value": c["s_per_step"], "unit": "s/step", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(),
               "sample": f"oracle multifrontal supernodal Cholesky (relaxed supernodes, per-supernode POTRF/TRSM/"
                         f"SYRK from scipy's OpenBLAS, tree-parallel levels over {threads} OpenMP threads, threaded "
                         f"BLAS at the top levels), same ordering as the GPU path, on the {args.cpu_n}^3 cube: "
                         f"assembly {c['t_assembly']:.2f}s (serial) + injection and factor {c['t_factor']:.2f}s "
                         f"({c['factor_gflops']:.0f} GFLOP/s) + solve {c['t_solve']:.2f}s = {c['s_sample']:.2f}s, "
                         f"scaled x{c['scale']:.2f} by the factorization flop ratio to the full workload",
               "s_per_step_16cube": c16["s_sample"]}
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
                   "free_dofs": nred, "tets": len(tets), "factor_nnz": f.getFactorNNZ(), "factor_flops": flops,
                   "supernodes": st["num_supernodes"], "levels": st["num_levels"]},
        "factor_tflops": factor_tflops, "t_assembly_s": t_asm, "t_factor_s": t_fac, "t_solve_s": t_sol,
        "refine_iters": iters, "relres_max": max(rrs), "forward_error_vs_fp64": fwd_err,
        "symbolic_s": t_sym, "kernel_ms_per_step": kernel_ms,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
