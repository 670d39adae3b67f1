#!/usr/bin/env python3
"""Benchmark: pseudo-transient 3D viscoelastic Stokes at 255³, FP64, one MI355X.

A step is one PT iteration (examples/stokes_3d_inc_ve_T.jl:164-165) on
fields resident in HBM.  Iterations run two per pass: the two-iteration kernel
(update_stress!, update_velocity! and the free-slip V batch of the first
iteration evaluated in-kernel, then the second iteration) followed by the V
batch, so K steps are K/2 passes.  Reported value: T_eff = algorithmic bytes
per iteration / time per iteration (GB/s), with algorithmic bytes = 27 array
passes over exact interior counts (updated Vx,Vy,Vz,Pr,τ×6 twice; read
τ_old×6 and η once).  Setup: [-0.5,0.5]³, background η = 1 with a sphere of
radius 0.1 at η = 1e-3, times exp(-T) with T = 0.1·exp(-|x|²/0.05²); G = 1;
dt from the reference's diffusion limit; pure-shear free-slip walls ±1.

The same line carries the other two targets: the time to an FP64 residual of
1e-8 with FP32 sweeps + FP64 defect correction against plain FP64 (2D at
2047², 3D at 255³), and the SIA adjoint/forward cost with its finite
difference check (512², 300 steps, 30 checkpoints).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "chmy.jl_amd"))

PEAK_HBM_GBS = 8000.0


def counter_traffic(n, dtype, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE x 2 + WRITE_SIZE x 1, the scale
    factors of a 1 GiB device copy), or None when the summary does not cover
    (kernel, n, dtype)."""
    path = os.path.join(ROOT, "profiles", "fused3d_stokes_traffic.txt")
    try:
        for line in open(path):
            line = line.strip()
            if line.startswith("{"):
                d = json.loads(line)
                if d.get("n") == n and d.get("dtype") == dtype and d.get("kernel", "fused3d_kernel") == kernel:
                    return d["hbm_bytes_per_launch"]
    except OSError:
        pass
    return None


def algorithmic_bytes(n, itemsize=8):
    C = n ** 3
    Vf = (n + 1) * n * n
    E = (n + 1) * (n + 1) * n
    updated = 3 * Vf + 4 * C + 3 * E
    read = 3 * C + 3 * E + C
    return (2 * updated + read) * itemsize


def reference_bytes(n, itemsize=8):
    """nIO = 2·10 + 17 arrays × nx·ny·nz (stokes_3d_inc_ve_T_mpi_perf.jl:210-213)."""
    return 37 * itemsize * n ** 3


def setup(n, dtype, exact=False):
    import torch
    import chmy
    from chmy import configs
    from chmy.stokes import PTStokes, pure_shear_initial, pure_shear_values
    arch = chmy.Arch(chmy.MI355XBackend())
    g, eta, dt = configs.stokes3d_inclusion(arch, n)   # dt_diff (stokes_3d_inc_ve_T.jl:157), λ_ρCp = 1e-4
    S = PTStokes(arch, g, dtype=dtype, eta_field=eta, G=1.0, dt=dt, normal_values=pure_shear_values(3),
                 exact=exact)
    pure_shear_initial(g, S.V)
    S.apply_bc()
    S.sync_buffers()
    torch.cuda.synchronize()
    return S


def cpu_baseline(n, iters, threads=None, zslab=None):
    """Time the oracle (C++ restatement of the reference's CPU path: unfused
    update_stress!, update_velocity!, bc! per iteration, OpenMP) on a bounded
    sample of the same workload: one warm-up iteration, then the median of
    `iters` iterations, on `threads` OpenMP threads (default: all).  With
    `zslab` the two kernels run over the Launcher box's first zslab planes
    only (the bc! batch as always); returns (seconds, threads, fraction of the
    Launcher box's points covered)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    from chmy.stokes import inclusion_setup_3d
    import chmy
    g = chmy.UniformGrid(origin=(-0.5,) * 3, extent=(1.0,) * 3, dims=(n,) * 3)
    og = O.Grid(g.origin, g.extent, g.dims)
    C, V = O.CENTER, O.VERTEX
    tl = [(C, C, C)] * 3 + [(V, V, C), (V, C, V), (C, V, V)]
    vl = [(V, C, C), (C, V, C), (C, C, V)]
    Vf = [O.HField(og, l) for l in vl]
    rV = [O.HField(og, l) for l in vl]
    tau = [O.HField(og, l) for l in tl]
    told = [O.HField(og, l) for l in tl]
    Pr, dV, eta, eml = (O.HField(og, C) for _ in range(4))
    eta.arr[...] = inclusion_setup_3d(g)
    O.maxloc(og, eml, eta)
    p = O.pt_params(og)
    dx = min(og.spacing)
    dt = dx ** 2 / 1e-4 / 3 / 2.1
    spec = []
    vals = {0: (1.0, -1.0), 1: (0.0, 0.0), 2: (-1.0, 1.0)}
    for D in range(3):
        s = {d: O.Neumann() for d in range(3)}
        s[D] = (O.Dirichlet(vals[D][0]), O.Dirichlet(vals[D][1]))
        spec.append((Vf[D], s))

    b = O.box((0, 0, 0), (n + 2, n + 2, zslab)) if zslab else None
    frac = zslab / (n + 2) if zslab else 1.0

    def step():
        O.update_stress(og, tau, Pr, dV, Vf, told, eta, 0.0, 1.0, dt, p["dtau_Pr"], p["dtau_r"], b=b)
        O.update_velocity(og, Vf, rV, Pr, tau, None, 0.0, p["nudtau"], eta_ml=eml, G=1.0, dt=dt, b=b)
        O.bc_apply(og, spec)

    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if threads is not None:
        O.set_threads(threads)
        cores = threads
    try:
        step()  # warm-up
        ts = []
        for _ in range(iters):
            t0 = time.perf_counter()
            step()
            ts.append(time.perf_counter() - t0)
    finally:
        if threads is not None:
            O.set_threads(0)
    el = sorted(ts)[len(ts) // 2]   # median
    return el, cores, frac


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_to_converge(args):
    """Time to an f64 residual below eps: plain f64 PT (fused kernel) versus
    f32 PT sweeps with f64 defect correction every `inner` iterations, on the
    same GPU and setup; also the per-field error of the mixed solution
    against the f64 one (max and L2, relative)."""
    import numpy as np
    import torch
    import chmy
    from chmy import configs
    from chmy.stokes import MixedPTStokes, PTStokes, pure_shear_initial, pure_shear_values
    arch = chmy.Arch(chmy.MI355XBackend())
    if args.ttc == "2d":
        g, eta, dt = configs.stokes2d_inclusions(arch, args.n)
    else:
        g, eta, dt = configs.stokes3d_inclusion(arch, args.n)
    nd = g.nd
    log = lambda *a: print("#", *a, file=sys.stderr, flush=True)
    res = {}
    S = PTStokes(arch, g, eta_field=eta, G=1.0, dt=dt, normal_values=pure_shear_values(nd))
    pure_shear_initial(g, S.V)
    S.apply_bc()
    S.sync_buffers()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it64, err64 = S.solve(eps=args.eps, max_iter=args.max_iter, ncheck=args.ncheck,
                          callback=lambda it, e: log("f64", it, [f"{x:.2e}" for x in e]) if it % (20 * args.ncheck) == 0 else None)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    res["f64"] = {"seconds": round(t64, 3), "iterations": it64, "residual": err64,
                  "true_residual_incl_tau": S.true_residual()}
    log("f64 done", it64, t64, err64)
    M = MixedPTStokes(arch, g, eta_field=eta, G=1.0, dt=dt, normal_values=pure_shear_values(nd), inner=args.inner)
    pure_shear_initial(g, M.x.V)
    M.x.apply_bc()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sw, errm = M.solve(eps=args.eps, max_cycles=args.max_iter // args.inner + 1,
                       callback=lambda s, e: log("mixed", s, [f"{x:.2e}" for x in e]) if (s // args.inner) % 20 == 0 else None)
    torch.cuda.synchronize()
    tm = time.perf_counter() - t0
    res["f32_fp64dc"] = {"seconds": round(tm, 3), "f32_iterations": sw, "cycles": M.cycles, "residual": errm}
    log("mixed done", sw, tm, errm)
    errs = {}
    names = list(g.axes_names())
    pairs = [("V" + names[D], S.V[D], M.x.V[D]) for D in range(nd)] + [("Pr", S.Pr, M.x.Pr)]
    pairs += [("tau_" + k, a, b) for k, a, b in zip(S.tau._fields, S.tau, M.x.tau)]
    for name, a, b in pairs:
        ai = a.interior().double().cpu().numpy()
        bi = b.interior().double().cpu().numpy()
        d = ai - bi
        errs[name] = {"max_rel": float(np.max(np.abs(d)) / max(np.max(np.abs(ai)), 1e-300)),
                      "l2_rel": float(np.linalg.norm(d) / max(np.linalg.norm(ai), 1e-300))}
    out = {"metric": "time-to-1e-8 FP32+fix vs FP64", "value": round(t64 / tm, 3), "unit": "x speedup",
           "higher_is_better": True, "n_gpus": 1, "dtype": "f32 sweeps + f64 residual",
           "config": {"workload": f"{nd}D VE Stokes PT, {args.n}^{nd} cells, inclusions, eps {args.eps}",
                      "inner": args.inner, "ncheck": args.ncheck},
           "runs": res, "field_errors_vs_f64": errs}
    return out


def algorithmic_bytes_2d(n, itemsize=8):
    """Updated {Vx, Vy, Pr, tau_xx, tau_yy, tau_xy} x2 + read {tau_old x3, eta}
    over the exact interior counts of the 2D staggered fields."""
    C, Vf, E = n * n, (n + 1) * n, (n + 1) * (n + 1)
    updated = 2 * Vf + 3 * C + E
    read = 2 * C + E + C
    return (2 * updated + read) * itemsize


def stokes2d_rates(args):
    """The 2D fused PT iteration at n^2 (inclusions setup): FP64 and FP32
    single-iteration kernel times (HIP events on the launch stream, median of
    `steps`); the time per iteration when iterations run in pairs (one pass
    per two iterations, HIP events around `steps` iterations, V batch
    included) and its T_eff; and the wall time per sweep of the FP32 + FP64
    defect-correction solver (200-sweep cycles with their FP64 residual
    refresh)."""
    import torch
    import chmy
    from chmy import configs
    from chmy.stokes import MixedPTStokes, PTStokes, pure_shear_initial, pure_shear_values
    arch = chmy.Arch(chmy.MI355XBackend())
    n = args.n
    g, eta, dt = configs.stokes2d_inclusions(arch, n)
    out = {"metric": "2D fused PT iteration rate", "unit": "GB/s", "n_gpus": 1, "higher_is_better": True,
           "config": {"workload": f"2D VE Stokes PT, {n}^2 cells, 32 inclusions", "grid": [n, n]}}
    for dt_, name, isz in ((torch.float64, "f64", 8), (torch.float32, "f32", 4)):
        S = PTStokes(arch, g, dtype=dt_, eta_field=eta, G=1.0, dt=dt, normal_values=pure_shear_values(2))
        pure_shear_initial(g, S.V)
        S.apply_bc()
        S.sync_buffers()
        for _ in range(args.warmup):
            S.iteration()
        torch.cuda.synchronize()
        K = args.steps
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
        t0 = time.perf_counter()
        for q in range(K):
            ev[q][0].record()
            S._iter_call(False)
            ev[q][1].record()
            S.apply_bc(S.V2)
            S._swap()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / K
        k_ms = sorted(a.elapsed_time(b) for a, b in ev)[K // 2]
        A = algorithmic_bytes_2d(n, isz)
        # iterations in pairs (chmy_stokes_iterate: one pass per two iterations,
        # the free-slip batch in-kernel between them, the batch launch after)
        S.iterate(20)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        S.iterate(K)
        e1.record()
        torch.cuda.synchronize()
        pair_ms = e0.elapsed_time(e1) / K
        out[name] = {"kernel_ms_median": round(k_ms, 5), "kernel_GBs": round(A / k_ms / 1e6, 1),
                     "step_ms": round(wall * 1e3, 5), "algorithmic_bytes": A,
                     "paired_iteration_ms": round(pair_ms, 5), "paired_teff_GBs": round(A / pair_ms / 1e6, 1)}
        del S
    M = MixedPTStokes(arch, g, eta_field=eta, G=1.0, dt=dt, normal_values=pure_shear_values(2), inner=200)
    pure_shear_initial(g, M.x.V)
    M.x.apply_bc()
    M.refresh()
    M.cycle()
    M.refresh()
    torch.cuda.synchronize()
    cycles = max(1, args.steps // 200)
    t0 = time.perf_counter()
    for _ in range(cycles):
        M.cycle()
        M.refresh()
    torch.cuda.synchronize()
    out["mixed_sweep_ms"] = round((time.perf_counter() - t0) / (200 * cycles) * 1e3, 5)
    out["value"] = out["f64"]["paired_teff_GBs"]
    print(json.dumps(out))


def diffusion_c1(args):
    """2D linear diffusion (examples/diffusion_2d.jl form, 256^2, 1000 steps, FP64) on the
    GPU: the reference's launch sequence (compute_q!, update_C! with the
    Neumann batch, through the Launcher), the fused step (chmy_diffusion_step
    into a second buffer + the batch, through the Launcher) and the whole loop
    in one native call (chmy_diffusion_iterate), each checked bitwise against
    the oracle's 1000 steps, which are also timed on the host (the reference
    runs this workload on its CPU backend)."""
    import numpy as np
    import torch
    import chmy
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n, steps = 256, 1000
    arch = chmy.Arch(chmy.MI355XBackend())
    g = chmy.UniformGrid(arch, origin=(-1.0, -1.0), extent=(2.0, 2.0), dims=(n, n))
    og = O.Grid(g.origin, g.extent, g.dims)
    rng = np.random.default_rng(1234)
    xs = og.coords(O.CENTER, 0)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    C0 = np.exp(-(X ** 2 + Y ** 2) / 0.1) + 1e-3 * rng.random((n, n))
    dt = min(chmy.spacing(g)) ** 2 / 1.0 / 4.1
    # oracle (host)
    hC = O.HField(og, O.CENTER)
    hq = [O.HField(og, (O.VERTEX, O.CENTER)), O.HField(og, (O.CENTER, O.VERTEX))]
    hC.set_interior(C0)
    O.bc_apply(og, [(hC, O.Neumann())])
    t0 = time.perf_counter()
    for _ in range(steps):
        O.compute_q(og, hq, hC, 1.0)
        O.update_C(og, hC, hq, dt)
        O.bc_apply(og, [(hC, O.Neumann())])
    t_cpu = time.perf_counter() - t0
    L = chmy.Launcher(arch, g, sync=False)
    res = {}
    for mode in ("reference_launches", "fused_step", "native_loop"):
        Cf = chmy.Field(arch, g, chmy.Center())
        C2 = chmy.Field(arch, g, chmy.Center())
        q = chmy.VectorField(arch, g)
        chmy.set_(Cf, C0)
        bcs = chmy.batch(g, (Cf, chmy.Neumann()))
        bcs2 = chmy.batch(g, (C2, chmy.Neumann()))
        chmy.bc_(arch, g, bcs)
        C2.data.copy_(Cf.data)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        if mode == "reference_launches":
            for _ in range(steps):
                L(arch, g, (chmy.kernels.compute_q, (q, Cf, 1.0, g)))
                L(arch, g, (chmy.kernels.update_C, (Cf, q, dt, g)), bc=bcs)
            out = Cf
        elif mode == "fused_step":
            cur, nxt = (Cf, bcs), (C2, bcs2)
            for _ in range(steps):
                L(arch, g, (chmy.kernels.diffusion_step, (cur[0], nxt[0], q, 1.0, dt, g)), bc=nxt[1])
                cur, nxt = nxt, cur
            out = cur[0]
        else:   # the whole loop in one native call
            import ctypes as C
            from chmy import _lib
            from chmy.kernels import _arr
            counts = (C.c_int32 * 6)(1, 1, 1, 1, 0, 0)
            conds = (_lib.ChmyBC * 4)(*[_lib.ChmyBC(_lib.NEUMANN, _lib.VALUE_ZERO, 0.0, None) for _ in range(4)])
            par = C.c_int(0)
            _lib.call("chmy_diffusion_iterate", C.byref(g.desc()), C.byref(Cf.desc()), C.byref(C2.desc()), _arr(q),
                      counts, conds, 1.0, dt, steps, C.byref(par), chmy.current_stream())
            out = (Cf, C2)[par.value]
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        bitwise = bool(np.array_equal(out.data.cpu().numpy(), hC.arr))
        res[mode] = {"ms_1000_steps": round(ms, 3), "bitwise_vs_oracle": bitwise}
    out = {"metric": "2D diffusion 256^2 wall time, 1000 steps", "value": res["native_loop"]["ms_1000_steps"], "unit": "ms",
           "higher_is_better": False, "n_gpus": 1, "dtype": "f64",
           "config": {"workload": "2D linear diffusion, 256^2 cells, 1000 steps, Neumann(0)"},
           "runs": res, "cpu_oracle_ms_1000_steps": round(t_cpu * 1e3, 1),
           "cpu_cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))}
    print(json.dumps(out))


def sia_gradient(args):
    """SIA forward run versus gradient (checkpointed forward + reverse sweep)
    wall time at 512² (300 steps, 30 checkpoints, fixed Δt), and the gradient
    against central differences of J along 16 seeded Gaussian directions."""
    import numpy as np
    import torch
    import chmy
    from chmy import configs
    from chmy.sia import SIA
    arch = chmy.Arch(chmy.MI355XBackend())
    g, B, la0, la1, mb, dt = configs.sia_ice_cap(arch, args.n)
    nsteps, ncheck = args.sia_steps, args.sia_ckpt
    dts = [dt] * nsteps
    S = SIA(arch, g, B, logA=la1, As=0.0, npow=3, mb=mb)
    S.reset()
    S.forward(nsteps, dts)
    H_obs = S.H.interior().clone()
    dt_lim = S.dt_rule()
    S.set_logA(la0)
    log = lambda *a: print("#", *a, file=sys.stderr, flush=True)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), ts

    def fwd():
        S.reset()
        S.forward(nsteps, dts)
        return S.misfit(H_obs)

    t_f, _ = timed(fwd, args.sia_reps)
    out = {}
    t_g, _ = timed(lambda: out.__setitem__("r", S.gradient(None, H_obs, nsteps, dts, ncheck)), args.sia_reps)
    J, grad = out["r"]
    log(f"forward {t_f * 1e3:.3f} ms, gradient {t_g * 1e3:.3f} ms, J = {J:.6e}")
    gn = grad.double().cpu().numpy()
    rng = np.random.default_rng(args.sia_seed)
    rows = []
    for d in range(args.sia_dirs):
        v = np.zeros(la0.shape)                      # cells only: the gradient covers no halo
        v[2:-2, 2:-2] = rng.standard_normal((la0.shape[0] - 4, la0.shape[1] - 4))
        h = args.sia_h
        S.set_logA(la0 + h * v)
        Jp = fwd()
        S.set_logA(la0 - h * v)
        Jm = fwd()
        fd = (Jp - Jm) / (2 * h)
        ad = float(np.sum(gn * v[2:-2, 2:-2].T))
        rows.append({"fd": fd, "adjoint": ad, "rel_err": abs(fd - ad) / abs(fd)})
    S.set_logA(la0)
    worst = max((r["rel_err"] for r in rows), default=None)
    log(f"FD directions: worst rel err {worst}")
    res = {"metric": "adjoint/fwd cost", "value": round(t_g / t_f, 3), "unit": "x forward wall time",
           "higher_is_better": False, "n_gpus": 1, "dtype": "f64",
           "config": {"workload": f"2D SIA ice cap, {args.n}^2 cells, {nsteps} steps of fixed dt {dt} a, "
                                  f"{ncheck} checkpoints, gradient of J = 1/2|H - H_obs|^2 wrt log A",
                      "forward_ms": round(t_f * 1e3, 3), "gradient_ms": round(t_g * 1e3, 3),
                      "repeats": args.sia_reps, "J": J, "dt_limit_at_end": dt_lim},
           "fd_check": {"directions": args.sia_dirs, "h": args.sia_h, "worst_rel_err": worst, "rows": rows}}
    return res


def stokes_gradient(args):
    """Steady 3D Stokes gradient dJ/dlog(eta), J the top-layer tangential
    velocity misfit against the solution for a perturbed viscosity: forward
    and adjoint PT solves as FP32 sweeps with FP64 defect correction, then the
    gradient pass; cost ratio (forward + adjoint + gradient) / forward.  The
    gradient is checked against central differences of J along seeded
    Gaussian directions at a smaller grid with f64 solves."""
    import numpy as np
    import torch
    import chmy
    from chmy import configs
    from chmy.stokes import pure_shear_values
    from chmy.stokes_adjoint import StokesAdjoint, smooth_log_field
    arch = chmy.Arch(chmy.MI355XBackend())
    log = lambda *a: print("#", *a, file=sys.stderr, flush=True)

    def problem(n, mixed):
        g, eta, dt = configs.stokes3d_inclusion(arch, n)
        la0 = np.log(eta)
        A = StokesAdjoint(arch, g, la0, G=1.0, dt=dt, normal_values=pure_shear_values(3), mixed=mixed,
                          inner=args.inner)
        return g, la0, A

    g, la0, A = problem(args.n, True)
    A.set_logeta(la0 + smooth_log_field(g, 0.5, seed=args.sia_seed, corr=8))
    sw_obs, _ = A.solve_forward(eps=args.eps, max_iter=args.max_iter)
    V_obs = A.observe()
    A.set_logeta(la0)
    log(f"observation solve: {sw_obs} sweeps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sw_f, err_f = A.solve_forward(eps=args.eps, max_iter=args.max_iter)
    torch.cuda.synchronize()
    t_f = time.perf_counter() - t0
    J = A.misfit(V_obs)
    t0 = time.perf_counter()
    sw_a, err_a = A.solve_adjoint(eps=args.eps, max_iter=args.max_iter)
    torch.cuda.synchronize()
    t_a = time.perf_counter() - t0
    t0 = time.perf_counter()
    A.gradient()
    torch.cuda.synchronize()
    t_g = time.perf_counter() - t0
    log(f"forward {t_f:.2f} s ({sw_f} sweeps), adjoint {t_a:.2f} s ({sw_a}), gradient pass {t_g * 1e3:.2f} ms")
    del A
    torch.cuda.empty_cache()
    # central differences at the check size, f64 solves
    gf, laf, B = problem(args.fd_n, False)
    B.set_logeta(laf + smooth_log_field(gf, 0.5, seed=args.sia_seed, corr=8))
    B.solve_forward(eps=args.fd_eps, max_iter=args.max_iter)
    Vo = B.observe()
    B.set_logeta(laf)
    Jf, grad = B.J_and_gradient(Vo, eps=args.fd_eps, max_iter=args.max_iter)
    gn = grad.xyz().cpu().numpy()
    rng = np.random.default_rng(args.sia_seed)
    rows = []
    for _ in range(args.fd_dirs):
        v = rng.standard_normal(laf.shape)
        Js = []
        for sgn in (1, -1):
            B.set_logeta(laf + sgn * args.fd_h * v)
            B.solve_forward(eps=args.fd_eps, max_iter=args.max_iter)
            Js.append(B.misfit(Vo))
        fd = (Js[0] - Js[1]) / (2 * args.fd_h)
        ad = float(np.sum(gn * np.transpose(v, (2, 1, 0))))
        rows.append({"fd": fd, "adjoint": ad, "rel_err": abs(fd - ad) / abs(fd)})
        log(f"FD direction: fd {fd:.10e} adjoint {ad:.10e}")
    worst = max((r["rel_err"] for r in rows), default=None)
    res = {"metric": "adjoint/fwd cost", "value": round((t_f + t_a + t_g) / t_f, 3),
           "unit": "x forward wall time (forward + adjoint + gradient over forward)", "higher_is_better": False,
           "n_gpus": 1, "dtype": "f32 sweeps + f64 residual",
           "config": {"workload": f"3D VE Stokes, {args.n}^3 cells, inclusion, J = top-layer tangential velocity "
                                  f"misfit, gradient wrt log eta, eps {args.eps}",
                      "forward_s": round(t_f, 3), "adjoint_s": round(t_a, 3), "gradient_pass_ms": round(t_g * 1e3, 3),
                      "forward_sweeps": sw_f, "adjoint_sweeps": sw_a, "residual_forward": err_f,
                      "residual_adjoint": err_a, "J": J},
           "fd_check": {"grid": [args.fd_n] * 3, "directions": args.fd_dirs, "h": args.fd_h, "solve_eps": args.fd_eps,
                        "J": Jf, "worst_rel_err": worst, "rows": rows}}
    print(json.dumps(res))


def thermal_loop(args):
    """The physical time loop's PT iterations with the thermal update inside
    (examples/stokes_3d_inc_ve_T.jl:161-169, time steps after the first): per
    iteration update_stress!, update_velocity! + V batch, update_thermal_flux!,
    update_thermal! + T batch.  3D FP64 at --n, rho g_z and T inclusions of
    radius 0.1 L (the example's setup), eta = 10, --steps iterations with no
    residual check (the example checks every 2 n_x iterations in its second
    step).  Timed two ways from the same state: one fused iteration per
    launch, and two iterations per pass with the first iteration's V kept for
    its thermal update (chmy_stokes_iter2_first); the final V, Pr, tau and T
    of the two are compared bitwise."""
    import numpy as np
    import torch
    import chmy
    from chmy import Center as Cc, Vertex as Vv
    from chmy.stokes_thermal import StokesThermal
    arch = chmy.Arch(chmy.MI355XBackend())
    n, K = args.n, args.steps
    L = 1.0
    g = chmy.UniformGrid(arch, origin=(-L / 2,) * 3, extent=(L,) * 3, dims=(n, n, n))

    def incl(locs, inside, outside):
        xs = [g.coords(Vv() if c == "V" else Cc(), d, first=1, count=n + (1 if c == "V" else 0))
              for d, c in enumerate(locs)]
        Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
        arr = np.zeros(tuple(len(x) + 4 for x in xs[::-1]))
        arr[2:-2, 2:-2, 2:-2] = np.where(X * X + Y * Y + Z * Z < (0.1 * L) ** 2, inside, outside)
        return arr

    rg_arr, T_arr = incl("CCV", 1.0, 0.0), incl("CCC", 1.0, 0.1)
    res, keep = {}, None
    for mode in ("single", "pairs"):
        rhog = chmy.Field(arch, g, (Cc(), Cc(), Vv()))
        chmy.set_storage(rhog, rg_arr)
        M = StokesThermal(arch, g, eta=10.0, G=1.0, rhog=rhog, exact=args.exact, pairs=mode == "pairs")
        chmy.set_storage(M.T, T_arr)
        M.apply_bc()
        M.time_step(2, 4, 2 * n, 0.0)   # warm-up: one step of 4 iterations
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        M.time_step(2, K, 2 * n, 0.0)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        fields = [f.data for f in list(M.S.V) + [M.S.Pr] + list(M.S.tau) + [M.T]]
        if keep is None:
            keep = [f.cpu() for f in fields]
            same = None
        else:
            same = all(torch.equal(a, b.cpu()) for a, b in zip(keep, fields))
        res[mode] = {"ms_per_iteration": round(sec * 1e3 / K, 4), "seconds": round(sec, 3)}
        print(f"# thermal loop {mode}: {res[mode]['ms_per_iteration']} ms per iteration", file=sys.stderr, flush=True)
        del M, rhog, fields
        torch.cuda.empty_cache()
    res["pairs"]["bitwise_equal_to_single"] = same
    speed = res["single"]["ms_per_iteration"] / res["pairs"]["ms_per_iteration"]
    print(json.dumps({"metric": "3D PT Stokes + thermal loop, ms per PT iteration", "value": res["pairs"]["ms_per_iteration"],
                      "unit": "ms", "higher_is_better": False, "n_gpus": 1, "dtype": "f64",
                      "config": {"workload": f"3D VE Stokes PT with the thermal update per iteration, {n}^3 cells, "
                                             f"{K} iterations, {'exact' if args.exact else 'fast'} form"},
                      "runs": res, "speedup_pairs_vs_single": round(speed, 3)}))


def pairs_rate(S, args, isz):
    """3D PT iterations two per pass (chmy_stokes_iterate with the pairs flag:
    fused3d_tb_kernel, then the V batch); one step = one pass = two
    iterations.  Prints one JSON line with the time per iteration."""
    import torch
    n = args.n
    K = max(1, args.steps // 2)
    for _ in range(max(1, args.warmup // 2)):
        S.iterate(2, pairs=True)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(K):
        S.iterate(2, pairs=True)
    t1.record()
    torch.cuda.synchronize()
    ms_pass = t0.elapsed_time(t1) / K
    A = algorithmic_bytes(n, isz)
    print(json.dumps({"metric": "T_eff GB/s, two PT iterations per pass", "value": round(2 * A / (ms_pass * 1e-3) / 1e9, 2),
                      "unit": "GB/s", "passes": K, "ms_per_pass": round(ms_pass, 4),
                      "ms_per_iteration": round(ms_pass / 2, 4), "dtype": args.dtype, "exact": bool(args.exact),
                      "config": {"workload": f"3D PT viscoelastic Stokes, {n}^3 cells, fused3d_tb_kernel + V batch",
                                 "algorithmic_bytes_per_iteration": A}}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=255)
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32"])
    ap.add_argument("--cpu-iters", type=int, default=5)
    ap.add_argument("--cpu-iters-1t", type=int, default=5)
    ap.add_argument("--no-targets", action="store_true",
                    help="the T_eff line only (no time-to-1e-8 and SIA adjoint parts)")
    ap.add_argument("--no-ttc3d", action="store_true", help="skip the 3D time-to-1e-8 part (about 2 minutes)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--exact", action="store_true",
                    help="the reference's expression order (bitwise equal to update_stress! + update_velocity!)")
    ap.add_argument("--mode", default="pass2", choices=["pass2", "fused", "unfused", "pairs"],
                    help="pass2: two iterations per pass (the free-slip batch in-kernel between them); "
                         "fused: one iteration per launch; unfused: the reference's update_stress! + "
                         "update_velocity! kernels; pairs: chmy_stokes_iterate two at a time, timed per pass")
    ap.add_argument("--ttc", choices=["2d", "3d"], default=None,
                    help="time-to-convergence comparison instead of the T_eff line")
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--inner", type=int, default=200)
    ap.add_argument("--ncheck", type=int, default=200)
    ap.add_argument("--max-iter", type=int, default=400000)
    ap.add_argument("--sia", action="store_true",
                    help="SIA gradient cost and finite-difference check instead of the T_eff line")
    ap.add_argument("--sia-steps", type=int, default=300)
    ap.add_argument("--sia-ckpt", type=int, default=30)
    ap.add_argument("--sia-dirs", type=int, default=16)
    ap.add_argument("--sia-h", type=float, default=1e-5)
    ap.add_argument("--sia-seed", type=int, default=16)
    ap.add_argument("--sia-reps", type=int, default=5)
    ap.add_argument("--stokes-adjoint", action="store_true",
                    help="3D Stokes gradient cost and finite-difference check instead of the T_eff line")
    ap.add_argument("--fd-n", type=int, default=63)
    ap.add_argument("--fd-dirs", type=int, default=8)
    ap.add_argument("--fd-h", type=float, default=1e-4)
    ap.add_argument("--fd-eps", type=float, default=1e-11)
    ap.add_argument("--stokes2d", action="store_true",
                    help="2D fused PT iteration rates at --n (default 2047): FP64, FP32, defect-correction sweep")
    ap.add_argument("--thermal-loop", action="store_true",
                    help="3D PT iterations with the thermal update inside (the time loop after its first step): "
                         "one iteration per launch vs two per pass, --steps iterations at --n")
    ap.add_argument("--diffusion", action="store_true",
                    help="2D diffusion 256^2 x 1000 steps: reference launches vs fused step vs native loop, vs the oracle")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU halo exchange is out of scope)", file=sys.stderr)
        sys.exit(2)
    if args.ttc:
        if args.ttc == "2d" and args.n == 255:
            args.n = 2047   # the 2D workload's grid
        print(json.dumps(time_to_converge(args)))
        return
    if args.stokes_adjoint:
        stokes_gradient(args)
        return
    if args.diffusion:
        diffusion_c1(args)
        return
    if args.thermal_loop:
        thermal_loop(args)
        return
    if args.stokes2d:
        if args.n == 255:
            args.n = 2047
        stokes2d_rates(args)
        return
    if args.sia:
        if args.n == 255:
            args.n = 512
        print(json.dumps(sia_gradient(args)))
        return

    import torch
    dtype = torch.float64 if args.dtype == "f64" else torch.float32
    isz = 8 if args.dtype == "f64" else 4
    n = args.n
    S = setup(n, dtype, exact=args.exact)
    print(f"# setup done: {n}^3 {args.dtype}", file=sys.stderr, flush=True)
    if args.mode == "pairs":
        pairs_rate(S, args, isz)
        return
    A = algorithmic_bytes(n, isz)
    if args.mode == "pass2":
        line = pass2_line(S, args, n, isz, A)
    else:
        line = single_line(S, args, n, isz, A)
    if not args.no_cpu and args.mode != "unfused":
        line["cpu_baseline"] = cpu_line(n, A, args)
    if args.mode == "pass2" and not args.no_targets:
        del S
        torch.cuda.empty_cache()
        line["targets"] = other_targets(args)
    print(json.dumps(line))


def timed_steps(S, K, call, use_events):
    """K calls of `call` (each followed by the V batch and the buffer swap),
    bracketed by a synchronize on both sides; HIP events on the launch stream
    around each kernel call.  Returns (total ms, mean kernel ms)."""
    import torch
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record()
    for q in range(K):
        ev[q][0].record()
        call()
        ev[q][1].record()
        S.apply_bc(S.V2)
        S._swap()
    t_end.record()
    torch.cuda.synchronize()
    return t_start.elapsed_time(t_end), sum(a.elapsed_time(b) for a, b in ev) / K


def pass2_line(S, args, n, isz, A):
    """The default line: K PT iterations as K/2 two-iteration passes."""
    import torch
    for _ in range(max(1, args.warmup // 2)):
        S.pass2()
    K = max(2, args.steps - args.steps % 2)
    total_ms, k_ms = timed_steps(S, K // 2, S._pass2_call, True)
    ms_step = total_ms / K
    teff = A / (ms_step * 1e-3) / 1e9
    achieved = 2 * A / (k_ms * 1e-3) / 1e9   # two iterations' algorithmic bytes per launch
    # the same pass in the reference's expression order (bitwise equal to
    # update_stress! + update_velocity!), and one iteration per launch
    S.exact = True
    for _ in range(2):
        S.pass2()
    ex_total, ex_k = timed_steps(S, 20, S._pass2_call, True)
    S.exact = False
    for _ in range(2):
        S.iteration()
    one_total, one_k = timed_steps(S, 40, lambda: S._iter_call(False), True)
    # residual sanity (not timed): one iteration writing the residual fields
    S.iteration(write_res=True)
    err = S.residual()
    return {
        "metric": "T_eff GB/s (% HBM peak); time-to-1e-8 FP32+fix vs FP64; adjoint/fwd cost",
        "value": round(teff, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seedless analytic inclusion setup)",
        "config": {"workload": f"3D PT viscoelastic Stokes, {n}^3 cells, one PT iteration per step "
                               "(two-iteration pass: fused stress+velocity, free-slip V batch in-kernel, "
                               "stress+velocity; then the V batch)",
                   "grid": [n, n, n], "teff_pct_of_peak": round(100 * teff / PEAK_HBM_GBS, 2),
                   "algorithmic_bytes_per_step": A,
                   "teff_reference_convention_GBs": round(reference_bytes(n, isz) / (ms_step * 1e-3) / 1e9, 2),
                   "exact_form": {"kernel_ms_per_iteration": round(ex_k / 2, 4),
                                  "ms_per_step": round(ex_total / 40, 4),
                                  "teff_GBs": round(A / (ex_total / 40 * 1e-3) / 1e9, 2)},
                   "one_iteration_per_launch": {"kernel": "fused3d_kernel", "kernel_ms": round(one_k, 4),
                                                "ms_per_step": round(one_total / 40, 4),
                                                "teff_GBs": round(A / (one_total / 40 * 1e-3) / 1e9, 2)},
                   "residual_after_timing": err},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": counter_traffic(n, args.dtype, "fused3d_tb_kernel"),
                     "kernel": "fused3d_tb_kernel", "kernel_ms": round(k_ms, 4),
                     "algorithmic_bytes_per_launch": 2 * A},
    }


def single_line(S, args, n, isz, A):
    """--mode fused / unfused: one iteration per launch (the fused kernel, or
    the reference's update_stress! + update_velocity! kernels)."""
    fused = args.mode == "fused"
    for _ in range(args.warmup):
        S.iteration() if fused else S.iteration_unfused()
    K = args.steps
    if fused:
        total_ms, k_ms = timed_steps(S, K, lambda: S._iter_call(False), True)
    else:
        import torch
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0.record()
        for q in range(K):
            ev[q][0].record()
            S.iteration_unfused_kernels()
            ev[q][1].record()
            S.apply_bc(S.V)
        t1.record()
        torch.cuda.synchronize()
        total_ms, k_ms = t0.elapsed_time(t1), sum(a.elapsed_time(b) for a, b in ev) / K
    ms_step = total_ms / K
    teff = A / (ms_step * 1e-3) / 1e9
    achieved = A / (k_ms * 1e-3) / 1e9
    if fused:
        S.iteration(write_res=True)
    else:
        S.iteration_unfused()
    err = S.residual()
    kernel = "fused3d_kernel" if fused else "stress_kernel + velocity_kernel"
    return {
        "metric": "T_eff GB/s (% HBM peak); time-to-1e-8 FP32+fix vs FP64; adjoint/fwd cost",
        "value": round(teff, 2), "unit": "GB/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.dtype, "data": "synthetic (seedless analytic inclusion setup)",
        "config": {"workload": f"3D PT viscoelastic Stokes, {n}^3 cells, one PT iteration per step "
                               f"({'fused stress+velocity kernel' if fused else 'reference kernels'} + free-slip V batch)",
                   "grid": [n, n, n], "teff_pct_of_peak": round(100 * teff / PEAK_HBM_GBS, 2),
                   "algorithmic_bytes_per_step": A, "residual_after_timing": err},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": counter_traffic(n, args.dtype, "fused3d_kernel") if fused else None,
                     "kernel": kernel, "kernel_ms": round(k_ms, 4)},
    }


def cpu_line(n, A, args):
    """The oracle on the host's cores: all threads (median of --cpu-iters
    iterations) and one thread (median of --cpu-iters-1t)."""
    c_s, cores, _ = cpu_baseline(n, args.cpu_iters)
    zs = 32
    c1_s, _, frac = cpu_baseline(n, args.cpu_iters_1t, threads=1, zslab=zs)
    return {"value": round(A / c_s / 1e9, 3), "unit": "GB/s", "cores": cores, "kind": "port",
            "sample": f"median of {args.cpu_iters} PT iterations at {n}^3 FP64 after one warm-up (oracle: unfused "
                      f"update_stress!, update_velocity!, bc!, OpenMP over {cores} threads; {c_s:.3f} s/iteration; "
                      f"CPU {cpu_model()})",
            "one_thread": {"value": round(A * frac / c1_s / 1e9, 3), "unit": "GB/s", "cores": 1,
                           "sample": f"median of {args.cpu_iters_1t} iterations after one warm-up, the kernels over "
                                     f"the first {zs} of {n + 2} planes of the Launcher box (bytes scaled by "
                                     f"{frac:.4f}), {c1_s:.3f} s/iteration"}}


def other_targets(args):
    """The metric's other two parts on their own workloads: the time to an
    FP64 residual of 1e-8 (FP32 sweeps + FP64 defect correction vs plain FP64)
    in 2D at 2047² and in 3D at 255³, and the SIA adjoint/forward cost at 512²
    with its finite-difference check."""
    import copy
    import torch
    out = {}
    log = lambda *a: print("#", *a, file=sys.stderr, flush=True)
    a2 = copy.copy(args)
    a2.ttc, a2.n = "2d", 2047
    r = time_to_converge(a2)
    out["time_to_1e-8_2d"] = {"value": r["value"], "unit": r["unit"], "config": r["config"], "runs": r["runs"],
                             "field_errors_vs_f64": r["field_errors_vs_f64"]}
    log("2D time to 1e-8:", r["value"])
    torch.cuda.empty_cache()
    a3 = copy.copy(args)
    a3.n = 512
    r = sia_gradient(a3)
    out["sia_adjoint"] = {"value": r["value"], "unit": r["unit"], "config": r["config"],
                          "fd_worst_rel_err": r["fd_check"]["worst_rel_err"], "fd_directions": r["fd_check"]["directions"],
                          "fd_h": r["fd_check"]["h"]}
    log("SIA gradient cost:", r["value"])
    torch.cuda.empty_cache()
    if not args.no_ttc3d:
        a4 = copy.copy(args)
        a4.ttc, a4.n = "3d", 255
        r = time_to_converge(a4)
        out["time_to_1e-8_3d"] = {"value": r["value"], "unit": r["unit"], "config": r["config"], "runs": r["runs"],
                                 "field_errors_vs_f64": r["field_errors_vs_f64"]}
        log("3D time to 1e-8:", r["value"])
    return out


if __name__ == "__main__":
    main()
