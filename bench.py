"""Benchmark: forward + adjoint gradient evaluations of the reaction-diffusion model on one MI355X.

Workload (default): the 256^3 members of the batching configuration -- an ensemble of 8 members at
256^3, each with its own seeded tissue maps, kappa ~ U(0.01, 0.1), rho ~ U(4, 15), 8-sparse p over
64 Gaussian bases, nt = 50 steps of dt = 0.02, observed data thresholded from a seeded
ground-truth forward run.  One step = one batched evaluation of J and dJ = [g_p | g_kappa | g_rho]
for every member (FP32 bulk, FP64 reductions).  value = gradient evaluations per second of the
whole ensemble, timed on the product path (graph replay, no per-kernel events).  Also reported:
per-kernel durations from one further evaluation with HIP events around each CG kernel launch;
the same evaluation run one member at a time ("plain"); a plain FP32 GPU path built the way the
reference's CUDA backend evaluates a gradient, timed over one whole evaluation
("refstyle_plain"); the relative errors of c(T), J and dJ against the FP64 oracle on the benched
grid at a short horizon (nt = 2), from the same oracle evaluation that gives the CPU baseline
(all host threads), plus a one-thread sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "grad evals/s at 256^3 (batched FP32 vs plain), rel err vs FP64; HBM GB/s"


def traffic_from_profiles(kernel, n, members):
    """HBM bytes per launch of `kernel` (all members active), from a committed counter summary.

    The summaries under profiles/ named hbm_traffic_*.txt hold lines
    `traffic <kernel> <n> <members> <bytes per launch>`, computed from rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes scaled by the same counters read for a device copy of known size."""
    import glob
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "hbm_traffic_*.txt"))):
        for line in open(f):
            p = line.split()
            if len(p) == 5 and p[0] == "traffic" and p[1] == kernel and int(p[2]) == n and int(p[3]) == members:
                return float(p[4])
    return None


def hbm_copy_rate(mib=1024, reps=20):
    """What a plain device-to-device copy of mib MiB moves per second (read + write bytes), timed
    with HIP events after three warm-up copies: the practical HBM rate beside the published peak."""
    import torch
    a = torch.ones(mib * 1024 * 1024 // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return 2.0 * mib * 1024 * 1024 / (ms * 1e-3) / 1e9


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_members(n, count, nbases, nnz, seed0):
    from glia_amd import synth
    kap, rho = synth.ensemble_params(count, seed0)
    out = []
    for m in range(count):
        pr = synth.problem(n, seed0 + m, nbases=nbases, nnz=nnz, kappa=float(kap[m]), rho=float(rho[m]),
                           sigma=2 * np.pi / 64)
        pr["pstar"] = synth.sparse_p(nbases, nnz, seed0 + 1000 + m)
        out.append(pr)
    return out


def build_solver(members, n, nt, dt, checkpoint_slots=0, adjoint_store=False):
    from glia_amd import rd
    s = rd.RDSolver(n, nmembers=len(members), dt=dt, nt=nt, k_gm_wm=0.2, r_gm_wm=0.2,
                    checkpoint_slots=checkpoint_slots, adjoint_store=adjoint_store)
    for m, pr in enumerate(members):
        s.set_tissue(m, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
        s.set_coefficients(m, pr["kappa"], pr["rho"])
        s.set_basis(m, pr["centers"], pr["sigma"])
    return s


def make_data(s, members):
    """d1 = 1[c*(T) > 0.5] from the ground-truth coefficients (setup, not timed)."""
    s.phi_apply([pr["pstar"] for pr in members])
    s.precFactor()
    s.solve_state(0)
    for m, pr in enumerate(members):
        ct = s.get_field(m, "ct")
        pr["d1"] = (ct > 0.5).astype(np.float32)
        s.set_data(m, pr["d1"], None, -0.99)


def xvec(pr):
    return np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]])


def accuracy_and_cpu(n, dt, pr, d1, nt_short):
    """Relative errors of c(T), J and dJ against the FP64 oracle on the benched grid, and the
    oracle's own speed, from one whole evaluation (forward + adjoint + gradient, mode 2: p, kappa
    and rho) of one member at nt_short steps on both sides, with the same d1.  The oracle runs on
    all host threads; evals/s at nt steps scale by nt_short / nt (the per-step work is the same at
    every step, setup is excluded on both sides)."""
    from glia_amd import rd
    from oracle import oracle
    x = xvec(pr)
    s = rd.RDSolver(n, dt=dt, nt=nt_short, k_gm_wm=0.2, r_gm_wm=0.2)
    s.set_tissue(0, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    s.set_coefficients(0, pr["kappa"], pr["rho"])
    s.set_basis(0, pr["centers"], pr["sigma"])
    s.set_data(0, d1, None, -0.99)
    J, g = s.objective_and_gradient([x])
    ct = s.get_field(0, "ct")
    s.close()
    threads = oracle.max_threads()
    o = oracle.Oracle(n)
    o.set_tissue(pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    o.set_time(dt, nt_short)
    o.set_k(pr["kappa"], 0.2, 0.0)
    o.set_rho(pr["rho"], 0.2, 0.0)
    o.prec_factor()
    o.set_phi(pr["centers"], pr["sigma"], 1.0, False)
    o.set_data(d1.astype(np.float64), None, -0.99)
    o.set_time(dt, nt_short)
    log("oracle: one %d^3 evaluation at nt=%d on %d threads" % (n, nt_short, threads))
    oracle.set_verbose(True)
    t0 = time.perf_counter()
    Jo, go = o.objective_and_gradient(x, mode=2)
    t_eval = time.perf_counter() - t0
    oracle.set_verbose(False)
    ksp = o.ksp()
    cto = o.ct()
    del o
    rl = lambda a, b: float(np.linalg.norm(np.ravel(a).astype(np.float64) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))
    acc = {"grid": "%d^3, nt=%d, dt=%g, %d bases" % (n, nt_short, dt, len(pr["p"])), "cT": rl(ct, cto),
           "J": abs(J[0] - Jo) / abs(Jo), "dJ": rl(g[0], go)}
    return acc, t_eval, threads, ksp


def cpu_one_thread(dt):
    """The oracle on one thread over one whole evaluation of a 64^3 member at nt = 2 (a bounded
    sample: the 256^3 evaluation would take tens of minutes on one core)."""
    from glia_amd import synth
    from oracle import oracle
    n, nt_s = 64, 2
    pr = synth.problem(n, 2, nbases=8, kappa=0.05, rho=8.0)
    d1 = (pr["wm"] > 0.5).astype(np.float64)
    prev = oracle.max_threads()
    oracle.set_threads(1)
    try:
        o = oracle.Oracle(n)
        o.set_tissue(pr["wm"], pr["gm"], pr["csf"], pr["vt"])
        o.set_time(dt, nt_s)
        o.set_k(pr["kappa"], 0.2, 0.0)
        o.set_rho(pr["rho"], 0.2, 0.0)
        o.prec_factor()
        o.set_phi(pr["centers"], pr["sigma"], 1.0, False)
        o.set_data(d1, None, -0.99)
        o.set_time(dt, nt_s)
        t0 = time.perf_counter()
        o.objective_and_gradient(xvec(pr), mode=2)
        t = time.perf_counter() - t0
    finally:
        oracle.set_threads(prev)
    return t, n, nt_s


def refstyle_gradient(nt, dt, k, rho, mask, wm, gm, c0, d1, rtol=1e-6, dev="cuda"):
    """The plain FP32 GPU path built the way the reference's CUDA backend evaluates a gradient,
    timed on the same GPU as a comparison leg (torch FFTs; not the engine).  Real-space vectors;
    PETSc-KSPCG preconditioned CG with three host-synchronized reductions per iteration; a matvec
    of eight full 3D transforms (computeGradient: 1 R2C + 3 C2R; computeDivergence: 3 R2C + 1 C2R;
    src/mat/DiffCoef.cpp:255-263, src/grad/SpectralOperators.cpp) and a preconditioner of two
    (applyPC, src/pde/DiffusionSolver.cpp:182).  One whole evaluation of one member: the Strang
    forward storing c_[i] (src/pde/PdeOperators.cpp:235-316), the misfit, the adjoint with the
    half-step state recomputed by a CN solve (adjoint_store_ = false, :318-420) and the
    gradDiffusion / gradReaction quadratures (src/grad/DerivativeOperators.cpp:189-321).  c(0) is
    given; Phi^T (two transforms per basis) is left out, which only flatters this leg.  Returns
    the timing and (J, g_kappa, g_rho) for a consistency check against the engine."""
    import torch
    f32 = torch.float32
    sync = torch.cuda.synchronize if dev == "cuda" else (lambda: None)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=f32, device=dev)
    kd, rd_, wmd, gmd = T(k), T(rho), T(wm), T(gm)
    shape = kd.shape
    h3 = (2 * np.pi) ** 3 / float(np.prod(shape))

    def wn(m, half=False):
        w = torch.fft.rfftfreq(m, 1.0 / m) if half else torch.fft.fftfreq(m, 1.0 / m)
        w = w.clone()
        w[w.abs() == m // 2] = 0.0  # multiplyXWaveNumber zeroes the Nyquist mode
        return w.to(device=dev, dtype=f32)
    wx = wn(shape[0])[:, None, None]
    wy = wn(shape[1])[None, :, None]
    wz = wn(shape[2], True)[None, None, :]
    w2 = wx * wx + wy * wy + wz * wz
    # kxx_avg_ = sum(k) / sum(brain mask) (src/mat/DiffCoef.cpp:103-111); factor 1 / (1 + dt/4 kbar |w|^2)
    kbar = float(np.sum(k, dtype=np.float64) / np.sum(mask, dtype=np.float64))
    its = [0]

    def grad(x):
        X = torch.fft.rfftn(x)
        return [torch.fft.irfftn(1j * w * X, s=shape) for w in (wx, wy, wz)]

    def div(a, b, c):
        return torch.fft.irfftn(1j * wx * torch.fft.rfftn(a) + 1j * wy * torch.fft.rfftn(b) + 1j * wz * torch.fft.rfftn(c),
                                s=shape)

    def applyD(x):
        gx, gy, gz = grad(x)
        return div(kd * gx, kd * gy, kd * gz)

    def cn(c, tau):
        A = lambda v: v - 0.5 * tau * applyD(v)
        P = 1.0 / (1.0 + 0.5 * tau * kbar * w2)
        B = lambda v: torch.fft.irfftn(torch.fft.rfftn(v) * P, s=shape)
        b = c + 0.5 * tau * applyD(c)
        x = c.clone()
        r = b - A(x)
        z = B(r)
        dp = float(torch.linalg.vector_norm(z))
        # KSPConvergedDefault with a nonzero initial guess: ttol = rtol ||B b||
        rnorm0 = float(torch.linalg.vector_norm(B(b)))
        beta = float(torch.sum(z * r))
        p = z.clone()
        for _ in range(1000):
            if dp <= rtol * rnorm0:
                break
            w = A(p)
            alpha = beta / float(torch.sum(p * w))
            x.add_(p, alpha=alpha)
            r.sub_(w, alpha=alpha)
            z = B(r)
            dp = float(torch.linalg.vector_norm(z))
            its[0] += 1
            if dp <= rtol * rnorm0:
                break
            betaold, beta = beta, float(torch.sum(z * r))
            p = z + (beta / betaold) * p
        return x

    e = torch.exp(rd_ * dt)

    def evaluate(c0d, d1d, nsteps):
        C = [c0d.clone()]
        c = c0d.clone()
        for _ in range(nsteps):
            c = cn(c, 0.5 * dt)
            c = c * e / (c * e + 1.0 - c)
            c = cn(c, 0.5 * dt)
            C.append(c)
        t = c - d1d
        J = h3 * 0.5 * float(torch.sum(t.double() * t.double()))
        a = -t
        A = [None] * (nsteps + 1)
        A[nsteps] = a
        for i in range(nsteps):
            j = nsteps - i - 1
            a = cn(a, 0.5 * dt)
            ch = cn(C[j].clone(), 0.5 * dt)
            den = ch * e + 1.0 - ch
            a = a * e / (den * den)
            a = cn(a, 0.5 * dt)
            A[j] = a
        Q = torch.zeros_like(c0d)
        P = torch.zeros_like(c0d)
        for i in range(nsteps + 1):
            wgt = dt * (0.5 if i in (0, nsteps) else 1.0)
            gc, ga = grad(C[i]), grad(A[i])
            Q += wgt * (gc[0] * ga[0] + gc[1] * ga[1] + gc[2] * ga[2])
            P += wgt * (C[i] * C[i] - C[i]) * A[i]
        gk = h3 * (float(torch.sum(wmd.double() * Q.double())) + 0.2 * float(torch.sum(gmd.double() * Q.double())))
        gr = h3 * (float(torch.sum(wmd.double() * P.double())) + 0.2 * float(torch.sum(gmd.double() * P.double())))
        return J, gk, gr

    c0d, d1d = T(c0), T(d1)
    evaluate(c0d, d1d, 1)  # warm-up (plans, allocator)
    its[0] = 0
    sync()
    t0 = time.perf_counter()
    J, gk, gr = evaluate(c0d, d1d, nt)
    sync()
    t_eval = time.perf_counter() - t0
    return {"value": 1.0 / t_eval, "unit": "grad evals/s",
            "sample": "reference-structured plain FP32 GPU path (torch FFTs, real-space PETSc-style CG, 10 full 3D "
                      "FFTs and 3 host-synchronized reductions per CG iteration): one whole gradient evaluation "
                      "(forward, adjoint with recomputed half steps, kappa/rho quadratures; Phi^T not timed) of "
                      "one %d^3 member at nt=%d, %d CG iterations: %.2f s" % (shape[0], nt, its[0], t_eval)}, (J, gk, gr)


def batch_config(args):
    """The batching configuration: 16 x 128^3 + 8 x 192^3 + 8 x 256^3 members, each with its own
    tissue, kappa, rho and 8-sparse p over 64 bases, nt = 50 at dt = 0.02; the three grids are three
    handles evaluated in lockstep from one host thread (time step i of the three handles is one
    graph launch).  One step = one evaluation of J and dJ for all 32 members.  The same handles are
    then timed evaluated concurrently by three host threads (one stream each) and one handle after
    another from one thread, and the lockstep results are checked bit for bit against both."""
    import torch
    from glia_amd import rd
    groups = [(128, 16, 300), (192, 8, 400), (256, 8, 500)]
    t0 = time.perf_counter()
    solvers, xs, names = [], [], []
    for n, count, seed in groups:
        members = make_members(n, count, args.nbases, args.nnz, seed)
        s = build_solver(members, n, args.nt, args.dt, args.checkpoint_slots)
        make_data(s, members)
        solvers.append(s)
        xs.append([xvec(pr) for pr in members])
        names.append("%dx%d^3" % (count, n))
        log("group %s ready (%.1f s)" % (names[-1], time.perf_counter() - t0))
    total = sum(c for _, c, _ in groups)

    def timed(fn, label):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res = None
        for _ in range(args.steps):
            res = fn()
        torch.cuda.synchronize()
        dt_ = time.perf_counter() - t1
        log("%s: %.3f s per evaluation of the 32 members" % (label, dt_ / args.steps))
        return dt_, res

    def one_after_another():
        Js, gs = [], []
        for s, x in zip(solvers, xs):
            J, g = s.objective_and_gradient(x)
            Js.append(J)
            gs.append(g)
        return Js, gs, 0

    t_lock, (Jl, gl, spanned) = timed(lambda: rd.objective_and_gradient_lockstep(solvers, xs), "lockstep")
    t_thr, (Jt, gt, _) = timed(lambda: rd.objective_and_gradient_group(solvers, xs), "threads")
    t_seq, (Jq, gq, _) = timed(one_after_another, "one handle after another")

    def same_as(Ja, ga):
        return all(np.array_equal(a, b) for a, b in zip(Jl, Ja)) and all(
            np.array_equal(u, v) for gx, gy in zip(gl, ga) for u, v in zip(gx, gy))
    same = same_as(Jt, gt)
    same_seq = same_as(Jq, gq)
    value = total * args.steps / t_lock
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "grad evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_lock * 1000.0 / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched ensemble of uneven grids, 16x128^3 + 8x192^3 + 8x256^3, nt=%d, dt=%g, "
                               "%d bases (%d nonzero), three handles in lockstep (one graph per time step)"
                               % (args.nt, args.dt, args.nbases, args.nnz),
                   "groups": names, "nt": args.nt, "dt": args.dt},
        "steps_spanning_handles_last_eval": spanned,
        "threaded_value": round(total * args.steps / t_thr, 4),
        "sequential_value": round(total * args.steps / t_seq, 4),
        "lockstep_equals_threaded_bitwise": bool(same),
        "lockstep_equals_sequential_bitwise": bool(same_seq),
    }
    for s in solvers:
        s.close()
    print(json.dumps(out), flush=True)
    return 0


def diff_config(args):
    """The differentiation configuration at 256^3: nt = 100 steps of dt = 0.01, 128 Gaussian bases
    (8 nonzero), gradient w.r.t. p, kappa and rho through the torch.autograd binding; binomial
    checkpointing with 10 stored states against store-all (time, device memory, bitwise equality);
    central finite differences along 10 random unit directions (h = 1e-4 by default: along the p
    block the objective's curvature makes the O(h^2) truncation of h = 1e-3 reach 16%, which the
    step-size sweep reported beside it shows falling to 0.1% at h = 1e-4)."""
    import torch
    from glia_amd import rd
    from glia_amd.autograd import rd_objective
    n, nt, dt, nb = 256, 100, 0.01, 128
    members = make_members(n, 1, nb, 8, 700)
    out = {}
    grads = {}
    for slots in (0, 10):
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        s = build_solver(members, n, nt, dt, slots)
        make_data(s, members)
        free1 = torch.cuda.mem_get_info()[0]
        x0 = xvec(members[0])
        s.objective([x0])  # warm-up; later evaluations share the solver-context dt
        x = torch.tensor(x0[None, :], dtype=torch.float64, requires_grad=True)
        s.sync()
        t0 = time.perf_counter()
        J = rd_objective(x, s)
        J.sum().backward()
        s.sync()
        t_grad = time.perf_counter() - t0
        g = x.grad.numpy()[0].copy()
        grads[slots] = (float(J.detach()[0]), g)
        rec = {"grad_eval_s": round(t_grad, 3), "device_bytes": int(free0 - free1)}
        if slots == 10:
            # J carries FP32 and PCG-tolerance (rtol 1e-6) noise that 1/(2h) amplifies, so the
            # error is reported both against |<dJ, v>| and against ||dJ||
            rng = np.random.default_rng(3)
            fds, ads = [], []
            t0 = time.perf_counter()
            for _ in range(10):
                v = rng.standard_normal(x0.size)
                v /= np.linalg.norm(v)
                h = args.fd_h
                fds.append((s.objective([x0 + h * v])[0] - s.objective([x0 - h * v])[0]) / (2 * h))
                ads.append(float(g @ v))
            fds, ads = np.array(fds), np.array(ads)
            gn = float(np.linalg.norm(g))
            rec["fd"] = [float("%.6g" % f) for f in fds]
            rec["adjoint"] = [float("%.6g" % a) for a in ads]
            rec["grad_norm"] = float("%.6g" % gn)
            rec["fd_err_rel_directional"] = [float("%.3g" % e) for e in np.abs(fds - ads) / np.abs(ads)]
            rec["fd_err_over_grad_norm"] = [float("%.3g" % e) for e in np.abs(fds - ads) / gn]
            # the bar of the 64^3 finite-difference test: |fd - ad| <= 2e-2 |ad| + 1e-3 ||dJ||
            rec["fd_pass"] = [bool(abs(f - a) <= 2e-2 * abs(a) + 1e-3 * gn) for f, a in zip(fds, ads)]
            # where a disagreement comes from: a direction inside the p block (there the gradient
            # is the discrete adjoint up to the CG tolerance) and the kappa and rho axes alone (the
            # reference's trapezoid quadratures match the discrete derivative to O(dt) only), each
            # at three step sizes (truncation error falls as h^2, noise in J grows as 1/h)
            vp = rng.standard_normal(x0.size)
            vp[-2:] = 0.0
            vp /= np.linalg.norm(vp)
            ek = np.zeros(x0.size)
            ek[-2] = 1.0
            er = np.zeros(x0.size)
            er[-1] = 1.0
            sweeps = {}
            for name, v, hs in (("p_block", vp, (1e-2, 1e-3, 1e-4)), ("kappa", ek, (1e-3, 1e-4, 1e-5)),
                                ("rho", er, (1e-2, 1e-3, 1e-4))):
                ad = float(g @ v)
                fdh = {}
                for h in hs:
                    fdh["%g" % h] = float("%.7g" % ((s.objective([x0 + h * v])[0] - s.objective([x0 - h * v])[0]) / (2 * h)))
                sweeps[name] = {"adjoint": float("%.7g" % ad), "fd": fdh,
                                "rel_err": {k: float("%.3g" % (abs(f - ad) / abs(ad))) for k, f in fdh.items()}}
            rec["fd_sweeps"] = sweeps
            rec["fd_s"] = round(time.perf_counter() - t0, 2)
        out["checkpoint_%d" % slots if slots else "store_all"] = rec
        s.close()
        del s
    out["bitwise_equal"] = bool(grads[0][0] == grads[10][0] and np.array_equal(grads[0][1], grads[10][1]))
    line = {
        "metric": METRIC,
        "value": round(1.0 / out["checkpoint_10"]["grad_eval_s"], 4),
        "unit": "grad evals/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(out["checkpoint_10"]["grad_eval_s"] * 1000, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "differentiation: one 256^3 member, nt=100, dt=0.01, 128 bases (8 nonzero), "
                               "torch.autograd, 10 checkpoint slots vs store-all, FD h=%g on 10 directions" % args.fd_h,
                   "n": n, "nt": nt, "dt": dt, "nbases": nb},
        "diff": out,
    }
    print(json.dumps(line), flush=True)
    return 0


def invert_config(args):
    """The inversion loop (glia_rd_invert, RDOptimizer semantics) on the 128^3 part of the batching
    configuration: 16 members, each with its own tissue, kappa* ~ U(0.01, 0.1), rho* ~ U(4, 15) and
    c0 = Phi p* from an 8-sparse p* over 64 bases; exact data d = c*(T); every member starts from
    (kappa, rho) = (0.05, 8) with the reference's bounds [1e-3, 1] x [1, 15], the reference's
    optimizer defaults (More-Thuente line search) except gttol = 1e-5 and max_it = 100 (the default
    relative tolerance 1e-3 can stop on a flat stretch far from the minimizer).  One step = one
    batched inversion of all members."""
    import torch
    from glia_amd import opt
    n, M = 128, 16
    members = make_members(n, M, args.nbases, args.nnz, 300)
    s = build_solver(members, n, args.nt, args.dt, args.checkpoint_slots)
    s.phi_apply([pr["pstar"] for pr in members])
    s.precFactor()
    s.solve_state(0)
    for m in range(M):
        s.set_data(m, s.get_field(m, "ct"), None, -0.99)
    nb = args.nbases
    x0 = [np.concatenate([pr["pstar"], [0.05, 8.0]]) for pr in members]
    lb = np.concatenate([np.full(nb, -np.inf), [1e-3, 1.0]])
    ub = np.concatenate([np.full(nb, np.inf), [1.0, 15.0]])
    torch.cuda.synchronize()
    s.sync()
    t0 = time.perf_counter()
    xs, res = opt.invert(s, x0, lb, ub, mode=1, gttol=1e-5, max_it=100)
    s.sync()
    t_inv = time.perf_counter() - t0
    s.close()
    fevals = max(r["fevals"] for r in res)  # batched rounds: every round evaluates all members
    ek = [abs(x[-2] - pr["kappa"]) / pr["kappa"] for x, pr in zip(xs, members)]
    er = [abs(x[-1] - pr["rho"]) / pr["rho"] for x, pr in zip(xs, members)]
    reasons = {}
    for r in res:
        reasons[r["reason"]] = reasons.get(r["reason"], 0) + 1
    line = {
        "metric": METRIC,
        "value": round(M * fevals / t_inv, 4),
        "unit": "grad evals/s",
        "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(t_inv * 1000, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched kappa/rho inversion (bounded L-BFGS, More-Thuente, reference stopping test "
                               "with gttol 1e-5), 16 members at 128^3, nt=%d, dt=%g, exact data, start (0.05, 8)"
                               % (args.nt, args.dt),
                   "n": n, "members": M, "nt": args.nt, "dt": args.dt},
        "inversions_per_s": round(M / t_inv, 4),
        "evaluation_rounds": fevals,
        "iterations": [r["iterations"] for r in res],
        "reasons": reasons,
        "kappa_rel_err_max": float("%.3g" % max(ek)),
        "rho_rel_err_max": float("%.3g" % max(er)),
    }
    print(json.dumps(line), flush=True)
    return 0


def advect_config(args):
    """The semi-Lagrangian transport step of the mass-effect models
    (src/pde/PdeOperatorsMassEffect.cpp:605-620) at 256^3: a new velocity (the reference's test field,
    src/test/helper.cpp:47-80), RK2 trajectories, div v, and six fields (gm, wm, vt, csf, c, mri with
    modes 1, 1, 2, 2, 1, 2) advected in one pass.  The advection kernel is also timed alone (fresh
    trajectories) for its roofline: per point it reads the offsets (12 B), div v at the grid and
    departure points (8 B) and each field (4 B), and writes each field (4 B)."""
    import torch
    from glia_amd import adv
    from glia_amd.synth import reference_velocity
    n, dt = 256, 0.02
    ng = n ** 3
    torch.manual_seed(11)
    v = [torch.from_numpy(a.astype(np.float32).reshape(-1)).cuda() for a in reference_velocity(n)]
    fields = [torch.rand(ng, device="cuda") for _ in range(6)]
    outs = [torch.empty_like(f) for f in fields]
    modes = [1, 1, 2, 2, 1, 2]
    torch.cuda.synchronize()
    # the solver's work and the timing events share one stream (default priority)
    stream = torch.cuda.Stream()
    s = adv.SemiLagrangianSolver(n, stream=stream.cuda_stream)

    def step():
        s.set_velocity(*v)
        s.advect(fields, modes, dt, out=outs)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    # the advection kernel alone: trajectories and div v fresh, so each call is one launch
    reps = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        s.advect(fields, modes, dt, out=outs)
    e1.record(stream)
    torch.cuda.synchronize()
    k_ms = e0.elapsed_time(e1) / reps
    s.close()
    bytes_launch = ng * (12 + 8 + 8 * len(fields))
    achieved = bytes_launch / (k_ms / 1000.0) / 1e9
    line = {
        "metric": "advection steps/s at 256^3 (6 fields, RK2 trajectories, cubic Lagrange)",
        "value": round(1.0 / t_step, 3),
        "unit": "steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_step * 1000, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "semi-Lagrangian step of the mass-effect model at 256^3: velocity update, RK2 "
                               "trajectories, spectral div v, 6 fields advected in one pass (dt 0.02)",
                   "n": n, "fields": len(fields), "modes": modes},
        "roofline": {"bound": "hbm", "kernel": "sl_apply", "achieved": round(achieved, 1), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(achieved / 8000.0, 4),
                     "algorithmic_bytes_per_launch": bytes_launch, "traffic": None},
        "kernel_ms": round(k_ms, 4),
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def elastic_config(args):
    """The linear-elasticity solve of the mass-effect model (VariableLinearElasticitySolver,
    src/pde/ElasticitySolver.cpp:471-533) at 160^3, the mass-effect grid: synthetic tissue with a
    background, a Gaussian tumor, material properties from them, and the force along the tumor
    gradient; one step = one CG solve from u = 0 (rtol 1e-3 in the preconditioned norm) on
    device-resident vectors.
    Algorithmic bytes per CG iteration: 167 fields of 4 N^3 bytes (operator 110: 20 transforms,
    the spectral combinations, 7 products and 3 screened updates, with the CG direction transformed
    in place; preconditioner 18: 3 + 3 transforms in place and the spectral inverse; <p, Ap> and the
    fused <z, z>, <z, r> 12; the CG updates 27)."""
    import torch
    from glia_amd import elasticity
    from glia_amd.synth import tissue_and_tumor, tumor_force
    n = 160
    fields = [a.astype(np.float32) for a in tissue_and_tumor(n)]
    f = tumor_force(fields[0].astype(np.float64)).astype(np.float32)
    s = elasticity.ElasticitySolver(n)
    s.set_material(*fields)
    # inputs resident in HBM: the timed solves read f and write u on the device
    fd = torch.from_numpy(f.reshape((3,) + (n, n, n)).copy()).cuda()
    ud = torch.zeros_like(fd)
    for _ in range(args.warmup):
        ud.zero_()
        s.solve_device(fd, ud)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = []
    for _ in range(args.steps):
        ud.zero_()
        k, reason, rn = s.solve_device(fd, ud)
        its.append(k)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    # the same solve with host vectors (PCIe staging of f, u in and u out), reported beside
    t1 = time.perf_counter()
    s.solve(f)
    t_host = time.perf_counter() - t1
    s.close()
    R = 4.0 * n ** 3
    bytes_it = 167 * R
    achieved = bytes_it * np.mean(its) / t / 1e9
    line = {
        "metric": "elasticity solves/s at 160^3 (CG, spectral preconditioner)",
        "value": round(1.0 / t, 3), "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t * 1000, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "linear elasticity of the mass-effect model at 160^3: material properties from "
                               "synthetic tissue and a Gaussian tumor, force along the tumor gradient, KSPCG "
                               "rtol 1e-3 from u = 0", "n": n},
        "cg_iterations": its, "reason": reason, "ms_per_iteration": round(t * 1000 / np.mean(its), 3),
        "ms_per_solve_host_vectors": round(t_host * 1000, 2),
        "roofline": {"bound": "hbm", "kernel": "whole CG iteration", "achieved": round(achieved, 1), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(achieved / 8000.0, 4), "algorithmic_bytes_per_iteration": bytes_it,
                     "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def masseffect_config(args):
    """The mass-effect forward of model 4 (PdeOperatorsMassEffect::solveState,
    src/pde/PdeOperatorsMassEffect.cpp:421-743) at 160^3 with 40 steps of dt = 0.025: synthetic
    tissue with all four classes, a Gaussian tumor seed, k = 0.01, rho = 8, forcing factor 6e4 (the
    reference's value for the SRI atlas; at 1.2e5 this tumor reaches the CFL limit of 14 after 31
    steps and the loop stops); one step = one whole forward (nt + 1 = 41 time steps, 42 elasticity
    solves).
    Algorithmic bytes per time step, in fields of 4 N^3 bytes: 167 per elasticity CG iteration,
    55 + 36 m per diffusion solve of m CG iterations, and 270 for the rest (elasticity set-up
    with the initial residual 150: force 10, material 10, operator and preconditioner 130; CFL 3,
    coefficients 6, velocity staging 6, trajectories and div v 14, advection of five fields 14,
    copies back 10, transforms around the diffusion solve 4, reaction 3, velocity 12, its
    smoothing 24, 24 for the engine's staging)."""
    from glia_amd.mass_effect import MassEffectSolver
    from glia_amd.synth import mass_effect_case
    n, dt, nt, FF = 160, 0.025, 40, 6e4
    c0, wm, gm, vt, csf, bg = [a.astype(np.float32) for a in mass_effect_case(n)]
    s = MassEffectSolver(n, dt, nt, k=0.01, rho=8.0)
    s.set_tissue(wm, gm, csf, vt)
    for _ in range(args.warmup):
        s.solve_state(c0, wm, gm, vt, csf, bg, forcing_factor=FF)
    times, infos = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        out = s.solve_state(c0, wm, gm, vt, csf, bg, forcing_factor=FF)
        times.append(time.perf_counter() - t0)
        infos.append(out["info"])
        print("forward %.3f s info %s" % (times[-1], out["info"]), file=sys.stderr, flush=True)
    s.close()
    t = float(np.mean(times))
    info = infos[-1]
    steps = info["steps"]
    R = 4.0 * n ** 3
    m_diff = info["diff_iterations"] / max(steps, 1)
    bytes_fwd = R * (167.0 * info["el_iterations"] + steps * (55.0 + 36.0 * m_diff + 270.0))
    achieved = bytes_fwd / t / 1e9
    line = {
        "metric": "mass-effect forward time steps/s at 160^3 (model 4)",
        "value": round(steps / t, 3), "unit": "time steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t * 1000, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "mass-effect forward of model 4 at 160^3: 40 steps of dt 0.025 (41 with the "
                               "reference's nt + 1 loop), k 0.01, rho 8, forcing factor 6e4, synthetic tissue "
                               "and Gaussian tumor; one step here = one whole forward", "n": n, "nt": nt,
                   "dt": dt},
        "time_steps": steps, "elasticity_cg_iterations": info["el_iterations"],
        "diffusion_cg_iterations": info["diff_iterations"], "cfl_max": round(info["cfl_max"], 4),
        "roofline": {"bound": "hbm", "kernel": "whole forward", "achieved": round(achieved, 1), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(achieved / 8000.0, 4),
                     "algorithmic_bytes_per_forward": bytes_fwd, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def multispecies_config(args):
    """The multispecies mass-effect forward of model 5 (PdeOperatorsMultiSpecies::solveState,
    src/pde/PdeOperatorsMultiSpecies.cpp:235-508) at 160^3 with 40 steps of dt = 0.025: the tissue,
    tumor seed and coefficients of the model-4 configuration, the reference's species parameters,
    forcing factor 6e4 and the model's CFL limit 8; one step = one whole forward (41 time steps).
    Algorithmic bytes per time step, in fields of 4 N^3 bytes: 167 per elasticity CG iteration,
    55 + 36 m per diffusion solve of m CG iterations, and 330 for the rest (the 270 of the model-4
    step, two more advected fields 6, k and its mean 10, clipping 14, the smoothing of i 8, D i 22)."""
    from glia_amd.mass_effect import MassEffectSolver
    from glia_amd.synth import mass_effect_case
    n, dt, nt, FF = 160, 0.025, 40, 6e4
    c0, wm, gm, vt, csf, bg = [a.astype(np.float32) for a in mass_effect_case(n)]
    s = MassEffectSolver(n, dt, nt, k=0.01, rho=8.0)
    s.set_tissue(wm, gm, csf, vt)
    for _ in range(args.warmup):
        s.solve_state_multispecies(c0, wm, gm, vt, csf, bg, forcing_factor=FF)
    times, infos = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        out = s.solve_state_multispecies(c0, wm, gm, vt, csf, bg, forcing_factor=FF)
        times.append(time.perf_counter() - t0)
        infos.append(out["info"])
        print("forward %.3f s info %s" % (times[-1], out["info"]), file=sys.stderr, flush=True)
    s.close()
    t = float(np.mean(times))
    info = infos[-1]
    steps = info["steps"]
    R = 4.0 * n ** 3
    m_diff = info["diff_iterations"] / max(steps, 1)
    bytes_fwd = R * (167.0 * info["el_iterations"] + steps * (55.0 + 36.0 * m_diff + 330.0))
    achieved = bytes_fwd / t / 1e9
    line = {
        "metric": "multispecies mass-effect forward time steps/s at 160^3 (model 5)",
        "value": round(steps / t, 3), "unit": "time steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t * 1000, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "multispecies mass-effect forward of model 5 at 160^3: 40 steps of dt 0.025 (41 "
                               "with the reference's nt + 1 loop), k 0.01, rho 8, forcing factor 6e4, default "
                               "species parameters, synthetic tissue and Gaussian tumor; one step here = one "
                               "whole forward", "n": n, "nt": nt, "dt": dt},
        "time_steps": steps, "elasticity_cg_iterations": info["el_iterations"],
        "diffusion_cg_iterations": info["diff_iterations"], "cfl_max": round(info["cfl_max"], 4),
        "roofline": {"bound": "hbm", "kernel": "whole forward", "achieved": round(achieved, 1), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(achieved / 8000.0, 4),
                     "algorithmic_bytes_per_forward": bytes_fwd, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--members", type=int, default=8)
    ap.add_argument("--nt", type=int, default=50)
    ap.add_argument("--dt", type=float, default=0.02)
    ap.add_argument("--nbases", type=int, default=64)
    ap.add_argument("--nnz", type=int, default=8)
    ap.add_argument("--checkpoint-slots", type=int, default=0)
    ap.add_argument("--fd-h", type=float, default=1e-4, help="finite-difference step of --config diff")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-store-leg", action="store_true",
                    help="skip the evaluations with the forward's half-step states kept for the adjoint")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--nt-oracle", type=int, default=2,
                    help="time steps of the oracle evaluation that gives rel_err_fp64 and cpu_baseline")
    ap.add_argument("--config", choices=["ensemble256", "batch", "diff", "invert", "advect", "elastic",
                                         "masseffect", "multispecies"],
                    default="ensemble256",
                    help="ensemble256: the 256^3 members (default, the headline metric); batch: the "
                         "uneven-grid batching configuration; diff: the 256^3 differentiation "
                         "configuration (checkpointing, finite differences); invert: a batched kappa/rho "
                         "inversion of 16 members at 128^3 through the inversion loop; advect: the "
                         "semi-Lagrangian step of the mass-effect models at 256^3; elastic: the "
                         "elasticity solve of the mass-effect model at 160^3; masseffect: the whole "
                         "mass-effect forward of model 4 at 160^3, 40 steps; multispecies: the "
                         "multispecies forward of model 5 at 160^3, 40 steps")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB device-to-device copy first (the reference dispatch for "
                         "scaling FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (ensemble members "
              "would shard across GPUs with no communication in the hot loop); run with --gpus 1",
              file=sys.stderr)
        return 2

    import torch
    from glia_amd import _lib
    _lib.load()
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 3
    if args.config == "batch":
        return batch_config(args)
    if args.config == "diff":
        return diff_config(args)
    if args.config == "invert":
        return invert_config(args)
    if args.config == "advect":
        return advect_config(args)
    if args.config == "elastic":
        return elastic_config(args)
    if args.config == "masseffect":
        return masseffect_config(args)
    if args.config == "multispecies":
        return multispecies_config(args)
    n, M, nt, dt = args.n, args.members, args.nt, args.dt
    if args.copy_calibration:
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    t0 = time.perf_counter()
    members = make_members(n, M, args.nbases, args.nnz, 100)
    log("synthetic inputs: %.1f s" % (time.perf_counter() - t0))
    s = build_solver(members, n, nt, dt, args.checkpoint_slots)
    make_data(s, members)
    log("setup + data: %.1f s" % (time.perf_counter() - t0))
    xs = [xvec(pr) for pr in members]

    for _ in range(args.warmup):
        s.objective_and_gradient(xs)
    # timed region: the product path (CG iterations replayed from captured graphs, no per-kernel
    # events)
    s.reset_stats()
    torch.cuda.synchronize()
    s.sync()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        J, g = s.objective_and_gradient(xs)
    s.sync()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    st = [s.stats(m) for m in range(M)]
    ms_step = (t2 - t1) * 1000.0 / args.steps
    value = M * args.steps / (t2 - t1)
    its = sum(x["pcg_iterations"] for x in st)
    solves = sum(x["cn_solves"] for x in st)
    mbar = its / max(solves, 1)
    log("batched: %.1f ms/step, %.3f grad evals/s, %.2f CG its per CN solve" % (ms_step, value, mbar))

    # one more evaluation with HIP events around every CG kernel launch on the solver stream
    # (graph replay off): the per-kernel durations.  A launch covers every member, but converged
    # members exit at once, so the bytes of one kernel are (fields per member-iteration) * (CG
    # iterations summed over members).
    s.reset_stats()
    s.profile(True)
    s.objective_and_gradient(xs)
    s.sync()
    prof = s.profile_get()
    s.profile(False)
    its_p = sum(s.stats(m)["pcg_iterations"] for m in range(M))
    field_bytes = 4.0 * n ** 3
    dom = max(prof, key=lambda k: prof[k][1])
    cnt, ms = prof[dom]
    achieved = _lib.KERNEL_FIELDS[dom] * field_bytes * its_p / (ms / 1000.0) / 1e9
    kernels = {k: {"launches": v[0], "ms": round(v[1], 3), "member_iterations": its_p,
                   "GBs": round(_lib.KERNEL_FIELDS[k] * field_bytes * its_p / (v[1] / 1000.0) / 1e9, 1)}
               for k, v in prof.items() if v[0] > 0}
    # whole-evaluation model of algorithmic bytes: B = R [nt (230 + 144 m) + 20 (nt + 1)] per member
    b_eval = field_bytes * (nt * (230 + 144 * mbar) + 20 * (nt + 1))
    model_gbs = b_eval * M * args.steps / (t2 - t1) / 1e9
    c0_member0 = s.get_field(0, "c0").astype(np.float64)
    k0, rho0, mask0 = s.get_field(0, "k"), s.get_field(0, "rho"), s.get_field(0, "filter")
    del s

    plain = None
    eng0 = None
    if not args.no_plain:
        # the same evaluations, one member at a time (a one-member solver per member)
        tsum = 0.0
        for m in range(M):
            sp = build_solver(members[m:m + 1], n, nt, dt, args.checkpoint_slots)
            sp.set_data(0, members[m]["d1"], None, -0.99)
            sp.objective_and_gradient(xs[m:m + 1])  # warm-up (graph capture)
            sp.sync()
            tp0 = time.perf_counter()
            Jm, gm_ = sp.objective_and_gradient(xs[m:m + 1])
            sp.sync()
            tsum += time.perf_counter() - tp0
            if m == 0:
                eng0 = (Jm[0], gm_[0][-2], gm_[0][-1])
            sp.close()
        plain = M / tsum
        log("plain (one member at a time): %.3f grad evals/s" % plain)

    store = None
    if not args.no_store_leg:
        # the same evaluations with the forward's half-step states kept for the adjoint instead of
        # recomputed by a CN solve per step (adjoint_store_, set in the reference's shipped configs
        # and scripts; the library default recomputes); the results are the same bit for bit
        ss = build_solver(members, n, nt, dt, args.checkpoint_slots, adjoint_store=True)
        for m in range(M):
            ss.set_data(m, members[m]["d1"], None, -0.99)
        for _ in range(max(args.warmup, 1)):
            Js, gs = ss.objective_and_gradient(xs)
        ss.sync()
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        for _ in range(args.steps):
            Js, gs = ss.objective_and_gradient(xs)
        ss.sync()
        torch.cuda.synchronize()
        t4 = time.perf_counter()
        same = bool(np.array_equal(np.asarray(Js), np.asarray(J)) and
                    all(np.array_equal(np.asarray(a), np.asarray(b)) for a, b in zip(gs, g)))
        store = {"value": round(M * args.steps / (t4 - t3), 4), "ms_per_step": round((t4 - t3) * 1000.0 / args.steps, 2),
                 "same_result_bitwise": same}
        ss.close()
        del ss
        log("adjoint_store leg: %s" % store)

    ref = None
    if not args.no_plain:
        pr0 = members[0]
        ref, (Jr, gkr, grr) = refstyle_gradient(nt, dt, k0, rho0, mask0, pr0["wm"], pr0["gm"],
                                                c0_member0.astype(np.float32), pr0["d1"])
        # the comparison leg computes the same quantities as the engine (FP32 both, different CG
        # iterates): J, g_kappa and g_rho of member 0
        ref["vs_engine_rel"] = {"J": float("%.3g" % (abs(Jr - eng0[0]) / abs(eng0[0]))),
                                "g_kappa": float("%.3g" % (abs(gkr - eng0[1]) / abs(eng0[1]))),
                                "g_rho": float("%.3g" % (abs(grr - eng0[2]) / abs(eng0[2])))}
        log("reference-structured plain path:", ref)

    copy_gbs = hbm_copy_rate()
    log("device copy of 1 GiB: %.1f GB/s (read + write)" % copy_gbs)

    acc, cpu = None, None
    if not (args.no_accuracy and args.no_cpu):
        acc, t_orc, threads, ksp = accuracy_and_cpu(n, dt, members[0], members[0]["d1"], args.nt_oracle)
        log("accuracy vs FP64 oracle:", acc)
        cpu = {"value": args.nt_oracle / (t_orc * nt), "unit": "grad evals/s", "cores": threads, "kind": "port",
               "sample": "oracle FP64 (the reference path restated), one whole gradient evaluation (forward, adjoint "
                         "with recomputed half steps, Phi^T, kappa/rho quadratures) of one %d^3 member at nt=%d "
                         "(%d + %d CG iterations): %.1f s on %d threads, scaled to nt=%d by the step count"
                         % (n, args.nt_oracle, ksp["state"], ksp["adjoint"], t_orc, threads, nt)}
        if not args.no_cpu:
            t1t, n1, nt1 = cpu_one_thread(dt)
            cpu["one_thread"] = {"seconds": round(t1t, 2), "sample": "one whole evaluation of a %d^3 member at nt=%d "
                                 "on 1 thread (grad evals/s there: %.4g)" % (n1, nt1, 1.0 / t1t)}
        log("cpu baseline:", cpu)

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "grad evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched forward+adjoint gradient, %d members at %d^3, nt=%d, dt=%g, %d bases "
                               "(%d nonzero), FP32 bulk + FP64 reductions" % (M, n, nt, dt, args.nbases, args.nnz),
                   "n": n, "members": M, "nt": nt, "dt": dt, "nbases": args.nbases,
                   "checkpoint_slots": args.checkpoint_slots},
        "timing": "value: product path (graph replay, no per-kernel events); roofline and kernels: one further "
                  "evaluation with HIP events around each CG kernel launch (graph replay off)",
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes_per_launch": _lib.KERNEL_FIELDS[dom] * field_bytes * M,
                     "traffic": traffic_from_profiles(dom, n, M)},
        "cpu_baseline": cpu,
        "plain_value": None if plain is None else round(plain, 4),
        "batched_vs_plain": None if plain is None else round(value / plain, 3),
        "refstyle_plain": ref,
        "batched_vs_refstyle_plain": None if ref is None else round(value / ref["value"], 2),
        "adjoint_store": store,
        "cg_iterations_per_solve": round(mbar, 3),
        "model_hbm_GBs": round(model_gbs, 1),
        "model_hbm_frac": round(model_gbs / HBM_PEAK_GBS, 4),
        "hbm_copy_GBs": round(copy_gbs, 1),
        "kernels": kernels,
        "rel_err_fp64": acc,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
