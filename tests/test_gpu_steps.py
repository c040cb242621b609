"""Time steps replayed from captured graphs, and the variant autotuner, through the C ABI.

Each forward and adjoint step of a store-all evaluation runs as one captured graph whose CG loops
are unrolled to the last iteration counts plus one; converged members' kernels exit at once, so
the result is the eager evaluation's bit for bit.  When a solve needs more iterations than the
capture holds, the step runs again eagerly from its stored start (GLIA_STEP_MARGIN = -1 forces
that path on every step).  The autotuner times the variants of the forward divergence passes;
they differ only in FP32 rounding."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NT, DT = 64, 6, 0.02


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def members():
    from glia_amd import synth
    out = []
    for seed, (kap, rho) in zip((41, 42), ((0.03, 7.0), (0.07, 11.0))):
        pr = synth.problem(N, seed, nbases=6, nnz=4, kappa=kap, rho=rho)
        pr["d1"] = (pr["wm"] > 0.6).astype(np.float32)
        out.append(pr)
    return out


def make(members, env=None, **kw):
    from glia_amd import rd
    old = {}
    for k, v in (env or {}).items():
        old[k] = os.environ.get(k)
        os.environ[k] = v
    try:
        s = rd.RDSolver(N, nmembers=len(members), dt=DT, nt=NT, k_gm_wm=0.2, r_gm_wm=0.2, beta=1e-3, **kw)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    for m, pr in enumerate(members):
        s.set_tissue(m, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
        s.set_coefficients(m, pr["kappa"], pr["rho"])
        s.set_basis(m, pr["centers"], pr["sigma"])
        s.set_data(m, pr["d1"], None, -0.99)
    return s


def xs(members):
    return [np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]]) for pr in members]


@pytest.mark.parametrize("adjoint_store", [False, True])
def test_captured_steps_equal_eager_steps_bitwise(members, adjoint_store):
    """Evaluations (and a Gauss-Newton Hessian matvec, whose incremental adjoint runs captured
    steps without quadratures) with captured steps, with every captured step run again eagerly,
    and with graphs off: J, dJ and H v bit for bit equal."""
    print("case captured steps, adjoint_store", adjoint_store, flush=True)
    x = xs(members)
    v = [np.concatenate([np.random.default_rng(3 + m).standard_normal(6), [0.0, 0.0]]) for m in range(2)]
    res = {}
    for name, env, graphs in (("eager", None, False), ("captured", None, True), ("rerun", {"GLIA_STEP_MARGIN": "-1"}, True)):
        s = make(members, env, adjoint_store=adjoint_store)
        s.use_graphs(graphs)
        out = [s.objective_and_gradient(x) for _ in range(3)]
        hv = s.hessian(v)
        launched, reruns, cached = s.step_info()
        print(name, "steps launched", launched, "reruns", reruns, "graphs", cached, flush=True)
        if name == "eager":
            assert launched == 0
        elif name == "captured":
            assert launched > 0 and cached > 0
        else:
            assert reruns > 0
        res[name] = (out, hv)
        s.close()
    for name in ("captured", "rerun"):
        for (J0, g0), (J1, g1) in zip(res["eager"][0], res[name][0]):
            assert np.array_equal(J0, J1), name
            for a, b in zip(g0, g1):
                assert np.array_equal(a, b), name
        for a, b in zip(res["eager"][1], res[name][1]):
            assert np.array_equal(a, b), name


def test_autotune_choice_cache_and_results(members, tmp_path, oracle_mod):
    """The autotuner picks a variant per forward divergence pass, writes it to the cache, and a
    later handle reads the cache instead of timing again; every variant combination evaluates to
    the same J and dJ within FP32 rounding (1e-6), and within the oracle bars."""
    print("case autotune", flush=True)
    cache = tmp_path / "tune.txt"
    x = xs(members)
    s = make(members)
    choice = s.autotune(str(cache))
    print("choice", choice, open(cache).read(), flush=True)
    assert choice[0] in (0, 1, 2) and choice[1] in (0, 1, 2, 3)
    lines = open(cache).read().splitlines()
    assert len(lines) == 1 and len(lines[0].split()) == 14
    Jt, gt = s.objective_and_gradient(x)
    s.close()
    s = make(members)
    assert s.autotune(str(cache)) == choice
    assert len(open(cache).read().splitlines()) == 1
    s.close()
    by_split = {}
    for split in ("none", "xf", "xf3", "yf", "both", "yfb", "xf,yfb2"):
        s = make(members, {"GLIA_SPLIT_DIV": split})
        J, g = s.objective_and_gradient(x)
        s.close()
        by_split[split] = (J, g)
        for m in range(2):
            assert abs(J[m] - Jt[m]) <= 1e-6 * abs(Jt[m]), split
            assert rel(g[m], gt[m]) < 1e-6, split
    # the X pass in three groups takes the operations of the two-group pass: the same bits
    for m in range(2):
        assert by_split["xf3"][0][m] == by_split["xf"][0][m]
        assert np.array_equal(by_split["xf3"][1][m], by_split["xf"][1][m])
    # member 0 against the oracle
    pr = members[0]
    o = oracle_mod.Oracle(N)
    o.set_tissue(pr["wm"], pr["gm"], pr["csf"], pr["vt"])
    o.set_time(DT, NT)
    o.set_k(pr["kappa"], 0.2, 0.0)
    o.set_rho(pr["rho"], 0.2, 0.0)
    o.prec_factor()
    o.set_phi(pr["centers"], pr["sigma"], 1.0, False)
    o.set_data(pr["d1"].astype(np.float64), None, -0.99)
    o.set_opt(beta=1e-3)
    o.set_time(DT, NT)
    Jo, go = o.objective_and_gradient(x[0], mode=2)
    assert abs(Jt[0] - Jo) / abs(Jo) < 1e-5
    assert rel(gt[0], go) < 1e-4


def test_lockstep_group_equals_separate_bitwise():
    """Handles of different grids, member counts, nt and history modes evaluated in lockstep
    (glia_rd_objective_and_gradient_lockstep: time step i of every handle in one graph) give J and
    dJ bit for bit equal to each handle evaluated alone, on three evaluations in a row; from the
    second evaluation on every time step runs as one graph across handles.  A handle whose captured
    steps all run again eagerly (GLIA_STEP_MARGIN = -1) keeps the equality, and a checkpointed
    handle makes the call evaluate one handle after another."""
    print("case lockstep", flush=True)
    from glia_amd import rd, synth
    specs = (  # n, seed, members, nt, adjoint_store, env
        (32, 5, 2, 6, False, None),
        (64, 6, 1, 6, True, None),
        (96, 7, 1, 4, False, {"GLIA_STEP_MARGIN": "-1"}),
    )
    problems = [[synth.problem(n, seed + m, nbases=4, nnz=None) for m in range(nm)] for n, seed, nm, *_ in specs]
    xs = [[np.concatenate([pr["p"], [pr["kappa"], pr["rho"]]]) for pr in members] for members in problems]

    def handles(checkpoint_last=False):
        out = []
        for k, ((n, _, nm, nt, ast, env), members) in enumerate(zip(specs, problems)):
            old = {}
            for key, v in (env or {}).items():
                old[key] = os.environ.get(key)
                os.environ[key] = v
            try:
                cp = 3 if (checkpoint_last and k == len(specs) - 1) else 0
                s = rd.RDSolver(n, nmembers=nm, dt=0.05, nt=nt, k_gm_wm=0.2, r_gm_wm=0.2, adjoint_store=ast,
                                checkpoint_slots=cp)
            finally:
                for key, v in old.items():
                    if v is None:
                        os.environ.pop(key, None)
                    else:
                        os.environ[key] = v
            for m, pr in enumerate(members):
                s.set_tissue(m, pr["wm"], pr["gm"], pr["csf"], pr["vt"])
                s.set_coefficients(m, pr["kappa"], pr["rho"])
                s.set_basis(m, pr["centers"], pr["sigma"])
                s.set_data(m, (pr["wm"] > 0.5).astype(np.float32), None, -0.99)
            out.append(s)
        return out

    def check(ref, got):
        for (J0, g0), (J1, g1) in zip(ref, got):
            assert np.array_equal(J0, J1)
            for a, b in zip(g0, g1):
                assert np.array_equal(a, b)

    # each handle alone, three evaluations (x moved a little each time)
    def x_at(t):
        return [[x * (1.0 + 0.01 * t) for x in xl] for xl in xs]

    alone = [[] for _ in specs]
    for k, s in enumerate(handles()):
        for t in range(3):
            alone[k].append(s.objective_and_gradient(x_at(t)[k]))
        s.close()
    group = handles()
    spans = []
    for t in range(3):
        Jl, gl, spanned = rd.objective_and_gradient_lockstep(group, x_at(t))
        spans.append(spanned)
        print("evaluation", t, "steps spanning handles", spanned, flush=True)
        for k in range(len(specs)):
            check([alone[k][t]], [(Jl[k], gl[k])])
    launched, reruns, _ = group[2].step_info()
    print("margin -1 handle: launched", launched, "reruns", reruns, flush=True)
    assert reruns > 0
    assert spans[0] < 2 * 6 and spans[1] == 2 * 6 and spans[2] == 2 * 6
    for s in group:
        s.close()
    # a checkpointed handle: one handle after another, the same results
    group = handles(checkpoint_last=True)
    Jl, gl, spanned = rd.objective_and_gradient_lockstep(group, x_at(0))
    assert spanned == 0
    for k in range(2):
        check([alone[k][0]], [(Jl[k], gl[k])])
    for s in group:
        s.close()
