"""The linear-elasticity solver on the GPU (include/glia_el.h) against the FP64 oracle: material
fields, the operator, the preconditioner and the CG solve (VariableLinearElasticitySolver,
src/pde/ElasticitySolver.cpp:6-533)."""
import numpy as np
import pytest

from el_cases import tissue_and_tumor, tumor_force

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def case(oracle_mod):
    from glia_amd import elasticity
    n = 32
    print("case elasticity 32^3", flush=True)
    fields = [a.astype(np.float32).astype(np.float64) for a in tissue_and_tumor(n)]
    o = oracle_mod.Oracle(n)
    mu, lam, scr, avg = o.el_material(*fields)
    s = elasticity.ElasticitySolver(n)
    s.set_material(*fields)
    yield n, o, s, fields, (mu, lam, scr, avg)
    s.close()


def test_material_fields_match_the_oracle(case):
    n, o, s, fields, (mu, lam, scr, avg) = case
    gmu, glam, gscr, gavg = s.material()
    print("rel mu", rel(gmu, mu), "lam", rel(glam, lam), "avg", gavg, avg, flush=True)
    assert rel(gmu, mu) < 1e-6 and rel(glam, lam) < 1e-6 and np.array_equal(gscr, scr.astype(np.float32))
    assert np.allclose(gavg, avg, rtol=1e-6)


def test_operator_and_preconditioner_match_the_oracle(case):
    n, o, s, fields, (mu, lam, scr, avg) = case
    rng = np.random.default_rng(2)
    X = np.arange(n) * 2 * np.pi / n
    u = np.stack([np.sin(X + a)[:, None, None] * np.cos(2 * X + b)[None, :, None] * np.sin(X)[None, None, :]
                  for a, b in rng.uniform(0, 3, (3, 2))]) + 0.1 * rng.standard_normal((3, n, n, n))
    u = u.astype(np.float32).astype(np.float64)
    f = s.apply(u)
    fo = o.el_matvec(mu.astype(np.float32).astype(np.float64), lam.astype(np.float32).astype(np.float64),
                     scr, u)
    z = s.precondition(fo.astype(np.float32))
    zo = o.el_prec(np.asarray(avg), fo.astype(np.float32).astype(np.float64))
    print("rel A u", rel(f, fo), "rel B f", rel(z, zo), flush=True)
    assert rel(f, fo) < 1e-5
    assert rel(z, zo) < 1e-5


def test_cg_solve_matches_the_oracle(case):
    n, o, s, fields, (mu, lam, scr, avg) = case
    f = tumor_force(fields[0]).astype(np.float32).astype(np.float64)
    u, its, reason, rn = s.solve(f)
    uo, itso, reasono, rno = o.el_solve(mu, lam, scr, avg, f, np.zeros((3, n, n, n)))
    # the GPU's solution meets the stopping bound measured with the FP64 operator
    res = o.el_prec(avg, f - o.el_matvec(mu, lam, scr, u.astype(np.float64)))
    bf = o.el_prec(avg, f)
    print("its", its, itso, reason, reasono, "rnorm", rn, rno, "rel u", rel(u, uo),
          "||B(f - A u)|| / ||B f||", np.linalg.norm(res) / np.linalg.norm(bf), flush=True)
    assert reason == "rtol" and reasono == 2
    assert abs(its - itso) <= 1
    assert np.linalg.norm(res) <= 2e-3 * np.linalg.norm(bf)
    assert rel(u, uo) < 1e-5


def test_tumor_force_matches_the_oracle(case):
    """Tumor::computeForce (src/mat/Tumor.cpp:114-158): the smoothed tumor's spectral gradient times
    the forcing factor, with and without the tanh scaling."""
    n, o, s, fields, _ = case
    c = fields[0]
    for use_tanh in (True, False):
        f = s.force(c, 1e5, use_tanh)
        fo = o.el_force(c, 1e5, use_tanh)
        print("force tanh", use_tanh, "rel", rel(f, fo), flush=True)
        assert rel(f, fo) < 1e-5


def test_cg_state_resets_between_solves(case):
    """The CG's stopping tests run on the device with iterations enqueued ahead; a solve that stops
    before its first iteration (zero force: the preconditioned residual is 0, KSP_CONVERGED_ATOL with
    0 iterations and u = 0, as KSPSolve_CG leaves it) and the solves around it leave no state behind:
    the same solve before and after gives the same iterations, reason and solution bit for bit, and an
    operator application after a solve is unaffected."""
    n, o, s, fields, (mu, lam, scr, avg) = case
    f = tumor_force(fields[0]).astype(np.float32).astype(np.float64)
    u1, its1, why1, rn1 = s.solve(f)
    u0, its0, why0, rn0 = s.solve(np.zeros_like(f))
    print("zero force:", its0, why0, rn0, flush=True)
    assert its0 == 0 and why0 == "atol" and rn0 == 0.0 and not np.any(u0)
    u2, its2, why2, rn2 = s.solve(f)
    assert (its2, why2) == (its1, why1) and rn2 == rn1
    assert np.array_equal(u1, u2)
    w = s.apply(u1)
    wo = o.el_matvec(mu, lam, scr, u1.astype(np.float64))
    assert rel(w, wo) < 1e-5
