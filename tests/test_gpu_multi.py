"""Several right-hand sides: the product of Q interleaved vectors with one
pass over the blocks (VALU and FP64 matrix-core kernels), and GMRES in
lockstep over them with graph-captured Arnoldi columns, against the
single-vector product and solver."""
import numpy as np
import pytest

from helpers import sphere_problem, upload_problem

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def L3(O, meshes):
    return sphere_problem(O, meshes, level=3, leaf=32, eta=2.0, freq=500.0)


@pytest.mark.parametrize("prec", [0, 1])
def test_multi_product_equals_single(bem, L3, prec):
    """Y = H X for 2, 3, 4 and 8 right-hand sides, VALU and matrix cores,
    FP64 and FP32 storage: every column equals the single-vector product to
    FP64 round-off (the same stored operator, sums in another order)."""
    ctx = upload_problem(bem, L3)
    G = ctx.assemble(1, 1e-6, 0, precision=prec, seed=3)
    rng = np.random.default_rng(2)
    n = L3.n
    for q in (2, 3, 4, 8):
        X = rng.standard_normal((n, q)) + 1j * rng.standard_normal((n, q))
        ref = np.stack([G.mvm(X[:, k]) for k in range(q)], 1)
        for mf in (True, False):
            ctx.set_mfma(mf)
            print("multi product: prec", prec, "nrhs", q, "mfma", mf, flush=True)
            Y = G.mvm_multi(X)
            assert rel(Y, ref) < 1e-13, (q, mf, rel(Y, ref))
    ctx.set_mfma(True)
    G.close()
    ctx.close()


@pytest.mark.parametrize("graphs", [True, False])
def test_lockstep_gmres_equals_single_solves(bem, L3, graphs):
    """Lockstep GMRES over 3 right-hand sides of the FP64 DPhi (the boundary
    data, a random vector and a third of another scale) to 1e-12, columns
    replayed as a graph or launched directly: each column equals the
    single-right-hand-side GMRES solution to 1e-10, in as many columns."""
    ctx = upload_problem(bem, L3)
    ctx.set_graphs(graphs)
    G = ctx.assemble(1, 1e-6, 0, seed=3)
    n = L3.n
    rng = np.random.default_rng(4)
    f = L3.rhs_parts()[3]                                        # the boundary data (g = 1)
    B = np.stack([f, rng.standard_normal(n) + 1j * rng.standard_normal(n), 1e-3 * np.exp(1j * np.arange(n))], 1)
    X, its, rr, st, rc = G.gmres_multi(B, 1e-12)
    assert rc == 0 and np.all(rr < 1e-12)
    for k in range(3):
        x1, it1, r1, rc1 = G.gmres(B[:, k], 1e-12)
        assert rc1 == 0
        assert rel(X[:, k], x1) < 1e-10, (k, rel(X[:, k], x1))
        assert abs(int(its[k]) - it1) <= 1, (k, its[k], it1)
    G.close()
    ctx.close()


def test_graph_columns_equal_direct_launches(bem, L3):
    """Single-right-hand-side GMRES on the lockstep solver with each Arnoldi
    column a replayed graph, and on bem_gmres_prec's own per-column launches
    (graphs off): the same iterates, bit for bit."""
    ctx = upload_problem(bem, L3)
    G = ctx.assemble(1, 1e-6, 0, seed=3)
    s = np.exp(0.37j * np.arange(L3.n)) + 0.5
    ctx.set_graphs(True)
    xg, itg, rg, rcg = G.gmres(s, 1e-12)
    ctx.set_graphs(False)
    xd, itd, rd, rcd = G.gmres(s, 1e-12)
    ctx.set_graphs(True)
    assert rcg == 0 and rcd == 0 and itg == itd
    assert np.array_equal(xg.view(np.uint64), xd.view(np.uint64))
    G.close()
    ctx.close()


def test_lockstep_refinement_fp32(bem, L3):
    """FP32 DPhi with its FP64 near field: lockstep refinement over two
    right-hand sides reaches |b - A_res x| < 1e-12 |b| for both."""
    ctx = upload_problem(bem, L3)
    ctx.set_near_fp64(True)
    G = ctx.assemble(1, 1e-6, 0, precision=bem.FP32, seed=3)
    s = np.exp(0.37j * np.arange(L3.n)) + 0.5
    B = np.stack([s, 2j * s[::-1]], 1)
    X, its, rr, st, rc = G.gmres_multi(B, 1e-12, max_steps=10)
    assert rc == 0 and np.all(rr < 1e-12) and st >= 1, (rr, st)
    for k in range(2):
        assert rel(G.mvm_refine(X[:, k]), B[:, k]) < 2e-12
    G.close()
    ctx.close()
