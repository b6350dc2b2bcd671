"""Leaf-block LU preconditioner (bem_precond_*, bem_gmres_prec): M^-1 is
the exact inverse of the diagonal leaf blocks, preconditioned GMRES reaches
the same solution, and it cuts the iterations of the impedance capsule."""
import numpy as np
import pytest

from helpers import sphere_problem, upload_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L3(O, meshes):
    return sphere_problem(O, meshes, level=3, leaf=32, eta=2.0, freq=500.0)


@pytest.mark.parametrize("prec", [0, 1])
def test_apply_is_block_inverse(bem, L3, prec):
    ctx = upload_problem(bem, L3)
    G = ctx.assemble(1, 1e-6, 0, precision=prec, seed=3)
    P = G.preconditioner()
    rng = np.random.default_rng(0)
    v = rng.standard_normal(L3.n) + 1j * rng.standard_normal(L3.n)
    z = P.apply(v)
    covered = 0
    for b, (r0, m, c0, n, adm) in enumerate(L3.blocks):
        if adm or r0 != c0 or m != n:
            continue
        A = G.block(b)
        ref = np.linalg.solve(A, v[r0:r0 + m])
        assert np.linalg.norm(z[r0:r0 + m] - ref) / np.linalg.norm(ref) < 1e-12
        covered += m
    assert covered == L3.n


def test_preconditioned_gmres_same_solution(bem, L3):
    ctx = upload_problem(bem, L3)
    G = ctx.assemble(1, 1e-6, 0, seed=3)
    rng = np.random.default_rng(2)
    b = rng.standard_normal(L3.n) + 1j * rng.standard_normal(L3.n)
    x0, it0, rel0, rc0 = G.gmres(b, 1e-11)
    x1, it1, rel1, rc1 = G.gmres(b, 1e-11, precond=G.preconditioner())
    assert rc0 == 0 and rc1 == 0 and rel1 < 1e-11
    assert np.linalg.norm(x1 - x0) / np.linalg.norm(x0) < 1e-9
    r = G.mvm(x1) - b
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-10


def test_capsule_iterations(bem, meshes):
    """The 76440-triangle impedance capsule at 1500 Hz (FP32, ACA 1e-4,
    GMRES 1e-8): unpreconditioned GMRES(300) needs five restarts, the
    leaf-block LU far fewer iterations, to the same pressure."""
    f, c, rho = 1500.0, 343.0, 1.21
    k = 2 * np.pi * f / c
    tri = meshes.capsule(2.0, 0.25, 0.0107)
    n = len(tri)
    bc = meshes.impedance_bc(n, k, 0.3, 4)
    d = meshes.unit_direction(4)
    S = bem.Solver(tri, bc, leaf=64, eta=1.5)
    S.set_sources(plane=np.array([d[0], d[1], d[2], 1.0, 0.0]))
    _, _, p0, st0 = S.solve(f, c=c, rho=rho, aca_tol=1e-4, gmres_tol=1e-8, seed=1, precision=bem.FP32)
    S.set_preconditioner(1)
    _, _, p1, st1 = S.solve(f, c=c, rho=rho, aca_tol=1e-4, gmres_tol=1e-8, seed=1, precision=bem.FP32)
    rel = np.linalg.norm(p1 - p0) / np.linalg.norm(p0)
    print("capsule iterations", st0["gmres_iters"], "->", st1["gmres_iters"], "time", st0["t_gmres"], "->",
          st1["t_gmres"], "ms, rel", rel)
    assert st1["residual"] < 1e-7
    assert st1["gmres_iters"] < 0.3 * st0["gmres_iters"]
    assert rel < 1e-6
