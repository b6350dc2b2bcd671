"""FP32 assembly against the FP64 restatement: regular quadrature and ACA in
complex FP32, self / near-singular entries in FP64 rounded to FP32 storage,
FP64 GMRES (restart residual recomputed with FP64 accumulation) over the
FP32 operator.  Tolerances are stated per test; the FP32 path is not part of
the bit-exact contract."""
import numpy as np
import pytest

from helpers import near_pairs, sphere_problem, upload_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L3(O, meshes):
    return sphere_problem(O, meshes, level=3, leaf=32, eta=2.0, freq=500.0)


@pytest.fixture(scope="module")
def ctx3(bem, L3):
    return upload_problem(bem, L3)


@pytest.mark.parametrize("kind", [0, 1])
def test_fp32_near_field_blocks(bem, L3, ctx3, kind):
    """Dense blocks: relative Frobenius error < 2e-6 per block against the
    FP64 entries; self and near-singular entries are the FP64 values rounded
    once to FP32."""
    G = ctx3.assemble(kind, 1e-6, 0, precision=bem.FP32, seed=7)
    dense = L3.dense(kind)
    nodes, nrm, col, area, hbar = L3.geometry()
    worst = 0.0
    for b, (r0, m, c0, n, adm) in enumerate(L3.blocks):
        if adm:
            continue
        A = dense[r0:r0 + m, c0:c0 + n]
        B = G.block(b)
        worst = max(worst, np.linalg.norm(B - A) / np.linalg.norm(A))
        d = np.linalg.norm(col[r0:r0 + m, None, :] - col[None, c0:c0 + n, :], axis=2)
        sing = d < 4 * hbar
        assert np.array_equal(B[sing], A[sing].astype(np.complex64).astype(np.complex128)), b
    print("FP32 near-field worst block rel err", worst)
    assert worst < 2e-6


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("tol", [1e-4, 1e-6])
def test_fp32_hmatrix_error(bem, L3, ctx3, kind, tol):
    """Whole-matrix error of the FP32 H-matrix within the ACA criterion
    (2 tol, Tests/tests.cpp:27-43) plus the FP32 entry error."""
    G = ctx3.assemble(kind, tol, 0, precision=bem.FP32, seed=11)
    dense = L3.dense(kind)
    full = G.to_full()
    err = np.linalg.norm(full - dense) / np.linalg.norm(dense)
    print("FP32 H-matrix kind", kind, "tol", tol, "rel err", err, "stats", G.stats())
    assert err < 2 * tol + 2e-6
    st = G.stats()
    H64 = ctx3.assemble(kind, tol, 0, seed=11)
    assert st["bytes"] < 0.6 * H64.stats()["bytes"]


def test_fp32_mvm_is_exact_product_of_stored_operator(bem, L3, ctx3):
    """The MVM accumulates in FP64: it equals the dense product of the stored
    FP32 operator to FP64 round-off."""
    G = ctx3.assemble(1, 1e-6, 0, precision=bem.FP32, seed=3)
    full = G.to_full()
    rng = np.random.default_rng(5)
    x = rng.standard_normal(L3.n) + 1j * rng.standard_normal(L3.n)
    y = G.mvm(x)
    assert np.linalg.norm(y - full @ x) / np.linalg.norm(full @ x) < 1e-13


def test_fp32_solve_sphere1280(bem, O, meshes):
    """Pulsating sphere L3 (1280 triangles, 500 Hz) in FP32:
    surface pressure within 1e-6 relative L2 of the FP64 CPU solve, GMRES
    residual of the FP32 operator < 1e-9."""
    tri = meshes.icosphere(3, 0.5)
    n = len(tri)
    bc = meshes.robin_bc(n, 0, 1, 1)
    S = bem.Solver(tri, bc, leaf=32, eta=2.0)
    _, _, pr, st = S.solve(500.0, c=343.0, rho=1.21, coupling=1, aca_tol=1e-6, gmres_tol=1e-10, seed=1,
                           precision=bem.FP32)
    P = O.Problem(tri, bc, leaf=32, eta=2.0)
    P.set_frequency(500.0, 343.0, 1)
    _, _, pro, _ = P.solve(1e-6, gmres_tol=1e-10, seed=1, rho=1.21, freq=500.0)
    rel = np.linalg.norm(pr - pro) / np.linalg.norm(pro)
    print("FP32 sphere-1280 pressure rel-L2 vs FP64 CPU", rel, "iters", st["gmres_iters"])
    assert rel < 1e-6
    assert st["residual"] < 1e-9


@pytest.mark.parametrize("freq", [100.0, 1500.0])
def test_fp32_solve_vs_dense_oracle(bem, O, meshes, freq):
    """Random Robin data on an L2 sphere: FP32 H-matrix solve against the FP64
    dense LU solve of the restatement (regularSolveWithLU,
    boundaryelementsolver.cpp:140-331), within the ACA tolerance."""
    tri = meshes.icosphere(2, 0.5)
    n = len(tri)
    rng = np.random.default_rng(8)
    bc = meshes.robin_bc(n, 0.3 + 0.1j, 1.0, rng.standard_normal(n))
    P = O.Problem(tri, bc, leaf=16, eta=2.0)
    P.set_frequency(freq, 343.0, 1)
    _, _, pd = P.dense_solve(1.21, freq)
    S = bem.Solver(tri, bc, leaf=16, eta=2.0)
    _, _, pr, st = S.solve(freq, aca_tol=1e-7, gmres_tol=1e-11, seed=2, precision=bem.FP32)
    rel = np.linalg.norm(pr - pd) / np.linalg.norm(pd)
    print("FP32 vs dense FP64 oracle", freq, rel)
    assert rel < 2e-6
