"""The host H-LU (bem_hlu_create; HArithm::LUDecomposition,
HMatrix/harithm.cpp:1192-1296, on the truncated copy of hmatrix.cpp:770-776)
on H-matrices assembled by the CPU restatement: without truncation the
factors multiply back to the matrix and solve it to round-off; with a
relative tolerance the factorisation error follows the tolerance; with the
reference's default (rank 2, no tolerance) it is a preconditioner that cuts
GMRES iterations several fold."""
import numpy as np
import pytest
from scipy.sparse.linalg import LinearOperator, gmres


def problem(O, meshes, level, leaf, freq=500.0):
    tri = meshes.icosphere(level, 0.5)
    n = len(tri)
    P = O.Problem(tri, meshes.robin_bc(n, 0, 1, 1), leaf=leaf, eta=2.0)
    P.set_frequency(freq, 343.0, 1)
    H = P.assemble(1, 1e-6, 0, 3)
    return P, H, [H.block(b) for b in range(len(P.blocks))]


@pytest.fixture(scope="module")
def L2(O, meshes):
    return problem(O, meshes, 2, 16)


@pytest.fixture(scope="module")
def L3(O, meshes):
    return problem(O, meshes, 3, 32)


def test_exact_lu(bem, L2):
    """no truncation: L U = A and the solve is exact to round-off (the
    diagonal leaves are QR-factored, L leaf = Q, U leaf = R)"""
    P, H, blocks = L2
    A = H.to_full()
    F = bem.HLU(P.n, P.clusters, P.blocks, blocks, max_rank=0, rel_err=0.0)
    assert np.linalg.norm(F.dense() - A) / np.linalg.norm(A) < 1e-13
    rng = np.random.default_rng(1)
    b = rng.standard_normal(P.n) + 1j * rng.standard_normal(P.n)
    x = F.solve(b)
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-12
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(x) < 1e-11


@pytest.mark.parametrize("tol", [1e-8, 1e-5, 1e-3])
def test_truncation_follows_tolerance(bem, L2, tol):
    P, H, blocks = L2
    A = H.to_full()
    F = bem.HLU(P.n, P.clusters, P.blocks, blocks, max_rank=0, rel_err=tol)
    e = np.linalg.norm(F.dense() - A) / np.linalg.norm(A)
    print("rel_err", tol, "||LU - A|| / ||A||", e, F.stats())
    assert e < 10 * tol
    if tol >= 1e-5:
        assert e > 1e-3 * tol        # it does truncate


def test_rank_two_preconditioner(bem, L3):
    """the reference's defaults (preconditionerRank 2, preconditionerRelError
    0) on the 1280-element operator: GMRES to 1e-10 with the H-LU needs far
    fewer iterations and reaches the same solution"""
    P, H, blocks = L3
    A = H.to_full()
    F = bem.HLU(P.n, P.clusters, P.blocks, blocks, max_rank=2, rel_err=0.0)
    assert F.stats()["max_rank"] <= 2
    rng = np.random.default_rng(2)
    b = rng.standard_normal(P.n) + 1j * rng.standard_normal(P.n)
    counts = []

    def run(M):
        it = [0]

        def cb(_):
            it[0] += 1

        x, info = gmres(A, b, rtol=1e-10, restart=300, maxiter=10, M=M, callback=cb, callback_type="pr_norm")
        assert info == 0
        counts.append(it[0])
        return x

    x0 = run(None)
    x1 = run(LinearOperator(A.shape, matvec=F.solve, dtype=np.complex128))
    ref = np.linalg.solve(A, b)
    print("GMRES iterations without / with the H-LU", counts)
    assert np.linalg.norm(x1 - ref) / np.linalg.norm(ref) < 1e-8
    assert np.linalg.norm(x0 - ref) / np.linalg.norm(ref) < 1e-8
    assert counts[1] * 2 < counts[0]


@pytest.mark.parametrize("rank,tol", [(0, 0.0), (0, 1e-5), (2, 0.0)])
def test_flat_schedule_equals_recursive_solve(bem, L2, L3, rank, tol):
    """the leaf-by-leaf substitution (the device preconditioner's order,
    Q^H and R^-1 of each leaf as products) gives the recursive
    forward / backward substitution's result to round-off, on the 320- and
    1280-element operators, exact and truncated"""
    for P, H, blocks in (L2, L3):
        F = bem.HLU(P.n, P.clusters, P.blocks, blocks, max_rank=rank, rel_err=tol)
        rng = np.random.default_rng(7)
        b = rng.standard_normal(P.n) + 1j * rng.standard_normal(P.n)
        x0, x1 = F.solve(b), F.solve(b, flat=True)
        err = np.linalg.norm(x1 - x0) / np.linalg.norm(x0)
        print(P.n, rank, tol, "flat vs recursive", err)
        assert err < 1e-12
        F.close()


def test_bad_partition_is_an_error(bem, L2):
    P, H, blocks = L2
    keep = np.ones(len(P.blocks), bool)
    keep[len(keep) // 2] = False
    with pytest.raises(bem.BemError):
        bem.HLU(P.n, P.clusters, P.blocks[keep], [b for b, k in zip(blocks, keep) if k])
