"""H-matrix layer of the CPU restatement: the reference's own invariants
(Tests/tests.cpp: ACA error <= 2 tol, MVM equal to the dense product to
round-off) plus partition invariants, a dense-LU cross check and the
pulsating-sphere closed form."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def L3(O, meshes):
    tri = meshes.icosphere(3, 0.5)
    P = O.Problem(tri, meshes.robin_bc(len(tri), 0, 1, 1), leaf=32, eta=2.0)
    P.set_frequency(500.0, 343.0, 1)
    return P


def test_cluster_tree_and_partition_invariants(O, L3):
    n = L3.n
    C = L3.clusters
    leaf = 32
    assert C[0, 0] == 0 and C[0, 1] == n
    leaves = C[C[:, 3] == 1]
    assert np.all(leaves[:, 1] < leaf) or np.all((leaves[:, 1] < leaf) | (leaves[:, 1] <= 1))
    # leaves tile [0, n) in order
    order = np.argsort(leaves[:, 0])
    starts, sizes = leaves[order, 0], leaves[order, 1]
    assert np.array_equal(starts[1:], (starts + sizes)[:-1]) and starts[0] == 0 and (starts + sizes)[-1] == n
    # permutation
    assert np.array_equal(np.sort(L3.perm), np.arange(n))
    # every matrix entry in exactly one block
    cover = np.zeros((n, n), np.int8)
    nodes, nrm, col, area, hbar = L3.geometry()
    for r0, m, c0, nn, adm in L3.blocks:
        cover[r0:r0 + m, c0:c0 + nn] += 1
    assert np.all(cover == 1)
    # admissible blocks satisfy min(diam)/dist < eta with box distance > 0
    for r0, m, c0, nn, adm in L3.blocks:
        if not adm:
            continue
        a, b = col[r0:r0 + m], col[c0:c0 + nn]
        da = np.linalg.norm(a.max(0) - a.min(0))
        db = np.linalg.norm(b.max(0) - b.min(0))
        gap = np.maximum(0, np.maximum(a.min(0) - b.max(0), b.min(0) - a.max(0)))
        dist = np.linalg.norm(gap)
        assert dist > 0 and min(da, db) / dist < 2.0
    B = L3.blocks
    assert (B[:, 4] == 0).sum() == 463 and (B[:, 4] == 1).sum() == 912


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("tol", [1e-3, 1e-6])
def test_aca_error_bound(L3, kind, tol):
    """|A - H|_F <= 2 tol |A|_F over the whole matrix (the reference's testACA
    criterion, Tests/tests.cpp:27-43); dense blocks exact; the heuristic
    stopping rule may under-resolve single blocks, bounded loosely."""
    H = L3.assemble(kind, tol, 0, 11)
    dense = L3.dense(kind)
    errs = []
    for b, (r0, m, c0, n, adm) in enumerate(L3.blocks):
        blk = H.block(b)
        A = dense[r0:r0 + m, c0:c0 + n]
        if not adm:
            assert np.array_equal(blk, A)
            continue
        U, al, V = blk
        errs.append(np.linalg.norm(A - (U * al) @ V) / np.linalg.norm(A))
    errs = np.array(errs)
    assert errs.max() < 20 * tol and np.mean(errs > 2 * tol) < 0.05, (errs.max(), np.mean(errs > 2 * tol))
    full = H.to_full()
    assert np.linalg.norm(full - dense) / np.linalg.norm(dense) < 2 * tol


def test_max_rank_cap(L3):
    H = L3.assemble(1, 1e-8, 3, 1)
    for b, (r0, m, c0, n, adm) in enumerate(L3.blocks):
        if adm:
            assert 1 <= H.rank(b) <= 3


def test_mvm_equals_full_product(L3):
    """HArithm::MVM against toFullMat() * x to round-off (Tests/tests.cpp:45-58)."""
    H = L3.assemble(1, 1e-4, 0, 5)
    full = H.to_full()
    rng = np.random.default_rng(0)
    x = rng.standard_normal(L3.n) + 1j * rng.standard_normal(L3.n)
    y0 = rng.standard_normal(L3.n) + 1j * rng.standard_normal(L3.n)
    y = H.mvm(x, y0)
    ref = y0 + full @ x
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-14


def test_gmres_converges(L3):
    H = L3.assemble(1, 1e-6, 0, 5)
    rng = np.random.default_rng(1)
    b = rng.standard_normal(L3.n) + 1j * rng.standard_normal(L3.n)
    x, its, est = H.gmres(b, 1e-11, restart=10)   # several restarts
    r = H.mvm(x) - b
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-11
    x2, its2, _ = H.gmres(b, 1e-11, restart=300)
    assert np.linalg.norm(x - x2) / np.linalg.norm(x2) < 1e-9
    assert its2 <= its


def test_pulsating_sphere_closed_form(O, meshes):
    """|p(a)| = |i w rho a / (i k a - 1)| on the sphere; first-order
    convergence of the collocation error with the mesh level."""
    errs = []
    for level in (2, 3):
        tri = meshes.icosphere(level, 0.5)
        P = O.Problem(tri, meshes.robin_bc(len(tri), 0, 1, 1), leaf=32, eta=2.0)
        P.set_frequency(500.0, 343.0, 1)
        _, _, pr, st = P.solve(1e-6, gmres_tol=1e-10, rho=1.21, freq=500.0)
        k = 2 * np.pi * 500 / 343
        pex = 1j * 2 * np.pi * 500 * 1.21 * 0.5 / (1j * k * 0.5 - 1)
        errs.append(np.linalg.norm(pr - pex) / (abs(pex) * np.sqrt(len(tri))))
        assert st["residual"] < 1e-9
    assert errs[1] < 0.03 and 1.6 < errs[0] / errs[1] < 2.5, errs


def test_hmatrix_solve_matches_dense_lu(O, meshes):
    """H-matrix + GMRES solve against the dense LU path
    (regularSolveWithLU, boundaryelementsolver.cpp:140-331)."""
    tri = meshes.icosphere(2, 0.5)
    n = len(tri)
    rng = np.random.default_rng(8)
    bc = meshes.robin_bc(n, 0.0, 1.0, rng.standard_normal(n))
    P = O.Problem(tri, bc, leaf=16, eta=2.0)
    P.set_frequency(900.0, 343.0, 1)
    _, _, pd = P.dense_solve(1.21, 900.0)
    for tol in (1e-4, 1e-8):
        _, _, ph, st = P.solve(tol, gmres_tol=1e-12, rho=1.21, freq=900.0)
        assert np.linalg.norm(ph - pd) / np.linalg.norm(pd) < 20 * tol


def test_impedance_and_plane_wave_rhs(O, meshes):
    """Rigid sphere scattering of a plane wave: total surface potential
    against the Mie series (the plane-wave RHS s = phi_inc + cc dphi_inc/dn)."""
    from scipy.special import spherical_jn, spherical_yn, eval_legendre
    tri = meshes.icosphere(3, 0.5)
    n = len(tri)
    P = O.Problem(tri, meshes.robin_bc(n, 0.0, 1.0, 0.0), leaf=32, eta=2.0)
    f = 300.0
    P.set_frequency(f, 343.0, 1)
    d = np.array([0.0, 0.0, 1.0])
    P.set_sources(plane=np.array([d[0], d[1], d[2], 1.0, 0.0]))
    phi, dphi, pr, st = P.solve(1e-6, gmres_tol=1e-10, rho=1.21, freq=f)
    assert np.abs(dphi).max() < 1e-12
    k = 2 * np.pi * f / 343
    a = 0.5
    nodes, nrm, col, area, hbar = P.geometry()
    cos_t = (col @ d) / np.linalg.norm(col, axis=1)
    ka = k * a
    tot = np.zeros(n, complex)
    for m in range(30):
        jp = spherical_jn(m, ka, derivative=True)
        hp = jp + 1j * spherical_yn(m, ka, derivative=True)
        # total field on a rigid sphere: phi = (i/(ka)^2) sum (2m+1) i^m P_m / h_m'(ka)
        tot += (2 * m + 1) * (1j ** m) * eval_legendre(m, cos_t) * (1j / (ka ** 2 * hp))
    err = np.linalg.norm(phi - tot) / np.linalg.norm(tot)
    assert err < 0.03, err
