"""GPU path vs the CPU restatement on the same inputs.  FP64 entries, dense
blocks and ACA factors must be bit-identical (the arithmetic contract of
include/bem_detmath.h); products and solves agree to round-off."""
import numpy as np
import pytest

from helpers import near_pairs, sphere_problem, upload_problem

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128))
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.fixture(scope="module")
def L3(O, meshes):
    return sphere_problem(O, meshes, level=3, leaf=32, eta=2.0, freq=500.0)


@pytest.fixture(scope="module")
def ctx3(bem, L3):
    return upload_problem(bem, L3)


def test_entries_bit_exact(L3, ctx3):
    n = L3.n
    rng = np.random.default_rng(0)
    ij = np.concatenate([rng.integers(0, n, size=(4000, 2)), near_pairs(L3), np.stack([np.arange(n)] * 2, 1)])
    ref = L3.entries4(ij)
    got = ctx3.entries4(ij)
    assert same_bits(ref, got), np.abs(ref - got).max()
    for kind in (0, 1):
        assert same_bits(L3.entries(kind, ij), ctx3.entries(kind, ij))


@pytest.mark.parametrize("kind", [0, 1])
def test_assembly_bit_exact(L3, ctx3, kind):
    tol, seed = 1e-6, 7
    H = L3.assemble(kind, tol, 0, seed)
    G = ctx3.assemble(kind, tol, 0, seed=seed)
    for b in range(len(L3.blocks)):
        rb, rg = H.rank(b), G.rank(b)
        assert rb == rg, (b, rb, rg)
        A, B = H.block(b), G.block(b)
        if rb < 0:
            assert same_bits(A, B), b
        else:
            for x, y in zip(A, B):
                assert same_bits(x, y), b
    st = G.stats()
    assert st["evaluated"] == H.evaluated


def test_assembly_bit_exact_with_rank_overflow(L3, ctx3):
    """ACA 1e-10 on the L3 blocks: ranks up to 55 / 60 exceed the first-pass
    capacity of 48, so the overflowing blocks are re-run with their full
    capacity and every block is packed (k_compact); the result is still
    bit-identical to the restatement."""
    for kind in (0, 1):
        H = L3.assemble(kind, 1e-10, 0, 5)
        G = ctx3.assemble(kind, 1e-10, 0, seed=5)
        st = G.stats()
        assert st["overflow_blocks"] > 0 and st["max_rank"] > 48
        for b in range(len(L3.blocks)):
            assert H.rank(b) == G.rank(b), (kind, b)
            A, B = H.block(b), G.block(b)
            if H.rank(b) < 0:
                assert same_bits(A, B), (kind, b)
            else:
                for x, y in zip(A, B):
                    assert same_bits(x, y), (kind, b)
        G.close()


def test_mvm_matches_oracle(L3, ctx3):
    H = L3.assemble(1, 1e-6, 0, 3)
    G = ctx3.assemble(1, 1e-6, 0, seed=3)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(L3.n) + 1j * rng.standard_normal(L3.n)
    y0 = rng.standard_normal(L3.n) + 1j * rng.standard_normal(L3.n)
    ref = H.mvm(x, y0)
    got = G.mvm(x, y0, accumulate=True)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-14
    full = H.to_full()
    assert np.linalg.norm(G.mvm(x) - full @ x) / np.linalg.norm(full @ x) < 1e-13


def test_gmres_matches_oracle(L3, ctx3):
    H = L3.assemble(1, 1e-6, 0, 3)
    G = ctx3.assemble(1, 1e-6, 0, seed=3)
    rng = np.random.default_rng(2)
    b = rng.standard_normal(L3.n) + 1j * rng.standard_normal(L3.n)
    xr, itr, _ = H.gmres(b, 1e-12)
    xg, itg, rel, rc = G.gmres(b, 1e-12)
    assert rc == 0 and rel < 1e-12
    assert abs(itr - itg) <= 1
    assert np.linalg.norm(xg - xr) / np.linalg.norm(xr) < 1e-10


def test_solver_sphere1280_parity(bem, O, meshes):
    """Pulsating sphere L3, 500 Hz, leaf 32, eta 2, ACA 1e-6, GMRES 1e-10."""
    tri = meshes.icosphere(3, 0.5)
    n = len(tri)
    bc = meshes.robin_bc(n, 0, 1, 1)
    S = bem.Solver(tri, bc, leaf=32, eta=2.0)
    phi, dphi, pr, st = S.solve(500.0, c=343.0, rho=1.21, coupling=1, aca_tol=1e-6, gmres_tol=1e-10, seed=1)
    P = O.Problem(tri, bc, leaf=32, eta=2.0)
    P.set_frequency(500.0, 343.0, 1)
    phio, dphio, pro, sto = P.solve(1e-6, gmres_tol=1e-10, seed=1, rho=1.21, freq=500.0)
    perm, blocks = S.layout()
    assert np.array_equal(perm, P.perm) and np.array_equal(blocks, P.blocks)
    rel = np.linalg.norm(pr - pro) / np.linalg.norm(pro)
    print("sphere-1280 pressure rel-L2 GPU vs CPU restatement", rel, "iters", st["gmres_iters"], sto["gmres_iters"])
    assert rel < 1e-10
    k = 2 * np.pi * 500 / 343
    pex = 1j * 2 * np.pi * 500 * 1.21 * 0.5 / (1j * k * 0.5 - 1)
    assert np.linalg.norm(pr - pex) / (abs(pex) * np.sqrt(n)) < 0.03
    assert st["residual"] < 1e-9


@pytest.mark.parametrize("level,leaf,eta", [(0, 50, 3.0), (1, 8, 2.0), (2, 50, 3.0)])
def test_small_meshes(bem, O, meshes, level, leaf, eta):
    tri = meshes.icosphere(level, 1.0)
    n = len(tri)
    bc = meshes.robin_bc(n, 0, 1, 1)
    S = bem.Solver(tri, bc, leaf=leaf, eta=eta)
    _, _, pr, st = S.solve(200.0, aca_tol=1e-8, gmres_tol=1e-12)
    P = O.Problem(tri, bc, leaf=leaf, eta=eta)
    P.set_frequency(200.0, 343.0, 1)
    _, _, pro, _ = P.solve(1e-8, gmres_tol=1e-12, rho=1.21, freq=200.0)
    assert np.linalg.norm(pr - pro) / np.linalg.norm(pro) < 1e-10


def test_golden_entries_on_device(bem, O, meshes):
    """The committed fixture's entries reproduced bit for bit by the device."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz"),
                allow_pickle=False)
    P = O.Problem(g["tri"], meshes.robin_bc(len(g["tri"]), 0, 1, 1), leaf=32, eta=2.0)
    P.set_frequency(500.0, 343.0, 1)
    ctx = upload_problem(bem, P)
    assert same_bits(ctx.entries4(g["ij"]), g["entries4"])


def test_near_pairing_bit_exact(bem, O, meshes):
    """near-field pairing (the DPhi near field formed by the Phi assembly and
    taken by the next DPhi assembly): every block still equals the
    restatement's, also after a frequency change between the two assemblies
    (the kept near field must not be used), and FP32 blocks equal the
    unpaired FP32 assembly"""
    from helpers import sphere_problem, upload_problem
    P = sphere_problem(O, meshes, level=3, leaf=32, eta=2.0, freq=600.0)
    ctx = upload_problem(bem, P)
    ctx.set_near_pairing(True)

    def check(kind, G, H):
        for b in range(len(P.blocks)):
            A, B = H.block(b), G.block(b)
            if isinstance(A, tuple):
                for x, y in zip(A, B):
                    assert same_bits(x, y), (kind, b)
            else:
                assert same_bits(A, B), (kind, b)

    G0 = ctx.assemble(0, 1e-6, 0, seed=3)
    G1 = ctx.assemble(1, 1e-6, 0, seed=3)       # takes the kept near field
    check(0, G0, P.assemble(0, 1e-6, 0, 3))
    check(1, G1, P.assemble(1, 1e-6, 0, 3))
    G0.close(); G1.close()
    # Phi at 600 Hz, then a new frequency: DPhi must be formed afresh
    G0 = ctx.assemble(0, 1e-6, 0, seed=3)
    P.set_frequency(750.0, 343.0, 1)
    prm = P.params()
    ctx.set_frequency(prm["k"], prm["cc"])
    G1 = ctx.assemble(1, 1e-6, 0, seed=3)
    check(1, G1, P.assemble(1, 1e-6, 0, 3))
    G0.close(); G1.close()
    # FP32: paired and unpaired give the same blocks
    F0 = ctx.assemble(0, 1e-6, 0, precision=bem.FP32, seed=3)
    F1 = ctx.assemble(1, 1e-6, 0, precision=bem.FP32, seed=3)
    ctx.set_near_pairing(False)
    U1 = ctx.assemble(1, 1e-6, 0, precision=bem.FP32, seed=3)
    for b in range(len(P.blocks)):
        A, B = F1.block(b), U1.block(b)
        if isinstance(A, tuple):
            for x, y in zip(A, B):
                assert np.array_equal(x, y), b
        else:
            assert np.array_equal(A, B), b
    for G in (F0, F1, U1):
        G.close()
    ctx.close()
