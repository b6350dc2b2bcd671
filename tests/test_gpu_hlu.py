"""The H-LU preconditioner fed with GPU-assembled blocks
(bem_precond_create_hlu, solver preconditioner 2): the 1280-triangle
pulsating sphere with H-LU preconditioned GMRES to 1e-10 against the CPU restatement and the
monopole, its application against the host H-LU built from the
restatement's own blocks, and the FP32 operator preconditioned the same
way."""
import numpy as np
import pytest

from helpers import sphere_problem, upload_problem

pytestmark = pytest.mark.gpu


def test_sphere1280_hlu_preconditioned(bem, O, meshes):
    """Pulsating sphere L3 (1280), 500 Hz, Burton-Miller, FP64, leaf 32,
    eta 2, ACA 1e-6, H-LU (rank 2, the reference's default) preconditioned
    GMRES to 1e-10"""
    tri = meshes.icosphere(3, 0.5)
    n = len(tri)
    bc = meshes.robin_bc(n, 0, 1, 1)
    S = bem.Solver(tri, bc, leaf=32, eta=2.0)
    _, _, p0, st0 = S.solve(500.0, c=343.0, rho=1.21, aca_tol=1e-6, gmres_tol=1e-10, seed=1)
    S.set_preconditioner(2)
    _, _, p1, st1 = S.solve(500.0, c=343.0, rho=1.21, aca_tol=1e-6, gmres_tol=1e-10, seed=1)
    P = O.Problem(tri, bc, leaf=32, eta=2.0)
    P.set_frequency(500.0, 343.0, 1)
    _, _, pro, _ = P.solve(1e-6, gmres_tol=1e-10, seed=1, rho=1.21, freq=500.0)
    rel = np.linalg.norm(p1 - pro) / np.linalg.norm(pro)
    print("sphere-1280 H-LU: iterations", st1["gmres_iters"], "(plain", st0["gmres_iters"], ") vs restatement", rel,
          "residual", st1["residual"])
    assert rel < 1e-9 and st1["residual"] < 1e-9
    assert st1["gmres_iters"] < st0["gmres_iters"]
    k = 2 * np.pi * 500 / 343
    pex = 1j * 2 * np.pi * 500 * 1.21 * 0.5 / (1j * k * 0.5 - 1)
    assert np.linalg.norm(p1 - pex) / (abs(pex) * np.sqrt(n)) < 0.03
    S.close()


def test_apply_equals_host_hlu(bem, O, meshes):
    """the device-fed preconditioner applies the same H-LU as the host
    factorisation of the restatement's blocks (the FP64 blocks are
    bit-identical)"""
    P = sphere_problem(O, meshes, level=3, leaf=32, eta=2.0, freq=700.0)
    ctx = upload_problem(bem, P)
    G = ctx.assemble(1, 1e-6, 0, seed=5)
    H = P.assemble(1, 1e-6, 0, 5)
    for rank, tol in ((2, 0.0), (0, 1e-6)):
        M = G.hlu(P.clusters, rank, tol)
        F = bem.HLU(P.n, P.clusters, P.blocks, [H.block(b) for b in range(len(P.blocks))], rank, tol)
        try:
            rng = np.random.default_rng(rank)
            v = rng.standard_normal(P.n) + 1j * rng.standard_normal(P.n)
            a, b = M.apply(v), F.solve(v)
            err = np.linalg.norm(a - b) / np.linalg.norm(b)
        finally:
            M.close()
            F.close()
        assert err < 1e-14, (rank, err)
    G.close()
    ctx.close()


def test_device_substitution_wide_leaves(bem, O, meshes):
    """diagonal leaves of 65-127 rows (leaf 128: each lane holds two rows of
    the leaf) apply the same H-LU as the host; leaves above 128 rows are
    refused before anything runs on the device"""
    P = sphere_problem(O, meshes, level=3, leaf=128, eta=2.0, freq=700.0)
    assert max(b[1] for b in P.blocks if b[0] == b[2]) > 64
    ctx = upload_problem(bem, P)
    G = ctx.assemble(1, 1e-6, 0, seed=5)
    H = P.assemble(1, 1e-6, 0, 5)
    M = G.hlu(P.clusters, 2, 0.0)
    F = bem.HLU(P.n, P.clusters, P.blocks, [H.block(b) for b in range(len(P.blocks))], 2, 0.0)
    try:
        rng = np.random.default_rng(3)
        v = rng.standard_normal(P.n) + 1j * rng.standard_normal(P.n)
        a, b = M.apply(v), F.solve(v)
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
    finally:
        M.close()
        F.close()
    print("leaf 128: device vs host H-LU", err)
    assert err < 1e-14
    G.close()
    ctx.close()
    P2 = sphere_problem(O, meshes, level=3, leaf=256, eta=2.0, freq=700.0)
    ctx2 = upload_problem(bem, P2)
    G2 = ctx2.assemble(1, 1e-6, 0, seed=5)
    try:
        with pytest.raises(bem.BemError, match="at most 128"):
            G2.hlu(P2.clusters, 2, 0.0)
    finally:
        G2.close()
        ctx2.close()


def test_fp32_operator_with_hlu(bem, meshes):
    """L4 (5120), 800 Hz, FP32 blocks with the H-LU built from them: same
    pressure as the FP64 solve, fewer iterations"""
    tri = meshes.icosphere(4, 0.5)
    n = len(tri)
    S = bem.Solver(tri, meshes.robin_bc(n, 0, 1, 1), leaf=64, eta=2.0)
    _, _, p64, _ = S.solve(800.0, aca_tol=1e-5, gmres_tol=1e-9, seed=1)
    _, _, p32, st0 = S.solve(800.0, aca_tol=1e-5, gmres_tol=1e-9, seed=1, precision=bem.FP32)
    S.set_preconditioner(2)
    _, _, p32h, st1 = S.solve(800.0, aca_tol=1e-5, gmres_tol=1e-9, seed=1, precision=bem.FP32)
    print("FP32 iterations plain / H-LU", st0["gmres_iters"], st1["gmres_iters"],
          "FP32+H-LU vs FP64", np.linalg.norm(p32h - p64) / np.linalg.norm(p64))
    assert np.linalg.norm(p32h - p32) / np.linalg.norm(p32) < 1e-7
    assert np.linalg.norm(p32h - p64) / np.linalg.norm(p64) < 1e-5
    assert st1["gmres_iters"] <= st0["gmres_iters"]
    S.close()


def test_left_preconditioning_as_reference(bem, O, meshes):
    """The reference's preconditioned GMRES (gmresLUPreconditionedSolve,
    gmres.h:200-216): GMRES on M^-1 A x = M^-1 b from x = M^-1 b, stopped
    on the preconditioned residual (bem_set_precond_side).  The reported
    residual is |M^-1 (b - A x)| / |M^-1 b| recomputed on the host, the
    solution agrees with the right-preconditioned one, and the solver form
    (bem_solver_set_precond_side) reproduces the restatement's pressure."""
    P = sphere_problem(O, meshes, level=3, leaf=32, eta=2.0, freq=700.0)
    ctx = upload_problem(bem, P)
    G = ctx.assemble(1, 1e-6, 0, seed=5)
    M = G.hlu(P.clusters, 2, 0.0)
    rng = np.random.default_rng(11)
    b = rng.standard_normal(P.n) + 1j * rng.standard_normal(P.n)
    x_r, its_r, rel_r, _ = G.gmres(b, 1e-10, precond=M)
    _, its_0, _, _ = G.gmres(b, 1e-10)
    ctx.set_precond_side(True)
    pb = M.apply(b)
    x_l, its_l, rel_l, rc = G.gmres(b, 1e-10, x0=pb, precond=M)
    ctx.set_precond_side(False)
    assert rc == 0 and rel_l < 1e-10
    r = M.apply(b - G.mvm(x_l))
    rel_host = np.linalg.norm(r) / np.linalg.norm(pb)
    print("left-preconditioned GMRES: iterations", its_l, "(right", its_r, ", none", its_0, ") residual", rel_l,
          "host", rel_host, "vs right", np.linalg.norm(x_l - x_r) / np.linalg.norm(x_r))
    assert rel_host < 1e-10 and abs(rel_host - rel_l) < 1e-3 * rel_l + 1e-14
    assert np.linalg.norm(x_l - x_r) / np.linalg.norm(x_r) < 1e-8
    assert its_l < its_0
    M.close()
    G.close()
    ctx.close()
    tri = meshes.icosphere(3, 0.5)
    bc = meshes.robin_bc(len(tri), 0, 1, 1)
    S = bem.Solver(tri, bc, leaf=32, eta=2.0)
    S.set_preconditioner(2)
    S.set_precond_side(True)
    _, _, p1, st1 = S.solve(500.0, c=343.0, rho=1.21, aca_tol=1e-6, gmres_tol=1e-10, seed=1)
    Q = O.Problem(tri, bc, leaf=32, eta=2.0)
    Q.set_frequency(500.0, 343.0, 1)
    _, _, pro, _ = Q.solve(1e-6, gmres_tol=1e-10, seed=1, rho=1.21, freq=500.0)
    rel = np.linalg.norm(p1 - pro) / np.linalg.norm(pro)
    print("solver, left H-LU: iterations", st1["gmres_iters"], "vs restatement", rel, "true residual", st1["residual"])
    assert rel < 1e-9 and st1["residual"] < 1e-9
    S.close()
