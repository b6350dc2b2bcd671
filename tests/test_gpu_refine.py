"""FP64-residual iterative refinement of the FP32 form (bem_set_near_fp64,
bem_hmvm_refine, bem_gmres_ir, bem_solver_set_refinement): the refinement
operator is the FP64 near field plus the stored low-rank factors (mode 1)
or the FP64 H-matrix (mode 2); GMRES iterates on the FP32 operator.
Checked against the restatement's FP64 blocks and solves."""
import numpy as np
import pytest

from helpers import sphere_problem, upload_problem

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def L3(O, meshes):
    return sphere_problem(O, meshes, level=3, leaf=32, eta=2.0, freq=500.0)


@pytest.mark.parametrize("kind", [0, 1])
def test_refine_operator_is_fp64_near_field_plus_factors(bem, L3, kind):
    """FP32 assembly with the FP64 near field kept: its FP32 dense blocks are
    the restatement's FP64 blocks rounded once, and bem_hmvm_refine equals
    (FP64 dense blocks + the stored FP32 factors) x to FP64 round-off."""
    ctx = upload_problem(bem, L3)
    ctx.set_near_pairing(True)
    ctx.set_near_fp64(True)
    G0 = ctx.assemble(0, 1e-6, 0, precision=bem.FP32, seed=7)
    G = G0 if kind == 0 else ctx.assemble(1, 1e-6, 0, precision=bem.FP32, seed=7)   # DPhi adopts the paired field
    H = L3.assemble(kind, 1e-6, 0, 7)
    n = L3.n
    A = np.zeros((n, n), complex)
    for b, (r0, m, c0, nn, adm) in enumerate(L3.blocks):
        if adm:
            U, al, V = G.block(b)
            A[r0:r0 + m, c0:c0 + nn] = (U * al) @ V
        else:
            F64 = H.block(b)
            assert np.array_equal(G.block(b), F64.astype(np.complex64).astype(np.complex128)), b
            A[r0:r0 + m, c0:c0 + nn] = F64
    rng = np.random.default_rng(kind)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = G.mvm_refine(x)
    assert rel(y, A @ x) < 1e-13
    st = G.stats()
    assert st["resident_bytes"] > st["bytes"]
    G.close()
    G0.close()
    ctx.close()


def test_gmres_ir_converges_to_refinement_operator(bem, L3):
    """GMRES-IR on the FP32 DPhi against its FP64 refinement operator reaches
    |b - A_res x| < 1e-12 |b| (plain GMRES on the FP32 operator leaves the
    A_res residual at the FP32 level), in a few refinement steps."""
    ctx = upload_problem(bem, L3)
    ctx.set_near_fp64(True)
    G = ctx.assemble(1, 1e-6, 0, precision=bem.FP32, seed=3)
    s = np.exp(0.37j * np.arange(L3.n)) + 0.5
    x, its, r, steps, rc = G.gmres_ir(s, 1e-12)
    assert rc == 0 and r < 1e-12 and 1 <= steps <= 6, (r, steps)
    assert rel(G.mvm_refine(x), s) < 2e-12
    x0, _, _, _ = G.gmres(s, 1e-12)
    assert rel(G.mvm_refine(x0), s) > 1e-10                       # FP32 operator alone: not FP64-accurate
    G.close()
    ctx.close()


def test_solver_refinement_sphere1280(bem, O, meshes):
    """Pulsating sphere L3, 500 Hz, ACA 1e-6: FP32 with refinement mode 2
    (FP64 H-matrix residuals, GMRES 1e-12) equals the FP64 CPU restatement to
    1e-9; modes 0 and 1 stay within the FP32 form's 1e-6 of it."""
    tri = meshes.icosphere(3, 0.5)
    n = len(tri)
    bc = meshes.robin_bc(n, 0, 1, 1)
    S = bem.Solver(tri, bc, leaf=32, eta=2.0)
    P = O.Problem(tri, bc, leaf=32, eta=2.0)
    P.set_frequency(500.0, 343.0, 1)
    _, _, pref, _ = P.solve(1e-6, gmres_tol=1e-12, seed=1, rho=1.21, freq=500.0)
    out = {}
    for mode in (0, 1, 2):
        S.set_refinement(mode)
        _, _, pr, st = S.solve(500.0, rho=1.21, aca_tol=1e-6, gmres_tol=1e-12, seed=1, precision=bem.FP32)
        rs = S.refine_stats()
        out[mode] = rel(pr, pref)
        print("refine mode", mode, "vs FP64 restatement", out[mode], "stats", rs, "iters", st["gmres_iters"])
        if mode:
            assert rs["steps"] >= 1 and rs["residual"] < 1e-12
    assert out[2] < 1e-9
    assert out[0] < 1e-6 and out[1] < 1e-6
