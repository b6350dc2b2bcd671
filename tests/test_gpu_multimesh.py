"""Several meshes in one ACA launch (bem_assemble_multi): three contexts of
different size, shape and frequency; each context's operator from the
multi-mesh launch equals its own bem_assemble bit for bit, FP64 and FP32,
Phi and DPhi, with one workgroup per item and with persistent workgroups."""
import numpy as np
import pytest

from helpers import upload_problem

pytestmark = pytest.mark.gpu


def same_blocks(G, H, nb):
    for b in range(nb):
        A, B = G.block(b), H.block(b)
        if isinstance(A, tuple):
            if not isinstance(B, tuple) or len(A[1]) != len(B[1]):
                return False, b
            for x, y in zip(A, B):
                if not np.array_equal(np.ascontiguousarray(x).view(np.uint64), np.ascontiguousarray(y).view(np.uint64)):
                    return False, b
        elif not np.array_equal(np.ascontiguousarray(A).view(np.uint64), np.ascontiguousarray(B).view(np.uint64)):
            return False, b
    return True, -1


def test_multi_mesh_equals_single(bem, O, meshes):
    cases = [(meshes.icosphere(2, 0.4), 400.0), (meshes.ellipsoid(3, (0.6, 0.35, 0.25)), 700.0),
             (meshes.icosphere(3, 0.3), 1200.0)]
    probs, ctxs = [], []
    for tri, f in cases:
        P = O.Problem(tri, meshes.robin_bc(len(tri), 0, 1, 1), leaf=32, eta=2.0)
        P.set_frequency(f, 343.0, 1)
        probs.append(P)
        ctxs.append(upload_problem(bem, P))
    for prec in (bem.FP64, bem.FP32):
        for kind in (0, 1):
            single = [c.assemble(kind, 1e-6, 0, precision=prec, seed=5) for c in ctxs]
            for queue in (0, 1):
                ctxs[0].set_aca_schedule(prec, 384, 2048 if prec == bem.FP32 else 1 << 40, queue)
                print("multi-mesh: prec", prec, "kind", kind, "queue", queue, flush=True)
                multi = bem.assemble_multi(ctxs, kind, 1e-6, precision=prec, seed=5)
                for q, (G, H) in enumerate(zip(multi, single)):
                    ok, b = same_blocks(G, H, len(probs[q].blocks))
                    assert ok, (prec, kind, queue, q, b)
                    assert G.stats()["evaluated"] == H.stats()["evaluated"]
                for G in multi:
                    G.close()
            for H in single:
                H.close()
    for c in ctxs:
        c.close()


@pytest.mark.parametrize("prec,refine", [(0, 0), (1, 1)])
def test_solve_multi_equals_single_solves(bem, meshes, prec, refine):
    """bem_solvers_solve_multi on three meshes at their own frequencies:
    every pressure equals that solver's own bem_solver_solve bit for bit
    (FP64; FP32 with refinement mode 1)."""
    cases = [(meshes.icosphere(3, 0.5), 500.0), (meshes.ellipsoid(3, (0.5, 0.3, 0.2)), 900.0),
             (meshes.icosphere(2, 0.3), 1300.0)]
    solvers = [bem.Solver(t, meshes.robin_bc(len(t), 0, 1, 1), leaf=32, eta=2.0) for t, _ in cases]
    for S in solvers:
        S.set_refinement(refine)
    freqs = [f for _, f in cases]
    pm, its, st = bem.solve_multi(solvers, freqs, aca_tol=1e-6, gmres_tol=1e-10, seed=3, precision=prec, threads=2)
    for q, S in enumerate(solvers):
        _, _, ps, st1 = S.solve(freqs[q], aca_tol=1e-6, gmres_tol=1e-10, seed=3, precision=prec)
        assert np.array_equal(ps.view(np.uint64), pm[q].view(np.uint64)), q
        assert its[q] == st1["gmres_iters"]
    for S in solvers:
        S.close()
