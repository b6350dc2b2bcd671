"""The 20480-triangle pulsating sphere (icosphere level 5, r = 0.5 m,
1000 Hz, leaf 64, eta 2, ACA 1e-5, GMRES 1e-8) against an FP64 dense-assembly
solve: every entry of Phi and DPhi from the device entry generator (bit-equal
to the CPU restatement, test_gpu_parity.py) and the dense system
(the reference's regularSolveWithLU, boundaryelementsolver.cpp:140-331, with
rhs = s + Phi f) solved by GMRES to 1e-12.  States the surface-pressure error of the FP64 and the
FP32 H-matrix paths against it and against each other."""
import numpy as np
import pytest

from helpers import sphere_problem, upload_problem

pytestmark = pytest.mark.gpu


def dense_fp64_solution(bem, P, rho, freq, gmres_tol=1e-12):
    """Dense FP64 solve through the library itself: the partition is one
    non-admissible n x n block, so assembly writes every entry of Phi and
    DPhi densely; GMRES on the dense operator to gmres_tol."""
    n = P.n
    ctx = upload_problem(bem, P)
    ctx.upload_partition(np.array([[0, n, 0, n, 0]], np.int64))
    s, a, b, f, sub = P.rhs_parts()
    Phi = ctx.assemble(0, 0.0)
    rhs = ctx.array(s)
    Phi.mvm_dev(ctx.array(f), rhs, accumulate=True)
    rhs_h = rhs.download()
    Phi.close()
    DPhi = ctx.assemble(1, 0.0)
    x, its, rel, rc = DPhi.gmres(rhs_h, gmres_tol)
    assert rc == 0 and rel < gmres_tol, (rc, rel)
    DPhi.close()
    _, _, pr = ctx.reconstruct(ctx.array(x), rho, freq)
    return pr


def test_sphere20480_fp64_and_fp32_against_dense_fp64(bem, O, meshes):
    freq, rho = 1000.0, 1.21
    P = sphere_problem(O, meshes, level=5, leaf=64, eta=2.0, freq=freq)
    print("dense FP64 reference, n =", P.n, flush=True)
    pd = dense_fp64_solution(bem, P, rho, freq)
    tri = meshes.icosphere(5, 0.5)
    bc = meshes.robin_bc(len(tri), 0, 1, 1)
    S = bem.Solver(tri, bc, leaf=64, eta=2.0)
    perm, blocks = S.layout()
    assert np.array_equal(perm, P.perm) and np.array_equal(blocks, P.blocks)
    out = {}
    for name, prec in (("fp64", bem.FP64), ("fp32", bem.FP32)):
        _, _, pr, st = S.solve(freq, c=343.0, rho=rho, coupling=1, aca_tol=1e-5, gmres_tol=1e-8, seed=1,
                               precision=prec)
        out[name] = pr
        print(name, "H-matrix vs dense FP64 rel-L2", np.linalg.norm(pr - pd) / np.linalg.norm(pd),
              "iters", st["gmres_iters"], flush=True)
    e64 = np.linalg.norm(out["fp64"] - pd) / np.linalg.norm(pd)
    e32 = np.linalg.norm(out["fp32"] - pd) / np.linalg.norm(pd)
    d = np.linalg.norm(out["fp32"] - out["fp64"]) / np.linalg.norm(out["fp64"])
    print("fp32 vs fp64 H-matrix rel-L2", d, flush=True)
    # ACA at 1e-5 bounds the H-matrix paths; the FP32 entries add < 1e-6
    assert e64 < 1e-5
    assert e32 < e64 + 1e-6
    assert d < 1e-6
