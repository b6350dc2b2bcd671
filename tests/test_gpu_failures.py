"""Failure detection, forced from the host so that every kernel still ends
normally.  The reference prints the blocks that hold NaN and continues
(hmatrix.cpp:812-820); this library refuses boundary data without finite
column coefficients before any kernel runs, and stops an assembly that
meets a non-finite value with BEM_ERR_NAN.  GMRES that reaches its
iteration limit above the tolerance is reported (BEM_ERR_NOCONV) with the
iterations and residual it reached."""
import numpy as np
import pytest

from helpers import sphere_problem, upload_problem

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="the one GPU run of an earlier form of these tests ended with the card faulted "
                            "after both had passed; the faulting kernel is not identified, so they stay off "
                            "the GPU until it is"),
]


@pytest.mark.parametrize("precision", ["FP64", "FP32"])
def test_non_finite_boundary_coefficients(bem, meshes, precision):
    """a = b = 0 on one element leaves its column without finite
    coefficients: the solve is refused (BEM_ERR_ARG) before any kernel
    runs, and a valid problem solves afterwards in the same process."""
    prec = getattr(bem, precision)
    tri = meshes.icosphere(3, 0.5)
    n = len(tri)
    for bad in (17, 600, n - 1):
        bc = meshes.robin_bc(n, 0.0, 1.0, 1.0)
        bc[bad, 2:4] = 0.0
        S = bem.Solver(tri, bc, leaf=32, eta=2.0)
        with pytest.raises(bem.BemError) as e:
            S.solve(500.0, aca_tol=1e-6, gmres_tol=1e-10, precision=prec)
        assert e.value.code == bem.BEM_ERR_ARG, (bad, str(e.value))
        S.close()
    S = bem.Solver(tri, meshes.robin_bc(n, 0.0, 1.0, 1.0), leaf=32, eta=2.0)
    _, _, p, st = S.solve(500.0, aca_tol=1e-6, gmres_tol=1e-10, precision=prec)
    assert np.all(np.isfinite(p)) and st["residual"] < 1e-9
    S.close()


def test_gmres_iteration_limit_is_reported(bem, O, meshes):
    """GMRES with a limit of three iterations and a tolerance it cannot reach
    in three returns BEM_ERR_NOCONV with at most three iterations, the
    residual it reached (above the tolerance) and a finite iterate; without
    the limit the same system converges."""
    P = sphere_problem(O, meshes, level=3, leaf=32, eta=2.0, freq=500.0)
    ctx = upload_problem(bem, P)
    H = ctx.assemble(bem.DPHI, 1e-6)
    rng = np.random.default_rng(3)
    b = (rng.standard_normal(P.n) + 1j * rng.standard_normal(P.n)).reshape(-1, 1)
    x, its, rel, _, rc = H.gmres_multi(b, 1e-12, maxit=3)
    assert rc == bem.BEM_ERR_NOCONV
    assert 1 <= its[0] <= 3 and rel[0] > 1e-12
    assert np.all(np.isfinite(x))
    x2, its2, rel2, _, rc2 = H.gmres_multi(b, 1e-12)
    assert rc2 == bem.BEM_OK and its2[0] > 3 and rel2[0] < 1e-12
    H.close()
    ctx.close()
