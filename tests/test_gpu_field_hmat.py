"""The field through the observation x boundary H-matrix on the device
(bem_field_hmat, bem_solver_field_hmat; calculateFieldSolutionFast,
boundaryelementsolver.cpp:2397-2476) against the restatement on the same
partition and pivots (1e-12: the same ACA, products summed in another
order), against the direct sum (full rank: round-off; tolerance t: O(t)),
and against the pulsating sphere's closed form."""
import numpy as np
import pytest

from helpers import sphere_problem, upload_problem

pytestmark = pytest.mark.gpu


def points(m, seed, rmin=0.55, rmax=4.0):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((m, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return rng.uniform(rmin, rmax, m)[:, None] * d


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("tol", [0.0, 1e-2, 1e-5])
def test_field_hmat_matches_restatement(bem, O, meshes, tol):
    level = 2 if tol == 0.0 else 4
    P = sphere_problem(O, meshes, level=level, leaf=32, eta=2.0, freq=700.0)
    ctx = upload_problem(bem, P)
    rng = np.random.default_rng(3)
    n = P.n
    phi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dphi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    obs = points(400 if tol == 0.0 else 3000, 4)
    perm, b5 = P.field_partition(obs, leaf=50, eta=3.0)
    print("field hmat: n", n, "points", len(obs), "blocks", len(b5), "admissible", int(b5[:, 4].sum()), "tol", tol,
          flush=True)
    got_c, st = ctx.field_hmat(obs[perm], b5, ctx.array(phi), ctx.array(dphi), tol, seed=7)
    got = np.empty_like(got_c)
    got[perm] = got_c
    ref, (nadm, maxr, ev) = P.field_hmat(obs, phi, dphi, tol, seed=7, leaf=50, eta=3.0)
    assert int(st[1]) == nadm and int(st[2]) == maxr and int(st[3]) == ev, (st, nadm, maxr, ev)
    assert rel(got, ref) < 1e-12, rel(got, ref)
    direct = ctx.field(obs, ctx.array(phi), ctx.array(dphi))
    assert rel(got, direct) < (1e-12 if tol == 0.0 else 10 * tol), rel(got, direct)
    ctx.close()


def test_solver_field_hmat_pulsating_sphere(bem, meshes):
    """L4 sphere, 800 Hz: bem_solver_field_hmat against the direct sum of
    the same solution (ACA 1e-6: 1e-5; the reference default 0.01: 0.05)
    and against p = i w rho a^2 / (i k a - 1) e^{ik(r-a)} / r (2 %); the
    points come back in their input order."""
    a, f, c, rho = 0.5, 800.0, 343.0, 1.21
    k = 2 * np.pi * f / c
    w = 2 * np.pi * f
    tri = meshes.icosphere(4, a)
    n = len(tri)
    S = bem.Solver(tri, meshes.robin_bc(n, 0, 1, 1), leaf=64, eta=2.0)
    S.solve(f, c=c, rho=rho, aca_tol=1e-6, gmres_tol=1e-10)
    obs = points(2000, 9, 0.8, 5.0)
    direct = S.field(obs)
    fast, st = S.field_hmat(obs, aca_tol=1e-6)
    print("solver field hmat:", st, flush=True)
    assert st["admissible_blocks"] > 0
    assert rel(fast, direct) < 1e-5, rel(fast, direct)
    fast2, _ = S.field_hmat(obs)
    assert rel(fast2, direct) < 0.05, rel(fast2, direct)
    r = np.linalg.norm(obs, axis=1)
    pex = 1j * w * rho * a * a / (1j * k * a - 1) * np.exp(1j * k * (r - a)) / r
    assert np.max(np.abs(fast - pex) / np.abs(pex)) < 0.02
    S.set_impedance_plane(np.array([0, 0, -1.0]), np.array([0, 0, 1.0]))
    with pytest.raises(bem.BemError):
        S.field_hmat(obs)
    S.close()
