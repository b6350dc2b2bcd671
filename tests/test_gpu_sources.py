"""Point sources (sourceTerm, boundaryelementsolver.cpp:3431-3441, the
reference's only excitation type) through the device solver: Burton-Miller
solves with point sources equal the CPU restatement (so the kept division
by (ikr - 1) runs on both sides), and a sweep carries the solver's point
sources into every right-hand side."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SOURCES = np.array([[0.3, -0.4, 1.1, 1.0, 0.0, 0.0],
                    [-0.9, 0.2, -0.1, 0.4, -0.7, 0.0]])


@pytest.mark.parametrize("bc", [(0, 1, 0), (0, 1, 1)])
def test_point_sources_burton_miller(bem, O, meshes, bc):
    """L3 sphere, 700 Hz, rigid (g = 0) and pulsating (g = 1) with two
    exterior point sources, Burton-Miller, ACA 1e-6, GMRES 1e-10: pressure
    equal to the restatement's at 1e-10."""
    tri = meshes.icosphere(3, 0.5)
    n = len(tri)
    S = bem.Solver(tri, meshes.robin_bc(n, *bc), leaf=32, eta=2.0)
    S.set_sources(points=SOURCES)
    phi, dphi, pr, st = S.solve(700.0, c=343.0, rho=1.21, coupling=1, aca_tol=1e-6, gmres_tol=1e-10, seed=1)
    P = O.Problem(tri, meshes.robin_bc(n, *bc), leaf=32, eta=2.0)
    P.set_frequency(700.0, 343.0, 1)
    P.set_sources(points=SOURCES)
    phio, _, pro, _ = P.solve(1e-6, gmres_tol=1e-10, seed=1, rho=1.21, freq=700.0)
    rel = np.linalg.norm(pr - pro) / np.linalg.norm(pro)
    print("point sources", bc, "rel-L2 vs restatement", rel)
    assert rel < 1e-10
    assert np.linalg.norm(phi - phio) / np.linalg.norm(phio) < 1e-10


def test_sweep_keeps_point_sources(bem, O, meshes):
    """A two-frequency sweep on two lanes with point sources set on the
    solver (and a plane wave for the second right-hand side) equals the
    restatement's single solves with the same excitation, 1e-10."""
    tri = meshes.icosphere(3, 0.5)
    n = len(tri)
    d = meshes.unit_direction(5)
    S = bem.Solver(tri, meshes.robin_bc(n, 0, 1, 0), leaf=32, eta=2.0)
    S.set_sources(points=SOURCES)
    freqs = np.array([350.0, 650.0])
    planes = np.array([[0, 0, 0, 0, 0], [d[0], d[1], d[2], 1.0, 0.0]])
    g = np.zeros((2, n), complex)
    pr, its, _ = S.sweep(freqs, g=g, planes=planes, aca_tol=1e-6, gmres_tol=1e-10, seed=1, lanes=2)
    P = O.Problem(tri, meshes.robin_bc(n, 0, 1, 0), leaf=32, eta=2.0)
    for fi, f in enumerate(freqs):
        for k in range(2):
            P.set_frequency(f, 343.0, 1)
            P.set_sources(points=SOURCES, plane=planes[1] if k == 1 else None)
            _, _, ref, _ = P.solve(1e-6, gmres_tol=1e-10, seed=1, rho=1.21, freq=f)
            rel = np.linalg.norm(pr[fi, k] - ref) / np.linalg.norm(ref)
            assert rel < 1e-10, (f, k, rel)
