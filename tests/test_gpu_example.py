"""The reference's example project end to end (Example/Sphere: the mesh
Sphere.msh kept under tests/golden, SolvingSphere.txt's set-up): read,
shifted, refined to a maximum edge length (SolvingScriptReader::
refineElements), a rigid radius-5 sphere with a point source and an
infinite baffle, solved on the device and by the CPU restatement."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPHERE_MSH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "Sphere.msh")


def test_example_sphere_point_source_and_baffle(bem, O, meshes):
    """Sphere.msh (110 triangles, radius 5 m) with SolvingSphere.txt's shift
    (6.1, 1.5, 1.5) m, refined to edges of at most 0.8 m (2780 triangles);
    rigid elements (the default boundary condition); the pressure point of
    weight 5000 at (20, 0, 0) m; the infinite baffle through (0, 0, -10) with
    normal (0, 0, 1), which the reader turns into (0, 0, -1); 100 Hz,
    c = 344 m/s.  Burton-Miller, leaf 50, eta 3 (the reference's
    constants), ACA 1e-6, GMRES 1e-10: pressure equal to the restatement's
    at 1e-10."""
    tri = bem.read_msh(SPHERE_MSH)
    tri = tri + np.tile([6.1, 1.5, 1.5], 3)
    tri = bem.refine_triangles(tri, 0.8)
    n = len(tri)
    assert n == 2780
    bc = meshes.robin_bc(n, 0, 1, 0)
    src = np.array([[20.0, 0.0, 0.0, 5000.0, 0.0, 0.0]])
    pt, nrm = np.array([0.0, 0.0, -10.0]), np.array([0.0, 0.0, -1.0])
    S = bem.Solver(tri, bc, leaf=50, eta=3.0)
    S.set_sources(points=src)
    S.set_impedance_plane(pt, nrm)
    print("example sphere: n", n, flush=True)
    _, _, pr, st = S.solve(100.0, c=344.0, rho=1.21, coupling=1, aca_tol=1e-6, gmres_tol=1e-10, seed=1)
    P = O.Problem(tri, bc, leaf=50, eta=3.0)
    P.set_frequency(100.0, 344.0, 1)
    P.set_sources(points=src)
    P.set_plane(pt, nrm)
    _, _, pro, _ = P.solve(1e-6, gmres_tol=1e-10, seed=1, rho=1.21, freq=100.0)
    rel = np.linalg.norm(pr - pro) / np.linalg.norm(pro)
    print("example sphere: GMRES iterations", st["gmres_iters"], "rel-L2 vs restatement", rel, flush=True)
    assert rel < 1e-10
