"""The impedance plane of the CPU restatement (the single-plane branch of
hMatrixSolve, boundaryelementsolver.cpp:549-557, 594-604, with the reflected
operators :1516-1643 and the field terms :2305-2311) pinned by the method of
images: a mesh O over a rigid plane is, entry for entry, the free-space
problem of O together with its mirror image O' carrying the same boundary
data.  The mirror image is built here independently (x - 2 ((x - p0).u) u,
nodes reversed), so a wrong reflection (sign of the offset, a missing factor
2, the node order and with it the normal's direction), a jump term on the
mirrored diagonal or a wrong boundary-condition coefficient on the image
column each break one of the identities."""
import numpy as np
import pytest

P0 = np.array([0.3, -0.2, 0.1])
NRM = np.array([0.3, 0.4, 0.866])          # not unit on purpose


def mirror(tri):
    u = NRM / np.linalg.norm(NRM)
    v = tri.reshape(-1, 3, 3)
    h = (v - P0) @ u
    m = v - 2.0 * h[..., None] * u
    return m[:, ::-1, :].reshape(-1, 9)     # nodes c, b, a


@pytest.fixture(scope="module")
def case(O, meshes):
    """L2 sphere (320 elements, r 0.4) with its centre 1 m from the plane,
    impedance patchwork (30 %) so both substituted and plain columns occur"""
    u = NRM / np.linalg.norm(NRM)
    centre = P0 + 1.0 * u + np.array([0.1, -0.05, 0.0])
    tri = meshes.icosphere(2, 0.4, centre)
    n = len(tri)
    f, c = 400.0, 343.0
    k = 2 * np.pi * f / c
    bc = meshes.impedance_bc(n, k, 0.3, 11)
    bc[:, 4] = np.random.default_rng(3).standard_normal(n)      # g: random normal velocity
    bc[bc[:, 0] != 0, 4] = 0.0
    P1 = O.Problem(tri, bc, leaf=16, eta=2.0)
    P1.set_frequency(f, c, 1)
    P1.set_plane(P0, NRM)
    tri2 = np.concatenate([tri, mirror(tri)])
    P2 = O.Problem(tri2, np.concatenate([bc, bc]), leaf=16, eta=2.0)
    P2.set_frequency(f, c, 1)
    return P1, P2, n, f


def positions(P):
    """cluster position of every input element"""
    return np.argsort(P.perm)


def test_mirror_geometry_is_outside_near_field(case):
    """the identity below is exact only if no element-image pair is
    near-singular in the doubled free-space mesh"""
    P1, P2, n, _ = case
    _, _, col, _, hbar = P2.geometry()
    pos = positions(P2)
    a, b = col[pos[:n]], col[pos[n:]]
    d = np.linalg.norm(a[:, None, :] - b[None, :, :], axis=2)
    assert d.min() > 4 * hbar


@pytest.mark.parametrize("kind", [0, 1])
def test_entries_equal_image_sum(case, kind):
    """Phi / DPhi with the plane = free-space block (O, O) + block (O, O')"""
    P1, P2, n, _ = case
    D1 = P1.dense(kind)
    D2 = P2.dense(kind)
    p1, p2 = positions(P1), positions(P2)
    A = D1[np.ix_(p1, p1)]
    B = D2[np.ix_(p2[:n], p2[:n])] + D2[np.ix_(p2[:n], p2[n:])]
    assert np.linalg.norm(A - B) / np.linalg.norm(B) < 1e-13
    # the image block is not negligible: dropping it fails
    C = D2[np.ix_(p2[:n], p2[:n])]
    assert np.linalg.norm(A - C) / np.linalg.norm(B) > 1e-3


def test_solution_and_field_equal_image_problem(case):
    """dense solves: the plane problem's phi, dphi equal the doubled free-space
    problem's on O (which is symmetric), and so do the fields"""
    P1, P2, n, f = case
    phi1, dphi1, pr1 = P1.dense_solve(1.21, f)
    phi2, dphi2, pr2 = P2.dense_solve(1.21, f)
    p1, p2 = positions(P1), positions(P2)
    a, b = phi1[p1], phi2[p2[:n]]
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-11
    assert np.linalg.norm(phi2[p2[n:]] - b) / np.linalg.norm(b) < 1e-11
    assert np.linalg.norm(dphi1[p1] - dphi2[p2[:n]]) / np.linalg.norm(dphi2) < 1e-11
    u = NRM / np.linalg.norm(NRM)
    rng = np.random.default_rng(5)
    obs = P0 + rng.uniform(-2, 2, (60, 3))
    obs = obs[(obs - P0) @ u > 0.05]                   # the half-space of the mesh
    f1 = P1.field(obs, phi1, dphi1)
    f2 = P2.field(obs, phi2, dphi2)
    assert np.linalg.norm(f1 - f2) / np.linalg.norm(f2) < 1e-11
    # the plane changes the answer
    P1.set_plane(None)
    try:
        phi0, _, _ = P1.dense_solve(1.21, f)
    finally:
        P1.set_plane(P0, NRM)
    assert np.linalg.norm(phi0[p1] - b) / np.linalg.norm(b) > 1e-2


def test_hmatrix_solve_with_plane(case):
    """the H-matrix path (ACA on the sum, same partition as without the
    plane) converges to the dense solution within the ACA tolerance"""
    P1, _, n, f = case
    phi_d, _, _ = P1.dense_solve(1.21, f)
    phi, dphi, pr, st = P1.solve(1e-6, gmres_tol=1e-10, seed=2, rho=1.21, freq=f)
    assert st["residual"] < 1e-9
    assert np.linalg.norm(phi - phi_d) / np.linalg.norm(phi_d) < 1e-5
