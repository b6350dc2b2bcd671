"""Field evaluation of the CPU restatement (BemOperatorField /
calculateFieldSolutionRegular, boundaryelementsolver.cpp:2262-2353) pinned
by the Kirchhoff-Helmholtz identity on a sphere: with the exact boundary
data of a pulsating sphere the integral reproduces the radiated field
outside and vanishes inside, and the BEM solution radiates the closed-form
field."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def sphere(O, meshes):
    tri = meshes.icosphere(3, 0.5)
    P = O.Problem(tri, meshes.robin_bc(len(tri), 0, 1, 1), leaf=32, eta=2.0)
    P.set_frequency(500.0, 343.0, 1)
    return P


def exact_phi(k, a, r):
    """pulsating sphere, dphi/dn = 1 at r = a: phi = a^2 / (i k a - 1) e^{ik(r - a)} / r"""
    return a * a / (1j * k * a - 1) * np.exp(1j * k * (r - a)) / r


def directions(m, seed):
    d = np.random.default_rng(seed).standard_normal((m, 3))
    return d / np.linalg.norm(d, axis=1)[:, None]


def test_field_identity_with_exact_boundary_data(sphere):
    """Exact phi(a), dphi/dn = 1 as boundary data: the field equals the
    radiated field at r = 1, 2, 4 m (to the 1-2 % discretisation error of the
    faceted sphere) and is ~0 at interior points."""
    P = sphere
    k = P.params()["k"]
    a = 0.5
    n = P.n
    phi = np.full(n, exact_phi(k, a, a), complex)
    dphi = np.ones(n, complex)
    for r in (1.0, 2.0, 4.0):
        obs = r * directions(20, 1)
        f = P.field(obs, phi, dphi)
        ref = exact_phi(k, a, r)
        assert np.max(np.abs(f - ref)) / abs(ref) < 0.03, r
    inner = 0.2 * directions(20, 2)
    f = P.field(inner, phi, dphi)
    assert np.max(np.abs(f)) < 0.01 * abs(exact_phi(k, a, a))
    # a sign error in either term fails: the M term alone is far from the field
    f_m = P.field(2.0 * directions(5, 3), phi, np.zeros(n, complex))
    assert np.max(np.abs(f_m - exact_phi(k, a, 2.0))) / abs(exact_phi(k, a, 2.0)) > 0.1


def test_field_of_bem_solution(sphere):
    P = sphere
    k = P.params()["k"]
    phi, dphi, pr, st = P.solve(1e-6, gmres_tol=1e-10, rho=1.21, freq=500.0)
    obs = 1.5 * directions(30, 4)
    f = P.field(obs, phi, dphi)
    ref = exact_phi(k, 0.5, 1.5)
    assert np.max(np.abs(f - ref)) / abs(ref) < 0.03
