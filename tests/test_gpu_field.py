"""Field evaluation on the device (bem_field, bem_solver_field) against the
CPU restatement (bit for bit) and against the closed forms of the
pulsating and the rigid sphere."""
import numpy as np
import pytest
from scipy.special import eval_legendre, spherical_jn, spherical_yn

from helpers import sphere_problem, upload_problem

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128))
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def directions(m, seed):
    d = np.random.default_rng(seed).standard_normal((m, 3))
    return d / np.linalg.norm(d, axis=1)[:, None]


def test_field_bit_exact(bem, O, meshes):
    """Random boundary data, observation points from just outside the
    surface to far away (n = 5120: five 1024-element chunks)."""
    P = sphere_problem(O, meshes, level=4, leaf=64, eta=2.0, freq=700.0)
    ctx = upload_problem(bem, P)
    rng = np.random.default_rng(3)
    n = P.n
    phi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dphi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    r = np.concatenate([np.full(50, 0.52), rng.uniform(0.6, 5.0, 250)])
    obs = r[:, None] * directions(300, 4)
    ref = P.field(obs, phi, dphi)
    got = ctx.field(obs, ctx.array(phi), ctx.array(dphi))
    assert same_bits(ref, got), np.abs(ref - got).max()


def test_solver_field_pulsating_and_rigid_sphere(bem, meshes):
    """L4 sphere, 800 Hz: radiated pressure of the pulsating sphere at
    r = 1 and 3 m against p = i w rho a^2 / (i k a - 1) e^{ik(r-a)} / r;
    total pressure of a plane wave scattered by the rigid sphere at r = 1 m
    against the Mie series."""
    a, f, c, rho = 0.5, 800.0, 343.0, 1.21
    k = 2 * np.pi * f / c
    w = 2 * np.pi * f
    tri = meshes.icosphere(4, a)
    n = len(tri)
    S = bem.Solver(tri, meshes.robin_bc(n, 0, 1, 1), leaf=64, eta=2.0)
    S.solve(f, c=c, rho=rho, aca_tol=1e-6, gmres_tol=1e-10)
    for r in (1.0, 3.0):
        obs = r * directions(40, 5)
        p = S.field(obs)
        pex = 1j * w * rho * a * a / (1j * k * a - 1) * np.exp(1j * k * (r - a)) / r
        assert np.max(np.abs(p - pex)) / abs(pex) < 0.02, r
    d = meshes.unit_direction(6)
    S2 = bem.Solver(tri, meshes.robin_bc(n, 0, 1, 0), leaf=64, eta=2.0)
    S2.set_sources(plane=np.array([d[0], d[1], d[2], 1.0, 0.0]))
    S2.solve(f, c=c, rho=rho, aca_tol=1e-6, gmres_tol=1e-10)
    r = 1.0
    obs = r * directions(40, 7)
    p = S2.field(obs)
    cos_t = obs @ d / r
    ka, kr = k * a, k * r
    phi = np.zeros(len(obs), complex)
    for m in range(60):
        jp = spherical_jn(m, ka, derivative=True)
        hp = jp + 1j * spherical_yn(m, ka, derivative=True)
        hr = spherical_jn(m, kr) + 1j * spherical_yn(m, kr)
        phi += (1j ** m) * (2 * m + 1) * (spherical_jn(m, kr) - jp / hp * hr) * eval_legendre(m, cos_t)
    pm = 1j * w * rho * phi
    err = np.linalg.norm(p - pm) / np.linalg.norm(pm)
    print("rigid sphere field vs Mie", err)
    assert err < 0.02
