"""The restatement's field through the observation x boundary H-matrix
(calculateFieldSolutionFast, boundaryelementsolver.cpp:2397-2476) pinned by
what the mathematics fixes: the partition covers every (point, element)
pair once; with ACA tolerance 0 (every admissible block to full rank, an
exact factorisation) it reproduces the direct sum to round-off, so a wrong
sign, kernel or index in the blocks' terms shows; with tolerance t the
error stays at the order of t; the field pivots stay in their groups."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def sphere(O, meshes):
    tri = meshes.icosphere(2, 0.5)
    P = O.Problem(tri, meshes.robin_bc(len(tri), 0, 1, 1), leaf=16, eta=2.0)
    P.set_frequency(700.0, 343.0, 1)
    return P


def points(m, seed, rmin=0.6, rmax=3.0):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((m, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return rng.uniform(rmin, rmax, m)[:, None] * d


def data(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n), rng.standard_normal(n) + 1j * rng.standard_normal(n))


def test_partition_covers_every_pair_once(sphere):
    obs = points(300, 1)
    perm, b5 = sphere.field_partition(obs, leaf=20, eta=2.0)
    assert sorted(perm.tolist()) == list(range(300))
    cnt = np.zeros((300, sphere.n), np.int32)
    for r0, m, c0, nn, adm in b5:
        cnt[r0:r0 + m, c0:c0 + nn] += 1
    assert np.all(cnt == 1)
    assert b5[:, 4].sum() > 0 and (b5[:, 4] == 0).sum() > 0   # both kinds of block occur


def test_full_rank_aca_is_the_direct_sum(sphere):
    """tol = 0: every admissible block factorised exactly, so the H-matrix
    field equals orc_field's direct sum to round-off (1e-12); swapping the
    kernels or dropping the minus sign would show at O(1)."""
    obs = points(300, 2)
    phi, dphi = data(sphere.n, 3)
    ref = sphere.field(obs, phi, dphi)
    got, (nadm, maxr, ev) = sphere.field_hmat(obs, phi, dphi, 0.0, leaf=20, eta=2.0)
    assert nadm > 0
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12
    # each kernel alone: phi only and dphi only
    z = np.zeros(sphere.n, complex)
    for a, b in ((phi, z), (z, dphi)):
        r1 = sphere.field(obs, a, b)
        g1, _ = sphere.field_hmat(obs, a, b, 0.0, leaf=20, eta=2.0)
        assert np.linalg.norm(g1 - r1) / np.linalg.norm(r1) < 1e-12


@pytest.mark.parametrize("tol", [1e-2, 1e-4, 1e-6])
def test_error_follows_the_aca_tolerance(sphere, tol):
    """The approximation error is of the order of the ACA tolerance (10 tol)
    and the ranks grow as it tightens."""
    obs = points(300, 4)
    phi, dphi = data(sphere.n, 5)
    ref = sphere.field(obs, phi, dphi)
    got, (nadm, maxr, ev) = sphere.field_hmat(obs, phi, dphi, tol, leaf=20, eta=2.0)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 10 * tol, err
    assert 1 <= maxr


def test_deterministic_and_seed_keyed(sphere):
    obs = points(200, 6)
    phi, dphi = data(sphere.n, 7)
    a, _ = sphere.field_hmat(obs, phi, dphi, 1e-3, seed=5, leaf=20, eta=2.0)
    b, _ = sphere.field_hmat(obs, phi, dphi, 1e-3, seed=5, leaf=20, eta=2.0)
    assert np.array_equal(a, b)
