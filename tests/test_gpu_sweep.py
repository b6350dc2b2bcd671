"""Frequency sweeps with several right-hand sides (bem_solver_sweep): every
(frequency, right-hand side) pressure equals the single-frequency solve of
the CPU restatement; the 64-frequency L4 sweep (pulsating + plane-wave
scattering, FP32) against the closed forms."""
import numpy as np
import pytest
from scipy.special import eval_legendre, spherical_jn, spherical_yn

pytestmark = pytest.mark.gpu


def mie_rigid_surface_phi(k, a, cos_t, terms=60):
    """Total potential on a rigid sphere for a unit plane wave."""
    ka = k * a
    tot = np.zeros(len(cos_t), complex)
    for m in range(terms):
        hp = spherical_jn(m, ka, derivative=True) + 1j * spherical_yn(m, ka, derivative=True)
        tot += (2 * m + 1) * (1j ** m) * eval_legendre(m, cos_t) * (1j / (ka ** 2 * hp))
    return tot


def test_sweep_equals_single_solves(bem, O, meshes):
    """L3 sphere, three frequencies on two lanes, two right-hand sides
    (pulsating g = 1; rigid scattering of a plane wave, g = 0): FP64 sweep
    against the restatement's single solves, 1e-10."""
    tri = meshes.icosphere(3, 0.5)
    n = len(tri)
    d = meshes.unit_direction(3)
    S = bem.Solver(tri, meshes.robin_bc(n, 0, 1, 1), leaf=32, eta=2.0)
    freqs = np.array([300.0, 500.0, 800.0])
    g = np.stack([np.ones(n, complex), np.zeros(n, complex)])
    planes = np.array([[0, 0, 0, 0, 0], [d[0], d[1], d[2], 1.0, 0.0]])
    pr, its, st = S.sweep(freqs, g=g, planes=planes, aca_tol=1e-6, gmres_tol=1e-10, seed=1, lanes=2)
    assert pr.shape == (3, 2, n) and np.all(its > 0)
    Pp = O.Problem(tri, meshes.robin_bc(n, 0, 1, 1), leaf=32, eta=2.0)
    Ps = O.Problem(tri, meshes.robin_bc(n, 0, 1, 0), leaf=32, eta=2.0)
    for fi, f in enumerate(freqs):
        for k, P in enumerate((Pp, Ps)):
            P.set_frequency(f, 343.0, 1)
            if k == 1:
                P.set_sources(plane=np.array([d[0], d[1], d[2], 1.0, 0.0]))
            _, _, ref, _ = P.solve(1e-6, gmres_tol=1e-10, seed=1, rho=1.21, freq=f)
            rel = np.linalg.norm(pr[fi, k] - ref) / np.linalg.norm(ref)
            assert rel < 1e-10, (f, k, rel)


def test_sphere5120_sweep64_fp32(bem, meshes):
    """Icosphere L4 (5120 triangles), 64 log-spaced frequencies 100-2000 Hz,
    pulsating and plane-wave scattering right-hand sides (seeded direction),
    FP32 assembly, leaf 64, eta 2, ACA 1e-5, GMRES 1e-8, run as the batched
    form: chunks of 4 frequencies (one batched assembly per operator and one
    lockstep GMRES over the 8 (frequency, right-hand side) columns), FP64
    near-field refinement, the ACA schedule tuned on the first chunk, two
    lanes.  Pulsating pressure against p(a) = i w rho a / (i k a - 1),
    scattering against the Mie series of the rigid sphere, and four
    frequencies against the FP64 sweep."""
    a, c, rho = 0.5, 343.0, 1.21
    tri = meshes.icosphere(4, a)
    n = len(tri)
    d = meshes.unit_direction(2026)
    S = bem.Solver(tri, meshes.robin_bc(n, 0, 1, 1), leaf=64, eta=2.0)
    perm, _ = S.layout()
    freqs = 100.0 * 20.0 ** (np.arange(64) / 63.0)
    g = np.stack([np.ones(n, complex), np.zeros(n, complex)])
    planes = np.array([[0, 0, 0, 0, 0], [d[0], d[1], d[2], 1.0, 0.0]])
    S.set_sweep_batch(4)
    S.set_autotune(True)
    S.set_refinement(1)
    pr, its, st = S.sweep(freqs, g=g, planes=planes, c=c, rho=rho, aca_tol=1e-5, gmres_tol=1e-8, seed=1,
                          precision=bem.FP32, lanes=2)
    assert np.all(its > 0) and all(s["residual"] < 1e-7 for s in st)
    assert S.aca_schedule(bem.FP32) is not None
    print("tuned ACA schedule", S.aca_schedule(bem.FP32))
    S.set_sweep_batch(0)
    S.set_refinement(0)
    col = tri.reshape(n, 3, 3).mean(axis=1)[perm]          # centroids in cluster order
    cos_t = (col @ d) / np.linalg.norm(col, axis=1)
    worst_p, worst_s = 0.0, 0.0
    for fi, f in enumerate(freqs):
        k = 2 * np.pi * f / c
        pex = 1j * 2 * np.pi * f * rho * a / (1j * k * a - 1)
        worst_p = max(worst_p, np.linalg.norm(pr[fi, 0] - pex) / (abs(pex) * np.sqrt(n)))
        phi = pr[fi, 1] / (1j * 2 * np.pi * f * rho)
        mie = mie_rigid_surface_phi(k, a, cos_t)
        worst_s = max(worst_s, np.linalg.norm(phi - mie) / np.linalg.norm(mie))
    print("64-frequency sweep: worst pulsating error", worst_p, "worst scattering error", worst_s)
    assert worst_p < 0.02 and worst_s < 0.05
    sel = [0, 21, 42, 63]
    pr64, _, _ = S.sweep(freqs[sel], g=g, planes=planes, c=c, rho=rho, aca_tol=1e-5, gmres_tol=1e-8, seed=1,
                         precision=bem.FP64, lanes=2)
    for j, fi in enumerate(sel):
        for k in range(2):
            rel = np.linalg.norm(pr[fi, k] - pr64[j, k]) / np.linalg.norm(pr64[j, k])
            assert rel < 1e-5, (freqs[fi], k, rel)
