"""The large mixed-precision cases: the 76440-triangle capsule with a seeded
30 % impedance patchwork, and 16 seeded ellipsoids of uneven size each at
its own frequency.  The capsule's excitation is a seeded plane wave (the
case names none; without one the solution is zero)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_capsule76k_impedance(bem, meshes):
    """Capsule L = 2 m, r = 0.25 m, edge 1.07 cm (76440 triangles), 1500 Hz,
    Z = rho c (1.5 + 0.5i) on a seeded 30 % of the elements, rigid elsewhere,
    leaf 64, eta 1.5, ACA 1e-4, leaf-block LU preconditioner.  The FP32
    operator alone agrees with the FP64 path at the ACA tolerance (different
    pivots); with FP64-residual refinement (mode 2: residuals with the FP64
    H-matrices, GMRES on the FP32 one) it reaches the FP64 solution to 1e-8.
    The reconstructed phi, dphi satisfy the impedance / rigid boundary
    conditions element by element."""
    f, c, rho = 1500.0, 343.0, 1.21
    k = 2 * np.pi * f / c
    tri = meshes.capsule(2.0, 0.25, 0.0107)
    n = len(tri)
    assert 70000 < n < 90000
    bc = meshes.impedance_bc(n, k, 0.3, 4)
    d = meshes.unit_direction(4)
    S = bem.Solver(tri, bc, leaf=64, eta=1.5)
    S.set_sources(plane=np.array([d[0], d[1], d[2], 1.0, 0.0]))
    S.set_preconditioner(1)
    perm, _ = S.layout()
    kw = dict(c=c, rho=rho, aca_tol=1e-4, seed=1)
    phi, dphi, p32, st32 = S.solve(f, gmres_tol=1e-8, precision=bem.FP32, **kw)
    _, _, p64, st64 = S.solve(f, gmres_tol=1e-10, precision=bem.FP64, **kw)
    S.set_refinement(2)
    _, _, pref, stref = S.solve(f, gmres_tol=1e-10, precision=bem.FP32, **kw)
    rs = S.refine_stats()
    S.set_refinement(0)
    print("capsule n", n, "FP32 vs FP64", rel(p32, p64), "refined FP32 vs FP64", rel(pref, p64), "iters",
          st32["gmres_iters"], st64["gmres_iters"], stref["gmres_iters"], "refinement", rs)
    assert st32["residual"] < 1e-7 and st64["residual"] < 1e-9
    assert rel(p32, p64) < 1e-4
    assert rs["residual"] < 1e-10 and rs["steps"] >= 2
    assert rel(pref, p64) < 1e-8
    a = (bc[:, 0] + 1j * bc[:, 1])[perm]                  # cluster order
    imp = a != 0
    assert 0.25 < imp.mean() < 0.35
    # a phi + dphi/dn = 0 on every element (b = 1, g = 0)
    assert np.abs(a * phi + dphi).max() < 1e-12 * np.abs(dphi).max()
    assert np.all(np.abs(dphi[~imp]) < 1e-12 * np.abs(dphi).max())


def test_sixteen_ellipsoids(bem, O, meshes):
    """16 ellipsoids, semi-axes uniform in [0.2, 1.0] m (seeded), 5120 or
    20480 triangles (250880 in all), each at a frequency uniform in
    [200, 2000] Hz with k max(axes) <= 20, normal velocity N(0, 1) per
    element, Burton-Miller, FP32, solved together (bem_solvers_solve_multi:
    Phi and DPhi of all 16 meshes from one multi-mesh ACA launch per size
    class each): every solve converges, each pressure equals the same
    solver's own FP32 solve bit for bit, FP32 agrees with FP64 on four of
    them, and the FP64 path equals the CPU restatement on a 5120-triangle one.
    The first joint solve includes the solvers' one-time set-up (device
    buffers), so it is timed a second time against the single solves."""
    import time
    rng = np.random.default_rng(5)
    c, rho = 343.0, 1.21
    levels = rng.permutation([5] * 11 + [4] * 5)
    solvers, freqs, bcs, tris = [], [], [], []
    for q, level in enumerate(levels):
        axes = rng.uniform(0.2, 1.0, 3)
        f = rng.uniform(200.0, 2000.0)
        f = min(f, 20.0 * c / (2 * np.pi * axes.max()))
        tri = meshes.ellipsoid(int(level), axes)
        bc = meshes.robin_bc(len(tri), 0.0, 1.0, rng.standard_normal(len(tri)))
        solvers.append(bem.Solver(tri, bc, leaf=64, eta=2.0))
        freqs.append(f)
        bcs.append(bc)
        tris.append(tri)
    assert sum(len(t) for t in tris) == 11 * 20480 + 5 * 5120
    p32, its, st = bem.solve_multi(solvers, freqs, c=c, rho=rho, aca_tol=1e-5, gmres_tol=1e-8, seed=7,
                                   precision=bem.FP32, threads=4)
    assert all(x["residual"] < 1e-7 for x in st), [x["residual"] for x in st]
    t0 = time.perf_counter()
    p32b, _, _ = bem.solve_multi(solvers, freqs, c=c, rho=rho, aca_tol=1e-5, gmres_tol=1e-8, seed=7,
                                 precision=bem.FP32, threads=4)
    t_multi = time.perf_counter() - t0
    assert all(np.array_equal(a.view(np.uint64), b.view(np.uint64)) for a, b in zip(p32, p32b))
    t0 = time.perf_counter()
    for q in range(16):
        _, _, ps, _ = solvers[q].solve(freqs[q], c=c, rho=rho, aca_tol=1e-5, gmres_tol=1e-8, seed=7,
                                       precision=bem.FP32)
        assert np.array_equal(ps.view(np.uint64), p32[q].view(np.uint64)), q
    t_single = time.perf_counter() - t0
    print("16 meshes solved together: %.3f s; one after another: %.3f s" % (t_multi, t_single), flush=True)
    checked = 0
    for q in range(0, 16, 4):
        S = solvers[q]
        _, _, p64, _ = S.solve(freqs[q], c=c, rho=rho, aca_tol=1e-5, gmres_tol=1e-8, seed=7, precision=bem.FP64)
        e = rel(p32[q], p64)
        print("ellipsoid", q, "n", len(tris[q]), "f", round(freqs[q], 1), "FP32 vs FP64", e, flush=True)
        assert e < 1e-4
        if levels[q] == 4 and checked == 0:
            P = O.Problem(tris[q], bcs[q], leaf=64, eta=2.0)
            P.set_frequency(freqs[q], c, 1)
            _, _, pc, _ = P.solve(1e-5, gmres_tol=1e-8, seed=7, rho=rho, freq=freqs[q])
            assert rel(p64, pc) < 1e-10
            checked += 1
    for S in solvers:
        S.close()


def test_capsule76k_hlu_preconditioned(bem, meshes):
    """The same capsule (FP32, ACA 1e-4, GMRES 1e-8) preconditioned by the
    H-LU with the reference's defaults (rank 2, no tolerance): factorised on
    the host from the GPU-assembled blocks, substituted on the device
    (k_hlu).  Same pressure as with the leaf-block LU, a fraction of
    its iterations."""
    f, c, rho = 1500.0, 343.0, 1.21
    k = 2 * np.pi * f / c
    tri = meshes.capsule(2.0, 0.25, 0.0107)
    n = len(tri)
    bc = meshes.impedance_bc(n, k, 0.3, 4)
    d = meshes.unit_direction(4)
    S = bem.Solver(tri, bc, leaf=64, eta=1.5)
    S.set_sources(plane=np.array([d[0], d[1], d[2], 1.0, 0.0]))
    kw = dict(c=c, rho=rho, aca_tol=1e-4, gmres_tol=1e-8, seed=1, precision=bem.FP32)
    S.set_preconditioner(1)
    _, _, p1, st1 = S.solve(f, **kw)
    S.set_preconditioner(2)
    _, _, p2, st2 = S.solve(f, **kw)
    ps = S.precond_stats()
    print("capsule H-LU: iterations", st2["gmres_iters"], "(leaf-block LU", st1["gmres_iters"], ") vs leaf-block LU",
          rel(p2, p1), "set-up ms", ps["setup_ms"], "factorisation ms", ps["factor_ms"], "applications",
          ps["applies"], "device ms", ps["apply_ms"])
    assert st1["residual"] < 1e-7 and st2["residual"] < 1e-7
    assert rel(p2, p1) < 1e-6
    assert st2["gmres_iters"] * 3 < st1["gmres_iters"]
    assert ps["applies"] >= st2["gmres_iters"] and ps["apply_ms"] > 0
    S.close()
