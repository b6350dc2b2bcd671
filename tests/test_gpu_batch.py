"""Batched assembly (bem_assemble_batch): the operators of several
frequencies from one ACA launch per size class over the flattened
(frequency, block) list equal the per-frequency assemblies bit for bit, in
FP64 and FP32, with one workgroup per item and with persistent workgroups on
a work queue, and under another size-class schedule."""
import numpy as np
import pytest

from helpers import sphere_problem, upload_problem

pytestmark = pytest.mark.gpu

FREQS = (300.0, 500.0, 820.0)


def same_blocks(G, H, nb):
    for b in range(nb):
        A, B = G.block(b), H.block(b)
        if isinstance(A, tuple):
            if not isinstance(B, tuple) or len(A[1]) != len(B[1]):
                return False, b
            for x, y in zip(A, B):
                if not np.array_equal(np.ascontiguousarray(x).view(np.uint64), np.ascontiguousarray(y).view(np.uint64)):
                    return False, b
        elif not np.array_equal(np.ascontiguousarray(A).view(np.uint64), np.ascontiguousarray(B).view(np.uint64)):
            return False, b
    return True, -1


@pytest.mark.parametrize("level,leaf", [(3, 32), (4, 64)])
def test_batch_equals_single(bem, O, meshes, level, leaf):
    P = sphere_problem(O, meshes, level=level, leaf=leaf, eta=2.0, freq=FREQS[0])
    ctx = upload_problem(bem, P)
    nb = len(P.blocks)
    kcc = []
    for f in FREQS:
        P.set_frequency(f, 343.0, 1)
        prm = P.params()
        kcc.append([prm["k"], prm["cc"].real, prm["cc"].imag])
    kcc = np.array(kcc)
    for prec in (bem.FP64, bem.FP32):
        for kind in (0, 1):
            single = []
            for f, (k, cr, ci) in enumerate(kcc):
                ctx.set_frequency(k, complex(cr, ci))
                single.append(ctx.assemble(kind, 1e-6, 0, precision=prec, seed=3))
            for queue in (0, 1):
                ctx.set_aca_schedule(prec, 384, 2048 if prec == bem.FP32 else 1 << 40, queue)
                print("batch: level", level, "prec", prec, "kind", kind, "queue", queue, flush=True)
                batch = ctx.assemble_batch(kind, kcc, 1e-6, precision=prec, seed=3)
                for f in range(len(FREQS)):
                    ok, b = same_blocks(batch[f], single[f], nb)
                    assert ok, (prec, kind, queue, f, b)
                    assert batch[f].stats()["evaluated"] == single[f].stats()["evaluated"]
                for G in batch:
                    G.close()
            for G in single:
                G.close()
    # another schedule (every block in the one-wavefront class): the same FP64 blocks
    ctx.set_aca_schedule(bem.FP64, 1 << 40, 1 << 40, 1)
    batch = ctx.assemble_batch(1, kcc[:1], 1e-6, seed=3)
    ctx.set_frequency(kcc[0, 0], complex(kcc[0, 1], kcc[0, 2]))
    ctx.set_aca_schedule(bem.FP64, 384, 1 << 40, 0)
    ref = ctx.assemble(1, 1e-6, 0, seed=3)
    ok, b = same_blocks(batch[0], ref, nb)
    assert ok, b
    ctx.close()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("prec", [0, 1])
def test_gmres_batch_equals_per_operator(bem, O, meshes, prec):
    """Lockstep GMRES over the DPhi of three frequencies with two right-hand
    sides each (six columns, bem_gmres_batch) against bem_gmres_multi on each
    operator alone: the same solutions to 1e-10 in as many columns (+-1).
    FP32: with the FP64 near field kept and refinement, every column reaches
    |b - A_res x| < 1e-12 |b| against its own operator."""
    P = sphere_problem(O, meshes, level=3, leaf=32, eta=2.0, freq=FREQS[0])
    ctx = upload_problem(bem, P)
    ctx.set_near_fp64(prec == bem.FP32)
    kcc = []
    for f in FREQS:
        P.set_frequency(f, 343.0, 1)
        prm = P.params()
        kcc.append([prm["k"], prm["cc"].real, prm["cc"].imag])
    Hs = ctx.assemble_batch(1, np.array(kcc), 1e-6, precision=prec, seed=3)
    n = P.n
    rng = np.random.default_rng(8)
    B = rng.standard_normal((n, 6)) + 1j * rng.standard_normal((n, 6))
    steps = 10 if prec == bem.FP32 else 0
    X, its, rr, st, rc = ctx.gmres_batch(Hs, B, 1e-12, max_steps=steps)
    assert rc == 0 and np.all(rr < 1e-12), rr
    for o, H in enumerate(Hs):
        cols = slice(2 * o, 2 * o + 2)
        X1, its1, rr1, st1, rc1 = H.gmres_multi(B[:, cols], 1e-12, max_steps=steps)
        assert rc1 == 0
        for r in range(2):
            assert rel(X[:, 2 * o + r], X1[:, r]) < 1e-10, (o, r, rel(X[:, 2 * o + r], X1[:, r]))
            if steps:
                assert rel(H.mvm_refine(X[:, 2 * o + r]), B[:, 2 * o + r]) < 2e-12
            else:
                assert abs(int(its[2 * o + r]) - int(its1[r])) <= 1, (o, r, its, its1)
    for H in Hs:
        H.close()
    ctx.close()


def test_batched_sweep(bem, O, meshes):
    """Three frequencies, two right-hand sides (pulsating g = 1; rigid
    scattering of a plane wave, g = 0) in one batch (bem_solver_set_sweep_batch):
    FP64 against the restatement's single solves, 1e-10; FP32 with
    refinement mode 1 against the lanes' FP32 sweep with the same mode, 1e-8."""
    tri = meshes.icosphere(3, 0.5)
    n = len(tri)
    d = meshes.unit_direction(3)
    S = bem.Solver(tri, meshes.robin_bc(n, 0, 1, 1), leaf=32, eta=2.0)
    freqs = np.array([300.0, 500.0, 800.0])
    g = np.stack([np.ones(n, complex), np.zeros(n, complex)])
    planes = np.array([[0, 0, 0, 0, 0], [d[0], d[1], d[2], 1.0, 0.0]])
    S.set_sweep_batch(4)
    pr, its, st = S.sweep(freqs, g=g, planes=planes, aca_tol=1e-6, gmres_tol=1e-10, seed=1)
    assert pr.shape == (3, 2, n) and np.all(its > 0)
    assert all(x["evaluated_phi"] > 0 and x["evaluated_dphi"] > 0 and x["max_rank_phi"] > 0 for x in st)
    Pp = O.Problem(tri, meshes.robin_bc(n, 0, 1, 1), leaf=32, eta=2.0)
    Ps = O.Problem(tri, meshes.robin_bc(n, 0, 1, 0), leaf=32, eta=2.0)
    for fi, f in enumerate(freqs):
        for k, Pr in enumerate((Pp, Ps)):
            Pr.set_frequency(f, 343.0, 1)
            if k == 1:
                Pr.set_sources(plane=np.array([d[0], d[1], d[2], 1.0, 0.0]))
            _, _, ref, _ = Pr.solve(1e-6, gmres_tol=1e-10, seed=1, rho=1.21, freq=f)
            assert rel(pr[fi, k], ref) < 1e-10, (f, k, rel(pr[fi, k], ref))
    S.set_refinement(1)
    pb, _, _ = S.sweep(freqs, g=g, planes=planes, aca_tol=1e-6, gmres_tol=1e-10, seed=1, precision=bem.FP32)
    S.set_sweep_batch(0)
    pl, _, _ = S.sweep(freqs, g=g, planes=planes, aca_tol=1e-6, gmres_tol=1e-10, seed=1, precision=bem.FP32, lanes=2)
    for fi in range(3):
        for k in range(2):
            assert rel(pb[fi, k], pl[fi, k]) < 1e-8, (fi, k, rel(pb[fi, k], pl[fi, k]))
    S.close()


@pytest.mark.parametrize("prec,near64", [(0, False), (1, False), (1, True)])
def test_batch_pairing(bem, O, meshes, prec, near64):
    """Near-field pairing across batches: a batched Phi assembly also forms
    the DPhi near fields, and the next batched DPhi assembly of the same
    frequency list takes them (no near-field tile launch of its own); both
    batches equal the unpaired single assemblies bit for bit."""
    import ctypes
    P = sphere_problem(O, meshes, level=3, leaf=32, eta=2.0, freq=FREQS[0])
    ctx = upload_problem(bem, P)
    ctx.set_near_fp64(near64)
    nb = len(P.blocks)
    kcc = []
    for f in FREQS:
        P.set_frequency(f, 343.0, 1)
        prm = P.params()
        kcc.append([prm["k"], prm["cc"].real, prm["cc"].imag])
    kcc = np.array(kcc)
    single = {}
    for f, (k, cr, ci) in enumerate(kcc):
        ctx.set_frequency(k, complex(cr, ci))
        for kind in (0, 1):
            single[f, kind] = ctx.assemble(kind, 1e-6, 0, precision=prec, seed=3)
    ctx.set_near_pairing(True)
    phi = ctx.assemble_batch(0, kcc, 1e-6, precision=prec, seed=3)
    prof = np.zeros(14)
    bem.lib().bem_profile(ctx.h, 1, prof.ctypes.data_as(ctypes.c_void_p))
    dphi = ctx.assemble_batch(1, kcc, 1e-6, precision=prec, seed=3)
    bem.lib().bem_profile(ctx.h, 0, prof.ctypes.data_as(ctypes.c_void_p))
    assert prof[7] == 0, prof            # no near-field tile launch in the DPhi batch
    for f in range(len(FREQS)):
        for kind, G in ((0, phi[f]), (1, dphi[f])):
            ok, b = same_blocks(G, single[f, kind], nb)
            assert ok, (f, kind, b)
        if near64:
            x = np.exp(0.3j * np.arange(P.n))
            assert np.array_equal(dphi[f].mvm_refine(x), single[f, 1].mvm_refine(x))
    for G in phi + dphi + list(single.values()):
        G.close()
    ctx.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_fused_product_equals_operator_loop(bem, O, meshes, prec):
    """bem_gmres_batch with every operator in one pass per kernel (fused)
    against one operator after another: bit for bit with one or two
    right-hand sides per operator (the same per-entry arithmetic), to 1e-12
    with four (the loop takes the matrix-core kernels there); FP32 with the
    refinement form too."""
    P = sphere_problem(O, meshes, level=3, leaf=32, eta=2.0, freq=FREQS[0])
    ctx = upload_problem(bem, P)
    ctx.set_near_fp64(prec == bem.FP32)
    kcc = []
    for f in FREQS:
        P.set_frequency(f, 343.0, 1)
        prm = P.params()
        kcc.append([prm["k"], prm["cc"].real, prm["cc"].imag])
    Hs = ctx.assemble_batch(1, np.array(kcc), 1e-6, precision=prec, seed=3)
    n = P.n
    rng = np.random.default_rng(9)
    steps = 10 if prec == bem.FP32 else 0
    for nops, nrhs in ((3, 1), (3, 2), (2, 4)):
        B = rng.standard_normal((n, nops * nrhs)) + 1j * rng.standard_normal((n, nops * nrhs))
        out = {}
        for fuse in (True, False):
            ctx.set_fuse_ops(fuse)
            print("fused product: prec", prec, "nops", nops, "nrhs", nrhs, "fuse", fuse, flush=True)
            out[fuse] = ctx.gmres_batch(Hs[:nops], B, 1e-11, max_steps=steps)
            assert out[fuse][4] == 0
        if nrhs < 4:
            assert np.array_equal(out[True][0].view(np.uint64), out[False][0].view(np.uint64)), (nops, nrhs)
            assert np.array_equal(out[True][1], out[False][1])
        else:
            assert rel(out[True][0], out[False][0]) < 1e-9, rel(out[True][0], out[False][0])
    ctx.set_fuse_ops(True)
    for H in Hs:
        H.close()
    ctx.close()


def test_autotune_keeps_results(bem, O, meshes):
    """bem_autotune_aca picks one of its candidate schedules and keeps it on
    the context; assemblies under it equal the built-in schedule's bit for
    bit.  A tuned batched sweep reports its schedule and gives the untuned
    sweep's pressures bit for bit, and so does a sweep with a schedule set
    by the caller."""
    P = sphere_problem(O, meshes, level=3, leaf=32, eta=2.0, freq=FREQS[0])
    ctx = upload_problem(bem, P)
    nb = len(P.blocks)
    kcc = []
    for f in FREQS:
        P.set_frequency(f, 343.0, 1)
        prm = P.params()
        kcc.append([prm["k"], prm["cc"].real, prm["cc"].imag])
    kcc = np.array(kcc)
    ref = ctx.assemble_batch(1, kcc, 1e-6, precision=bem.FP32, seed=3)
    (small, huge, queue), ms = ctx.autotune_aca(1, kcc, 1e-6, precision=bem.FP32, seed=3)
    print("tuned schedule", small, huge, queue, ms, flush=True)
    assert small in (128, 256, 384, 512) and huge in (1024, 2048, 4096, 1 << 40) and queue in (0, 1) and ms > 0
    tuned = ctx.assemble_batch(1, kcc, 1e-6, precision=bem.FP32, seed=3)
    for f in range(len(FREQS)):
        ok, b = same_blocks(tuned[f], ref[f], nb)
        assert ok, (f, b)
    for G in ref + tuned:
        G.close()
    ctx.close()
    tri = meshes.icosphere(3, 0.5)
    n = len(tri)
    S = bem.Solver(tri, meshes.robin_bc(n, 0, 1, 1), leaf=32, eta=2.0)
    S.set_sweep_batch(4)
    p0, _, _ = S.sweep(np.array(FREQS), aca_tol=1e-6, gmres_tol=1e-10, precision=bem.FP32)
    assert S.aca_schedule(bem.FP32) is None
    S.set_autotune(True)
    p1, _, _ = S.sweep(np.array(FREQS), aca_tol=1e-6, gmres_tol=1e-10, precision=bem.FP32)
    assert S.aca_schedule(bem.FP32) is not None
    assert np.array_equal(p0.view(np.uint64), p1.view(np.uint64))
    # a known schedule (bem_solver_set_aca_schedule): no tuning run, the same pressures
    S.set_autotune(False)
    S.set_aca_schedule(bem.FP32, 256, 1024, 1)
    p2, _, _ = S.sweep(np.array(FREQS), aca_tol=1e-6, gmres_tol=1e-10, precision=bem.FP32)
    sch = S.aca_schedule(bem.FP32)
    assert (sch["small_bound"], sch["huge_bound"], sch["queue"]) == (256, 1024, 1)
    assert np.array_equal(p0.view(np.uint64), p2.view(np.uint64))
    S.close()


def test_batched_sweep_impedance_ellipsoid(bem, O, meshes):
    """Batched sweep on a non-spherical body with mixed boundary conditions:
    ellipsoid (0.5, 0.3, 0.2) m, 1280 triangles, Z = rho c (1.5 + 0.5i) on a
    seeded 30 % of the elements (a != 0 there), rigid elsewhere; two
    frequencies, two seeded plane waves in one batch.  FP64 against the
    restatement's single solves, 1e-10; FP32 with refinement mode 1 against
    the lanes' FP32 sweep, 1e-8, and within 1e-6 of FP64 (mode 1 refines
    against FP64 near-field blocks and FP32 low-rank factors, so it converges
    to that operator's solution, about 1e-7 from the FP64 one)."""
    tri = meshes.ellipsoid(3, (0.5, 0.3, 0.2))
    n = len(tri)
    bc = meshes.impedance_bc(n, 2 * np.pi * 600.0 / 343.0, 0.3, 9)
    assert 0.25 < np.mean(bc[:, 0] ** 2 + bc[:, 1] ** 2 > 0) < 0.35
    d1, d2 = meshes.unit_direction(5), meshes.unit_direction(6)
    planes = np.array([[d1[0], d1[1], d1[2], 1.0, 0.0], [d2[0], d2[1], d2[2], 0.5, -0.25]])
    g = np.zeros((2, n), complex)
    freqs = np.array([450.0, 750.0])
    S = bem.Solver(tri, bc, leaf=32, eta=2.0)
    S.set_sweep_batch(2)
    pr, its, st = S.sweep(freqs, g=g, planes=planes, aca_tol=1e-6, gmres_tol=1e-10, seed=1)
    assert pr.shape == (2, 2, n) and np.all(its > 0)
    P = O.Problem(tri, bc, leaf=32, eta=2.0)
    for fi, f in enumerate(freqs):
        P.set_frequency(f, 343.0, 1)
        for k in range(2):
            P.set_sources(plane=planes[k])
            _, _, ref, _ = P.solve(1e-6, gmres_tol=1e-10, seed=1, rho=1.21, freq=f)
            assert rel(pr[fi, k], ref) < 1e-10, (f, k, rel(pr[fi, k], ref))
    S.set_refinement(1)
    pb, _, _ = S.sweep(freqs, g=g, planes=planes, aca_tol=1e-6, gmres_tol=1e-10, seed=1, precision=bem.FP32)
    S.set_sweep_batch(0)
    pl, _, _ = S.sweep(freqs, g=g, planes=planes, aca_tol=1e-6, gmres_tol=1e-10, seed=1, precision=bem.FP32, lanes=2)
    for fi in range(2):
        for k in range(2):
            assert rel(pb[fi, k], pl[fi, k]) < 1e-8, (fi, k, rel(pb[fi, k], pl[fi, k]))
            assert rel(pb[fi, k], pr[fi, k]) < 1e-6, (fi, k, rel(pb[fi, k], pr[fi, k]))
    S.close()
