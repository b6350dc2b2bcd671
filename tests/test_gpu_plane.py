"""The impedance plane on the device (bem_set_plane,
bem_solver_set_impedance_plane): entries, H-matrix blocks and the field bit
for bit equal to the CPU restatement, solves to round-off, the FP32 form
against FP64, and the method of images (the mesh over the plane = the mesh
and its mirror image in free space) through the whole FP32 pipeline."""
import numpy as np
import pytest

from helpers import near_pairs, upload_problem

pytestmark = pytest.mark.gpu

P0 = np.array([0.3, -0.2, 0.1])
NRM = np.array([0.3, 0.4, 0.866])          # not unit on purpose
UNIT = NRM / np.linalg.norm(NRM)


def same_bits(a, b):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128))
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def mirror(tri):
    v = tri.reshape(-1, 3, 3)
    h = (v - P0) @ UNIT
    return (v - 2.0 * h[..., None] * UNIT)[:, ::-1, :].reshape(-1, 9)


def scene(meshes, level, f, seed=11):
    """sphere r 0.4 centred 0.9 m from the plane; 30 % impedance, the rest
    driven with a seeded normal velocity"""
    tri = meshes.icosphere(level, 0.4, P0 + 0.9 * UNIT + np.array([0.1, -0.05, 0.0]))
    n = len(tri)
    bc = meshes.impedance_bc(n, 2 * np.pi * f / 343.0, 0.3, seed)
    rigid = bc[:, 0] == 0
    bc[rigid, 4] = np.random.default_rng(seed).standard_normal(rigid.sum())
    return tri, bc


@pytest.fixture(scope="module")
def L3(bem, O, meshes):
    tri, bc = scene(meshes, 3, 500.0)
    P = O.Problem(tri, bc, leaf=32, eta=2.0)
    P.set_frequency(500.0, 343.0, 1)
    P.set_plane(P0, NRM)
    ctx = upload_problem(bem, P)
    ctx.set_plane(P0, NRM)
    return tri, bc, P, ctx


def test_plane_entries_bit_exact(L3):
    _, _, P, ctx = L3
    n = P.n
    rng = np.random.default_rng(0)
    ij = np.concatenate([rng.integers(0, n, size=(4000, 2)), near_pairs(P), np.stack([np.arange(n)] * 2, 1)])
    for kind in (0, 1):
        ref, got = P.entries(kind, ij), ctx.entries(kind, ij)
        assert same_bits(ref, got), (kind, np.abs(ref - got).max())
    # the free-space operators are untouched by the plane
    assert same_bits(P.entries4(ij), ctx.entries4(ij))


@pytest.mark.parametrize("kind", [0, 1])
def test_plane_assembly_bit_exact(L3, kind):
    _, _, P, ctx = L3
    H = P.assemble(kind, 1e-6, 0, 7)
    G = ctx.assemble(kind, 1e-6, 0, seed=7)
    for b in range(len(P.blocks)):
        rb, rg = H.rank(b), G.rank(b)
        assert rb == rg, (b, rb, rg)
        A, B = H.block(b), G.block(b)
        if rb < 0:
            assert same_bits(A, B), b
        else:
            for x, y in zip(A, B):
                assert same_bits(x, y), b


def test_plane_field_bit_exact(L3):
    _, _, P, ctx = L3
    n = P.n
    rng = np.random.default_rng(3)
    phi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dphi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    obs = P0 + rng.uniform(-3, 3, (400, 3))
    obs = obs[(obs - P0) @ UNIT > 0.05]
    ref = P.field(obs, phi, dphi)
    got = ctx.field(obs, ctx.array(phi), ctx.array(dphi))
    assert same_bits(ref, got), np.abs(ref - got).max()


def test_plane_solver_parity_and_fp32(bem, O, L3):
    """FP64 solve with the plane equals the restatement's (1e-10); the FP32
    form agrees with it at the ACA tolerance; a sweep on two lanes (lanes
    made after the plane was set, and a lane updated when it changes) repeats
    the single solves"""
    tri, bc, P, _ = L3
    S = bem.Solver(tri, bc, leaf=32, eta=2.0)
    S.set_impedance_plane(P0, NRM)
    _, _, p64, st = S.solve(500.0, c=343.0, rho=1.21, aca_tol=1e-6, gmres_tol=1e-10, seed=1)
    _, _, pro, _ = P.solve(1e-6, gmres_tol=1e-10, seed=1, rho=1.21, freq=500.0)
    rel = np.linalg.norm(p64 - pro) / np.linalg.norm(pro)
    print("plane FP64 vs restatement", rel, "iters", st["gmres_iters"])
    assert rel < 1e-10 and st["residual"] < 1e-9
    _, _, p32, st32 = S.solve(500.0, c=343.0, rho=1.21, aca_tol=1e-6, gmres_tol=1e-10, seed=1, precision=bem.FP32)
    e32 = np.linalg.norm(p32 - p64) / np.linalg.norm(p64)
    print("plane FP32 vs FP64", e32)
    assert e32 < 1e-5 and st32["residual"] < 1e-9
    freqs = np.array([400.0, 500.0, 650.0])
    pr, _, _ = S.sweep(freqs, c=343.0, rho=1.21, aca_tol=1e-6, gmres_tol=1e-10, seed=1, lanes=2)
    assert np.linalg.norm(pr[1, 0] - p64) / np.linalg.norm(p64) < 1e-9
    S.set_impedance_plane(None)
    pr0, _, _ = S.sweep(freqs, c=343.0, rho=1.21, aca_tol=1e-6, gmres_tol=1e-10, seed=1, lanes=2)
    _, _, pfree, _ = S.solve(650.0, c=343.0, rho=1.21, aca_tol=1e-6, gmres_tol=1e-10, seed=1)
    assert np.linalg.norm(pr0[2, 0] - pfree) / np.linalg.norm(pfree) < 1e-9
    assert np.linalg.norm(pr0[1, 0] - p64) / np.linalg.norm(p64) > 1e-2
    S.close()


def same_blocks(G, H, nb):
    for b in range(nb):
        A, B = G.block(b), H.block(b)
        if isinstance(A, tuple):
            if not isinstance(B, tuple) or len(A) != len(B) or not all(same_bits(x, y) for x, y in zip(A, B)):
                return False, b
        elif not same_bits(A, B):
            return False, b
    return True, -1


def test_plane_batched_sweep(bem, O, L3):
    """Batched frequency chunks with the plane: two frequencies assembled in
    one launch per size class (every entry with its mirror image's
    quadrature) equal the one-by-one assemblies bit for bit in FP64 and FP32
    (a work-queue request is ignored with the plane); a batched FP64 sweep
    equals the restatement's single solves (1e-10), and the batched FP32
    sweep with refinement mode 1 the lanes' (1e-8)."""
    tri, bc, P, ctx = L3
    kcc = []
    for f in (400.0, 650.0):
        P.set_frequency(f, 343.0, 1)
        prm = P.params()
        kcc.append([prm["k"], prm["cc"].real, prm["cc"].imag])
    kcc = np.array(kcc)
    nb = len(P.blocks)
    for prec in (bem.FP64, bem.FP32):
        ctx.set_aca_schedule(prec, 384, 2048 if prec == bem.FP32 else 1 << 40, 1)
        for kind in (0, 1):
            single = []
            for k, cr, ci in kcc:
                ctx.set_frequency(k, complex(cr, ci))
                single.append(ctx.assemble(kind, 1e-6, 0, precision=prec, seed=3))
            batch = ctx.assemble_batch(kind, kcc, 1e-6, precision=prec, seed=3)
            for f in range(len(kcc)):
                ok, b = same_blocks(batch[f], single[f], nb)
                assert ok, (prec, kind, f, b)
            for G in batch + single:
                G.close()
        ctx.set_aca_schedule(prec, 384, 2048 if prec == bem.FP32 else 1 << 40, 0)
    P.set_frequency(500.0, 343.0, 1)
    prm = P.params()
    ctx.set_frequency(prm["k"], prm["cc"])
    S = bem.Solver(tri, bc, leaf=32, eta=2.0)
    S.set_impedance_plane(P0, NRM)
    freqs = np.array([400.0, 500.0, 650.0])
    S.set_sweep_batch(2)
    pb, its, _ = S.sweep(freqs, c=343.0, rho=1.21, aca_tol=1e-6, gmres_tol=1e-10, seed=1)
    assert pb.shape == (3, 1, len(tri)) and np.all(its > 0)
    for fi, f in enumerate(freqs):
        P.set_frequency(f, 343.0, 1)
        _, _, ref, _ = P.solve(1e-6, gmres_tol=1e-10, seed=1, rho=1.21, freq=f)
        e = np.linalg.norm(pb[fi, 0] - ref) / np.linalg.norm(ref)
        assert e < 1e-10, (f, e)
    P.set_frequency(500.0, 343.0, 1)
    S.set_refinement(1)
    p32b, _, _ = S.sweep(freqs, c=343.0, rho=1.21, aca_tol=1e-6, gmres_tol=1e-10, seed=1, precision=bem.FP32)
    S.set_sweep_batch(0)
    p32l, _, _ = S.sweep(freqs, c=343.0, rho=1.21, aca_tol=1e-6, gmres_tol=1e-10, seed=1, precision=bem.FP32,
                         lanes=2)
    for fi in range(3):
        e = np.linalg.norm(p32b[fi, 0] - p32l[fi, 0]) / np.linalg.norm(p32l[fi, 0])
        assert e < 1e-8, (fi, e)
    S.close()


def test_method_of_images_fp32(bem, meshes):
    """L4 sphere (5120 elements) over the plane vs the sphere and its mirror
    image (10240 elements) in free space, both FP32 H-matrix solves with ACA
    1e-5 (the doubled mesh's 5120 x 5120 sphere-to-image blocks do not reach
    1e-6 in FP32 within the rank cap): the pressures on the sphere agree to
    the ACA tolerance, and so do the fields above the plane"""
    f = 700.0
    tri, bc = scene(meshes, 4, f, seed=5)
    n = len(tri)
    S1 = bem.Solver(tri, bc, leaf=64, eta=2.0)
    S1.set_impedance_plane(P0, NRM)
    _, _, p1, st1 = S1.solve(f, c=343.0, rho=1.21, aca_tol=1e-5, gmres_tol=1e-9, seed=2, precision=bem.FP32)
    perm1, _ = S1.layout()
    S2 = bem.Solver(np.concatenate([tri, mirror(tri)]), np.concatenate([bc, bc]), leaf=64, eta=2.0)
    _, _, p2, st2 = S2.solve(f, c=343.0, rho=1.21, aca_tol=1e-5, gmres_tol=1e-9, seed=2, precision=bem.FP32)
    perm2, _ = S2.layout()
    a = np.empty(n, complex)
    a[perm1] = p1                       # input order
    b = np.empty(2 * n, complex)
    b[perm2] = p2
    e = np.linalg.norm(a - b[:n]) / np.linalg.norm(b[:n])
    sym = np.linalg.norm(b[n:] - b[:n]) / np.linalg.norm(b[:n])
    print("images: plane vs doubled mesh", e, "symmetry of the doubled solve", sym,
          "iters", st1["gmres_iters"], st2["gmres_iters"])
    assert e < 1e-4 and sym < 1e-4
    rng = np.random.default_rng(8)
    obs = P0 + rng.uniform(-3, 3, (300, 3))
    obs = obs[(obs - P0) @ UNIT > 0.05]
    fa, fb = S1.field(obs), S2.field(obs)
    ef = np.linalg.norm(fa - fb) / np.linalg.norm(fb)
    print("images: field", ef)
    assert ef < 1e-4
    S1.close()
    S2.close()
