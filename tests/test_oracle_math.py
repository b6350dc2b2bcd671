"""Pins of the CPU restatement (the oracle) against mathematics that does not
depend on the oracle's own formulas: extended-precision sin/cos, polynomial
exactness of the quadrature rules, brute-force nearest points, independent
numpy evaluations of the Helmholtz kernels, closed-form polar identities of
the self integrals, and the Laplace limit."""
import numpy as np
import pytest


def test_det_sincos_ulp(O):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(0, 300, 200000), rng.uniform(-2, 2, 50000), rng.uniform(-300, 0, 50000),
                        (np.arange(-400, 400) * (np.pi / 2)) + rng.uniform(-1e-6, 1e-6, 800)])
    s, c = O.sincos(x)
    xl = x.astype(np.longdouble)
    rs, rc = np.sin(xl), np.cos(xl)
    us = np.abs(s - rs) / np.spacing(np.abs(rs).astype(np.float64))
    uc = np.abs(c - rc) / np.spacing(np.abs(rc).astype(np.float64))
    assert us.max() < 2.0 and uc.max() < 2.0, (us.max(), uc.max())


def _tri_moment(a, b):
    """exact integral of x1^a x2^b over the reference triangle, area 1/2"""
    from math import factorial
    return factorial(a) * factorial(b) / factorial(a + b + 2)


def test_triangle_rule_exactness(O):
    tri, gl5, half = O.tables()
    x1, x2, w = tri[:, 0], tri[:, 1], tri[:, 2]
    assert abs(w.sum() - 1.0) < 1e-12
    # rule weights sum to 1 over a triangle of area 1/2: integral = w . f / 2
    for deg in range(9):
        for a in range(deg + 1):
            b = deg - a
            approx = 0.5 * np.sum(w * x1 ** a * x2 ** b)
            assert abs(approx - _tri_moment(a, b)) < 2e-13, (a, b)
    # degree 9 is not integrated exactly (order 8 rule)
    errs = [abs(0.5 * np.sum(w * x1 ** a * x2 ** (9 - a)) - _tri_moment(a, 9 - a)) for a in range(10)]
    assert max(errs) > 1e-9


def test_line_rules(O):
    tri, gl5, half = O.tables()
    x, w = np.polynomial.legendre.leggauss(5)
    assert np.allclose(gl5[:, 0], x, atol=1e-15, rtol=0) and np.allclose(gl5[:, 1], w, atol=1e-15, rtol=0)
    x32, w32 = np.polynomial.legendre.leggauss(32)
    assert np.allclose(half[:, 0], x32[16:], atol=1e-15, rtol=0)
    assert np.allclose(half[:, 1], w32[16:], atol=1e-15, rtol=0)


def _brute_nearest(tri9, p, m=400):
    a, b, c = tri9[0:3], tri9[3:6], tri9[6:9]
    u, v = np.meshgrid(np.linspace(0, 1, m + 1), np.linspace(0, 1, m + 1))
    keep = u + v <= 1
    u, v = u[keep], v[keep]
    pts = a + np.outer(u, b - a) + np.outer(v, c - a)
    d = np.linalg.norm(pts - p, axis=1)
    k = np.argmin(d)
    return pts[k], 1 - u[k] - v[k], u[k], v[k]


def test_projection_regions(O):
    rng = np.random.default_rng(3)
    for trial in range(60):
        tri9 = rng.standard_normal(9)
        a, b, c = tri9[0:3], tri9[3:6], tri9[6:9]
        p = (a + b + c) / 3 + 1.5 * rng.standard_normal(3)
        region, out = O.project(tri9, p)
        near, w1, w2, w3 = _brute_nearest(tri9, p)
        w = np.array([w1, w2, w3])
        big = w > 0.05 + 0.01
        small = w < 0.05 - 0.01
        if not (big | small).all():
            continue  # within the snapping band: classification ambiguous at this sampling
        if big.all():
            assert region == 0
            assert np.linalg.norm(out - near) < 5e-3 * np.linalg.norm(b - a)
        elif big.sum() == 2:
            assert region in (4, 5, 6)
        else:
            assert region in (1, 2, 3)
            assert np.allclose(out, [a, b, c][region - 1])


# ---- independent numpy evaluation of the four operators -------------------

TRI16 = None


def _ops_numpy(p, npv, nq, qpts, w, k):
    """sum_q w_q (G, dG/dn_q, dG/dn_p, d2G/dn_p dn_q) with G = e^{ikr}/(4 pi r),
    r = |p - q|, derivatives by complex arithmetic written from the formulas."""
    R = p - qpts
    r = np.linalg.norm(R, axis=1)
    G = np.exp(1j * k * r) / (4 * np.pi * r)
    dGdr = G * (1j * k - 1 / r)
    rnq = -(R @ nq) / r
    rnp = (R @ npv) / r
    d2Gdr2 = G * (-k * k - 2j * k / r + 2 / r ** 2)
    N = dGdr * (-(npv @ nq + rnp * rnq) / r) + d2Gdr2 * rnp * rnq
    return np.array([np.sum(w * G), np.sum(w * dGdr * rnq), np.sum(w * dGdr * rnp), np.sum(w * N)])


@pytest.fixture(scope="module")
def L3(O, meshes):
    tri = meshes.icosphere(3, 0.5)
    P = O.Problem(tri, meshes.robin_bc(len(tri), 0, 1, 1), leaf=32, eta=2.0)
    P.set_frequency(800.0, 343.0, 1)
    return P


def test_regular_entries_match_numpy_formula(O, L3):
    """Same 16 quadrature points, kernel formulas evaluated independently."""
    tri16, _, _ = O.tables()
    nodes, nrm, col, area, hbar = L3.geometry()
    k = L3.params()["k"]
    rng = np.random.default_rng(5)
    ij = rng.integers(0, L3.n, size=(300, 2))
    d = np.linalg.norm(col[ij[:, 0]] - col[ij[:, 1]], axis=1)
    ij = ij[d > 4 * hbar]
    got = L3.entries4(ij)
    for t, (i, j) in enumerate(ij):
        a, b, c = nodes[j, 0:3], nodes[j, 3:6], nodes[j, 6:9]
        x1, x2, w = tri16[:, 0], tri16[:, 1], tri16[:, 2]
        q = np.outer(x1, a) + np.outer(x2, b) + np.outer(1 - x1 - x2, c)
        ref = area[j] * _ops_numpy(col[i], nrm[i], nrm[j], q, w, k)
        assert np.allclose(got[t], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), (i, j)


def _fine_rule(tri9, levels):
    """midpoint subdivision of a triangle into 4^levels pieces, 16-point rule on
    each (an independent high-order composite quadrature)"""
    import oracle
    tri16, _, _ = oracle.tables()
    tris = [tri9.reshape(3, 3)]
    for _ in range(levels):
        nt = []
        for t in tris:
            a, b, c = t
            ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
            nt += [np.array([a, ab, ca]), np.array([ab, b, bc]), np.array([ca, bc, c]), np.array([ab, bc, ca])]
        tris = nt
    pts, wts = [], []
    for a, b, c in tris:
        x1, x2, w = tri16[:, 0], tri16[:, 1], tri16[:, 2]
        pts.append(np.outer(x1, a) + np.outer(x2, b) + np.outer(1 - x1 - x2, c))
        wts.append(w * 0.5 * np.linalg.norm(np.cross(b - a, c - a)))
    return np.concatenate(pts), np.concatenate(wts)


def test_near_singular_polar_decomposition(O, L3):
    """The near-singular branch splits the source triangle at the projection
    of the target and integrates in polar coordinates (5 x 5 Gauss-Legendre in
    the reference).  With the 5-point rule the forced polar evaluation is the
    production entry bit for bit; with 48 points the same decomposition
    (projection, sub-triangles, angles, rho_max, in-plane axes) must converge
    to an independent 4^5-piece composite rule of the same integrals, which a
    wrong index, sign or angle would break.  The 5-point scheme itself is
    coarse (the reference's accuracy), checked only loosely."""
    nodes, nrm, col, area, hbar = L3.geometry()
    k = L3.params()["k"]
    d = np.linalg.norm(col[:, None] - col[None], axis=2)
    i, j = np.nonzero((d < 4 * hbar) & (d > 0))
    sel = np.random.default_rng(2).choice(len(i), 20, replace=False)
    ij = np.stack([i[sel], j[sel]], 1)
    prod = L3.entries4(ij)
    _, gl5, _ = O.tables()   # the reference's printed 5-point rule
    x5, w5 = gl5[:, 0], gl5[:, 1]
    assert np.array_equal(L3.entries4_polar(ij, x5, w5).view(np.uint64), prod.view(np.uint64))
    x48, w48 = np.polynomial.legendre.leggauss(48)
    hi = L3.entries4_polar(ij, x48, w48)
    for t, (a, b) in enumerate(ij):
        q, w = _fine_rule(nodes[b], 5)
        ref = _ops_numpy(col[a], nrm[a], nrm[b], q, w, k)
        assert np.all(np.abs(hi[t] - ref) <= 1e-7 * np.abs(ref) + 1e-12), (a, b, hi[t], ref)
        assert abs(prod[t, 0] - ref[0]) < 0.2 * abs(ref[0])


def _polar_radius(tri9, p, theta, e1, e2):
    """distance from p (in the plane) to the triangle boundary along direction theta"""
    v = [tri9[0:3], tri9[3:6], tri9[6:9]]
    dirv = np.cos(theta)[:, None] * e1 + np.sin(theta)[:, None] * e2
    best = np.full(len(theta), np.inf)
    for a, b in ((v[0], v[1]), (v[1], v[2]), (v[2], v[0])):
        e = b - a
        # solve p + t dir = a + s e in the plane (least squares in 3D)
        for idx in range(len(theta)):
            M = np.stack([dirv[idx], -e], 1)
            sol, *_ = np.linalg.lstsq(M, a - p, rcond=None)
            t, s = sol
            if t > 0 and -1e-12 <= s <= 1 + 1e-12:
                best[idx] = min(best[idx], t)
    return best


def test_self_terms_closed_form(O, L3):
    """Self entries: L = (i/2k)(1 - (1/2pi) int e^{ikR(th)} dth) and
    N = ik/2 - (1/4pi) int e^{ikR}/R dth, with R(theta) from an independent
    ray-triangle intersection and a 4000-point angular rule."""
    nodes, nrm, col, area, hbar = L3.geometry()
    k = L3.params()["k"]
    for e in (0, 17, 400, 1279):
        tri9 = nodes[e]
        a, b = tri9[0:3], tri9[3:6]
        e1 = (b - a) / np.linalg.norm(b - a)
        e2 = np.cross(nrm[e], e1)
        x, w = np.polynomial.legendre.leggauss(200)
        # integrate piecewise between the vertex angles for accuracy
        ang = []
        for v in (tri9[0:3], tri9[3:6], tri9[6:9]):
            r = v - col[e]
            ang.append(np.arctan2(r @ e2, r @ e1) % (2 * np.pi))
        ang = np.sort(np.array(ang))
        edges = np.concatenate([ang, [ang[0] + 2 * np.pi]])
        IL, IN = 0j, 0j
        for lo, hi in zip(edges[:-1], edges[1:]):
            th = 0.5 * (hi - lo) * x + 0.5 * (hi + lo)
            R = _polar_radius(tri9, col[e], th, e1, e2)
            IL += 0.5 * (hi - lo) * np.sum(w * np.exp(1j * k * R))
            IN += 0.5 * (hi - lo) * np.sum(w * np.exp(1j * k * R) / R)
        Lref = (1j / (2 * k)) * (1 - IL / (2 * np.pi))
        Nref = 1j * k / 2 - IN / (4 * np.pi)
        got = L3.entries4([[e, e]])[0]
        assert abs(got[0] - Lref) < 1e-9 * abs(Lref), (got[0], Lref)
        assert abs(got[3] - Nref) < 1e-9 * abs(Nref), (got[3], Nref)
        assert got[1] == 0 and got[2] == 0


def _laplace_triangle_inplane(tri9, p):
    """int_T 1/|x - p| dA for p in the plane of T (edge-sum closed form)"""
    v = [tri9[0:3], tri9[3:6], tri9[6:9]]
    nrm = np.cross(v[1] - v[0], v[2] - v[0])
    nrm /= np.linalg.norm(nrm)
    tot = 0.0
    for a, b in ((v[0], v[1]), (v[1], v[2]), (v[2], v[0])):
        e = b - a
        L = np.linalg.norm(e)
        t = e / L
        m = np.cross(t, nrm)           # outward in-plane normal of the edge (CCW)
        d = (a - p) @ m                # signed distance of p to the edge line
        s1, s2 = (a - p) @ t, (b - p) @ t
        r1, r2 = np.linalg.norm(a - p), np.linalg.norm(b - p)
        tot += d * np.log((s2 + r2) / (s1 + r1))
    return tot


def test_self_single_layer_laplace_limit(O, meshes):
    tri = meshes.icosphere(1, 1.0)
    P = O.Problem(tri, meshes.robin_bc(len(tri), 0, 1, 1), leaf=50, eta=3.0)
    P.set_frequency(1e-3, 343.0, 1)   # k = 1.8e-5
    nodes, nrm, col, area, hbar = P.geometry()
    k = P.params()["k"]
    for e in range(0, len(tri), 7):
        I0 = _laplace_triangle_inplane(nodes[e], col[e])
        # L(k) = I0/(4 pi) + i k A/(4 pi) + O(k^2)
        ref = I0 / (4 * np.pi) + 1j * k * area[e] / (4 * np.pi)
        got = P.entries4([[e, e]])[0, 0]
        assert abs(got.real - ref.real) < 1e-9 * abs(ref.real)
        # Im L = (1 - Re S / 2 pi) / 2k loses ~10 digits to cancellation at this k
        assert abs(got.imag - ref.imag) < 1e-4 * abs(ref.imag)


def test_bc_substitution_formulas(O, meshes):
    """implicitPhiMatrix2 / implicitDPhiMatrix2 from the four operators, for
    both substitution branches (impedance |a| > |b| on some elements)."""
    tri = meshes.icosphere(2, 0.5)
    n = len(tri)
    rng = np.random.default_rng(4)
    a = np.where(rng.random(n) < 0.5, 3.0 + 1j, 0.2j)
    b = np.where(rng.random(n) < 0.5, 1.0, 0.5 - 0.25j)
    P = O.Problem(tri, meshes.robin_bc(n, a, b, 1.0), leaf=32, eta=2.0)
    P.set_frequency(700.0, 343.0, 1)
    s, al, be, f, sub = P.rhs_parts()
    cc = P.params()["cc"]
    assert np.array_equal(sub, (np.abs(al) < np.abs(be)).astype(np.int32))
    ij = np.concatenate([rng.integers(0, n, size=(200, 2)), np.stack([np.arange(20)] * 2, 1)])
    o4 = P.entries4(ij)
    phi, dphi = P.entries(0, ij), P.entries(1, ij)
    for t, (i, j) in enumerate(ij):
        L, M, Mt, N = o4[t]
        dlt = 1.0 if i == j else 0.0
        X = L + cc * Mt + dlt * cc / 2
        Y = M + cc * N - dlt * 0.5
        if sub[j]:
            ph, dp = -X / be[j], -al[j] / be[j] * X - Y
        else:
            ph, dp = Y / al[j], X + be[j] / al[j] * Y
        assert abs(phi[t] - ph) <= 1e-13 * abs(ph) + 1e-15
        assert abs(dphi[t] - dp) <= 1e-13 * abs(dp) + 1e-15
