"""ctypes binding of the CPU restatement (liboracle.so).  Test infrastructure
only: imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg as the checker, never by the product."""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_p = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        sig = {
            "orc_new": (_p, [_i64, _p, _p, _i64, _dbl]),
            "orc_free": (None, [_p]),
            "orc_n": (_i64, [_p]),
            "orc_nblocks": (_i64, [_p]),
            "orc_get_perm": (None, [_p, _p]),
            "orc_get_blocks": (None, [_p, _p]),
            "orc_get_geometry": (None, [_p, _p, _p, _p, _p, _p]),
            "orc_nclusters": (_i64, [_p]),
            "orc_get_clusters": (None, [_p, _p]),
            "orc_set_frequency": (ctypes.c_int, [_p, _dbl, _dbl, ctypes.c_int]),
            "orc_set_sources": (None, [_p, _i64, _p, ctypes.c_int, _p]),
            "orc_get_rhs_parts": (None, [_p, _p, _p, _p, _p, _p]),
            "orc_get_params": (None, [_p, _p]),
            "orc_entries4": (None, [_p, _i64, _p, _p]),
            "orc_entries": (None, [_p, ctypes.c_int, _i64, _p, _p]),
            "orc_entries4_polar": (None, [_p, _i64, _p, ctypes.c_int, _p, _p, _p]),
            "orc_dense_matrix": (None, [_p, ctypes.c_int, _p]),
            "orc_assemble": (_p, [_p, ctypes.c_int, _dbl, _i64, ctypes.c_uint64, ctypes.c_int]),
            "orc_hmat_free": (None, [_p]),
            "orc_hmat_rank": (_i64, [_p, _i64]),
            "orc_hmat_block": (None, [_p, _i64, _p, _p, _p]),
            "orc_hmvm": (None, [_p, _p, _p]),
            "orc_pivot_list": (_i64, [_p, _i64, ctypes.c_uint64, _p, _i64]),
            "orc_normal_groups": (_i64, [_p, _i64, _i64, _p]),
            "orc_hmat_evaluated": (_i64, [_p]),
            "orc_gmres": (_i64, [_p, _p, _p, _dbl, _i64, _i64, _p]),
            "orc_solve": (ctypes.c_int, [_p, _dbl, _i64, _dbl, _i64, ctypes.c_uint64, _dbl, _dbl, ctypes.c_int,
                                         _p, _p, _p, _p]),
            "orc_dense_solve": (ctypes.c_int, [_p, _dbl, _dbl, _p, _p, _p]),
            "orc_field": (None, [_p, ctypes.c_int64, _p, _p, _p, _p]),
            "orc_field_partition": (ctypes.c_int64, [_p, ctypes.c_int64, _p, ctypes.c_int64, ctypes.c_double, _p,
                                                     _p, ctypes.c_int64]),
            "orc_field_hmat": (None, [_p, ctypes.c_int64, _p, _p, _p, ctypes.c_double, ctypes.c_int64,
                                      ctypes.c_uint64, ctypes.c_int64, ctypes.c_double, _p, _p]),
            "orc_set_plane": (None, [_p, _p, _p]),
            "orc_tables": (None, [_p, _p, _p]),
            "orc_sincos": (None, [_i64, _p, _p, _p]),
            "orc_project": (ctypes.c_int, [_p, _p, _p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class Problem:
    """Cluster-ordered BEM problem (mesh, tree, partition) in the oracle."""

    def __init__(self, tri, bc, leaf=50, eta=3.0):
        self._tri = _c(tri, np.float64)
        self._bc = _c(bc, np.float64)
        self.h = lib().orc_new(len(self._tri), _ptr(self._tri), _ptr(self._bc), int(leaf), float(eta))
        if not self.h:
            raise ValueError("empty mesh")
        self.n = len(self._tri)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    @property
    def perm(self):
        out = np.empty(self.n, dtype=np.int64)
        lib().orc_get_perm(self.h, _ptr(out))
        return out

    @property
    def blocks(self):
        nb = lib().orc_nblocks(self.h)
        out = np.empty((nb, 5), dtype=np.int64)
        lib().orc_get_blocks(self.h, _ptr(out))
        return out

    @property
    def clusters(self):
        nc = lib().orc_nclusters(self.h)
        out = np.empty((nc, 10), dtype=np.float64)
        lib().orc_get_clusters(self.h, _ptr(out))
        return out

    def geometry(self):
        n = self.n
        nodes = np.empty((n, 9)); nrm = np.empty((n, 3)); col = np.empty((n, 3)); area = np.empty(n)
        hb = np.empty(1)
        lib().orc_get_geometry(self.h, _ptr(nodes), _ptr(nrm), _ptr(col), _ptr(area), _ptr(hb))
        return nodes, nrm, col, area, float(hb[0])

    def set_frequency(self, freq, c=343.0, coupling=1):
        lib().orc_set_frequency(self.h, float(freq), float(c), int(coupling))

    def set_sources(self, points=None, plane=None):
        pts = _c(points if points is not None else np.zeros((0, 6)), np.float64)
        pl = _c(plane if plane is not None else np.zeros(5), np.float64)
        lib().orc_set_sources(self.h, len(pts), _ptr(pts), 1 if plane is not None else 0, _ptr(pl))

    def set_plane(self, point=None, normal=None):
        """one impedance plane (rigid image) through point with normal; None removes it"""
        if normal is None:
            lib().orc_set_plane(self.h, None, None)
            return
        pt, nr = _c(point, np.float64), _c(normal, np.float64)
        lib().orc_set_plane(self.h, _ptr(pt), _ptr(nr))

    def rhs_parts(self):
        n = self.n
        s = np.empty(n, np.complex128); a = np.empty(n, np.complex128); b = np.empty(n, np.complex128)
        f = np.empty(n, np.complex128); sub = np.empty(n, np.int32)
        lib().orc_get_rhs_parts(self.h, _ptr(s), _ptr(a), _ptr(b), _ptr(f), _ptr(sub))
        return s, a, b, f, sub

    def params(self):
        o = np.empty(4)
        lib().orc_get_params(self.h, _ptr(o))
        return {"k": o[0], "cc": complex(o[1], o[2]), "hbar": o[3]}

    def entries4(self, ij):
        ij = _c(ij, np.int64).reshape(-1, 2)
        out = np.empty((len(ij), 4), np.complex128)
        lib().orc_entries4(self.h, len(ij), _ptr(ij), _ptr(out))
        return out

    def entries4_polar(self, ij, x, w):
        ij = _c(ij, np.int64).reshape(-1, 2)
        x, w = _c(x, np.float64), _c(w, np.float64)
        out = np.empty((len(ij), 4), np.complex128)
        lib().orc_entries4_polar(self.h, len(ij), _ptr(ij), len(x), _ptr(x), _ptr(w), _ptr(out))
        return out

    def entries(self, kind, ij):
        ij = _c(ij, np.int64).reshape(-1, 2)
        out = np.empty(len(ij), np.complex128)
        lib().orc_entries(self.h, int(kind), len(ij), _ptr(ij), _ptr(out))
        return out

    def dense(self, kind):
        out = np.empty((self.n, self.n), np.complex128, order="F")
        lib().orc_dense_matrix(self.h, int(kind), _ptr(out))
        return out

    def assemble(self, kind, tol, maxrank=0, seed=1, nthreads=0):
        return HMat(self, lib().orc_assemble(self.h, int(kind), float(tol), int(maxrank), int(seed), int(nthreads)))

    def pivot_list(self, b, seed):
        cap = 4096
        out = np.empty((cap, 2), np.int64)
        c = lib().orc_pivot_list(self.h, int(b), int(seed), _ptr(out), cap)
        return out[:c].copy()

    def normal_groups(self, start, count):
        """group index of each element start .. start + count - 1 (cluster order)"""
        gid = np.empty(int(count), np.int32)
        ng = lib().orc_normal_groups(self.h, int(start), int(count), _ptr(gid))
        return gid, int(ng)

    def solve(self, acatol, maxrank=0, gmres_tol=None, restart=300, seed=1, rho=1.21, freq=None, nthreads=0):
        n = self.n
        phi = np.empty(n, np.complex128); dphi = np.empty(n, np.complex128); pr = np.empty(n, np.complex128)
        st = np.zeros(16)
        gt = 1e-3 * acatol if gmres_tol is None else gmres_tol
        lib().orc_solve(self.h, float(acatol), int(maxrank), float(gt), int(restart), int(seed), float(rho),
                        float(freq), int(nthreads), _ptr(phi), _ptr(dphi), _ptr(pr), _ptr(st))
        keys = ["t_phi_asm", "t_rhs_mvm", "t_dphi_asm", "t_gmres", "t_total", "gmres_iters", "residual",
                "max_rank_phi", "max_rank_dphi", "evaluated_phi", "evaluated_dphi", "near_blocks", "far_blocks",
                "gmres_est"]
        return phi, dphi, pr, dict(zip(keys, st[:len(keys)].tolist()))

    def field(self, obs, phi, dphi):
        """boundary term of the field potential at observation points (nobs, 3)"""
        obs = _c(obs, np.float64).reshape(-1, 3)
        phi, dphi = _c(phi, np.complex128), _c(dphi, np.complex128)
        out = np.empty(len(obs), np.complex128)
        lib().orc_field(self.h, len(obs), _ptr(obs), _ptr(phi), _ptr(dphi), _ptr(out))
        return out

    def field_hmat(self, obs, phi, dphi, tol, maxrank=0, seed=1, leaf=50, eta=3.0):
        """boundary term through the observation x boundary H-matrix (input
        order of obs); returns (values, (admissible blocks, max rank, evaluated))"""
        obs = _c(obs, np.float64).reshape(-1, 3)
        phi, dphi = _c(phi, np.complex128), _c(dphi, np.complex128)
        out = np.empty(len(obs), np.complex128)
        st = np.zeros(3)
        lib().orc_field_hmat(self.h, len(obs), _ptr(obs), _ptr(phi), _ptr(dphi), float(tol), int(maxrank), int(seed),
                             int(leaf), float(eta), _ptr(out), _ptr(st))
        return out, tuple(int(v) for v in st)

    def field_partition(self, obs, leaf=50, eta=3.0):
        """(perm, blocks (nb, 5)) of the points' tree and the observation x
        boundary partition"""
        obs = _c(obs, np.float64).reshape(-1, 3)
        perm = np.empty(len(obs), np.int64)
        cap = 1 << 16
        b5 = np.empty((cap, 5), np.int64)
        nb = lib().orc_field_partition(self.h, len(obs), _ptr(obs), int(leaf), float(eta), _ptr(perm), _ptr(b5), cap)
        if nb > cap:
            raise RuntimeError("too many field blocks")
        return perm, b5[:nb].copy()

    def dense_solve(self, rho, freq):
        n = self.n
        phi = np.empty(n, np.complex128); dphi = np.empty(n, np.complex128); pr = np.empty(n, np.complex128)
        lib().orc_dense_solve(self.h, float(rho), float(freq), _ptr(phi), _ptr(dphi), _ptr(pr))
        return phi, dphi, pr


class HMat:
    def __init__(self, prob, h):
        self.prob = prob
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_hmat_free(self.h)
            self.h = None

    def rank(self, b):
        return lib().orc_hmat_rank(self.h, int(b))

    @property
    def evaluated(self):
        return lib().orc_hmat_evaluated(self.h)

    def block(self, b):
        row0, m, col0, n, adm = self.prob.blocks[b]
        k = self.rank(b)
        if k < 0:
            F = np.empty((m, n), np.complex128, order="F")
            lib().orc_hmat_block(self.h, int(b), _ptr(F), None, None)
            return F
        U = np.empty((m, k), np.complex128, order="F")
        al = np.empty(k, np.complex128)
        V = np.empty((k, n), np.complex128)
        lib().orc_hmat_block(self.h, int(b), _ptr(U), _ptr(al), _ptr(V))
        return U, al, V

    def mvm(self, x, y=None):
        x = _c(x, np.complex128)
        y = np.zeros_like(x) if y is None else _c(y, np.complex128).copy()
        lib().orc_hmvm(self.h, _ptr(x), _ptr(y))
        return y

    def gmres(self, b, tol, restart=300, maxit=100000, x0=None):
        b = _c(b, np.complex128)
        x = np.zeros_like(b) if x0 is None else _c(x0, np.complex128).copy()
        est = np.zeros(1)
        its = lib().orc_gmres(self.h, _ptr(b), _ptr(x), float(tol), int(restart), int(maxit), _ptr(est))
        return x, its, float(est[0])

    def to_full(self):
        n = self.prob.n
        A = np.zeros((n, n), np.complex128)
        for b, (r0, m, c0, nn, adm) in enumerate(self.prob.blocks):
            blk = self.block(b)
            if isinstance(blk, tuple):
                U, al, V = blk
                A[r0:r0 + m, c0:c0 + nn] = (U * al) @ V
            else:
                A[r0:r0 + m, c0:c0 + nn] = blk
        return A


def tables():
    tri = np.empty((16, 3)); gl5 = np.empty((5, 2)); half = np.empty((16, 2))
    lib().orc_tables(_ptr(tri), _ptr(gl5), _ptr(half))
    return tri, gl5, half


def sincos(x):
    x = _c(x, np.float64)
    s = np.empty_like(x); c = np.empty_like(x)
    lib().orc_sincos(len(x), _ptr(x), _ptr(s), _ptr(c))
    return s, c


def project(tri9, p):
    t = _c(tri9, np.float64); pp = _c(p, np.float64); out = np.empty(3)
    r = lib().orc_project(_ptr(t), _ptr(pp), _ptr(out))
    return r, out


def refine_triangles(tri, edge_length):
    """Restatement of ElementSection::refineTriangles / recursiveRefineTriangle
    (SolvingScript/elementSection.cpp:688-800) with lineLength (:1001-1005),
    middleOfLine (:993-999) and global::indexOfMaxOfThree (global.cpp:593-617),
    in plain Python floats (the same IEEE double operations): every triangle
    whose longest edge is above edge_length is split at that edge's midpoint,
    depth first."""
    import math

    def length(p, q):
        dx, dy, dz = p[0] - q[0], p[1] - q[1], p[2] - q[2]
        return math.sqrt(dx * dx + dy * dy + dz * dz)

    def imax(a, b, c):
        if a > b:
            return 1 if a > c else 3
        return 2 if b > c else 3

    def mid(p, q):
        return ((p[0] + q[0]) / 2, (p[1] + q[1]) / 2, (p[2] + q[2]) / 2)

    out = []

    def rec(n1, n2, n3, longest):
        e = {1: (n1, n2), 2: (n2, n3), 3: (n3, n1)}[longest]
        if length(*e) <= edge_length:
            out.append(n1 + n2 + n3)
            return
        m = mid(*e)
        if longest == 1:
            kids = ((n1, m, n3), (m, n2, n3))
        elif longest == 2:
            kids = ((n1, n2, m), (m, n3, n1))
        else:
            kids = ((n1, n2, m), (m, n2, n3))
        for a, b, c in kids:
            rec(a, b, c, imax(length(a, b), length(b, c), length(c, a)))

    for t in np.asarray(tri, dtype=np.float64).reshape(-1, 9).tolist():
        n1, n2, n3 = tuple(t[0:3]), tuple(t[3:6]), tuple(t[6:9])
        rec(n1, n2, n3, imax(length(n1, n2), length(n2, n3), length(n3, n1)))
    return np.array(out, dtype=np.float64).reshape(-1, 9)
