"""MI355X (gfx950) H-matrix BEM hot path: Python binding of the C ABI in
include/bem_hip.h (lib3dbem_hip.so next to this file).

There is no CPU fallback: every call goes to the HIP library, and importing
the binding raises if the library has not been built (``make -C 3d_bem_amd``
or ``__graft_entry__.build()``).  The package directory name starts with a
digit, so import it with ``importlib.import_module("3d_bem_amd")``.
"""
import ctypes
import os
import weakref

import numpy as np

from . import meshes  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib3dbem_hip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "bem_hip.h")

BEM_OK, BEM_ERR_ARG, BEM_ERR_HIP, BEM_ERR_STATE, BEM_ERR_NAN, BEM_ERR_NOCONV = 0, -1, -2, -3, -4, -5
PHI, DPHI = 0, 1
FP64, FP32 = 0, 1

_i64, _dbl, _p, _int, _u64 = ctypes.c_int64, ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64
_SIGS = {
    "bem_create": (_int, [ctypes.POINTER(_p), _int]),
    "bem_destroy": (_int, [_p]),
    "bem_last_error": (ctypes.c_char_p, [_p]),
    "bem_upload_mesh": (_int, [_p, _i64, _p, _p, _p, _p, _dbl]),
    "bem_set_bc": (_int, [_p, _p, _p, _p]),
    "bem_set_plane": (_int, [_p, _p, _p]),
    "bem_set_near_pairing": (_int, [_p, _int]),
    "bem_hlu_create": (_int, [_i64, _i64, _p, _i64, _p, _p, _p, _i64, _dbl, ctypes.POINTER(_p)]),
    "bem_hlu_solve": (_int, [_p, _p, _p]),
    "bem_hlu_solve_flat": (_int, [_p, _p, _p]),
    "bem_hlu_dense": (_int, [_p, _p]),
    "bem_hlu_stats": (_int, [_p, _p]),
    "bem_hlu_destroy": (_int, [_p]),
    "bem_hlu_error": (ctypes.c_char_p, []),
    "bem_solver_set_impedance_plane": (_int, [_p, _p, _p]),
    "bem_upload_partition": (_int, [_p, _i64, _p, _p, _p, _p, _p]),
    "bem_set_frequency": (_int, [_p, _dbl, _dbl, _dbl]),
    "bem_assemble": (_int, [_p, _int, _dbl, _i64, _int, _u64, ctypes.POINTER(_p)]),
    "bem_hmat_destroy": (_int, [_p]),
    "bem_assemble_batch": (_int, [_p, _int, _i64, _p, _dbl, _i64, _int, _u64, _p]),
    "bem_set_aca_schedule": (_int, [_p, _int, _i64, _i64, _int]),
    "bem_hmvm": (_int, [_p, _p, _p, _p, _int]),
    "bem_block_info": (_int, [_p, _i64, ctypes.POINTER(_i64)]),
    "bem_block_download": (_int, [_p, _i64, _p, _p, _p]),
    "bem_hmat_stats": (_int, [_p, _p]),
    "bem_gmres": (_int, [_p, _p, _p, _p, _dbl, _i64, _i64, ctypes.POINTER(_i64), ctypes.POINTER(_dbl)]),
    "bem_reconstruct": (_int, [_p, _p, _dbl, _dbl, _p, _p, _p]),
    "bem_reconstruct_g": (_int, [_p, _p, _p, _dbl, _dbl, _p, _p, _p]),
    "bem_field": (_int, [_p, _i64, _p, _p, _p, _p]),
    "bem_field_noalloc": (_int, [_p, _i64, _p, _p, _p, _p]),
    "bem_set_near_fp64": (_int, [_p, _int]),
    "bem_hmvm_refine": (_int, [_p, _p, _p, _p, _int]),
    "bem_gmres_ir": (_int, [_p, _p, _p, _p, _p, _p, _int, _dbl, _i64, _i64, _i64, ctypes.POINTER(_i64),
                            ctypes.POINTER(_dbl), ctypes.POINTER(_i64)]),
    "bem_solver_set_refinement": (_int, [_p, _int, _i64]),
    "bem_solver_set_sweep_batch": (_int, [_p, _int]),
    "bem_set_fuse_ops": (_int, [_p, _int]),
    "bem_set_precond_side": (_int, [_p, _int]),
    "bem_solver_set_precond_side": (_int, [_p, _int]),
    "bem_solver_set_aca_schedule": (_int, [_p, _int, _i64, _i64, _int]),
    "bem_solver_profile": (_int, [_p, _int, _p]),
    "bem_assemble_multi": (_int, [_p, _i64, _int, _dbl, _i64, _int, _u64, _p]),
    "bem_solvers_solve_multi": (_int, [_p, _i64, _p, _dbl, _dbl, _int, _dbl, _i64, _dbl, _i64, _u64, _int, _int, _p, _p,
                                       _p]),
    "bem_autotune_aca": (_int, [_p, _int, _int, _i64, _p, _dbl, _i64, _u64, _p, _p]),
    "bem_solver_set_autotune": (_int, [_p, _int]),
    "bem_solver_aca_schedule": (_int, [_p, _int, _p]),
    "bem_gmres_batch": (_int, [_p, _i64, _p, _p, _i64, _p, _p, _int, _dbl, _i64, _i64, _i64, _p, _p,
                               ctypes.POINTER(_i64)]),
    "bem_set_mfma": (_int, [_p, _int]),
    "bem_set_graphs": (_int, [_p, _int]),
    "bem_hmvm_multi": (_int, [_p, _p, _i64, _p, _p, _int, _int]),
    "bem_gmres_multi": (_int, [_p, _p, _p, _i64, _p, _p, _int, _dbl, _i64, _i64, _i64, _p, _p,
                               ctypes.POINTER(_i64)]),
    "bem_solver_solve_dense": (_int, [_p, _dbl, _dbl, _dbl, _int, _dbl, _i64, _p, _p, _p, _p]),
    "bem_solver_refine_stats": (_int, [_p, _p]),
    "bem_read_msh": (_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, _i64, _p, ctypes.POINTER(_i64)]),
    "bem_read_msh_error": (ctypes.c_char_p, []),
    "bem_refine_triangles": (_int, [_i64, _p, _dbl, _i64, _p, ctypes.POINTER(_i64)]),
    "bem_solver_field": (_int, [_p, _i64, _p, _p]),
    "bem_solver_field_hmat": (_int, [_p, _i64, _p, _dbl, _i64, _i64, _dbl, _u64, _p, _p]),
    "bem_field_hmat": (_int, [_p, _i64, _p, _i64, _p, _dbl, _i64, _u64, _p, _p, _p, _p]),
    "bem_solver_set_preconditioner": (_int, [_p, _int]),
    "bem_solver_set_hlu": (_int, [_p, _i64, _dbl]),
    "bem_solver_precond_stats": (_int, [_p, _p]),
    "bem_precond_create_hlu": (_int, [_p, _p, _i64, _p, _i64, _dbl, ctypes.POINTER(_p)]),
    "bem_precond_stats": (_int, [_p, _p]),
    "bem_precond_create": (_int, [_p, _p, ctypes.POINTER(_p)]),
    "bem_precond_destroy": (_int, [_p]),
    "bem_precond_apply": (_int, [_p, _p, _p, _p]),
    "bem_gmres_prec": (_int, [_p, _p, _p, _p, _p, _dbl, _i64, _i64, ctypes.POINTER(_i64), ctypes.POINTER(_dbl)]),
    "bem_entries4": (_int, [_p, _i64, _p, _p]),
    "bem_entries": (_int, [_p, _int, _i64, _p, _p]),
    "bem_malloc": (_int, [_p, ctypes.c_size_t, ctypes.POINTER(_p)]),
    "bem_free": (_int, [_p, _p]),
    "bem_memcpy_htod": (_int, [_p, _p, _p, ctypes.c_size_t]),
    "bem_memcpy_dtoh": (_int, [_p, _p, _p, ctypes.c_size_t]),
    "bem_synchronize": (_int, [_p]),
    "bem_profile": (_int, [_p, _int, _p]),
    "bem_solver_create": (_int, [ctypes.POINTER(_p), _int, _i64, _p, _p, _i64, _dbl]),
    "bem_solver_destroy": (_int, [_p]),
    "bem_solver_last_error": (ctypes.c_char_p, [_p]),
    "bem_solver_context": (_p, [_p]),
    "bem_solver_layout": (_int, [_p, _p, ctypes.POINTER(_i64), _p, _i64]),
    "bem_solver_set_sources": (_int, [_p, _i64, _p, _int, _p]),
    "bem_solver_solve": (_int, [_p, _dbl, _dbl, _dbl, _int, _dbl, _i64, _dbl, _i64, _u64, _int, _p, _p, _p, _p]),
    "bem_solver_sweep": (_int, [_p, _i64, _p, _dbl, _dbl, _int, _dbl, _i64, _dbl, _i64, _u64, _int, _i64, _p, _p,
                                _int, _p, _p, _p]),
}
EXPORTS = tuple(_SIGS)

_LIB = None


class BemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


def lib():
    """The loaded HIP library; raises if it is missing (no fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C 3d_bem_amd` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class DeviceArray:
    """Complex128 vector in device memory owned by a Context."""

    def __init__(self, ctx, n):
        self.ctx, self.n = ctx, int(n)
        p = _p()
        ctx._check(lib().bem_malloc(ctx.h, 16 * self.n, ctypes.byref(p)))
        self.ptr = p

    def upload(self, x):
        x = _c(x, np.complex128)
        assert x.size == self.n
        self.ctx._check(lib().bem_memcpy_htod(self.ctx.h, self.ptr, _ptr(x), 16 * self.n))
        return self

    def download(self):
        out = np.empty(self.n, np.complex128)
        self.ctx._check(lib().bem_memcpy_dtoh(self.ctx.h, _ptr(out), self.ptr, 16 * self.n))
        return out

    def free(self):
        if self.ptr:
            lib().bem_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One HIP device + stream (bem_ctx_t)."""

    def __init__(self, device=0):
        h = _p()
        rc = lib().bem_create(ctypes.byref(h), int(device))
        if rc != BEM_OK:
            raise BemError(rc, "bem_create failed (no HIP device?)")
        self.h = h
        self.n = 0
        self.blocks = None
        self._children = weakref.WeakSet()   # H-matrices and preconditioners of this context

    def _check(self, rc):
        if rc != BEM_OK:
            raise BemError(rc, lib().bem_last_error(self.h).decode())
        return rc

    def close(self):
        """destroys the context; its H-matrices and preconditioners are closed first"""
        if getattr(self, "h", None):
            for child in list(getattr(self, "_children", ())):
                child.close()
            lib().bem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_mesh(self, nodes9, normal3, area, colloc3, hbar):
        nodes9, normal3 = _c(nodes9, np.float64), _c(normal3, np.float64)
        area, colloc3 = _c(area, np.float64), _c(colloc3, np.float64)
        self.n = len(area)
        self._check(lib().bem_upload_mesh(self.h, self.n, _ptr(nodes9), _ptr(normal3), _ptr(area), _ptr(colloc3),
                                          float(hbar)))

    def set_bc(self, alpha, beta, f):
        a, b, g = (_c(v, np.complex128) for v in (alpha, beta, f))
        self._check(lib().bem_set_bc(self.h, _ptr(a), _ptr(b), _ptr(g)))

    def set_near_pairing(self, on=True):
        """a Phi assembly also forms the DPhi near field for the next DPhi
        assembly (bem_set_near_pairing)"""
        self._check(lib().bem_set_near_pairing(self.h, int(bool(on))))

    def set_near_fp64(self, on=True):
        """later FP32 assemblies keep their near field in FP64 too (bem_set_near_fp64)"""
        self._check(lib().bem_set_near_fp64(self.h, int(bool(on))))

    def set_mfma(self, on=True):
        """matrix-core products for four or more right-hand sides (bem_set_mfma)"""
        self._check(lib().bem_set_mfma(self.h, int(bool(on))))

    def set_graphs(self, on=True):
        """graph-captured GMRES columns (bem_set_graphs)"""
        self._check(lib().bem_set_graphs(self.h, int(bool(on))))

    def set_plane(self, point=None, normal=None):
        """one impedance plane (rigid image) through point with normal; None removes it"""
        if normal is None:
            self._check(lib().bem_set_plane(self.h, None, None))
            return
        pt, nr = _c(point, np.float64), _c(normal, np.float64)
        self._check(lib().bem_set_plane(self.h, _ptr(pt), _ptr(nr)))

    def upload_partition(self, blocks5):
        B = _c(blocks5, np.int64)
        cols = [np.ascontiguousarray(B[:, q]) for q in range(4)]
        adm = np.ascontiguousarray(B[:, 4].astype(np.uint8))
        self._check(lib().bem_upload_partition(self.h, len(B), *[_ptr(c) for c in cols], _ptr(adm)))
        self.blocks = B.copy()

    def set_frequency(self, k, cc):
        self._check(lib().bem_set_frequency(self.h, float(k), float(np.real(cc)), float(np.imag(cc))))

    def assemble(self, kind, tol, maxrank=0, precision=FP64, seed=1):
        h = _p()
        self._check(lib().bem_assemble(self.h, int(kind), float(tol), int(maxrank), int(precision), int(seed),
                                       ctypes.byref(h)))
        return HMatrix(self, h)

    def assemble_batch(self, kind, kcc, tol, maxrank=0, precision=FP64, seed=1):
        """bem_assemble_batch: kcc (nf, 3) = (k, cc re, cc im) per frequency;
        one HMatrix per frequency"""
        kcc = _c(np.asarray(kcc).reshape(-1, 3), np.float64)
        outs = (_p * len(kcc))()
        self._check(lib().bem_assemble_batch(self.h, int(kind), len(kcc), _ptr(kcc), float(tol), int(maxrank),
                                             int(precision), int(seed), ctypes.cast(outs, _p)))
        return [HMatrix(self, _p(h)) for h in outs]

    def autotune_aca(self, kind, kcc, tol, maxrank=0, precision=FP64, seed=1):
        """bem_autotune_aca: (small bound, huge bound, queue), best ms; the
        schedule stays on the context"""
        kcc = _c(np.asarray(kcc).reshape(-1, 3), np.float64)
        ch = np.zeros(3, np.int64)
        ms = np.zeros(1)
        self._check(lib().bem_autotune_aca(self.h, int(kind), int(precision), len(kcc), _ptr(kcc), float(tol),
                                           int(maxrank), int(seed), _ptr(ch), _ptr(ms)))
        return tuple(int(v) for v in ch), float(ms[0])

    def field_hmat(self, obs, blocks5, phi, dphi, tol, maxrank=0, seed=1):
        """bem_field_hmat: boundary term at obs (nobs, 3, in the points'
        cluster order) through the partition blocks5 (nb, 5); phi / dphi
        device arrays; returns (values, stats4)"""
        obs = _c(obs, np.float64).reshape(-1, 3)
        b5 = _c(np.asarray(blocks5).reshape(-1, 5), np.int64)
        out = np.empty(len(obs), np.complex128)
        st = np.zeros(4)
        self._check(lib().bem_field_hmat(self.h, len(obs), _ptr(obs), len(b5), _ptr(b5), float(tol), int(maxrank),
                                         int(seed), phi.ptr, dphi.ptr, _ptr(out), _ptr(st)))
        return out, st

    def set_fuse_ops(self, on):
        """bem_set_fuse_ops: products of several operators in one pass (default on)"""
        self._check(lib().bem_set_fuse_ops(self.h, int(bool(on))))

    def set_precond_side(self, left):
        """bem_set_precond_side: preconditioned GMRES on M^-1 A (left, the
        reference's form) or on A M^-1 (right, the default)"""
        self._check(lib().bem_set_precond_side(self.h, int(bool(left))))

    def gmres_batch(self, hs, B, tol, hres=None, restart=300, maxit=100000, max_steps=0):
        """lockstep GMRES over several operators (bem_gmres_batch): B (n,
        nops * nrhs), column o * nrhs + r a right-hand side of hs[o]; X,
        iterations, relative residuals per column, refinement steps, status"""
        B = _c(B, np.complex128)
        n, q = B.shape
        nops = len(hs)
        if q % nops:
            raise ValueError("B needs nrhs columns per operator")
        hv = (_p * nops)(*[h.h for h in hs])
        rv = None if hres is None else (_p * nops)(*[h.h for h in hres])
        bd, xd = self.array(B.ravel(), n=n * q), self.array(n=n * q)
        its, rel, st = np.zeros(q, np.int64), np.zeros(q), _i64()
        rc = lib().bem_gmres_batch(self.h, nops, ctypes.cast(hv, _p), None if rv is None else ctypes.cast(rv, _p),
                                   q // nops, bd.ptr, xd.ptr, 1, float(tol), int(restart), int(maxit), int(max_steps),
                                   _ptr(its), _ptr(rel), ctypes.byref(st))
        if rc not in (BEM_OK, BEM_ERR_NOCONV):
            self._check(rc)
        return xd.download().reshape(n, q), its, rel, st.value, rc

    def set_aca_schedule(self, precision, small_bound, huge_bound, queue=-1):
        """ACA size-class bounds (m + n) for one precision and the batched
        launches' work queue (bem_set_aca_schedule)"""
        self._check(lib().bem_set_aca_schedule(self.h, int(precision), int(small_bound), int(huge_bound), int(queue)))

    def entries4(self, ij):
        ij = _c(ij, np.int64).reshape(-1, 2)
        out = np.empty((len(ij), 4), np.complex128)
        self._check(lib().bem_entries4(self.h, len(ij), _ptr(ij), _ptr(out)))
        return out

    def entries(self, kind, ij):
        ij = _c(ij, np.int64).reshape(-1, 2)
        out = np.empty(len(ij), np.complex128)
        self._check(lib().bem_entries(self.h, int(kind), len(ij), _ptr(ij), _ptr(out)))
        return out

    def array(self, x=None, n=None):
        d = DeviceArray(self, self.n if n is None else n)
        if x is not None:
            d.upload(x)
        return d

    def reconstruct(self, x_dev, rho, freq):
        phi, dphi, p = self.array(), self.array(), self.array()
        self._check(lib().bem_reconstruct(self.h, x_dev.ptr, float(rho), float(freq), phi.ptr, dphi.ptr, p.ptr))
        return phi.download(), dphi.download(), p.download()

    def field(self, obs, phi_dev, dphi_dev):
        """bem_field: boundary term of the field potential at obs (nobs, 3)"""
        obs = _c(obs, np.float64).reshape(-1, 3)
        out = np.empty(len(obs), np.complex128)
        self._check(lib().bem_field(self.h, len(obs), _ptr(obs), phi_dev.ptr, dphi_dev.ptr, _ptr(out)))
        return out

    def synchronize(self):
        self._check(lib().bem_synchronize(self.h))

    def profile(self, enable=-1):
        """bem_profile: average ms per kernel family [0..6], launches [7..13]"""
        out = np.zeros(14)
        self._check(lib().bem_profile(self.h, int(enable), _ptr(out)))
        return out


def assemble_multi(ctxs, kind, tol, maxrank=0, precision=FP64, seed=1):
    """bem_assemble_multi: one operator per context (each its own mesh and
    frequency) from one ACA launch per size class; HMatrix per context"""
    hv = (_p * len(ctxs))(*[c.h for c in ctxs])
    outs = (_p * len(ctxs))()
    rc = lib().bem_assemble_multi(ctypes.cast(hv, _p), len(ctxs), int(kind), float(tol), int(maxrank), int(precision),
                                  int(seed), ctypes.cast(outs, _p))
    if rc != BEM_OK:
        ctxs[0]._check(rc)
    return [HMatrix(c, _p(h)) for c, h in zip(ctxs, outs)]


def solve_multi(solvers, freqs, c=343.0, rho=1.21, coupling=1, aca_tol=1e-5, maxrank=0, gmres_tol=None, restart=300,
                seed=1, precision=FP64, threads=4):
    """bem_solvers_solve_multi: solver q at freqs[q], Phi and DPhi of all
    meshes from one multi-mesh assembly each; pressures (cluster order),
    GMRES iterations, stats per solver"""
    freqs = _c(freqs, np.float64)
    ns = len(solvers)
    hv = (_p * ns)(*[S.h for S in solvers])
    outs = [np.empty(S.n, np.complex128) for S in solvers]
    pv = (_p * ns)(*[o.ctypes.data for o in outs])
    its = np.zeros(ns, np.int64)
    st = np.zeros((ns, 16))
    gt = -1.0 if gmres_tol is None else float(gmres_tol)
    rc = lib().bem_solvers_solve_multi(ctypes.cast(hv, _p), ns, _ptr(freqs), float(c), float(rho), int(coupling),
                                       float(aca_tol), int(maxrank), gt, int(restart), int(seed), int(precision),
                                       int(threads), ctypes.cast(pv, _p), _ptr(its), _ptr(st))
    if rc != BEM_OK:
        raise BemError(rc, lib().bem_solver_last_error(solvers[0].h).decode())
    return outs, its, [dict(zip(STAT_KEYS, r.tolist())) for r in st]


class HMatrix:
    """Assembled H-matrix on the device (bem_hmat_t)."""

    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h
        ctx._children.add(self)

    def close(self):
        # a destroyed context has released the H-matrix's device memory with
        # its own (garbage collection may reach the two in either order)
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            lib().bem_hmat_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        s = np.zeros(9)
        lib().bem_hmat_stats(self.h, _ptr(s))
        keys = ["evaluated", "max_rank", "mean_rank", "bytes", "assembly_ms", "overflow_blocks", "dense_ms", "aca_ms",
                "resident_bytes"]
        return dict(zip(keys, s.tolist()))

    def rank(self, b):
        r = _i64()
        self.ctx._check(lib().bem_block_info(self.h, int(b), ctypes.byref(r)))
        return r.value

    def block(self, b):
        """Dense block (m x n) or (U, alpha, V) in the reference layout."""
        r0, m, c0, n, adm = self.ctx.blocks[b]
        k = self.rank(b)
        if k < 0:
            F = np.empty((m, n), np.complex128, order="F")
            self.ctx._check(lib().bem_block_download(self.h, int(b), _ptr(F), None, None))
            return F
        U = np.empty((m, k), np.complex128, order="F")
        al = np.empty(k, np.complex128)
        V = np.empty((k, n), np.complex128, order="F")
        self.ctx._check(lib().bem_block_download(self.h, int(b), _ptr(U), _ptr(al), _ptr(V)))
        return U, al, V

    def to_full(self):
        """Dense n x n matrix of the H-matrix (HMatrix::toFullMat, small n only)."""
        n = self.ctx.n
        F = np.zeros((n, n), np.complex128)
        for b, (r0, m, c0, nn, adm) in enumerate(self.ctx.blocks):
            blk = self.block(b)
            F[r0:r0 + m, c0:c0 + nn] = blk if not isinstance(blk, tuple) else (blk[0] * blk[1]) @ blk[2]
        return F

    def mvm(self, x, y=None, accumulate=False):
        """y (+)= H x with host arrays (copies through device buffers)."""
        xd = self.ctx.array(x)
        yd = self.ctx.array(np.zeros(self.ctx.n, np.complex128) if y is None else y)
        self.mvm_dev(xd, yd, accumulate)
        return yd.download()

    def mvm_dev(self, xd, yd, accumulate=False):
        self.ctx._check(lib().bem_hmvm(self.ctx.h, self.h, xd.ptr, yd.ptr, 1 if accumulate else 0))

    def mvm_multi(self, X, refine=False):
        """Y = H X for X (n, nrhs) (bem_hmvm_multi)"""
        X = _c(X, np.complex128)
        n, q = X.shape
        xd, yd = self.ctx.array(X.ravel(), n=n * q), self.ctx.array(n=n * q)
        self.ctx._check(lib().bem_hmvm_multi(self.ctx.h, self.h, q, xd.ptr, yd.ptr, 0, int(bool(refine))))
        return yd.download().reshape(n, q)

    def gmres_multi(self, B, tol, hres=None, restart=300, maxit=100000, max_steps=0):
        """lockstep GMRES over the columns of B (n, nrhs) (bem_gmres_multi):
        X, iterations, relative residuals, refinement steps, status"""
        B = _c(B, np.complex128)
        n, q = B.shape
        bd, xd = self.ctx.array(B.ravel(), n=n * q), self.ctx.array(n=n * q)
        its, rel, st = np.zeros(q, np.int64), np.zeros(q), _i64()
        rc = lib().bem_gmres_multi(self.ctx.h, self.h, None if hres is None else hres.h, q, bd.ptr, xd.ptr, 1,
                                   float(tol), int(restart), int(maxit), int(max_steps), _ptr(its), _ptr(rel),
                                   ctypes.byref(st))
        if rc not in (BEM_OK, BEM_ERR_NOCONV):
            self.ctx._check(rc)
        return xd.download().reshape(n, q), its, rel, st.value, rc

    def mvm_refine(self, x):
        """y = A_res x: the FP64 near field + stored low-rank factors (bem_hmvm_refine)"""
        xd, yd = self.ctx.array(x), self.ctx.array()
        self.ctx._check(lib().bem_hmvm_refine(self.ctx.h, self.h, xd.ptr, yd.ptr, 0))
        return yd.download()

    def gmres_ir(self, b, tol, hres=None, restart=300, maxit=100000, max_steps=10, precond=None):
        """iterative refinement around GMRES (bem_gmres_ir): x, inner iterations,
        |b - A_res x| / |b|, refinement steps, status"""
        bd, xd = self.ctx.array(b), self.ctx.array(np.zeros(self.ctx.n, np.complex128))
        its, rel, st = _i64(), _dbl(), _i64()
        rc = lib().bem_gmres_ir(self.ctx.h, self.h, None if hres is None else hres.h,
                                None if precond is None else precond.h, bd.ptr, xd.ptr, 1, float(tol), int(restart),
                                int(maxit), int(max_steps), ctypes.byref(its), ctypes.byref(rel), ctypes.byref(st))
        if rc not in (BEM_OK, BEM_ERR_NOCONV):
            self.ctx._check(rc)
        return xd.download(), its.value, rel.value, st.value, rc

    def preconditioner(self):
        """leaf-block LU preconditioner of this operator (bem_precond_create)"""
        p = _p()
        self.ctx._check(lib().bem_precond_create(self.ctx.h, self.h, ctypes.byref(p)))
        return Preconditioner(self.ctx, p)

    def hlu(self, clusters2, rank=2, rel_err=0.0):
        """host H-LU preconditioner of this operator (bem_precond_create_hlu);
        clusters2 (nc, 2) = (start, size) of every cluster"""
        cl = _c(np.asarray(clusters2)[:, :2], np.int64)
        p = _p()
        self.ctx._check(lib().bem_precond_create_hlu(self.ctx.h, self.h, len(cl), _ptr(cl), int(rank), float(rel_err),
                                                     ctypes.byref(p)))
        return Preconditioner(self.ctx, p)

    def gmres(self, b, tol, restart=300, maxit=100000, x0=None, precond=None):
        bd = self.ctx.array(b)
        xd = self.ctx.array(np.zeros(self.ctx.n, np.complex128) if x0 is None else x0)
        its, rel = _i64(), _dbl()
        rc = lib().bem_gmres_prec(self.ctx.h, self.h, None if precond is None else precond.h, bd.ptr, xd.ptr,
                                  float(tol), int(restart), int(maxit), ctypes.byref(its), ctypes.byref(rel))
        if rc not in (BEM_OK, BEM_ERR_NOCONV):
            self.ctx._check(rc)
        return xd.download(), its.value, rel.value, rc


class Preconditioner:
    """bem_prec_t: z = M^-1 v with M the diagonal leaf blocks"""

    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h
        ctx._children.add(self)

    def apply(self, v):
        vd, zd = self.ctx.array(v), self.ctx.array()
        self.ctx._check(lib().bem_precond_apply(self.ctx.h, self.h, vd.ptr, zd.ptr))
        return zd.download()

    def stats(self):
        o = np.zeros(5)
        lib().bem_precond_stats(self.h, _ptr(o))
        return {"setup_ms": o[0], "applies": int(o[1]), "apply_ms": o[2], "entries": o[3], "factor_ms": o[4]}

    def close(self):
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            lib().bem_precond_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HLU:
    """Host H-LU of an H-matrix given block by block (bem_hlu_create):
    clusters2 (nc, 2) = (start, size) of every cluster, blocks5 the
    partition, blocks[b] = dense (m, n) array or (U, alpha, V) with the block
    U diag(alpha) V; max_rank / rel_err the truncation (0 / 0: exact)."""

    def __init__(self, n, clusters2, blocks5, blocks, max_rank=2, rel_err=0.0):
        cl = _c(np.asarray(clusters2)[:, :2], np.int64)
        B = _c(blocks5, np.int64)
        rank = np.empty(len(B), np.int64)
        self._keep = []
        for b, blk in enumerate(blocks):
            if isinstance(blk, tuple):
                U, al, V = blk
                rank[b] = len(al)
                buf = np.concatenate([np.asarray(U, np.complex128).ravel(order="F"), np.asarray(al, np.complex128),
                                      np.asarray(V, np.complex128).ravel(order="F")])
            else:
                rank[b] = -1
                buf = np.asarray(blk, np.complex128).ravel(order="F")
            self._keep.append(np.ascontiguousarray(buf))
        ptrs = (ctypes.c_void_p * len(B))(*[a.ctypes.data for a in self._keep])
        h = _p()
        rc = lib().bem_hlu_create(int(n), len(cl), _ptr(cl), len(B), _ptr(B), _ptr(rank), ptrs, int(max_rank),
                                  float(rel_err), ctypes.byref(h))
        self._keep = None
        if rc != BEM_OK:
            raise BemError(rc, lib().bem_hlu_error().decode())
        self.h, self.n = h, int(n)

    def solve(self, b, flat=False):
        """x = U^-1 L^-1 b; flat: in the leaf-by-leaf order of the device
        preconditioner (bem_hlu_solve_flat)"""
        b = _c(b, np.complex128)
        x = np.empty_like(b)
        f = lib().bem_hlu_solve_flat if flat else lib().bem_hlu_solve
        rc = f(self.h, _ptr(b), _ptr(x))
        if rc != BEM_OK:
            raise BemError(rc, lib().bem_hlu_error().decode())
        return x

    def dense(self):
        """the product L U (n, n)"""
        out = np.empty((self.n, self.n), np.complex128, order="F")
        rc = lib().bem_hlu_dense(self.h, _ptr(out))
        if rc != BEM_OK:
            raise BemError(rc, lib().bem_hlu_error().decode())
        return out

    def stats(self):
        o = np.zeros(3)
        lib().bem_hlu_stats(self.h, _ptr(o))
        return {"entries": o[0], "max_rank": int(o[1]), "factor_s": o[2]}

    def close(self):
        if getattr(self, "h", None):
            lib().bem_hlu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


STAT_KEYS = ["t_phi_asm", "t_rhs_mvm", "t_dphi_asm", "t_gmres", "t_total", "gmres_iters", "residual",
             "max_rank_phi", "max_rank_dphi", "evaluated_phi", "evaluated_dphi", "near_blocks", "far_blocks",
             "precond_setup_ms", "precond_apply_ms", "dphi_bytes"]


def read_msh(path, include=None, exclude=None):
    """Triangles (n, 9) of a GMSH .msh file (bem_read_msh); include / exclude
    are lists of group ranges "a...b", group numbers or physical names."""
    enc = lambda v: None if not v else ",".join(str(x) for x in v).encode()
    cnt = _i64()
    rc = lib().bem_read_msh(str(path).encode(), enc(include), enc(exclude), 0, None, ctypes.byref(cnt))
    if rc != BEM_OK:
        raise BemError(rc, lib().bem_read_msh_error().decode())
    tri = np.empty((cnt.value, 9), np.float64)
    rc = lib().bem_read_msh(str(path).encode(), enc(include), enc(exclude), cnt.value, _ptr(tri),
                            ctypes.byref(cnt))
    if rc != BEM_OK:
        raise BemError(rc, lib().bem_read_msh_error().decode())
    return tri


def refine_triangles(tri, edge_length):
    """bisect every triangle (n, 9) until its longest edge is at most
    edge_length (bem_refine_triangles: ElementSection::refineTriangles)"""
    tri = _c(np.asarray(tri).reshape(-1, 9), np.float64)
    cnt = _i64()
    rc = lib().bem_refine_triangles(len(tri), _ptr(tri), float(edge_length), 0, None, ctypes.byref(cnt))
    if rc != BEM_OK:
        raise BemError(rc, "bem_refine_triangles")
    out = np.empty((cnt.value, 9), np.float64)
    rc = lib().bem_refine_triangles(len(tri), _ptr(tri), float(edge_length), cnt.value, _ptr(out), ctypes.byref(cnt))
    if rc != BEM_OK:
        raise BemError(rc, "bem_refine_triangles")
    return out


class Solver:
    """BoundaryElementSolver::hMatrixSolve on the GPU (bem_solver_t).

    tri: (n, 9) nodes in input order; bc: (n, 6) Robin a, b, g (re/im)."""

    def __init__(self, tri, bc, leaf=50, eta=3.0, device=0):
        tri, bc = _c(tri, np.float64), _c(bc, np.float64)
        h = _p()
        rc = lib().bem_solver_create(ctypes.byref(h), int(device), len(tri), _ptr(tri), _ptr(bc), int(leaf),
                                     float(eta))
        if rc != BEM_OK:
            raise BemError(rc, "bem_solver_create failed")
        self.h = h
        self.n = len(tri)

    def close(self):
        if getattr(self, "h", None):
            lib().bem_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def layout(self):
        perm = np.empty(self.n, np.int64)
        nb = _i64()
        lib().bem_solver_layout(self.h, _ptr(perm), ctypes.byref(nb), None, 0)
        blocks = np.empty((nb.value, 5), np.int64)
        lib().bem_solver_layout(self.h, None, ctypes.byref(nb), _ptr(blocks), nb.value)
        return perm, blocks

    def set_sources(self, points=None, plane=None):
        pts = _c(points if points is not None else np.zeros((0, 6)), np.float64)
        pl = _c(plane if plane is not None else np.zeros(5), np.float64)
        rc = lib().bem_solver_set_sources(self.h, len(pts), _ptr(pts), 1 if plane is not None else 0, _ptr(pl))
        if rc != BEM_OK:
            raise BemError(rc, "bem_solver_set_sources")

    def solve(self, freq, c=343.0, rho=1.21, coupling=1, aca_tol=1e-6, maxrank=0, gmres_tol=None, restart=300,
              seed=1, precision=FP64):
        n = self.n
        phi, dphi, pr = (np.empty(n, np.complex128) for _ in range(3))
        st = np.zeros(16)
        gt = -1.0 if gmres_tol is None else float(gmres_tol)
        rc = lib().bem_solver_solve(self.h, float(freq), float(c), float(rho), int(coupling), float(aca_tol),
                                    int(maxrank), gt, int(restart), int(seed), int(precision), _ptr(phi),
                                    _ptr(dphi), _ptr(pr), _ptr(st))
        if rc != BEM_OK:
            raise BemError(rc, lib().bem_solver_last_error(self.h).decode())
        return phi, dphi, pr, dict(zip(STAT_KEYS, st.tolist()))

    def set_autotune(self, on):
        """bem_solver_set_autotune: tune the ACA schedule on the first batched chunk"""
        rc = lib().bem_solver_set_autotune(self.h, int(bool(on)))
        if rc != BEM_OK:
            raise BemError(rc, "bem_solver_set_autotune")

    def set_aca_schedule(self, precision, small_bound, huge_bound, queue=0):
        """bem_solver_set_aca_schedule: a known ACA schedule, used without tuning"""
        rc = lib().bem_solver_set_aca_schedule(self.h, int(precision), int(small_bound), int(huge_bound), int(queue))
        if rc != BEM_OK:
            raise BemError(rc, "bem_solver_set_aca_schedule")

    def aca_schedule(self, precision):
        """the tuned ACA schedule of a precision (bem_solver_aca_schedule) or None"""
        out = np.zeros(4)
        rc = lib().bem_solver_aca_schedule(self.h, int(precision), _ptr(out))
        if rc != BEM_OK:
            raise BemError(rc, "bem_solver_aca_schedule")
        if out[0] < 0:
            return None
        return {"small_bound": int(out[0]), "huge_bound": int(out[1]), "queue": int(out[2]), "ms": float(out[3])}

    def set_sweep_batch(self, batch):
        """sweeps in frequency chunks of batch with batched assembly and
        lockstep GMRES on one context (bem_solver_set_sweep_batch; 0: lanes)"""
        rc = lib().bem_solver_set_sweep_batch(self.h, int(batch))
        if rc != BEM_OK:
            raise BemError(rc, "bem_solver_set_sweep_batch")

    def set_refinement(self, mode, max_steps=10):
        """FP32 solves: 0 no refinement, 1 residuals with the FP64 near field
        and the low-rank factors, 2 residuals with FP64 H-matrices
        (bem_solver_set_refinement)"""
        rc = lib().bem_solver_set_refinement(self.h, int(mode), int(max_steps))
        if rc != BEM_OK:
            raise BemError(rc, "bem_solver_set_refinement")

    def refine_stats(self):
        """the last solve's refinement: steps, inner iterations, final residual,
        FP64 assembly ms (bem_solver_refine_stats)"""
        o = np.zeros(4)
        lib().bem_solver_refine_stats(self.h, _ptr(o))
        return {"steps": int(o[0]), "inner_iters": int(o[1]), "residual": o[2], "fp64_assembly_ms": o[3]}

    def solve_dense(self, freq, c=343.0, rho=1.21, coupling=1, gmres_tol=1e-12, restart=300):
        """regularSolveWithLU on the device: dense FP64 Phi and DPhi, GMRES to
        gmres_tol (bem_solver_solve_dense); phi, dphi, pressure (cluster order), stats"""
        n = self.n
        phi, dphi, pr = (np.empty(n, np.complex128) for _ in range(3))
        st = np.zeros(16)
        rc = lib().bem_solver_solve_dense(self.h, float(freq), float(c), float(rho), int(coupling), float(gmres_tol),
                                          int(restart), _ptr(phi), _ptr(dphi), _ptr(pr), _ptr(st))
        if rc != BEM_OK:
            raise BemError(rc, lib().bem_solver_last_error(self.h).decode())
        return phi, dphi, pr, dict(zip(STAT_KEYS, st.tolist()))

    def precond_stats(self):
        """the last solve's preconditioner: set-up ms, applications, their ms,
        stored entries, factorisation ms (bem_solver_precond_stats)"""
        o = np.zeros(5)
        lib().bem_solver_precond_stats(self.h, _ptr(o))
        return {"setup_ms": o[0], "applies": int(o[1]), "apply_ms": o[2], "entries": o[3], "factor_ms": o[4]}

    def set_impedance_plane(self, point=None, normal=None):
        """the scene's impedance plane (a rigid image) through point with
        normal; None removes it (bem_solver_set_impedance_plane)"""
        if normal is None:
            rc = lib().bem_solver_set_impedance_plane(self.h, None, None)
        else:
            pt, nr = _c(point, np.float64), _c(normal, np.float64)
            rc = lib().bem_solver_set_impedance_plane(self.h, _ptr(pt), _ptr(nr))
        if rc != BEM_OK:
            raise BemError(rc, lib().bem_solver_last_error(self.h).decode())

    def set_preconditioner(self, kind, rank=2, rel_err=0.0):
        """0: none, 1: leaf-block LU, 2: host H-LU with the truncation rank /
        rel_err (bem_solver_set_preconditioner, bem_solver_set_hlu)"""
        if kind == 2:
            rc = lib().bem_solver_set_hlu(self.h, int(rank), float(rel_err))
        else:
            rc = lib().bem_solver_set_preconditioner(self.h, int(kind))
        if rc != BEM_OK:
            raise BemError(rc, "bem_solver_set_preconditioner")

    def set_precond_side(self, left):
        """bem_solver_set_precond_side: 1 the reference's left preconditioning
        from x = M^-1 b (tolerance on the preconditioned residual), 0 right"""
        rc = lib().bem_solver_set_precond_side(self.h, int(bool(left)))
        if rc != BEM_OK:
            raise BemError(rc, "bem_solver_set_precond_side")

    def field(self, obs):
        """sound pressure at observation points (nobs, 3) from the last solve"""
        obs = _c(obs, np.float64).reshape(-1, 3)
        out = np.empty(len(obs), np.complex128)
        rc = lib().bem_solver_field(self.h, len(obs), _ptr(obs), _ptr(out))
        if rc != BEM_OK:
            raise BemError(rc, lib().bem_solver_last_error(self.h).decode())
        return out

    def field_hmat(self, obs, aca_tol=0.01, maxrank=0, leaf=50, eta=3.0, seed=1):
        """sound pressure at observation points (nobs, 3) from the last solve
        through the observation x boundary H-matrix (bem_solver_field_hmat,
        the reference's calculateFieldSolutionFast); returns (pressure, stats)"""
        obs = _c(obs, np.float64).reshape(-1, 3)
        out = np.empty(len(obs), np.complex128)
        st = np.zeros(4)
        rc = lib().bem_solver_field_hmat(self.h, len(obs), _ptr(obs), float(aca_tol), int(maxrank), int(leaf),
                                         float(eta), int(seed), _ptr(out), _ptr(st))
        if rc != BEM_OK:
            raise BemError(rc, lib().bem_solver_last_error(self.h).decode())
        return out, {"ms": st[0], "admissible_blocks": int(st[1]), "max_rank": int(st[2]), "evaluated": int(st[3])}

    def sweep(self, freqs, g=None, planes=None, c=343.0, rho=1.21, coupling=1, aca_tol=1e-6, maxrank=0,
              gmres_tol=None, restart=300, seed=1, precision=FP64, lanes=1):
        """Frequency sweep (bem_solver_sweep): pressure (nf, nrhs, n) in
        cluster order, GMRES iterations (nf, nrhs), stats per frequency.
        g: (nrhs, n) complex boundary data in input order (None: the g given
        at creation, one right-hand side); planes: (nrhs, 5) plane waves."""
        freqs = _c(freqs, np.float64)
        nf = len(freqs)
        if g is not None:
            g = _c(np.atleast_2d(g), np.complex128)
            nrhs = len(g)
        else:
            nrhs = 1 if planes is None else len(planes)
        pl = None if planes is None else _c(np.atleast_2d(planes), np.float64)
        if pl is not None and len(pl) != nrhs:
            raise ValueError("planes must have one row per right-hand side")
        pr = np.empty((nf, nrhs, self.n), np.complex128)
        its = np.zeros((nf, nrhs), np.int64)
        st = np.zeros((nf, 16))
        gt = -1.0 if gmres_tol is None else float(gmres_tol)
        rc = lib().bem_solver_sweep(self.h, nf, _ptr(freqs), float(c), float(rho), int(coupling), float(aca_tol),
                                    int(maxrank), gt, int(restart), int(seed), int(precision), nrhs,
                                    None if g is None else _ptr(g), None if pl is None else _ptr(pl), int(lanes),
                                    _ptr(pr), _ptr(its), _ptr(st))
        if rc != BEM_OK:
            raise BemError(rc, lib().bem_solver_last_error(self.h).decode())
        return pr, its, [dict(zip(STAT_KEYS, r.tolist())) for r in st]
