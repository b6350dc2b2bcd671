"""bem_field_noalloc (the field evaluation with its partial sums in the
library's fixed device scratch, no allocation) against bem_field, through
the C ABI: bit-identical results on two calls in a row, the second with
fewer points, and a clear error for sizes beyond the scratch."""
import ctypes

import numpy as np
import pytest

from helpers import sphere_problem, upload_problem

pytestmark = pytest.mark.gpu


def test_field_noalloc_equals_field(bem, O, meshes):
    P = sphere_problem(O, meshes, level=3, leaf=32, eta=2.0, freq=500.0)
    ctx = upload_problem(bem, P)
    n = P.n
    rng = np.random.default_rng(11)
    phi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dphi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    phi_d, dphi_d = ctx.array(phi), ctx.array(dphi)
    L = bem.lib()
    for nobs in (3000, 700):
        print("bem_field_noalloc: L3 sphere, nobs", nobs, flush=True)
        d = rng.standard_normal((nobs, 3))
        obs = d / np.linalg.norm(d, axis=1)[:, None] * rng.uniform(0.6, 3.0, (nobs, 1))
        ref = ctx.field(obs, phi_d, dphi_d)
        soa = np.ascontiguousarray(obs.T).ravel()
        obs_d = ctx.array(soa.view(np.complex128), n=len(soa) // 2)   # 3 nobs doubles as raw device bytes
        out_d = ctx.array(n=nobs)
        ctx._check(L.bem_field_noalloc(ctx.h, nobs, obs_d.ptr, phi_d.ptr, dphi_d.ptr, out_d.ptr))
        ctx.synchronize()
        got = out_d.download()
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
        obs_d.free()
        out_d.free()
    big = (32 << 20) // 16 // ((n + 1023) // 1024) + 1              # one point beyond the scratch
    rc = L.bem_field_noalloc(ctx.h, big, phi_d.ptr, phi_d.ptr, dphi_d.ptr, phi_d.ptr)
    assert rc == bem.BEM_ERR_ARG
    assert b"scratch" in L.bem_last_error(ctx.h)
    ctx.close()
