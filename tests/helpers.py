"""Shared set-up for the parity tests: one oracle problem and the same
cluster-ordered data uploaded through the C ABI."""
import numpy as np


def sphere_problem(O, meshes, level=3, radius=0.5, leaf=32, eta=2.0, freq=500.0, c=343.0, bc=(0, 1, 1)):
    tri = meshes.icosphere(level, radius)
    n = len(tri)
    P = O.Problem(tri, meshes.robin_bc(n, *bc), leaf=leaf, eta=eta)
    P.set_frequency(freq, c, 1)
    return P


def upload_problem(bem, P):
    nodes, nrm, col, area, hbar = P.geometry()
    s, a, b, f, sub = P.rhs_parts()
    prm = P.params()
    ctx = bem.Context(0)
    ctx.upload_mesh(nodes, nrm, area, col, hbar)
    ctx.set_bc(a, b, f)
    ctx.upload_partition(P.blocks)
    ctx.set_frequency(prm["k"], prm["cc"])
    return ctx


def near_pairs(P, limit=None):
    nodes, nrm, col, area, hbar = P.geometry()
    d = np.linalg.norm(col[:, None, :] - col[None, :, :], axis=2)
    i, j = np.nonzero((d < 4 * hbar) & ~np.eye(len(col), dtype=bool))
    ij = np.stack([i, j], 1)
    return ij if limit is None else ij[:limit]
