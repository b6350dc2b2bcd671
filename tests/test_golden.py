"""The CPU restatement reproduces the committed fixtures (bit for bit for
the entries and the partition), and the fixtures' pressure agrees with the
pulsating-sphere closed form."""
import os

import numpy as np

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz")


def test_golden_sphere1280(O, meshes):
    g = np.load(G, allow_pickle=False)
    tri = meshes.icosphere(3, 0.5)
    assert np.array_equal(tri, g["tri"])
    P = O.Problem(tri, meshes.robin_bc(len(tri), 0, 1, 1), leaf=32, eta=2.0)
    P.set_frequency(500.0, 343.0, 1)
    assert np.array_equal(P.perm, g["perm"]) and np.array_equal(P.blocks, g["blocks"])
    assert np.array_equal(P.entries4(g["ij"]).view(np.uint64), g["entries4"].view(np.uint64))
    _, _, pr, st = P.solve(1e-6, gmres_tol=1e-10, seed=1, rho=1.21, freq=500.0)
    assert np.linalg.norm(pr - g["pressure"]) / np.linalg.norm(g["pressure"]) < 1e-12
    k = 2 * np.pi * 500 / 343
    pex = 1j * 2 * np.pi * 500 * 1.21 * 0.5 / (1j * k * 0.5 - 1)
    assert np.linalg.norm(g["pressure"] - pex) / (abs(pex) * np.sqrt(len(tri))) < 0.03
