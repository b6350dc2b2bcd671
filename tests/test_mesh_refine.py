"""Edge-length refinement (bem_refine_triangles, the reference's
ElementSection::refineTriangles) against the restatement in oracle.py, and
the restatement against a case worked out by hand and the invariants the
bisection keeps (area, orientation, edge bound)."""
import os

import numpy as np
import pytest

SPHERE_MSH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "Sphere.msh")


def areas_normals(tri):
    t = np.asarray(tri).reshape(-1, 3, 3)
    cr = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    a = 0.5 * np.linalg.norm(cr, axis=1)
    return a, cr / (2 * a[:, None])


def max_edges(tri):
    t = np.asarray(tri).reshape(-1, 3, 3)
    return np.max(np.stack([np.linalg.norm(t[:, k] - t[:, (k + 1) % 3], axis=1) for k in range(3)], 1), axis=1)


def test_hand_worked_right_triangle(O):
    """Legs 1 at the origin, edge length 0.75: the hypotenuse (edge 2) is cut
    at (0.5, 0.5); child 1 (n1, n2, m) has its edge 1 cut at (0.5, 0); child
    2 (m, n3, n1) has its edge 2 cut at (0, 0.5): four triangles in this
    order."""
    tri = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], float)
    got = O.refine_triangles(tri, 0.75)
    want = np.array([[0, 0, 0, .5, 0, 0, .5, .5, 0],
                     [.5, 0, 0, 1, 0, 0, .5, .5, 0],
                     [.5, .5, 0, 0, 1, 0, 0, .5, 0],
                     [0, .5, 0, 0, 0, 0, .5, .5, 0]])
    assert np.array_equal(got, want)
    assert len(O.refine_triangles(tri, 1.5)) == 1                     # nothing above the bound


@pytest.mark.parametrize("L", [0.9, 0.45, 0.2])
def test_invariants_and_product(bem, O, L):
    """The example sphere (radius 5, 110 triangles) refined to edges of at most
    L: every edge <= L, the total area and every child's normal equal the
    parent's, child counts double per split; the library's refinement equals
    the restatement bit for bit."""
    tri = bem.read_msh(SPHERE_MSH)
    ref = O.refine_triangles(tri, L)
    assert max_edges(ref).max() <= L
    a0, n0 = areas_normals(tri)
    a1, n1 = areas_normals(ref)
    assert abs(a1.sum() - a0.sum()) < 1e-12 * a0.sum()
    got = bem.refine_triangles(tri, L)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    # children of each parent (depth-first, in order): their normals are the parent's
    k = 0
    for p in range(len(tri)):
        cnt = len(O.refine_triangles(tri[p:p + 1], L))
        assert np.allclose(n1[k:k + cnt], n0[p], atol=1e-12)
        k += cnt
    assert k == len(ref)
