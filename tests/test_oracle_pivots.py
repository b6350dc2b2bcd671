"""Pins of the restatement's normal-filtered pivot grouping
(getNormalFilteredPivotIndices, boundaryelementsolver.cpp:2052-2172, called
with the filter angle 0.44 from partialPivotACAextra, hmatrix.cpp:624):
the groups against the properties the greedy sweep fixes, against an
independent numpy sweep written from the reference text, and on a
hand-built block whose groups are worked out by hand; the pivot pairs
against the groups."""
import numpy as np
import pytest

COS = np.cos(0.44)
MAXC = 20


def greedy_groups(normals, idx):
    """Independent statement of the greedy sweep: group g keeps its first
    element and every element within the cone |n . n_first| >= cos 0.44 of
    it; the rest, in order, form group g + 1; at most 20 groups (the last one
    keeps whatever is left when the cap is reached)."""
    groups = [list(idx)]
    while len(groups[-1]) > 1 and len(groups) < MAXC:
        cur = groups[-1]
        ref = normals[cur[0]]
        inside = [e for e in cur[1:] if abs(float(normals[e] @ ref)) >= COS]
        moved = [e for e in cur[1:] if abs(float(normals[e] @ ref)) < COS]
        groups[-1] = [cur[0]] + inside
        if not moved:
            break
        groups.append(moved)
    return groups


def groups_from_ids(gid, start):
    ng = int(gid.max()) + 1
    return [list(start + np.nonzero(gid == g)[0]) for g in range(ng)]


def check_group_properties(normals, groups, start, count):
    members = sorted(e for g in groups for e in g)
    assert members == list(range(start, start + count))          # a partition of the range
    assert 1 <= len(groups) <= MAXC
    for g, G in enumerate(groups):
        assert len(G) >= 1 and G == sorted(G)                      # input order kept
        first = normals[G[0]]
        capped = g == MAXC - 1                                     # the remainder group at the cap
        for e in G:
            if not capped:
                assert abs(normals[e] @ first) >= COS             # inside the cone of its group
            for h in range(g):                                     # moved out of every earlier group
                assert abs(normals[e] @ normals[groups[h][0]]) < COS
        if g + 1 < len(groups):                                    # group g + 1 starts at its first leaver
            assert groups[g + 1][0] == min(e for G2 in groups[g + 1:] for e in G2)


@pytest.fixture(scope="module")
def L3(O, meshes):
    tri = meshes.icosphere(3, 0.5)
    P = O.Problem(tri, meshes.robin_bc(len(tri), 0, 1, 1), leaf=32, eta=2.0)
    P.set_frequency(500.0, 343.0, 1)
    return P


def test_groups_on_every_admissible_block(L3):
    """L3 sphere (leaf 32, eta 2): the row and column groups of every
    admissible block satisfy the sweep's properties and equal the numpy
    sweep; the whole sphere in one range hits the 20-group cap."""
    nodes, nrm, col, area, hbar = L3.geometry()
    B = L3.blocks
    adm = B[B[:, 4] == 1]
    ranges = {(int(r0), int(m)) for r0, m, _, _, _ in adm} | {(int(c0), int(n)) for _, _, c0, n, _ in adm}
    assert len(ranges) > 50
    for start, count in sorted(ranges) + [(0, L3.n)]:
        gid, ng = L3.normal_groups(start, count)
        groups = groups_from_ids(gid, start)
        assert len(groups) == ng
        check_group_properties(nrm, groups, start, count)
        assert groups == greedy_groups(nrm, range(start, start + count)), (start, count)


def test_group_cap_on_full_sphere(O, meshes):
    """The L4 sphere's 5120 normals in one range need 26 groups uncapped:
    the sweep stops at 20, the last group holding the remainder."""
    tri = meshes.icosphere(4, 0.5)
    P = O.Problem(tri, meshes.robin_bc(len(tri), 0, 1, 1), leaf=64, eta=2.0)
    nodes, nrm, col, area, hbar = P.geometry()
    gid, ng = P.normal_groups(0, P.n)
    groups = groups_from_ids(gid, 0)
    assert ng == MAXC
    check_group_properties(nrm, groups, 0, P.n)
    assert groups == greedy_groups(nrm, range(P.n))
    last = groups[-1]
    assert any(abs(nrm[e] @ nrm[last[0]]) < COS for e in last)   # the remainder is not one cone


def test_pivot_pairs_follow_the_groups(L3):
    """pivot_list: |R| |C| pairs, pair a |C| + b = (a member of row group a,
    a member of column group b), block-relative indices in range."""
    nodes, nrm, col, area, hbar = L3.geometry()
    B = L3.blocks
    ids = np.nonzero(B[:, 4] == 1)[0]
    for b in ids[:: max(1, len(ids) // 60)]:
        r0, m, c0, n, _ = (int(v) for v in B[b])
        R = groups_from_ids(L3.normal_groups(r0, m)[0], r0)
        C = groups_from_ids(L3.normal_groups(c0, n)[0], c0)
        piv = L3.pivot_list(int(b), 1)
        assert len(piv) == len(R) * len(C)
        for p, (i, j) in enumerate(piv):
            assert 0 <= i < m and 0 <= j < n
            assert r0 + i in R[p // len(C)] and c0 + j in C[p % len(C)]
        assert np.array_equal(piv, L3.pivot_list(int(b), 1))        # deterministic for a seed


def _tri_with_normal(center, nvec, s=0.01):
    nvec = np.asarray(nvec, float) / np.linalg.norm(nvec)
    t = np.array([1.0, 0, 0]) if abs(nvec[0]) < 0.9 else np.array([0, 1.0, 0])
    u = np.cross(nvec, t)
    u /= np.linalg.norm(u)
    v = np.cross(nvec, u)                                          # (u, v, n) right-handed: (b-a) x (c-a) ~ n
    a = np.asarray(center, float)
    return np.concatenate([a, a + s * u, a + s * v])


def test_hand_built_block(O, meshes):
    """Twelve elements with chosen normals, groups worked out by hand:
    e0 +z; e2 -z (|dot| = 1) and e3 (20 deg from z, 0.349 rad < 0.44) join
    it; e1 +x with e6 -x and e7 (10 deg); e4 (30 deg from z toward y) alone;
    e5 +y with e11 (24 deg, 0.419 rad); e8 (30 deg toward -y) alone; e9 and
    e10 = -e9 together.  A missing |.| splits e2 and e10 off, an angle used
    as a cosine or a cone of 0.44 on the wrong side moves e3 / e11 / e4."""
    d = np.deg2rad
    normals = [
        (0, 0, 1), (1, 0, 0), (0, 0, -1), (np.sin(d(20)), 0, np.cos(d(20))),
        (0, np.sin(d(30)), np.cos(d(30))), (0, 1, 0), (-1, 0, 0), (np.cos(d(10)), np.sin(d(10)), 0),
        (0, -np.sin(d(30)), np.cos(d(30))), (1, 1, 1), (-1, -1, -1), (0, np.cos(d(24)), np.sin(d(24))),
    ]
    tri = np.array([_tri_with_normal((0.05 * i, 0.0, 0.0), nv) for i, nv in enumerate(normals)])
    P = O.Problem(tri, meshes.robin_bc(len(tri), 0, 1, 1), leaf=32, eta=2.0)
    perm = P.perm                                                  # perm[new] = input index
    gid, ng = P.normal_groups(0, len(tri))
    got = [sorted(int(perm[e]) for e in np.nonzero(gid == g)[0]) for g in range(ng)]
    assert got == [[0, 2, 3], [1, 6, 7], [4], [5, 11], [8], [9, 10]]
