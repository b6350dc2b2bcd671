"""Synthetic triangle meshes and boundary conditions for the BEM hot path.

The reference reads GMSH files (mshreader.cpp:10); the workloads here are
generated instead: icospheres (the pulsating-sphere and sweep workloads),
ellipsoids (the multi-mesh batch) and capsules (the elongated body).  All
meshes are returned as an (n, 9) float64 array of triangle nodes
(a.xyz, b.xyz, c.xyz) ordered counter-clockwise seen from outside, so that
(b - a) x (c - a) points into the exterior domain as the reference expects
(vectortriangle.h:90, global.cpp:903-908).
"""
import numpy as np

_PHI = (1.0 + 5.0 ** 0.5) / 2.0


def _icosahedron():
    v = np.array([[-1, _PHI, 0], [1, _PHI, 0], [-1, -_PHI, 0], [1, -_PHI, 0],
                  [0, -1, _PHI], [0, 1, _PHI], [0, -1, -_PHI], [0, 1, -_PHI],
                  [_PHI, 0, -1], [_PHI, 0, 1], [-_PHI, 0, -1], [-_PHI, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1)[:, None]
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v, f


def icosphere_vf(level):
    """Unit icosphere vertices/faces after `level` midpoint subdivisions."""
    v, f = _icosahedron()
    verts = [tuple(x) for x in v]
    for _ in range(level):
        cache = {}
        vl = verts

        def mid(i, j):
            key = (i, j) if i < j else (j, i)
            if key not in cache:
                p = np.array(vl[i]) + np.array(vl[j])
                p /= np.linalg.norm(p)
                cache[key] = len(vl)
                vl.append(tuple(p))
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = np.array(nf, dtype=np.int64)
        verts = vl
    return np.array(verts, dtype=np.float64), f


def tri_from_vf(v, f):
    return np.ascontiguousarray(np.concatenate([v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]], axis=1))


def icosphere(level, radius=1.0, center=(0.0, 0.0, 0.0)):
    """20 * 4**level triangles on a sphere of the given radius."""
    v, f = icosphere_vf(level)
    v = v * radius + np.asarray(center, dtype=np.float64)
    tri = tri_from_vf(v, f)
    return _orient_outward(tri, np.asarray(center, dtype=np.float64))


def ellipsoid(level, axes, center=(0.0, 0.0, 0.0)):
    """Icosphere scaled to semi-axes `axes`."""
    v, f = icosphere_vf(level)
    v = v * np.asarray(axes, dtype=np.float64) + np.asarray(center, dtype=np.float64)
    return _orient_outward(tri_from_vf(v, f), np.asarray(center, dtype=np.float64))


def capsule(length, radius, edge):
    """Cylinder of the given length along z with hemispherical caps, near
    uniform triangles of edge length about `edge`."""
    nphi = max(8, int(round(2 * np.pi * radius / edge)))
    nz = max(1, int(round(length / edge)))
    ncap = max(2, int(round(0.5 * np.pi * radius / edge)))
    rings = []  # (z, r) from bottom pole to top pole
    for i in range(1, ncap + 1):          # bottom cap (excluding pole)
        th = -0.5 * np.pi + 0.5 * np.pi * i / ncap
        rings.append((-0.5 * length + radius * np.sin(th), radius * np.cos(th)))
    for i in range(1, nz + 1):
        rings.append((-0.5 * length + length * i / nz, radius))
    for i in range(1, ncap):
        th = 0.5 * np.pi * i / ncap
        rings.append((0.5 * length + radius * np.sin(th), radius * np.cos(th)))
    verts = [(0.0, 0.0, -0.5 * length - radius)]
    ring_idx = []
    for ri, (z, r) in enumerate(rings):
        off = 0.5 * (ri % 2)
        idx = []
        for k in range(nphi):
            ph = 2 * np.pi * (k + off) / nphi
            idx.append(len(verts))
            verts.append((r * np.cos(ph), r * np.sin(ph), z))
        ring_idx.append(idx)
    top = len(verts)
    verts.append((0.0, 0.0, 0.5 * length + radius))
    faces = []
    first = ring_idx[0]
    for k in range(nphi):
        faces.append([0, first[(k + 1) % nphi], first[k]])
    for ri in range(len(ring_idx) - 1):
        lo, hi = ring_idx[ri], ring_idx[ri + 1]
        for k in range(nphi):
            a, b = lo[k], lo[(k + 1) % nphi]
            if ri % 2 == 0:
                c, d = hi[k], hi[(k + 1) % nphi]
                faces.append([a, b, c]); faces.append([b, d, c])
            else:
                c, d = hi[(k - 1) % nphi], hi[k]
                faces.append([a, b, d]); faces.append([a, d, c])
    last = ring_idx[-1]
    for k in range(nphi):
        faces.append([top, last[k], last[(k + 1) % nphi]])
    v = np.array(verts, dtype=np.float64)
    f = np.array(faces, dtype=np.int64)
    tri = tri_from_vf(v, f)
    # orient each face away from the axis (z axis segment)
    cen = tri.reshape(-1, 3, 3).mean(axis=1)
    axis_pt = np.zeros_like(cen)
    axis_pt[:, 2] = np.clip(cen[:, 2], -0.5 * length, 0.5 * length)
    return _orient_by(tri, cen - axis_pt)


def _orient_by(tri, outward):
    a, b, c = tri[:, 0:3], tri[:, 3:6], tri[:, 6:9]
    nrm = np.cross(b - a, c - a)
    flip = np.einsum("ij,ij->i", nrm, outward) < 0
    out = tri.copy()
    out[flip, 3:6], out[flip, 6:9] = tri[flip, 6:9], tri[flip, 3:6]
    return np.ascontiguousarray(out)


def _orient_outward(tri, center):
    cen = tri.reshape(-1, 3, 3).mean(axis=1)
    return _orient_by(tri, cen - center)


def robin_bc(n, a=0.0, b=1.0, g=1.0):
    """(n, 6) array of Robin coefficients a phi + b dphi/dn = g (re, im each).
    a, b, g may be scalars or length-n complex arrays."""
    out = np.zeros((n, 6), dtype=np.float64)
    for col, val in ((0, a), (2, b), (4, g)):
        v = np.broadcast_to(np.asarray(val, dtype=np.complex128), (n,))
        out[:, col] = v.real
        out[:, col + 1] = v.imag
    return out


def impedance_bc(n, k, fraction, seed, rho_c_ratio=1.5 + 0.5j):
    """Locally reacting impedance Z = rho c (1.5 + 0.5i) on a seeded random
    fraction of the elements (a = i k rho c / Z, b = 1, g = 0,
    wallimpedancesection.cpp:358-363), rigid (a = 0, b = 1, g = 0) elsewhere."""
    rng = np.random.default_rng(seed)
    a = np.zeros(n, dtype=np.complex128)
    sel = rng.permutation(n)[: int(round(fraction * n))]
    a[sel] = 1j * k / rho_c_ratio
    return robin_bc(n, a=a, b=1.0, g=0.0)


def unit_direction(seed):
    """Direction uniform on the unit sphere (normalised N(0, I3))."""
    d = np.random.default_rng(seed).standard_normal(3)
    return d / np.linalg.norm(d)


def write_msh(path, tri, groups=None, names=None, version="4.1"):
    """Write triangles (n, 9) as an ASCII GMSH mesh (format "4.1" or "2.2").
    groups: geometrical group (surface entity tag) per triangle, default 1;
    each group is also a physical group of the same number, named by
    names[group] if given.  Coincident vertices share one node."""
    tri = np.asarray(tri, dtype=np.float64).reshape(-1, 9)
    n = len(tri)
    groups = np.ones(n, np.int64) if groups is None else np.asarray(groups, np.int64)
    names = names or {}
    pts = tri.reshape(-1, 3)
    uniq, inv = np.unique(pts, axis=0, return_inverse=True)
    conn = inv.reshape(-1, 3) + 1
    tags = sorted(set(groups.tolist()))
    out = []
    out += ["$MeshFormat", f"{version} 0 8", "$EndMeshFormat"]
    if names:
        out += ["$PhysicalNames", str(len(names))]
        out += [f'2 {g} "{names[g]}"' for g in sorted(names)]
        out += ["$EndPhysicalNames"]
    if version.startswith("4"):
        out += ["$Entities", f"0 0 {len(tags)} 0"]
        lo, hi = uniq.min(0), uniq.max(0)
        for g in tags:
            out.append(f"{g} {lo[0]} {lo[1]} {lo[2]} {hi[0]} {hi[1]} {hi[2]} 1 {g} 0")
        out += ["$EndEntities", "$Nodes", f"1 {len(uniq)} 1 {len(uniq)}", f"2 {tags[0]} 0 {len(uniq)}"]
        out += [str(i + 1) for i in range(len(uniq))]
        out += [f"{x!r} {y!r} {z!r}" for x, y, z in uniq.tolist()]
        out += ["$EndNodes", "$Elements", f"{len(tags)} {n} 1 {n}"]
        eid = 1
        for g in tags:
            sel = np.nonzero(groups == g)[0]
            out.append(f"2 {g} 2 {len(sel)}")
            for t in sel:
                out.append(f"{eid} {conn[t, 0]} {conn[t, 1]} {conn[t, 2]}")
                eid += 1
        out += ["$EndElements"]
    else:
        out += ["$Nodes", str(len(uniq))]
        out += [f"{i + 1} {x!r} {y!r} {z!r}" for i, (x, y, z) in enumerate(uniq.tolist())]
        out += ["$EndNodes", "$Elements", str(n)]
        out += [f"{t + 1} 2 2 {groups[t]} {groups[t]} {conn[t, 0]} {conn[t, 1]} {conn[t, 2]}" for t in range(n)]
        out += ["$EndElements"]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
