"""GMSH reader (bem_read_msh, mshReader::readMsh, mshreader.cpp:10-428):
round trips in formats 2.2 and 4.1, group filters as the reference's
include / exclude lists, quadrilateral split, error paths."""
import os

import numpy as np
import pytest


def groups_by_hemisphere(tri):
    z = tri.reshape(-1, 3, 3)[:, :, 2].mean(axis=1)
    return np.where(z >= 0, 7, 9)


@pytest.mark.parametrize("version", ["2.2", "4.1"])
def test_round_trip(bem, meshes, tmp_path, version):
    tri = meshes.icosphere(2, 0.5)
    path = tmp_path / f"s{version}.msh"
    meshes.write_msh(path, tri, version=version)
    got = bem.read_msh(path)
    assert np.array_equal(got, tri)


@pytest.mark.parametrize("version", ["2.2", "4.1"])
def test_group_filters(bem, meshes, tmp_path, version):
    tri = meshes.icosphere(2, 0.5)
    g = groups_by_hemisphere(tri)
    path = tmp_path / f"g{version}.msh"
    meshes.write_msh(path, tri, groups=g, names={7: "north", 9: "South Cap"}, version=version)
    in_file = tri if version == "2.2" else np.concatenate([tri[g == 7], tri[g == 9]])   # 4.1: by entity
    assert np.array_equal(bem.read_msh(path, include=["all"]), in_file)
    assert np.array_equal(bem.read_msh(path, include=["7"]), tri[g == 7])
    assert np.array_equal(bem.read_msh(path, include=["8...9"]), tri[g == 9])
    assert np.array_equal(bem.read_msh(path, include=["...7"]), tri[g == 7])
    assert np.array_equal(bem.read_msh(path, include=["south cap"]), tri[g == 9])
    assert np.array_equal(bem.read_msh(path, exclude=["North"]), tri[g == 9])
    assert len(bem.read_msh(path, include=["11.."])) == 0


def test_quadrilaterals_split_and_errors(bem, tmp_path):
    path = tmp_path / "q.msh"
    path.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n4\n1 0 0 0\n2 1 0 0\n3 1 1 0\n4 0 1 0\n"
                    "$EndNodes\n$Elements\n1\n1 3 2 1 1 1 2 3 4\n$EndElements\n")
    tri = bem.read_msh(path)
    # equal diagonals: (4, 1, 3) + (1, 2, 3) (global::quadrilateralsToTriangles)
    assert tri.shape == (2, 9)
    assert np.array_equal(tri[0], [0, 1, 0, 0, 0, 0, 1, 1, 0]) and np.array_equal(tri[1], [0, 0, 0, 1, 0, 0, 1, 1, 0])
    # diagonal 1-3 longer than 2-4: (1, 2, 4) + (2, 3, 4)
    path2 = tmp_path / "q2.msh"
    path2.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n4\n1 0 0 0\n2 1 0 0\n3 2 1 0\n4 1 1 0\n"
                     "$EndNodes\n$Elements\n1\n1 3 2 1 1 1 2 3 4\n$EndElements\n")
    tri = bem.read_msh(path2)
    assert np.array_equal(tri[0], [0, 0, 0, 1, 0, 0, 1, 1, 0]) and np.array_equal(tri[1], [1, 0, 0, 2, 1, 0, 1, 1, 0])
    with pytest.raises(bem.BemError):
        bem.read_msh(tmp_path / "missing.msh")
    bad = tmp_path / "bad.msh"
    bad.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n1\n1 0 0 0\n$EndNodes\n$Elements\n1\n"
                   "1 2 2 1 1 1 2 3\n$EndElements\n")
    with pytest.raises(bem.BemError):
        bem.read_msh(bad)


REF_SPHERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "Sphere.msh")


def test_reference_example_sphere():
    """The reference's own example mesh (Example/Sphere/Sphere.msh, format
    4.1, kept as a fixture): a closed surface whose vertices lie on the
    radius-5 sphere."""
    import importlib
    bem = importlib.import_module("3d_bem_amd")
    tri = bem.read_msh(REF_SPHERE)
    assert len(tri) > 50
    r = np.linalg.norm(tri.reshape(-1, 3), axis=1)
    assert np.allclose(r, 5.0, rtol=1e-6)
    edges = {}
    for t in tri.reshape(-1, 3, 3):
        for a, b in ((0, 1), (1, 2), (2, 0)):
            key = tuple(sorted((tuple(np.round(t[a], 9)), tuple(np.round(t[b], 9)))))
            edges[key] = edges.get(key, 0) + 1
    assert set(edges.values()) == {2}
