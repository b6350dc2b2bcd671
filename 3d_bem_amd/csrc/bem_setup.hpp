// Host-side, per-mesh set-up of the H-matrix BEM path: element geometry,
// cluster tree, block partition, the frequency-independent geometry of the
// self and near-singular quadratures, and the normal-filtered ACA pivot
// lists.  All of it is integer/geometry work that runs once per mesh; the
// frequency-dependent integrals run on the GPU (bem_kernels.hip).
#pragma once
#include <cstdint>
#include <vector>

namespace bemsetup {

struct Vec3 { double x, y, z; };

// Per-element geometry in one order (cluster order after reordering).
struct Elements {
    int64_t n = 0;
    std::vector<double> nodes;   // n x 9: a.xyz b.xyz c.xyz
    std::vector<Vec3> normal, colloc;
    std::vector<double> area;
    double hbar = 0;             // mean(sqrt(area)) / 2
};

// Fills normal (normalised (b-a)x(c-a), global.cpp:903-908), centroid
// collocation point (global.cpp:910-913), area (global.cpp:1285-1290) and
// averageTriangleDim (boundaryelements.cpp:35-46).
void prepare_elements(Elements& E);

// Mirror image of every element in the impedance plane through `point` with
// normal `normal` (normalised here): MeshFunctions::reflectTriangles
// (meshfunctions.cpp:349-378) as called by BoundaryElements::reflectGeometry
// (boundaryelements.cpp:151-155).  x' = R [x; 1] with the rows of
// R = [I - 2 n n^T | -2 n d], d = -n.point, each row summed left to right;
// the node order is reversed (a' = R c, b' = R b, c' = R a), so the
// recomputed normal (rnormal, as prepare_elements) is the mirror image of the
// element's normal.  The element keeps its area (the reference copies the
// areas before reflecting).
void reflect_elements(int64_t n, const double* nodes9, const double point[3], const double normal[3],
                      std::vector<double>& rnodes9, std::vector<Vec3>& rnormal);

struct Cluster {
    int64_t start = 0, size = 0;
    int depth = 0;
    bool leaf = false;
    Vec3 bmin, bmax;
    int son1 = -1, son2 = -1;
};

struct Block {
    int64_t row0, rows, col0, cols;
    bool adm;
};

// ClusterTree(BoundaryElements*) (HMatrix/clustertree.h:77-86,215-310,
// clustertree.cpp:265-283): bisection of the centroid bounding box at the
// middle of its longest side, son1 = centroids with coordinate >= cut; leaf
// when |I| < leaf or |I| <= 1; depth-first reordering.  perm[new] = old.
struct ClusterTree {
    std::vector<Cluster> clusters;   // depth first, root = 0
    std::vector<int64_t> perm;
};
ClusterTree build_cluster_tree(int64_t n, const double* nodes9, int64_t leaf);
// the same over points (ClusterTree(QVector<Eigen::Vector3d>*) and the
// observation-triangle tree, whose members are their midpoints,
// clustertree.h:96-113; pts3: n x 3)
ClusterTree build_cluster_tree_points(int64_t n, const double* pts3, int64_t leaf);

// HMatrix::populateBlockClusterTree (hmatrix.cpp:922-1007) with the
// admissibility min(diam)/dist < eta (hmatrix.cpp:1052-1068).
std::vector<Block> build_partition(const ClusterTree& T, double eta);
// rows from TR, columns from TC (HMatrix(rowTree, columnTree, true), the
// field's observation x boundary partition, boundaryelementsolver.cpp:2425)
std::vector<Block> build_partition(const ClusterTree& TR, const ClusterTree& TC, double eta);

// Near-singular pairs |p_i - p_j| < 4 hbar, i != j
// (boundaryelementsolver.cpp:851), as CSR over rows with columns sorted, plus
// the per-pair polar quadrature geometry of BemOperatorsNearSing
// (boundaryelementsolver.cpp:1387-1514).  Per sub-triangle 25 doubles:
// node1.xyz, e1.xyz, e2.xyz (in-plane orthonormal axes), angle/2, then for
// the 5 angular Gauss points (rho_max, cos, sin).
constexpr int NEAR_SUB_DOUBLES = 25;
struct NearList {
    std::vector<int64_t> rowptr;     // n + 1
    std::vector<int32_t> col;        // nnz
    std::vector<int32_t> nsub;       // nnz
    std::vector<int64_t> geo_off;    // nnz (offset in doubles)
    std::vector<double> geo;
};
NearList build_near(const Elements& E);

// Frequency-independent radii of the polar self integral
// (boundaryelementsolver.cpp:896-1029): per element 3 sub-triangles x
// (lower, r1[16], r2[16]) = 99 doubles.
constexpr int SELF_DOUBLES = 99;
std::vector<double> build_self(const Elements& E);

// Pivot list of partialPivotACAextra for every admissible block
// (getNormalFilteredPivotIndices, boundaryelementsolver.cpp:2052-2172, cone
// 0.44 rad, at most 20 groups; hmatrix.cpp:624).  Pairs relative to the
// block origin; offsets per block (non-admissible blocks get empty lists).
void build_pivot_lists(const Elements& E, const std::vector<Block>& blocks, uint64_t seed,
                       std::vector<int64_t>& offsets, std::vector<int32_t>& pairs);

// Pivot lists of the field's partialPivotACAextra
// (getNormalFilteredPivotIndicesForField, boundaryelementsolver.cpp:
// 2721-2807): orientation groups of the block's boundary columns only (cone
// 0.44 rad, at most 20 groups); as many random observation rows as groups,
// and for every (row, group) pair a random member of the group.  Blocks
// index rows by observation point and columns by element of E.  Draws from
// stream kFieldPivotStream of the counter-based generator.
constexpr uint64_t kFieldPivotStream = 2;
void build_field_pivot_lists(const Elements& E, const std::vector<Block>& blocks, uint64_t seed,
                             std::vector<int64_t>& offsets, std::vector<int32_t>& pairs);

}  // namespace bemsetup
