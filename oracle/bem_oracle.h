/*
 * CPU restatement of the reference's H-matrix BEM hot path -- TEST
 * INFRASTRUCTURE ONLY.  Nothing in the product (3d_bem_amd/) links or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load it, as the checker.
 *
 * Each function cites the reference file:line it restates.  Deliberate
 * deviations from the reference (all round-off level or determinism):
 *   D1 pivot randomness: counter-based generator keyed by (seed, stream,
 *      block, draw) instead of QRandomGenerator::system()
 *      (HMatrix/hmatrix.cpp:508; boundaryelementsolver.cpp:2156-2158);
 *   D2 leaf size and admissibility eta are parameters (reference constexpr 50
 *      and 3: HMatrix/clustertree.h:146, HMatrix/hmatrix.h:240);
 *   D3 GMRES is restarted GMRES(m) with classical Gram-Schmidt applied twice
 *      and complex Givens rotations, start vector zero (reference:
 *      Eigen::GMRES, Householder Arnoldi, random start vector);
 *   D4 floating-point operation order is fixed (bem_detmath.h) so that the
 *      GPU path reproduces every entry bit for bit; sin/cos of k r use
 *      det_sincos (<= 1.5 ulp) instead of libm's complex exp;
 *   D5 ACA argmax compares |z|^2 (reference |z|), ties keep the first index
 *      as Eigen's maxCoeff does; reductions over a block row/column use a
 *      fixed 64-lane (one wavefront) strided partial sum followed by a pairwise tree.
 */
#ifndef BEM_ORACLE_H
#define BEM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_problem orc_problem;
typedef struct orc_hmat orc_hmat;

/* Build a problem from a triangle mesh in input order.
   nodes: n x 9 (a.xyz, b.xyz, c.xyz) row-major; bc: n x 6 (a, b, g complex:
   Robin a phi + b dphi = g).  The cluster tree reorders elements; all other
   accessors use cluster order, perm[new] = old. */
orc_problem* orc_new(int64_t n, const double* nodes, const double* bc, int64_t leaf, double eta);
void orc_free(orc_problem* p);
int64_t orc_n(const orc_problem* p);
int64_t orc_nblocks(const orc_problem* p);
void orc_get_perm(const orc_problem* p, int64_t* perm);
/* blocks: nblocks x 5 (row0, rows, col0, cols, admissible) in partition order */
void orc_get_blocks(const orc_problem* p, int64_t* blocks);
/* nodes9, normal3, colloc3 (n x 9 / x 3 / x 3) and area (n) in cluster order */
void orc_get_geometry(const orc_problem* p, double* nodes9, double* normal3, double* colloc3, double* area, double* hbar);
/* cluster tree in depth-first order: for each cluster (start, size, depth, is_leaf, bbox min xyz, bbox max xyz) */
int64_t orc_nclusters(const orc_problem* p);
void orc_get_clusters(const orc_problem* p, double* out10);

/* coupling: 0 none, 1 Burton-Miller i/k, 2 Kirkup i/(k+1) */
int orc_set_frequency(orc_problem* p, double freq, double c, int coupling);
/* add point sources (position xyz, weight re/im) and/or one plane wave
   (direction xyz, amplitude re/im) to the source term vector */
void orc_set_sources(orc_problem* p, int64_t npoint, const double* points6, int has_plane, const double* plane5);
void orc_get_rhs_parts(const orc_problem* p, double* s2, double* alpha2, double* beta2, double* f2, int32_t* subst);
void orc_get_params(const orc_problem* p, double* out4); /* k, cc.re, cc.im, hbar */

/* the four operators L, M, Mt, N of pairs (row, col): out npairs x 8 */
void orc_entries4(const orc_problem* p, int64_t npairs, const int64_t* ij, double* out8);
/* the near-singular polar decomposition of pairs (row, col) forced for every
   pair, with an nq-point line rule (x, w on [-1, 1]) in angle and radius */
void orc_entries4_polar(const orc_problem* p, int64_t npairs, const int64_t* ij, int nq, const double* x,
                        const double* w, double* out8);
/* kind 0 Phi, 1 DPhi: out npairs x 2 */
void orc_entries(const orc_problem* p, int kind, int64_t npairs, const int64_t* ij, double* out2);
/* dense n x n matrix (column-major complex) */
void orc_dense_matrix(const orc_problem* p, int kind, double* out);

/* H-matrix assembly.  kind 0: Phi with partialPivotACA; kind 1: DPhi with
   partialPivotACAextra + normal-filtered pivots. */
orc_hmat* orc_assemble(const orc_problem* p, int kind, double tol, int64_t maxrank, uint64_t seed, int nthreads);
void orc_hmat_free(orc_hmat* h);
int64_t orc_hmat_rank(const orc_hmat* h, int64_t b); /* -1 for dense blocks */
/* dense: m x n column-major into U; low rank: U m x k col-major, alpha k,
   V k x n row-major (V row l = l-th ACA residual row) */
void orc_hmat_block(const orc_hmat* h, int64_t b, double* U, double* alpha, double* V);
/* y += H x (complex interleaved, cluster order) */
void orc_hmvm(const orc_hmat* h, const double* x, double* y);
/* pivot list of partialPivotACAextra for one block: returns count, pairs relative */
int64_t orc_pivot_list(const orc_problem* p, int64_t b, uint64_t seed, int64_t* out_pairs, int64_t cap);
/* group of every element start .. start + count - 1 (cluster order) under the
   greedy normal-orientation grouping of getNormalFilteredPivotIndices
   (boundaryelementsolver.cpp:2052-2172); returns the number of groups */
int64_t orc_normal_groups(const orc_problem* p, int64_t start, int64_t count, int32_t* gid);
/* count of evaluated entries (dense + ACA rows/columns + pivot probes) of the last assemble */
int64_t orc_hmat_evaluated(const orc_hmat* h);

/* restarted GMRES: A x = b with A an H-matrix; x in/out; returns iterations */
int64_t orc_gmres(const orc_hmat* h, const double* b, double* x, double tol, int64_t restart, int64_t maxit, double* resid_est);

/* Full hMatrixSolve (boundaryelementsolver.cpp:526-744) + pressure.
   out: phi, dphi, pressure (n complex each, cluster order).
   stats (16 doubles): [0] t_phi_asm [1] t_rhs_mvm [2] t_dphi_asm [3] t_gmres
   [4] t_total [5] gmres iterations [6] true relative residual [7] max rank phi
   [8] max rank dphi [9] evaluated entries phi [10] evaluated entries dphi
   [11] near blocks [12] far blocks [13] gmres estimate. */
int orc_solve(orc_problem* p, double acatol, int64_t maxrank, double gmres_tol, int64_t restart,
              uint64_t seed, double rho, double freq, int nthreads,
              double* phi, double* dphi, double* pressure, double* stats);

/* dense path (boundaryelementsolver.cpp:140-331): dense assembly + partial
   pivoting LU; phi/dphi/pressure out */
int orc_dense_solve(orc_problem* p, double rho, double freq, double* phi, double* dphi, double* pressure);

/* One impedance plane through point3 with normal normal3 (a rigid image,
   the single-plane branch of hMatrixSolve, boundaryelementsolver.cpp:549-557,
   594-604): every Phi / DPhi entry gains the regular-rule term of the mirror
   image of its column element (implicitPhiMatrixReflected /
   implicitDPhiMatrixReflected, :1556-1643, added as in
   implicit*MatrixPlusReflected, :2042-2050), and the field gains the mirror
   images' terms (:2270-2312).  normal3 NULL or zero removes the plane.
   orc_entries4 keeps returning the free-space operators. */
void orc_set_plane(orc_problem* p, const double* point3, const double* normal3);

/* Field potential of the boundary solution phi, dphi (n complex each,
   cluster order) at nobs observation points (nobs x 3): out[i] =
   sum_j M_j(x_i) phi_j - L_j(x_i) dphi_j with the order-5 7-point rule, as
   calculateFieldSolutionRegular / BemOperatorField
   (boundaryelementsolver.cpp:2262-2353).  Summation order (fixed, the GPU's):
   chunks of 1024 elements in order; inside a chunk four partial sums over
   element index mod 4 combined (s0 + s1) + (s2 + s3).  Sources excluded. */
void orc_field(const orc_problem* p, int64_t nobs, const double* obs3, const double* phi2, const double* dphi2,
               double* out2);
/* The same boundary term through the observation x boundary H-matrix of
   calculateFieldSolutionFast (boundaryelementsolver.cpp:2397-2476): the
   points' cluster tree (leaf), the partition against the boundary tree
   (eta), ACA of the M kernel and (field pivots) of the L kernel per
   admissible block to tol / maxrank, dense blocks directly; obs3 and out2
   in input order.  stats3: admissible blocks, largest rank, evaluated
   entries. */
int64_t orc_field_partition(const orc_problem* p, int64_t nobs, const double* obs3, int64_t leaf, double eta,
                            int64_t* perm, int64_t* blocks5, int64_t cap);
void orc_field_hmat(const orc_problem* p, int64_t nobs, const double* obs3, const double* phi2, const double* dphi2,
                    double tol, int64_t maxrank, uint64_t seed, int64_t leaf, double eta, double* out2, double* stats3);

/* quadrature tables used (for tests): order-8 triangle rule 16x3 and the
   5-point Gauss-Legendre rule 5x2, the 16-point half Gauss-Legendre 16x2 */
void orc_tables(double* tri16x3, double* gl5x2, double* half16x2);
/* det_sincos exposed for the accuracy test */
void orc_sincos(int64_t n, const double* x, double* s, double* c);
/* nearest-point projection on a triangle: returns region, point out */
int orc_project(const double* tri9, const double* p3, double* out3);

#ifdef __cplusplus
}
#endif
#endif
