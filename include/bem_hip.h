/*
 * bem_hip.h -- C ABI of the MI355X (gfx950) H-matrix BEM path.
 *
 * The reference runs this path behind callback-parameterised C++ methods:
 *   HMatrix::assembleBlocks(rank, relErr, std::function<complex(long,long)>)
 *       HMatrix/hmatrix.h:200, HMatrix/hmatrix.cpp:793-823
 *   HMatrix::assembleBlocksExtra(rank, relErr, getPivotIndices, implicitMatrix)
 *       HMatrix/hmatrix.h:201, HMatrix/hmatrix.cpp:825-846
 *   HArithm::MVM(VectorXcd& y, HMatrix&, const VectorXcd& x)   (y += H x)
 *       HMatrix/harithm.h:32, HMatrix/harithm.cpp:3-28
 *   GMRES::gmresSolve(HMatrixWrapper, x0, b, tol, restart)
 *       HMatrix/gmres.h:170-180
 * A per-entry host callback cannot run on the GPU, so this ABI takes the
 * description of the entry generator instead (geometry, boundary
 * conditions, wavenumber, coupling, operator kind) and owns all device
 * memory.  Handles are opaque; every call returns 0 on success or a negative
 * BEM_ERR_* code and records a message for bem_last_error().  No call throws.
 * One context = one HIP stream; a context is not re-entrant.
 *
 * Numbers: complex values are interleaved (re, im) doubles (Eigen's
 * std::complex<double> layout).  Block storage follows the reference's
 * BlockCluster fields (hmatrix.h:130-133): dense blocks column-major;
 * low-rank blocks A ~ U diag(alpha) V with U m x k column-major, alpha the
 * ACA pivots' reciprocals (the reference's "singularValues") and V k x n
 * whose rows are the ACA residual rows (the reference's "VAdjMat").
 */
#ifndef BEM_HIP_H
#define BEM_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BEM_OK 0
#define BEM_ERR_ARG -1      /* bad argument or size */
#define BEM_ERR_HIP -2      /* HIP runtime error (message has the HIP string) */
#define BEM_ERR_STATE -3    /* call order (e.g. assemble before upload) */
#define BEM_ERR_NAN -4      /* NaN detected in an assembled operator */
#define BEM_ERR_NOCONV -5   /* GMRES stopped at maxit above tolerance */

#define BEM_KIND_PHI 0      /* implicitPhiMatrix2,  boundaryelementsolver.cpp:2004-2020 */
#define BEM_KIND_DPHI 1     /* implicitDPhiMatrix2, boundaryelementsolver.cpp:2022-2040 */
#define BEM_PREC_FP64 0
#define BEM_PREC_FP32 1     /* FP32 regular quadrature and block storage; FP64 singular/near-singular */

typedef struct bem_ctx* bem_ctx_t;
typedef struct bem_hmat* bem_hmat_t;
typedef struct bem_solver* bem_solver_t;
typedef struct bem_prec* bem_prec_t;

/* context on HIP device `device` (BoundaryElementSolver owns one per solve);
   destroy its H-matrices and preconditioners before it */
int bem_create(bem_ctx_t* out, int device);
int bem_destroy(bem_ctx_t ctx);
const char* bem_last_error(bem_ctx_t ctx);

/* Mesh in cluster order (after ClusterTree's reordering,
   HMatrix/clustertree.cpp:265-283): nodes n x 9 (a, b, c), normal n x 3,
   area n, colloc n x 3 (centroids), hbar = averageTriangleDim
   (boundaryelements.cpp:35-46).  Replaces the geometry read by
   BemOperatorsConst (boundaryelementsolver.cpp:844-894).  An element with
   a non-finite coordinate or an area not > 0 is refused (BEM_ERR_ARG,
   naming it) before the context changes. */
int bem_upload_mesh(bem_ctx_t ctx, int64_t n, const double* nodes9, const double* normal3, const double* area,
                    const double* colloc3, double hbar);
/* Host H-LU (the reference keeps the H-LU preconditioner on the host:
   HArithm::LUDecomposition / recursiveLUDecomposition,
   HMatrix/harithm.cpp:1192-1296, of the truncated copy
   HMatrix(const HMatrix&, rank, relError), hmatrix.cpp:770-776, as
   hMatrixSolve builds it, boundaryelementsolver.cpp:622-643).  clusters2:
   (start, size) of every cluster of the cluster tree; blocks5: the partition
   (bem_upload_partition's blocks); rank[b] = -1: data[b] is the dense block
   (m x n column-major complex), else data[b] holds U (m x k column-major),
   alpha (k), V (k x n column-major) one after the other (bem_block_download's
   three arrays).  max_rank / rel_err: preconditionerRank /
   preconditionerRelError (bemparameters.h:27-28; 0 / 0 = no truncation, an
   exact LU).  Factors on return.  bem_hlu_solve: x = U^-1 L^-1 b (host
   vectors, n complex; x may equal b): forwardSubstitution +
   backwardSubstitution (harithm.cpp:1314-1412).  bem_hlu_solve_flat: the
   same solve in the leaf-by-leaf order the device preconditioner of
   bem_precond_create_hlu runs (diagonal leaves in order, each after the row
   slices of the blocks left of it in L / right of it in U, Q^H and R^-1 of
   the leaf applied as products), on the host; equal to bem_hlu_solve up to
   the order of the additions.  bem_hlu_dense: L U as a
   dense n x n matrix (column-major complex).  bem_hlu_stats: stored complex
   entries, largest rank, factor seconds.  Errors: bem_hlu_error(). */
typedef struct bem_hlu* bem_hlu_t;
int bem_hlu_create(int64_t n, int64_t nclusters, const int64_t* clusters2, int64_t nb, const int64_t* blocks5,
                   const int64_t* rank, const double* const* data, int64_t max_rank, double rel_err, bem_hlu_t* out);
int bem_hlu_solve(bem_hlu_t h, const double* b2, double* x2);
int bem_hlu_solve_flat(bem_hlu_t h, const double* b2, double* x2);
int bem_hlu_dense(bem_hlu_t h, double* out2);
int bem_hlu_stats(bem_hlu_t h, double* out3);
int bem_hlu_destroy(bem_hlu_t h);
const char* bem_hlu_error(void);

/* One impedance plane, a rigid image: the plane through point3 with normal
   normal3 (normalised here); normal3 NULL or zero removes it.  Later
   assemblies add to every Phi / DPhi entry the regular-rule entry of the
   column element's mirror image (no jump term, the column's boundary
   condition), and bem_field adds the mirror images' terms.  Replaces the
   single-plane branch of hMatrixSolve (boundaryelementsolver.cpp:549-557,
   594-604): BoundaryElements::reflectGeometry (boundaryelements.cpp:151-155,
   MeshFunctions::reflectTriangles meshfunctions.cpp:349-378) and
   implicitPhiMatrixPlusReflected / implicitDPhiMatrixPlusReflected
   (boundaryelementsolver.cpp:2042-2050 over :1516-1643), and the plane terms
   of calculateFieldSolutionRegular (:2270-2312).  The plane is kept across
   bem_upload_mesh (rebuilt for the new mesh); already assembled H-matrices
   keep the operator they were assembled with. */
int bem_set_plane(bem_ctx_t ctx, const double* point3, const double* normal3);

/* Near-field pairing (off by default): a Phi assembly also forms the DPhi
   near field (dense blocks and self / near-singular entries) from the same
   quadrature and keeps it on the device; the next DPhi assembly with the
   same mesh, partition, boundary conditions, frequency, plane and precision
   takes it instead of recomputing.  The order of hMatrixSolve (Phi, then
   DPhi, boundaryelementsolver.cpp:546-604) is the one it serves; entries
   are unchanged. */
int bem_set_near_pairing(bem_ctx_t ctx, int on);
/* Robin coefficients a phi + b dphi = g per element (complex, n each):
   the alpha/beta/f vectors of calculateBoundaryConditionAndSourceTermVector
   (boundaryelementsolver.cpp:3386-3429). */
int bem_set_bc(bem_ctx_t ctx, const double* alpha_ri, const double* beta_ri, const double* f_ri);
/* Block partition in minPartition order (HMatrix::populateBlockClusterTree,
   hmatrix.cpp:922-1007). */
int bem_upload_partition(bem_ctx_t ctx, int64_t nblocks, const int64_t* row0, const int64_t* rows,
                         const int64_t* col0, const int64_t* cols, const uint8_t* admissible);
/* wavenumber k = 2 pi f / c and coupling parameter (boundaryelementsolver.cpp:30-59);
   also forms the column coefficients of the boundary conditions and refuses
   (BEM_ERR_ARG, naming the element) an element whose a and b are both zero
   or not finite, before any kernel sees it. */
int bem_set_frequency(bem_ctx_t ctx, double k, double cc_re, double cc_im);

/* assembleBlocks (kind PHI: dense + partialPivotACA) or assembleBlocksExtra
   (kind DPHI: dense + partialPivotACAextra with normal-filtered pivots).
   max_rank 0 = unlimited; seed keys the deterministic pivot generator. */
int bem_assemble(bem_ctx_t ctx, int kind, double aca_tol, int64_t max_rank, int precision, uint64_t seed,
                 bem_hmat_t* out);
int bem_hmat_destroy(bem_hmat_t h);
/* Batched assembly (no counterpart in the reference, whose frequency loop
   assembles one operator at a time, bemsolvercontroller.cpp:59-79): the
   operator `kind` at nf frequencies (kcc3: nf x (k, cc re, cc im)) on the
   context's mesh, partition and boundary conditions, returned as nf
   handles in outs, each equal to bem_assemble at that frequency (the FP64
   near field of bem_set_near_fp64 is kept).  With near-field pairing on, a
   batched Phi assembly also forms the DPhi near field of every frequency
   and the next batched DPhi assembly of the same state and frequency list
   takes them.  The ACA runs as one
   launch per size class over the flattened (frequency, block) list, the
   same-size blocks of all frequencies adjacent and the largest first; the
   self / near-singular tables and dense tiles per frequency.  With an
   impedance plane every entry carries the mirror image's quadrature as in
   bem_assemble, and the ACA runs one workgroup per item (a work-queue
   schedule is ignored); 1 <= nf <= 64.  The context's frequency is
   unchanged.
   bem_set_aca_schedule: the ACA size-class bounds on m + n for one
   precision (one wavefront up to small_bound, the middle workgroup shape
   up to huge_bound, the largest above) and, for queue >= 0, whether batched
   ACA launches run as persistent workgroups that take items from a device
   counter (1) or one workgroup per item (0).  Results do not depend on it. */
int bem_assemble_batch(bem_ctx_t ctx, int kind, int64_t nf, const double* kcc3, double aca_tol, int64_t max_rank,
                       int precision, uint64_t seed, bem_hmat_t* outs);
/* Several meshes at once (no counterpart in the reference, which solves
   one mesh after another): the operator `kind` of every context in ctxs
   (nctx, all on one device; each with its own mesh, partition, boundary
   conditions and frequency), returned in outs[q] for ctxs[q] and equal to
   bem_assemble on that context.  Near fields and singular tables per
   context; the ACA of every (context, block) item in one launch per size
   class on the first context's streams, sized by its schedule (and, with
   its work queue on, as persistent workgroups).  No impedance plane, no
   near-field pairing (a kept near field is dropped). */
int bem_assemble_multi(const bem_ctx_t* ctxs, int64_t nctx, int kind, double tol, int64_t max_rank, int precision,
                       uint64_t seed, bem_hmat_t* outs);
int bem_set_aca_schedule(bem_ctx_t ctx, int precision, int64_t small_bound, int64_t huge_bound, int queue);
/* ACA schedule autotuner (no counterpart in the reference, whose per-block
   tasks go to a thread pool, hmatrix.cpp:793-846): times the assembly of
   `kind` at the nf frequencies kcc3 (bem_assemble for nf = 1,
   bem_assemble_batch otherwise; the fastest of two runs each) under
   candidate schedules by coordinate descent -- the one-wavefront bound
   (128, 256, 384, 512), then for FP32 the huge bound (1024, 2048, 4096,
   none), then for nf > 1 the work queue, a candidate taking over only when
   3 % faster than the current choice -- and keeps the choice on the
   context.  choice3: small bound, huge bound, queue; best_ms: its time.
   Operators kept for near-field pairing are dropped. */
int bem_autotune_aca(bem_ctx_t ctx, int kind, int precision, int64_t nf, const double* kcc3, double tol,
                     int64_t maxrank, uint64_t seed, int64_t* choice3, double* best_ms);
/* HArithm::MVM: y (+)= H x on device pointers (complex double, cluster
   order); accumulate = 1 adds to y as the reference does. */
int bem_hmvm(bem_ctx_t ctx, bem_hmat_t h, const void* x_dev, void* y_dev, int accumulate);
/* rank of block b (-1 dense), and its factors in the reference layout (host
   buffers: dense m x n col-major into full_or_U; else U m x k, alpha k, V k x n
   col-major), for the host H-LU (HMatrix(const HMatrix&, rank, err),
   hmatrix.cpp:770-776) and for inspection. */
int bem_block_info(bem_hmat_t h, int64_t b, int64_t* rank);
int bem_block_download(bem_hmat_t h, int64_t b, double* full_or_U, double* alpha, double* V);
/* stats[0] evaluated entries, [1] max rank, [2] mean rank, [3] bytes of block
   storage at the final ranks (what one H-MVM reads), [4] assembly ms (host
   wall clock), [5] blocks re-run after rank-capacity overflow, [6] device ms
   of the singular, dense and first ACA pass, [7] device ms of the re-run and
   the packing, [8] device bytes the H-matrix holds (when no block
   overflowed, the first ACA pass's storage at its capacity of 48 per block
   is kept as is, so this exceeds [3]) */
int bem_hmat_stats(bem_hmat_t h, double* stats9);

/* GMRES::gmresSolve replacement: restarted GMRES(restart) with CGS2 Arnoldi
   on the device; x_dev holds the start vector on entry and the solution on
   return.  Returns BEM_ERR_NOCONV when maxit is reached above tol. */
int bem_gmres(bem_ctx_t ctx, bem_hmat_t h, const void* b_dev, void* x_dev, double tol, int64_t restart,
              int64_t maxit, int64_t* iters, double* rel_resid);
/* Mixed-precision refinement (no counterpart in the reference, whose solve
   is FP64 throughout).  bem_set_near_fp64(ctx, 1): later FP32 assemblies
   form their near field (dense blocks and self / near-singular entries) in
   FP64, keep it beside the FP32 copy they round from it, and so carry an
   FP64-accurate refinement operator: the FP64 near field plus the stored
   low-rank factors, products accumulated in FP64.  bem_hmvm_refine:
   y (+)= that operator times x (FP64 operators: bem_hmvm).
   bem_gmres_ir: iterative refinement around GMRES,
     r = b - A_res x (FP64), stop when |r| < tol |b|,
     d = GMRES(A, r) to the relative tolerance min(0.5, 0.5 tol |b| / |r|),
     x += d,
   with A = h (the FP32 operator the Krylov iterations use, preconditioner
   pc as bem_gmres_prec) and A_res = hres (an FP64 H-matrix, or an FP32 one
   with its FP64 near field; NULL: h itself).  x_is_zero: x holds 0 on entry
   (saves the first residual product).  iters: inner GMRES iterations;
   rel_resid: |b - A_res x| / |b|; steps: refinement steps.  The residual
   check replaces GMRES's own (gmres.h:170-180) and the true residual of
   hMatrixSolve (boundaryelementsolver.cpp:682-684). */
int bem_set_near_fp64(bem_ctx_t ctx, int on);
int bem_hmvm_refine(bem_ctx_t ctx, bem_hmat_t h, const void* x_dev, void* y_dev, int accumulate);
int bem_gmres_ir(bem_ctx_t ctx, bem_hmat_t h, bem_hmat_t hres, bem_prec_t pc, const void* b_dev, void* x_dev,
                 int x_is_zero, double tol, int64_t restart, int64_t maxit, int64_t max_steps, int64_t* iters,
                 double* rel_resid, int64_t* steps);

/* Several right-hand sides (no counterpart in the reference, which solves
   them one after another): vectors of nrhs right-hand sides (1 <= nrhs <= 8)
   interleaved, v[i * nrhs + q], in device memory.
   bem_hmvm_multi: Y (+)= H X for all nrhs with one pass over the blocks
   (refine = 1: with the refinement operator of bem_hmvm_refine); for four
   or more right-hand sides the near-field tiles and the U t phase of the
   low-rank blocks run on the FP64 matrix cores unless bem_set_mfma(ctx, 0).
   bem_gmres_multi: GMRES(restart) in lockstep over the nrhs right-hand
   sides B (start / solution X; x_is_zero: X is taken as 0), one product per
   Arnoldi column for all of them, each stopping at its own tolerance;
   hres != NULL or max_steps > 0: inside bem_gmres_ir's refinement loop with
   the residual operator hres (NULL: h).  iters / rel_resid: per right-hand
   side (columns used, final |b - A x| / |b|); steps: refinement steps.
   With bem_set_graphs(ctx, 1) (the default) the launches of one Arnoldi
   column are captured once per solve as a HIP graph and replayed per
   column; bem_gmres / bem_gmres_ir without a preconditioner run on the
   same solver with one right-hand side. */
int bem_set_mfma(bem_ctx_t ctx, int on);
int bem_set_graphs(bem_ctx_t ctx, int on);
int bem_hmvm_multi(bem_ctx_t ctx, bem_hmat_t h, int64_t nrhs, const void* x_dev, void* y_dev, int accumulate,
                   int refine);
int bem_gmres_multi(bem_ctx_t ctx, bem_hmat_t h, bem_hmat_t hres, int64_t nrhs, const void* b_dev, void* x_dev,
                    int x_is_zero, double tol, int64_t restart, int64_t maxit, int64_t max_steps, int64_t* iters,
                    double* rel_resid, int64_t* steps);

/* bem_gmres_multi over several operators at once (the frequencies of a
   batch): column q = o * nrhs + r of B / X (interleaved, nops * nrhs <= 8,
   nrhs 1, 2 or 4) is right-hand side r of operator hs[o]; every Arnoldi
   column applies each operator to its own columns inside the same replayed
   graph.  hres: NULL, or the residual operator of each hs[o] (refinement
   as bem_gmres_multi; max_steps > 0 with hres NULL: each operator's own
   refinement form).  iters / rel_resid per column. */
int bem_gmres_batch(bem_ctx_t ctx, int64_t nops, const bem_hmat_t* hs, const bem_hmat_t* hres, int64_t nrhs,
                    const void* b_dev, void* x_dev, int x_is_zero, double tol, int64_t restart, int64_t maxit,
                    int64_t max_steps, int64_t* iters, double* rel_resid, int64_t* steps);
/* bem_set_fuse_ops(ctx, 1) (the default): the product of bem_gmres_batch
   applies all its operators (of one storage precision) in one launch per
   kernel, operator o taking columns o * nrhs .. of the interleaved vectors;
   0: one operator after another through column gathers. */
int bem_set_fuse_ops(bem_ctx_t ctx, int on);

/* Leaf-block preconditioner of an assembled operator: LU with partial
   pivoting (FP64) of every diagonal leaf block, i.e. the leaf level of the
   reference's H-LU preconditioner (HMatrix(const HMatrix&, rank, err) +
   HArithm::LUDecomposition, hmatrix.cpp:770-776, harithm.cpp:1192-1296)
   without its low-rank coupling.  Needs diagonal blocks of at most 64 rows
   (leaf size <= 64).  bem_gmres_prec runs GMRES on A M^-1 (right
   preconditioning; x is the solution of A x = b); pc = NULL is bem_gmres.
   bem_precond_apply: z = M^-1 v (device vectors, n complex, z != v). */
int bem_precond_create(bem_ctx_t ctx, bem_hmat_t h, bem_prec_t* out);
int bem_precond_destroy(bem_prec_t p);
int bem_precond_apply(bem_ctx_t ctx, bem_prec_t p, const void* v_dev, void* z_dev);
/* The H-LU preconditioner of hMatrixSolve (boundaryelementsolver.cpp:622-643,
   usePreconditioner): the host H-LU of bem_hlu_create (truncation max_rank /
   rel_err = preconditionerRank / preconditionerRelError) fed with the blocks
   of the assembled H-matrix h (downloaded, FP32 blocks widened to FP64);
   clusters2 = (start, size) of every cluster of the cluster tree.  The
   factors are then flattened (bem_hlu_solve_flat's form: Q^H and R^-1 of
   every diagonal leaf, the off-diagonal blocks' row slices per leaf) and
   kept on the device; bem_precond_apply / bem_gmres_prec run the forward
   and backward substitution there (forwardSubstitution /
   backwardSubstitution, harithm.cpp:1314-1412; one workgroup walks the
   leaves), asynchronously on the context stream like the leaf-block LU.
   Diagonal leaves of more than 128 rows are refused (BEM_ERR_ARG). */
int bem_precond_create_hlu(bem_ctx_t ctx, bem_hmat_t h, int64_t nclusters, const int64_t* clusters2, int64_t max_rank,
                           double rel_err, bem_prec_t* out);
/* [0] set-up ms (host wall clock, incl. block download, factorisation and
   upload), [1] applications, [2] application ms in all (H-LU: device time
   of the substitutions; 0 for the leaf-block LU, whose applications are
   part of the GMRES iterations), [3] stored complex entries, [4]
   factorisation ms (host H-LU; 0 for the leaf-block LU) */
int bem_precond_stats(bem_prec_t p, double* out5);
/* Side of the preconditioning in bem_gmres_prec (and the inner solves of
   bem_gmres_ir) on this context: 0 (default) right, GMRES on A M^-1 with
   the true residual |b - A x| / |b| against tol; 1 left, the reference's
   gmresLUPreconditionedSolve (gmres.h:200-216): GMRES on M^-1 A x = M^-1 b
   from the given x, the preconditioned residual |M^-1 (b - A x)| /
   |M^-1 b| against tol (Eigen's criterion).  The reference starts it from
   x = M^-1 b (hMatrixSolve, boundaryelementsolver.cpp:689-692), as the
   solver does with bem_solver_set_precond_side. */
int bem_set_precond_side(bem_ctx_t ctx, int left);
int bem_gmres_prec(bem_ctx_t ctx, bem_hmat_t h, bem_prec_t p, const void* b_dev, void* x_dev, double tol,
                   int64_t restart, int64_t maxit, int64_t* iters, double* rel_resid);

/* boundary-condition reconstruction (boundaryelementsolver.cpp:729-740) and
   p = i rho 2 pi f phi (phiSolutionToSoundPressure, :2251-2261) */
int bem_reconstruct(bem_ctx_t ctx, const void* x_dev, double rho, double freq, void* phi_dev, void* dphi_dev,
                    void* pressure_dev);
/* bem_reconstruct with the boundary-condition right-hand side g taken from
   g_dev (n complex, device) instead of the one set by bem_set_bc: the
   reconstruction of a further right-hand side on the same a, b
   (boundaryelementsolver.cpp:729-740 with f = g_dev). */
int bem_reconstruct_g(bem_ctx_t ctx, const void* x_dev, const void* g_dev, double rho, double freq, void* phi_dev,
                      void* dphi_dev, void* pressure_dev);

/* Field potential of the boundary solution at observation points:
   out[i] = sum_j M_j(x_i) phi_j - L_j(x_i) dphi_j with the order-5 7-point
   rule (BemOperatorField and calculateFieldSolutionRegular,
   boundaryelementsolver.cpp:2262-2353; the H-matrix variant
   calculateFieldSolutionFast, :2397-2570, approximates this sum).  obs3:
   nobs x 3 host; phi_dev, dphi_dev: n complex on the device (cluster order);
   out2: nobs complex host.  Sources are not included. */
/* The boundary term of calculateFieldSolutionFast (boundaryelementsolver.cpp:
   2397-2476) on the GPU: obs3 (nobs x 3) the observation points in the
   order of their cluster tree, blocks5 (nblocks x (row0, rows, col0, cols,
   admissible)) the observation x element partition (minPartition order of
   HMatrix(fieldTree, boundaryTree, true)).  Admissible blocks: ACA of
   implicitPhiMatrixField (the M kernel, :2647-2683) and, with the pivots of
   getNormalFilteredPivotIndicesForField (:2721-2807, counter-based draws
   from `seed`), of implicitDPhiMatrixField (the L kernel, :2685-2719), to
   tol / max_rank; the block's term U A V phi - U' A' V' dphi right after
   (the blocks are not kept, as in the reference).  Dense blocks: M phi -
   L dphi directly.  Row sums in partition order; out2 in the points' order.
   No impedance plane (bem_field covers it).  stats4 as bem_solver_field_hmat. */
int bem_field_hmat(bem_ctx_t ctx, int64_t nobs, const double* obs3, int64_t nblocks, const int64_t* blocks5,
                   double tol, int64_t max_rank, uint64_t seed, const void* phi_dev, const void* dphi_dev,
                   double* out2, double* stats4);
int bem_field(bem_ctx_t ctx, int64_t nobs, const double* obs3, const void* phi_dev, const void* dphi_dev,
              double* out2);

/* No counterpart in the reference: bem_field for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc).  It allocates nothing: the
   per-chunk partial sums go to a fixed 32 MiB device array that the library
   holds on every GPU it runs on until the process ends, so nobs x
   ceil(n / 1024) x 16 bytes must fit in it (else BEM_ERR_ARG).  obs_dev:
   the points as three device rows (x[nobs], y[nobs], z[nobs]); phi_dev,
   dphi_dev: n complex; out_dev: nobs complex, all device memory.
   Asynchronous on the context stream (bem_synchronize waits for it).
   Every call on one GPU shares that array, so calls that overlap on one GPU
   (two contexts, two host threads) race; the array keeps the last call's
   contents. */
int bem_field_noalloc(bem_ctx_t ctx, int64_t nobs, const double* obs_dev, const void* phi_dev, const void* dphi_dev,
                      void* out_dev);

/* L, M, Mt, N (out npairs x 8) or Phi/DPhi entries (out npairs x 2) of
   (row, col) pairs given in host memory, evaluated on the device
   (BemOperatorsConst / implicitPhiMatrix2 / implicitDPhiMatrix2). */
int bem_entries4(bem_ctx_t ctx, int64_t npairs, const int64_t* ij, double* out8);
int bem_entries(bem_ctx_t ctx, int kind, int64_t npairs, const int64_t* ij, double* out2);

/* device memory helpers for callers without a HIP runtime binding */
int bem_malloc(bem_ctx_t ctx, size_t bytes, void** dev_ptr);
int bem_free(bem_ctx_t ctx, void* dev_ptr);
int bem_memcpy_htod(bem_ctx_t ctx, void* dst_dev, const void* src_host, size_t bytes);
int bem_memcpy_dtoh(bem_ctx_t ctx, void* dst_host, const void* src_dev, size_t bytes);
int bem_synchronize(bem_ctx_t ctx);
/* average device time (ms) of the last `bem_profile` window per kernel
   family: [0] dense assembly [1] singular/near-singular fix-up [2] ACA
   [3] MVM dense [4] MVM low-rank [5] MVM gather [6] field evaluation;
   counts in [7..13] */
int bem_profile(bem_ctx_t ctx, int enable, double* out14);

/* BoundaryElementSolver::hMatrixSolve (boundaryelementsolver.cpp:526-744) as
   one call: cluster tree + partition on the host, everything else on the
   device.  Mesh in input order: nodes n x 9, bc n x 6 (a, b, g complex).
   Non-finite vertex coordinates are refused (BEM_ERR_ARG) before the
   cluster tree is built or a device is touched. */
int bem_solver_create(bem_solver_t* out, int device, int64_t n, const double* nodes9, const double* bc6,
                      int64_t leaf, double eta);
int bem_solver_destroy(bem_solver_t s);
const char* bem_solver_last_error(bem_solver_t s);
bem_ctx_t bem_solver_context(bem_solver_t s);
/* perm[new] = old; blocks nb x 5 (row0, rows, col0, cols, admissible) */
int bem_solver_layout(bem_solver_t s, int64_t* perm, int64_t* nblocks, int64_t* blocks5, int64_t cap);
/* point sources (x, y, z, weight re, im) and an optional plane wave
   (dir x, y, z, amplitude re, im) for the source-term vector */
int bem_solver_set_sources(bem_solver_t s, int64_t npoint, const double* points6, int has_plane,
                           const double* plane5);
/* one frequency point.  coupling: 0 none, 1 Burton-Miller, 2 Kirkup.
   Outputs (cluster order, complex n each): phi, dphi, pressure.
   stats[0..15]: t_phi_asm, t_rhs, t_dphi_asm, t_gmres, t_total (ms),
   gmres iterations, true relative residual, max rank phi, max rank dphi,
   evaluated phi, evaluated dphi, near blocks, far blocks, preconditioner
   set-up ms, preconditioner application ms (all applications of the solve),
   bytes of the DPhi block storage. */
int bem_solver_solve(bem_solver_t s, double freq, double c, double rho, int coupling, double aca_tol,
                     int64_t max_rank, double gmres_tol, int64_t restart, uint64_t seed, int precision,
                     double* phi, double* dphi, double* pressure, double* stats16);

/* Several meshes solved together (no counterpart in the reference, which
   runs them one after another): solver ss[q] at freqs[q], every other
   parameter shared (and the first solver's preconditioner and refinement
   settings); Phi and DPhi of all meshes from one bem_assemble_multi each,
   then the GMRES solves on up to `threads` host threads (each solver its
   own context and streams).  pressure[q] (n_q complex, cluster order, as
   bem_solver_solve), iters[q], stats16 (16 per solver) optional.  No
   impedance plane.  Errors are reported by ss[0]. */
int bem_solvers_solve_multi(const bem_solver_t* ss, int64_t ns, const double* freqs, double c, double rho, int coupling,
                            double aca_tol, int64_t max_rank, double gmres_tol, int64_t restart, uint64_t seed,
                            int precision, int threads, double* const* pressure, int64_t* iters, double* stats16);

/* regularSolveWithLU (boundaryelementsolver.cpp:140-331), the reference's
   dense path: every entry of Phi and DPhi in FP64 (the same entry
   generator as the H-matrix assembly) on a context of its own, b = s + Phi f,
   the dense system solved by GMRES to gmres_tol (the reference factorises
   it by partial-pivoting LU), reconstruction and pressure as
   bem_solver_solve.  Needs 2 n^2 x 16 bytes of device memory (13.4 GB at
   n = 20480).  stats16: [0] assembly ms, [3] GMRES ms, [4] total ms,
   [5] iterations, [6] GMRES relative residual, the rest 0. */
int bem_solver_solve_dense(bem_solver_t s, double freq, double c, double rho, int coupling, double gmres_tol,
                           int64_t restart, double* phi, double* dphi, double* pressure, double* stats16);

/* GMRES preconditioner of bem_solver_solve / bem_solver_sweep: 0 none (the
   default), 1 the leaf-block LU of bem_precond_create, 2 the H-LU of
   bem_precond_create_hlu (usePreconditioner, bemparameters.h:26) */
int bem_solver_set_preconditioner(bem_solver_t s, int kind);
/* 1: GMRES with the preconditioner as the reference's
   gmresLUPreconditionedSolve (gmres.h:200-216): left preconditioning from
   x = M^-1 b, the tolerance on the preconditioned residual (the true
   residual is still reported in stats[6]); 0 (default): right
   preconditioning from 0, the tolerance on the true residual */
int bem_solver_set_precond_side(bem_solver_t s, int left);
/* the H-LU preconditioner with preconditionerRank / preconditionerRelError
   (bemparameters.h:27-28; defaults 2 / 0) */
int bem_solver_set_hlu(bem_solver_t s, int64_t rank, double rel_err);
/* bem_precond_stats of the last bem_solver_solve's preconditioner (zeros
   when it ran without one) */
int bem_solver_precond_stats(bem_solver_t s, double* out5);
/* Refinement of the FP32 form (no counterpart in the reference): mode 0
   none (the default), 1 residuals with the FP32 operators' own FP64 near
   field and their low-rank factors in FP64 accumulation (bem_set_near_fp64;
   b = s + Phi g with the same operator), 2 residuals with FP64 H-matrices
   assembled beside the FP32 ones (the plain path's Phi and DPhi); GMRES
   iterates on the FP32 DPhi inside bem_gmres_ir, at most max_steps
   refinement steps.  FP64 solves ignore it.  bem_solver_refine_stats of
   the last solve: [0] refinement steps, [1] inner GMRES iterations, [2]
   final |b - A_res x| / |b|, [3] ms assembling the FP64 operators of mode 2. */
int bem_solver_set_refinement(bem_solver_t s, int mode, int64_t max_steps);
int bem_solver_refine_stats(bem_solver_t s, double* out4);

/* The scene's impedance plane (BoundaryElements::impedancePlanes,
   boundaryelements.h:97, filled from an infinite-baffle section,
   solvingscriptreader.cpp:1070-1077): later solves and field evaluations use
   the mirror image as bem_set_plane describes.  normal3 NULL or zero removes
   it.  The sources are not mirrored (calculateBoundaryConditionAndSourceTermVector
   and sourcePhiTerm use the free-field incident field, as the reference). */
int bem_solver_set_impedance_plane(bem_solver_t s, const double* point3, const double* normal3);
/* Sound pressure at observation points (nobs x 3, host) from the last
   bem_solver_solve: the boundary term of bem_field plus the incident point
   sources and plane wave, times i rho 2 pi f (calculateFieldSolution with
   the direct sum, boundaryelementsolver.cpp:2192-2316).  pressure2: nobs
   complex. */
int bem_solver_field(bem_solver_t s, int64_t nobs, const double* obs3, double* pressure2);
/* calculateFieldSolutionFast (boundaryelementsolver.cpp:2397-2476, the
   reference's default field path, bemparameters.h:29-30): the sound
   pressure at nobs points (input order) from the last boundary solution,
   the boundary term through an observation x boundary H-matrix (cluster
   tree over the points with leaf size `leaf`, admissibility eta, ACA to
   aca_tol / max_rank; bem_field_hmat) plus the incident sources.  No
   impedance plane.  stats4 (optional): ms, admissible blocks, largest rank,
   evaluated entries. */
int bem_solver_field_hmat(bem_solver_t s, int64_t nobs, const double* obs3, double aca_tol, int64_t max_rank,
                          int64_t leaf, double eta, uint64_t seed, double* pressure2, double* stats4);

/* Boundary mesh from a GMSH .msh file (mshReader::readMsh,
   mshreader.cpp:10-428): ASCII format 2 (as the reference parses) or 4.1.
   Triangles, and quadrilaterals split along their shorter diagonal into
   (1, 2, 4) + (2, 3, 4) or (4, 1, 3) + (1, 2, 3) (global::
   quadrilateralsToTriangles; the reference's aspect-ratio splitting of
   quadrilaterals before it is not applied), of the
   included groups: include_csv / exclude_csv are comma-separated entries
   "a...b" (geometrical group range, either end may be empty), a number
   (geometrical group) or a physical group name; "all" or an empty include
   list includes everything except the excluded groups.  With nodes9 NULL
   only *ntri is set; otherwise nodes9 (cap x 9) receives the triangles in
   file order.  Errors: BEM_ERR_ARG (file, format, cap too small; the reason
   in bem_read_msh_error). */
int bem_read_msh(const char* path, const char* include_csv, const char* exclude_csv, int64_t cap, double* nodes9,
                 int64_t* ntri);
const char* bem_read_msh_error(void);
/* Edge-length refinement of a triangle mesh (ElementSection::refineTriangles
   / recursiveRefineTriangle, SolvingScript/elementSection.cpp:688-800, as
   SolvingScriptReader::refineElements applies it, solvingscriptreader.cpp:
   1104-1128, with EdgeLength = the section's edge length, or c / (2
   MeshFrequency)): every triangle whose longest edge exceeds edge_length is
   bisected at that edge's midpoint, recursively, the children in the
   parent's orientation and in depth-first order.  tri9: n x 9 input; out9:
   cap x 9 (NULL: only *nout is set). */
int bem_refine_triangles(int64_t n, const double* tri9, double edge_length, int64_t cap, double* out9, int64_t* nout);

/* Frequency sweep with several right-hand sides per frequency
   (BemSolverController::run's loop, bemsolvercontroller.cpp:59-79, with the
   per-frequency hMatrixSolve of boundaryelementsolver.cpp:526-744).  For
   every frequency Phi and DPhi are assembled once and GMRES runs once per
   right-hand side; right-hand side k has the boundary data g_k (nrhs x n
   complex, input element order; NULL: the g given at creation for every k)
   and the plane wave planes5[k] (dir x, y, z, amplitude re, im; zero
   amplitude: none; planes5 NULL: the plane wave of bem_solver_set_sources,
   if any) on the a, b given at creation; the point sources of
   bem_solver_set_sources belong to every right-hand side.  `lanes` solver copies on
   their own streams and host threads process the frequencies concurrently
   (frequency f on lane f % lanes); every frequency point is computed exactly
   as by bem_solver_solve.  Outputs: pressure nf x nrhs x n complex (cluster
   order), iters nf x nrhs, stats nf x 16 (bem_solver_solve's layout, for the
   last right-hand side). */
/* batch > 1: bem_solver_sweep processes the frequencies in chunks of
   min(batch, 8 / nrhs): Phi and DPhi of a chunk from one batched assembly
   each (bem_assemble_batch), the GMRES of all its (frequency, right-hand
   side) pairs in one lockstep solve (bem_gmres_batch); the sweep's lanes
   take whole chunks (one lane: no extra hardware queues needed); nrhs 1, 2
   or 4, no preconditioner (an impedance plane is taken).  0 or 1: one
   point per lane at a time. */
int bem_solver_set_sweep_batch(bem_solver_t s, int batch);
/* on: the first batched sweep of each precision first runs
   bem_autotune_aca on its first chunk (DPhi, the sweep's parameters, timed
   alone on the solver's context) and every lane then assembles with the
   chosen schedule; off: the built-in schedule, tuning forgotten.
   bem_solver_aca_schedule: small bound, huge bound, queue, tuned ms of a
   precision's chosen schedule (all -1 before tuning). */
int bem_solver_set_autotune(bem_solver_t s, int on);
int bem_solver_aca_schedule(bem_solver_t s, int precision, double* out4);
/* a known ACA schedule for a precision's sweeps (the bounds and queue of
   bem_set_aca_schedule), taken as tuned: no tuning run, every lane uses it
   (the tuned ms reads 0); bem_solver_set_autotune(s, 0) forgets it */
int bem_solver_set_aca_schedule(bem_solver_t s, int precision, int64_t small_bound, int64_t huge_bound, int queue);
/* bem_profile over the solver's context and every lane's: the windows of
   all of them merged (average ms and launch count per kernel family, the
   layout of bem_profile); enable as bem_profile, applied to each */
int bem_solver_profile(bem_solver_t s, int enable, double* out14);
int bem_solver_sweep(bem_solver_t s, int64_t nf, const double* freqs, double c, double rho, int coupling,
                     double aca_tol, int64_t max_rank, double gmres_tol, int64_t restart, uint64_t seed, int precision,
                     int64_t nrhs, const double* g, const double* planes5, int lanes, double* pressure,
                     int64_t* iters, double* stats16);

#ifdef __cplusplus
}
#endif
#endif
