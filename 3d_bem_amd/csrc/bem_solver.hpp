// C++ mirror of the reference's solver-side API for the H-matrix BEM path.
// Class and method names follow the reference so that its callers read the
// same (boundaryelementsolver.h, HMatrix/clustertree.h, HMatrix/hmatrix.h,
// HMatrix/harithm.h, HMatrix/gmres.h); underneath, every operator entry,
// block, product and Krylov step runs on the GPU through include/bem_hip.h.
#pragma once
#include <complex>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/bem_hip.h"
#include "bem_setup.hpp"

namespace bem {

using cplx = std::complex<double>;

// robinboundarycondition.h: a phi + b dphi/dn = g
struct RobinBoundaryCondition {
    cplx a{0.0, 0.0}, b{1.0, 0.0}, g{0.0, 0.0};
};

// impedanceplane.h:10-28: a plane (point, normal) the solve mirrors the mesh in
struct ImpedancePlane {
    double halfSpacePlanePoint[3] = {0, 0, 0};
    double halfSpacePlaneNormal[3] = {0, 0, 0};
};

// boundaryelements.h:35-170 (triangles + derived geometry), SoA here
struct BoundaryElements {
    int64_t n = 0;
    std::vector<double> nodes;                      // n x 9
    std::vector<RobinBoundaryCondition> bc;          // n
    std::vector<double> normal, colloc, area;        // n x 3, n x 3, n
    double averageTriangleDim = 0;
    std::vector<int64_t> elementIndex;               // input index of each element
    std::vector<ImpedancePlane> impedancePlanes;     // boundaryelements.h:97 (at most one on this path)
    void prepare();                                  // prepareBoundaryElements (:3463-3468)
};

// HMatrix/clustertree.h:77-86: builds the tree and reorders the elements
class ClusterTree {
public:
    explicit ClusterTree(BoundaryElements* elements, int64_t triangleThreshold = 50);
    const std::vector<bemsetup::Cluster>& clusters() const { return tree_.clusters; }
    const std::vector<int64_t>& permutation() const { return tree_.perm; }
    bemsetup::ClusterTree tree_;
};

// A device-resident complex vector (Eigen::VectorXcd's role in the reference)
class DeviceVector {
public:
    DeviceVector(bem_ctx_t ctx, int64_t n);
    ~DeviceVector();
    DeviceVector(const DeviceVector&) = delete;
    DeviceVector& operator=(const DeviceVector&) = delete;
    void upload(const std::vector<cplx>& h);
    std::vector<cplx> download() const;
    void* data() const { return p_; }
    int64_t size() const { return n_; }
private:
    bem_ctx_t ctx_;
    int64_t n_;
    void* p_ = nullptr;
};

// HMatrix/hmatrix.h:157-260 -- blocks live on the GPU
class HMatrix {
public:
    HMatrix(bem_ctx_t ctx) : ctx_(ctx) {}
    ~HMatrix() { clear(); }
    void setClusterTrees(std::shared_ptr<ClusterTree> rows, std::shared_ptr<ClusterTree> cols);
    // takes an operator assembled elsewhere on this context (bem_assemble_multi)
    void adopt(bem_hmat_t h);
    // HMatrix::populateBlockClusterTree (hmatrix.cpp:922-1007); admissibility
    // constant eta (reference constexpr 3, hmatrix.h:240)
    void populateBlockClusterTree(double eta = 3.0);
    // assembleBlocks with the BEM Phi operator (implicitPhiMatrix2)
    void assembleBlocks(long rank, double relativeError, uint64_t seed, int precision = BEM_PREC_FP64);
    // assembleBlocksExtra with the BEM DPhi operator (implicitDPhiMatrix2) and
    // getNormalFilteredPivotIndices
    void assembleBlocksExtra(long rank, double relativeError, uint64_t seed, int precision = BEM_PREC_FP64);
    void clear();
    long calculateMaxBlockRank() const;
    int64_t evaluatedEntries() const;
    const std::vector<bemsetup::Block>& getMinPartition() const { return partition_; }
    bem_hmat_t handle() const { return h_; }
    bem_ctx_t context() const { return ctx_; }
    double stat(int q) const;
private:
    void assemble(int kind, long rank, double relErr, uint64_t seed, int precision);
    bem_ctx_t ctx_;
    bem_hmat_t h_ = nullptr;
    std::shared_ptr<ClusterTree> rows_, cols_;
    std::vector<bemsetup::Block> partition_;
};

// HMatrix/harithm.h:32: y += H x
struct HArithm {
    static void MVM(DeviceVector& y, HMatrix& h, const DeviceVector& x);
};

// HMatrix/gmres.h:170-180
struct GMRES {
    static int64_t gmresSolve(HMatrix& A, DeviceVector& x, const DeviceVector& b, double tolerance = 0.01,
                              long iterations = 300, double* relResidual = nullptr, bem_prec_t precond = nullptr);
};

enum BemCoupling : int { NoCoupling = 0, BurtonMillerCoupling = 1, KirkupCoupling = 2 };

// bemparameters.h:21-34 (the fields this path uses)
struct BemSolverParameters {
    double acaRelativeError = 0.01;
    unsigned long acaMaxRank = 0;
    BemCoupling coupling = KirkupCoupling;
    double gmresTolerance = -1;   // < 0: 1e-3 * acaRelativeError (boundaryelementsolver.cpp:670)
    long gmresRestart = 300;
    uint64_t pivotSeed = 1;
    int precision = BEM_PREC_FP64;
    int preconditioner = 0;       // 0 none, 1 leaf-block LU (bem_precond_create), 2 H-LU (bem_precond_create_hlu)
    long preconditionerRank = 2;  // bemparameters.h:27
    double preconditionerRelError = 0;   // bemparameters.h:28
    // left preconditioning from x = M^-1 b, tolerance on the preconditioned
    // residual (gmresLUPreconditionedSolve, gmres.h:200-216); false: right
    // preconditioning from 0, tolerance on the true residual
    bool preconditionLeft = false;
    // FP32 operators only: iterative refinement of the GMRES solution against
    // an FP64-accurate operator (bem_gmres_ir): 0 none, 1 the FP32
    // operators' own FP64 near field + their low-rank factors
    // (bem_set_near_fp64), 2 FP64 H-matrices assembled beside them (the
    // plain path's operators)
    int refine = 0;
    long refineMaxSteps = 10;
    // 2..8 right-hand sides without a preconditioner: one lockstep GMRES
    // (bem_gmres_multi) instead of one GMRES per right-hand side
    bool lockstepRhs = true;
};

struct PointSource { double x, y, z; cplx weight; };
struct PlaneWave { double dx, dy, dz; cplx amplitude; };

// boundaryelementsolver.h: the H-matrix solve of one frequency point
class BoundaryElementSolver {
public:
    BoundaryElementSolver(int device, BoundaryElements elements, int64_t leaf, double eta);
    ~BoundaryElementSolver();
    void setFrequency(double f) { frequency = f; }
    // the context's image plane from boundaryElements.impedancePlanes (none or
    // one: hMatrixSolve's branches, boundaryelementsolver.cpp:549-557)
    void applyImpedancePlanes();
    // the GMRES preconditioner of bemSolverParameters.preconditioner for the
    // assembled DPhi (nullptr for none)
    bem_prec_t makePreconditioner();
    void setWaveSpeed(double c) { waveSpeed = c; }
    void setAirDensity(double rho) { airDensity = rho; }
    void calculateBoundarySolution();      // boundaryelementsolver.cpp:12-138 (H-matrix branch)
    // regularSolveWithLU (boundaryelementsolver.cpp:140-331): every entry of
    // Phi and DPhi assembled densely in FP64 on the device (a one-block
    // partition on a context of its own), b = s + Phi f, and the dense system
    // solved by GMRES to gmresTol instead of the reference's partial-pivoting
    // LU; fills phiSolution, dPhiSolution, soundPressure like
    // calculateBoundarySolution.  Holds two n x n complex matrices.
    void regularSolve(double gmresTol, long restart);
    // One right-hand side of a multi-right-hand-side solve on the same a, b:
    // boundary data g (cluster order) and its own sources.
    struct RightHandSide {
        std::vector<cplx> g;
        std::vector<PointSource> pointSources;
        std::vector<PlaneWave> planeWaves;
    };
    // calculateBoundarySolution for several right-hand sides: Phi and DPhi
    // assembled once, b_k = s_k + Phi g_k, GMRES and reconstruction per k;
    // each right-hand side's pressure is the same as calculateBoundarySolution
    // gives for it alone
    void calculateBoundarySolutions(const std::vector<RightHandSide>& rhs, std::vector<std::vector<cplx>>& pressure,
                                    std::vector<int64_t>& iterations);
    // calculateBoundarySolutions for several frequencies at once: Phi and DPhi
    // of all of them from one batched assembly each (bem_assemble_batch), the
    // GMRES of every (frequency, right-hand side) pair in one lockstep solve
    // (bem_gmres_batch); freqs.size() * rhs.size() <= 8, rhs.size() 1, 2 or 4,
    // no preconditioner, no impedance plane.  pressure[f][k], iterations[f][k];
    // stats (16 per frequency) carry the batch's times shared out evenly.
    void calculateBoundarySolutionsBatch(const std::vector<double>& freqs, const std::vector<RightHandSide>& rhs,
                                         std::vector<std::vector<std::vector<cplx>>>& pressure,
                                         std::vector<std::vector<int64_t>>& iterations, std::vector<double>& stats16);
    // bem_autotune_aca of the DPhi assembly at these frequencies with the
    // current parameters; the chosen schedule stays on this solver's context
    void tuneAca(const std::vector<double>& freqs, int64_t choice[3], double* ms);
    // calculateFieldSolution with the direct sum (calculateFieldSolutionRegular,
    // boundaryelementsolver.cpp:2192-2316): sound pressure at observation
    // points from the last boundary solution, incident sources included
    std::vector<cplx> calculateFieldSolution(const std::vector<double>& observationPoints3);
    // calculateFieldSolutionFast (boundaryelementsolver.cpp:2397-2476): the
    // boundary term through the observation x boundary H-matrix -- a cluster
    // tree over the observation points (leaf; the reference's is 50), the
    // admissible blocks as ACA of the M kernel times phi minus ACA (normal-
    // filtered pivots) of the L kernel times dphi, relativeError / maxRank
    // (fieldACARelativeError, default 0.01, bemparameters.h:29-30), the
    // dense blocks directly -- on the GPU (bem_field_hmat); incident sources
    // and pressure as calculateFieldSolution.  No impedance plane.
    // stats4: ms, admissible blocks, largest rank, evaluated entries.
    std::vector<cplx> calculateFieldSolutionFast(const std::vector<double>& observationPoints3, double relativeError,
                                                 long maxRank, int64_t leaf, double eta, uint64_t seed,
                                                 double* stats4);
    // the incident field of the sources added to the boundary term, times
    // i rho 2 pi f: the sound pressure at the points
    std::vector<cplx> fieldToPressure(const std::vector<double>& observationPoints3,
                                      const std::vector<cplx>& field) const;
    BoundaryElements boundaryElements;     // cluster order after construction
    BemSolverParameters bemSolverParameters;
    std::vector<PointSource> pointSources;
    std::vector<PlaneWave> planeWaves;
    std::vector<cplx> phiSolution, dPhiSolution, soundPressure;
    // The next solve's Phi / DPhi already assembled on this solver's context
    // (solveMany): taken instead of assembling, once.
    bem_hmat_t presetPhi = nullptr, presetDPhi = nullptr;
    // BemSolverController::run over several meshes: each solver at its own
    // frequency, Phi and DPhi of all of them from one multi-mesh assembly
    // each (bem_assemble_multi; the solvers' parameters, seed and precision
    // of the first), then the solves on up to `threads` host threads.
    // Refinement mode 2 assembles its FP64 operators per solver.
    static void solveMany(const std::vector<BoundaryElementSolver*>& solvers, const std::vector<double>& freqs,
                          int threads);
    double frequency = 0, waveSpeed = 343.0, airDensity = 1.2041;
    double stats[16] = {0};
    double precondStats[5] = {0};   // bem_precond_stats of the last solve's preconditioner
    // refinement of the last solve: steps, inner GMRES iterations, final
    // |b - A_res x| / |b|, ms spent assembling FP64 residual operators (mode 2)
    double refineStats[4] = {0};
    bem_ctx_t context() const { return ctx_; }
    const std::vector<bemsetup::Block>& partition() const { return phiHMat->getMinPartition(); }
    std::shared_ptr<ClusterTree> clusterTree;
private:
    void hMatrixSolve();
    void setWavenumberAndCoupling();
    void calculateBoundaryConditionAndSourceTermVector();
    bem_ctx_t ctx_ = nullptr;
    int device_ = 0;
    double wavenumber = 0;
    cplx couplingParameter{0, 0};
    std::vector<cplx> sourceTermVector;
    std::unique_ptr<HMatrix> phiHMat, dPhiHMat, dPhiResHMat;
    double eta_;
    // b (+)= Phi g with the operator the refinement measures residuals with
    void rhsProduct(DeviceVector& b, const DeviceVector& g);
    // the right-hand side operator and the residual DPhi of the refinement
    // mode (before the DPhi the GMRES iterations use)
    void assembleRhsOperator();
    void assembleResidualOperator();
    void solveLockstep(const std::vector<RightHandSide>& rhs, const std::vector<std::vector<cplx>>& s, bool anyG,
                       double t0, std::vector<std::vector<cplx>>& pressure, std::vector<int64_t>& iterations);
    // GMRES (refine 0) or GMRES with refinement; returns inner iterations
    int64_t solveSystem(DeviceVector& x, const DeviceVector& b, double tol, bem_prec_t pc, double* rel);
};

struct Error : std::exception {
    int code;
    std::string msg;
    Error(int c, std::string m) : code(c), msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

}  // namespace bem
