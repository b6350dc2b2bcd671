// Host mirror of the reference solver (see bem_solver.hpp) and the
// solver-level C ABI (bem_solver_*).  Compiled with -ffp-contract=off.
#include "bem_solver.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

#include "../../include/bem_detmath.h"

namespace bem {
namespace {

double bem_hmat_stats_value(bem_hmat_t h, int q)
{
    double s[9] = {0};
    if (h) bem_hmat_stats(h, s);
    return s[q];
}

void check(bem_ctx_t ctx, int rc, const char* what)
{
    if (rc != BEM_OK) throw Error(rc, std::string(what) + ": " + bem_last_error(ctx));
}

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

void BoundaryElements::prepare()
{
    bemsetup::Elements E;
    E.n = n;
    E.nodes = nodes;
    bemsetup::prepare_elements(E);
    normal.resize(3 * n);
    colloc.resize(3 * n);
    area = E.area;
    for (int64_t e = 0; e < n; e++) {
        normal[3 * e] = E.normal[e].x; normal[3 * e + 1] = E.normal[e].y; normal[3 * e + 2] = E.normal[e].z;
        colloc[3 * e] = E.colloc[e].x; colloc[3 * e + 1] = E.colloc[e].y; colloc[3 * e + 2] = E.colloc[e].z;
    }
    averageTriangleDim = E.hbar;
}

ClusterTree::ClusterTree(BoundaryElements* be, int64_t leaf)
{
    tree_ = bemsetup::build_cluster_tree(be->n, be->nodes.data(), leaf);
    // ClusterTree::recursiveReorder (clustertree.cpp:265-283)
    std::vector<double> nodes(be->nodes.size());
    std::vector<RobinBoundaryCondition> bc(be->n);
    std::vector<int64_t> idx(be->n);
    for (int64_t e = 0; e < be->n; e++) {
        int64_t o = tree_.perm[e];
        std::memcpy(&nodes[9 * e], &be->nodes[9 * o], 9 * sizeof(double));
        bc[e] = be->bc[o];
        idx[e] = be->elementIndex.empty() ? o : be->elementIndex[o];
    }
    be->nodes.swap(nodes);
    be->bc.swap(bc);
    be->elementIndex.swap(idx);
    be->prepare();
}

DeviceVector::DeviceVector(bem_ctx_t ctx, int64_t n) : ctx_(ctx), n_(n)
{
    check(ctx_, bem_malloc(ctx_, (size_t)n * sizeof(cplx), &p_), "DeviceVector");
}
DeviceVector::~DeviceVector()
{
    if (p_) bem_free(ctx_, p_);
}
void DeviceVector::upload(const std::vector<cplx>& h)
{
    if ((int64_t)h.size() != n_) throw Error(BEM_ERR_ARG, "DeviceVector::upload: size");
    check(ctx_, bem_memcpy_htod(ctx_, p_, h.data(), n_ * sizeof(cplx)), "DeviceVector::upload");
}
std::vector<cplx> DeviceVector::download() const
{
    std::vector<cplx> h(n_);
    check(ctx_, bem_memcpy_dtoh(ctx_, h.data(), p_, n_ * sizeof(cplx)), "DeviceVector::download");
    return h;
}

void HMatrix::setClusterTrees(std::shared_ptr<ClusterTree> rows, std::shared_ptr<ClusterTree> cols)
{
    rows_ = std::move(rows);
    cols_ = std::move(cols);
}

void HMatrix::populateBlockClusterTree(double eta)
{
    if (!rows_ || rows_ != cols_) throw Error(BEM_ERR_STATE, "populateBlockClusterTree: one cluster tree expected");
    partition_ = bemsetup::build_partition(rows_->tree_, eta);
    std::vector<int64_t> r0, rs, c0, cs;
    std::vector<uint8_t> adm;
    for (auto& B : partition_) {
        r0.push_back(B.row0); rs.push_back(B.rows); c0.push_back(B.col0); cs.push_back(B.cols);
        adm.push_back(B.adm ? 1 : 0);
    }
    check(ctx_, bem_upload_partition(ctx_, (int64_t)partition_.size(), r0.data(), rs.data(), c0.data(), cs.data(),
                                     adm.data()),
          "populateBlockClusterTree");
}

void HMatrix::assemble(int kind, long rank, double relErr, uint64_t seed, int precision)
{
    clear();
    check(ctx_, bem_assemble(ctx_, kind, relErr, rank, precision, seed, &h_), "assembleBlocks");
}
void HMatrix::assembleBlocks(long rank, double relErr, uint64_t seed, int precision)
{
    assemble(BEM_KIND_PHI, rank, relErr, seed, precision);
}
void HMatrix::assembleBlocksExtra(long rank, double relErr, uint64_t seed, int precision)
{
    assemble(BEM_KIND_DPHI, rank, relErr, seed, precision);
}
void HMatrix::adopt(bem_hmat_t h)
{
    clear();
    h_ = h;
}
void HMatrix::clear()
{
    if (h_) bem_hmat_destroy(h_);
    h_ = nullptr;
}
double HMatrix::stat(int q) const
{
    double s[9] = {0};
    if (h_) bem_hmat_stats(h_, s);
    return s[q];
}
long HMatrix::calculateMaxBlockRank() const { return (long)stat(1); }
int64_t HMatrix::evaluatedEntries() const { return (int64_t)stat(0); }

void HArithm::MVM(DeviceVector& y, HMatrix& h, const DeviceVector& x)
{
    check(h.context(), bem_hmvm(h.context(), h.handle(), x.data(), y.data(), 1), "HArithm::MVM");
    check(h.context(), bem_synchronize(h.context()), "HArithm::MVM");
}

int64_t GMRES::gmresSolve(HMatrix& A, DeviceVector& x, const DeviceVector& b, double tol, long restart, double* rel,
                          bem_prec_t precond)
{
    int64_t its = 0;
    double r = 0;
    int rc = bem_gmres_prec(A.context(), A.handle(), precond, b.data(), x.data(), tol, restart, 100000, &its, &r);
    if (rc != BEM_OK && rc != BEM_ERR_NOCONV) check(A.context(), rc, "GMRES::gmresSolve");
    if (rel) *rel = r;
    return its;
}

BoundaryElementSolver::BoundaryElementSolver(int device, BoundaryElements elements, int64_t leaf, double eta)
    : boundaryElements(std::move(elements)), device_(device), eta_(eta)
{
    int rc = bem_create(&ctx_, device);
    if (rc != BEM_OK) throw Error(rc, "bem_create failed (no HIP device?)");
    try {
        clusterTree = std::make_shared<ClusterTree>(&boundaryElements, leaf);
        auto& E = boundaryElements;
        check(ctx_, bem_upload_mesh(ctx_, E.n, E.nodes.data(), E.normal.data(), E.area.data(), E.colloc.data(),
                                    E.averageTriangleDim),
              "bem_upload_mesh");
        applyImpedancePlanes();
        check(ctx_, bem_set_near_pairing(ctx_, 1), "bem_set_near_pairing");
        std::vector<cplx> a(E.n), b(E.n), g(E.n);
        for (int64_t e = 0; e < E.n; e++) { a[e] = E.bc[e].a; b[e] = E.bc[e].b; g[e] = E.bc[e].g; }
        check(ctx_, bem_set_bc(ctx_, (const double*)a.data(), (const double*)b.data(), (const double*)g.data()),
              "bem_set_bc");
        phiHMat = std::make_unique<HMatrix>(ctx_);
        dPhiHMat = std::make_unique<HMatrix>(ctx_);
        phiHMat->setClusterTrees(clusterTree, clusterTree);
        dPhiHMat->setClusterTrees(clusterTree, clusterTree);
        phiHMat->populateBlockClusterTree(eta_);
    } catch (...) {
        phiHMat.reset();
        dPhiHMat.reset();
        bem_destroy(ctx_);
        ctx_ = nullptr;
        throw;
    }
}

void BoundaryElementSolver::applyImpedancePlanes()
{
    const auto& P = boundaryElements.impedancePlanes;
    if (P.size() > 1)
        throw Error(BEM_ERR_ARG, "more than one impedance plane: the solve handles none or one "
                                 "(hMatrixSolve, boundaryelementsolver.cpp:549-557)");
    if (P.empty()) check(ctx_, bem_set_plane(ctx_, nullptr, nullptr), "bem_set_plane");
    else check(ctx_, bem_set_plane(ctx_, P[0].halfSpacePlanePoint, P[0].halfSpacePlaneNormal), "bem_set_plane");
}

bem_prec_t BoundaryElementSolver::makePreconditioner()
{
    const auto& P = bemSolverParameters;
    bem_prec_t pc = nullptr;
    if (P.preconditioner == 1) {
        check(ctx_, bem_precond_create(ctx_, dPhiHMat->handle(), &pc), "bem_precond_create");
    } else if (P.preconditioner == 2) {
        // dPhiHMatCopy + HArithm::LUDecomposition (boundaryelementsolver.cpp:622-643)
        std::vector<int64_t> cl;
        for (const auto& C : clusterTree->clusters()) {
            cl.push_back(C.start);
            cl.push_back(C.size);
        }
        check(ctx_, bem_precond_create_hlu(ctx_, dPhiHMat->handle(), (int64_t)cl.size() / 2, cl.data(),
                                           (int64_t)P.preconditionerRank, P.preconditionerRelError, &pc),
              "bem_precond_create_hlu");
    }
    return pc;
}

BoundaryElementSolver::~BoundaryElementSolver()
{
    phiHMat.reset();
    dPhiHMat.reset();
    if (ctx_) bem_destroy(ctx_);
}

// calculateBoundaryConditionAndSourceTermVector + sourceTerm
// (boundaryelementsolver.cpp:3386-3441; the point-source term keeps the
// reference's division by (ikr - 1)); a plane wave A e^{ik d.x} contributes
// phi_inc + cc ik (d.n) phi_inc
void BoundaryElementSolver::calculateBoundaryConditionAndSourceTermVector()
{
    auto& E = boundaryElements;
    sourceTermVector.assign(E.n, cplx(0, 0));
    const bem_c64 cc = c64(couplingParameter.real(), couplingParameter.imag());
    for (int64_t i = 0; i < E.n; i++) {
        double x = E.colloc[3 * i], y = E.colloc[3 * i + 1], z = E.colloc[3 * i + 2];
        double nx = E.normal[3 * i], ny = E.normal[3 * i + 1], nz = E.normal[3 * i + 2];
        bem_c64 acc = c64(0, 0);
        for (auto& s : pointSources) {
            double Rx = x - s.x, Ry = y - s.y, Rz = z - s.z;
            double r = std::sqrt(fma(Rz, Rz, fma(Ry, Ry, Rx * Rx)));
            double rup = fma(Rz, nz, fma(Ry, ny, Rx * nx)) / r;
            bem_c64 G = c64_mul(c64(s.weight.real(), s.weight.imag()),
                                c64_scale(det_expi(wavenumber * r), 1.0 / (4.0 * BEM_PI * r)));
            bem_c64 dG = c64_div(c64_scale(G, 1.0 / r), c64(-1.0, wavenumber * r));
            acc = c64_add(acc, c64_add(G, c64_mul(cc, c64_scale(dG, rup))));
        }
        for (auto& w : planeWaves) {
            double dx = fma(w.dz, z, fma(w.dy, y, w.dx * x));
            double dn = fma(w.dz, nz, fma(w.dy, ny, w.dx * nx));
            bem_c64 phi = c64_mul(c64(w.amplitude.real(), w.amplitude.imag()), det_expi(wavenumber * dx));
            bem_c64 dphi = c64_mul(c64(0.0, wavenumber * dn), phi);
            acc = c64_add(acc, c64_add(phi, c64_mul(cc, dphi)));
        }
        sourceTermVector[i] = cplx(acc.re, acc.im);
    }
}

void BoundaryElementSolver::calculateBoundarySolution()
{
    setWavenumberAndCoupling();
    hMatrixSolve();
}

// boundaryelementsolver.cpp:12-59: k = 2 pi f / c and the coupling parameter
void BoundaryElementSolver::setWavenumberAndCoupling()
{
    if (!(frequency > 0)) throw Error(BEM_ERR_ARG, "frequency must be > 0 (the f = 0 Laplace branch is not on this path)");
    wavenumber = (2.0 * frequency * BEM_PI) / waveSpeed;
    switch (bemSolverParameters.coupling) {
    case NoCoupling: couplingParameter = cplx(0, 0); break;
    case BurtonMillerCoupling: couplingParameter = cplx(0.0, 1.0 / wavenumber); break;
    case KirkupCoupling: {
        bem_c64 q = c64_div(c64(0, 1), c64(wavenumber + 1.0, 0));
        couplingParameter = cplx(q.re, q.im);
        break;
    }
    }
    check(ctx_, bem_set_frequency(ctx_, wavenumber, couplingParameter.real(), couplingParameter.imag()),
          "bem_set_frequency");
}

std::vector<cplx> BoundaryElementSolver::calculateFieldSolution(const std::vector<double>& obs3)
{
    auto& E = boundaryElements;
    if ((int64_t)phiSolution.size() != E.n || (int64_t)dPhiSolution.size() != E.n)
        throw Error(BEM_ERR_STATE, "calculateFieldSolution: no boundary solution yet");
    const int64_t nobs = (int64_t)obs3.size() / 3;
    DeviceVector phi(ctx_, E.n), dphi(ctx_, E.n);
    phi.upload(phiSolution);
    dphi.upload(dPhiSolution);
    std::vector<cplx> field(nobs);
    check(ctx_, bem_field(ctx_, nobs, obs3.data(), phi.data(), dphi.data(), (double*)field.data()), "bem_field");
    return fieldToPressure(obs3, field);
}

std::vector<cplx> BoundaryElementSolver::calculateFieldSolutionFast(const std::vector<double>& obs3,
                                                                    double relativeError, long maxRank, int64_t leaf,
                                                                    double eta, uint64_t seed, double* stats4)
{
    auto& E = boundaryElements;
    if ((int64_t)phiSolution.size() != E.n || (int64_t)dPhiSolution.size() != E.n)
        throw Error(BEM_ERR_STATE, "calculateFieldSolutionFast: no boundary solution yet");
    if (!E.impedancePlanes.empty())
        throw Error(BEM_ERR_ARG, "calculateFieldSolutionFast: not with an impedance plane (use the direct sum)");
    if (leaf < 1 || !(eta > 0)) throw Error(BEM_ERR_ARG, "calculateFieldSolutionFast: leaf >= 1 and eta > 0");
    const int64_t nobs = (int64_t)obs3.size() / 3;
    if (nobs == 0) return {};
    // the observation points' cluster tree (ClusterTree(&triangles) on the
    // observation triangles' midpoints, :2420) against the boundary tree of
    // the solve (HMatrix(fieldTree, boundaryTree, true), :2425)
    bemsetup::ClusterTree T = bemsetup::build_cluster_tree_points(nobs, obs3.data(), leaf);
    std::vector<double> obsc(3 * nobs);
    for (int64_t i = 0; i < nobs; i++)
        for (int d = 0; d < 3; d++) obsc[3 * i + d] = obs3[3 * T.perm[i] + d];
    std::vector<bemsetup::Block> blocks = bemsetup::build_partition(T, clusterTree->tree_, eta);
    std::vector<int64_t> b5(5 * blocks.size());
    for (size_t b = 0; b < blocks.size(); b++) {
        b5[5 * b] = blocks[b].row0;
        b5[5 * b + 1] = blocks[b].rows;
        b5[5 * b + 2] = blocks[b].col0;
        b5[5 * b + 3] = blocks[b].cols;
        b5[5 * b + 4] = blocks[b].adm;
    }
    DeviceVector phi(ctx_, E.n), dphi(ctx_, E.n);
    phi.upload(phiSolution);
    dphi.upload(dPhiSolution);
    std::vector<cplx> fc(nobs), field(nobs);
    check(ctx_, bem_field_hmat(ctx_, nobs, obsc.data(), (int64_t)blocks.size(), b5.data(), relativeError,
                               (int64_t)maxRank, seed, phi.data(), dphi.data(), (double*)fc.data(), stats4),
          "bem_field_hmat");
    for (int64_t i = 0; i < nobs; i++) field[T.perm[i]] = fc[i];
    return fieldToPressure(obs3, field);
}

std::vector<cplx> BoundaryElementSolver::fieldToPressure(const std::vector<double>& obs3,
                                                         const std::vector<cplx>& field) const
{
    const int64_t nobs = (int64_t)field.size();
    // incident field of the sources (sourcePhiTerm, :3443-3450; plane waves A e^{ik d.x})
    const double k = wavenumber;
    const bem_c64 s = c64(0.0, airDensity * (2.0 * BEM_PI) * frequency);   // p = i rho 2 pi f phi (:2234-2240)
    std::vector<cplx> pressure(nobs);
    for (int64_t i = 0; i < nobs; i++) {
        const double x = obs3[3 * i], y = obs3[3 * i + 1], z = obs3[3 * i + 2];
        bem_c64 acc = c64(field[i].real(), field[i].imag());
        for (auto& src : pointSources) {
            double Rx = src.x - x, Ry = src.y - y, Rz = src.z - z;
            double r = std::sqrt(fma(Rz, Rz, fma(Ry, Ry, Rx * Rx)));
            bem_c64 G = c64_mul(c64(src.weight.real(), src.weight.imag()),
                                c64_scale(det_expi(k * r), 1.0 / (4.0 * BEM_PI * r)));
            acc = c64_add(acc, G);
        }
        for (auto& w : planeWaves) {
            double dx = fma(w.dz, z, fma(w.dy, y, w.dx * x));
            acc = c64_add(acc, c64_mul(c64(w.amplitude.real(), w.amplitude.imag()), det_expi(k * dx)));
        }
        bem_c64 pr = c64_mul(s, acc);
        pressure[i] = cplx(pr.re, pr.im);
    }
    return pressure;
}

// refinement mode of the current parameters (FP32 operators only)
static int refineMode(const BemSolverParameters& P)
{
    return P.precision == BEM_PREC_FP32 ? P.refine : 0;
}

// Phi for b = s + Phi g: the GMRES operators' precision, or with refinement
// the FP64-accurate operator the residuals are measured with (mode 1: the
// FP32 Phi with its FP64 near field; mode 2: Phi in FP64)
void BoundaryElementSolver::assembleRhsOperator()
{
    const auto& P = bemSolverParameters;
    const int R = refineMode(P);
    if (presetPhi) {
        phiHMat->adopt(presetPhi);
        presetPhi = nullptr;
        return;
    }
    double t0 = now_ms();
    phiHMat->assembleBlocks((long)P.acaMaxRank, P.acaRelativeError, P.pivotSeed, R == 2 ? BEM_PREC_FP64 : P.precision);
    if (R == 2) refineStats[3] += now_ms() - t0;
}

void BoundaryElementSolver::rhsProduct(DeviceVector& b, const DeviceVector& g)
{
    if (refineMode(bemSolverParameters) == 1)
        check(ctx_, bem_hmvm_refine(ctx_, phiHMat->handle(), g.data(), b.data(), 1), "b = s + Phi g");
    else
        check(ctx_, bem_hmvm(ctx_, phiHMat->handle(), g.data(), b.data(), 1), "b = s + Phi g");
}

// mode 2: DPhi in FP64 first (it takes the near field the FP64 Phi
// assembly paired with it), then the FP32 DPhi of the iterations
void BoundaryElementSolver::assembleResidualOperator()
{
    const auto& P = bemSolverParameters;
    dPhiResHMat.reset();
    if (refineMode(P) != 2) return;
    double t0 = now_ms();
    dPhiResHMat = std::make_unique<HMatrix>(ctx_);
    dPhiResHMat->setClusterTrees(clusterTree, clusterTree);
    dPhiResHMat->assembleBlocksExtra((long)P.acaMaxRank, P.acaRelativeError, P.pivotSeed, BEM_PREC_FP64);
    check(ctx_, bem_synchronize(ctx_), "sync");
    refineStats[3] += now_ms() - t0;
}

int64_t BoundaryElementSolver::solveSystem(DeviceVector& x, const DeviceVector& b, double tol, bem_prec_t pc,
                                           double* rel)
{
    const auto& P = bemSolverParameters;
    const int R = refineMode(P);
    check(ctx_, bem_set_precond_side(ctx_, P.preconditionLeft ? 1 : 0), "bem_set_precond_side");
    if (R == 0) {
        const bool left = pc && P.preconditionLeft;
        if (left)   // the reference's guess, LUSubstitutionSolve (boundaryelementsolver.cpp:689)
            check(ctx_, bem_precond_apply(ctx_, pc, b.data(), x.data()), "bem_precond_apply");
        else
            x.upload(std::vector<cplx>(x.size(), cplx(0, 0)));
        const int64_t its = GMRES::gmresSolve(*dPhiHMat, x, b, tol, P.gmresRestart, rel, pc);
        if (left && rel) {   // GMRES reports the preconditioned residual; the true one, as the reference prints (:696-698)
            DeviceVector y(ctx_, x.size());
            check(ctx_, bem_hmvm(ctx_, dPhiHMat->handle(), x.data(), y.data(), 0), "bem_hmvm");
            const std::vector<cplx> yv = y.download(), bv = b.download();
            double num = 0, den = 0;
            for (size_t i = 0; i < bv.size(); i++) {
                num += std::norm(bv[i] - yv[i]);
                den += std::norm(bv[i]);
            }
            *rel = den > 0 ? std::sqrt(num / den) : 0;
        }
        return its;
    }
    int64_t its = 0, steps = 0;
    double r = 0;
    int rc = bem_gmres_ir(ctx_, dPhiHMat->handle(), R == 2 ? dPhiResHMat->handle() : dPhiHMat->handle(), pc, b.data(),
                          x.data(), 1, tol, P.gmresRestart, 100000, P.refineMaxSteps, &its, &r, &steps);
    if (rc != BEM_OK && rc != BEM_ERR_NOCONV) check(ctx_, rc, "bem_gmres_ir");
    refineStats[0] += (double)steps;
    refineStats[1] += (double)its;
    refineStats[2] = r;
    if (rel) *rel = r;
    return its;
}

void BoundaryElementSolver::calculateBoundarySolutions(const std::vector<RightHandSide>& rhs,
                                                       std::vector<std::vector<cplx>>& pressure,
                                                       std::vector<int64_t>& iterations)
{
    setWavenumberAndCoupling();
    auto& E = boundaryElements;
    const auto& P = bemSolverParameters;
    const size_t nr = rhs.size();
    std::memset(stats, 0, sizeof stats);
    std::memset(precondStats, 0, sizeof precondStats);
    std::memset(refineStats, 0, sizeof refineStats);
    check(ctx_, bem_set_near_fp64(ctx_, refineMode(P) == 1), "bem_set_near_fp64");
    double t0 = now_ms();
    // source term of every right-hand side
    std::vector<std::vector<cplx>> s(nr);
    auto keepPoints = pointSources;
    auto keepPlanes = planeWaves;
    bool anyG = false;
    for (size_t k = 0; k < nr; k++) {
        if ((int64_t)rhs[k].g.size() != E.n) throw Error(BEM_ERR_ARG, "right-hand side of the wrong length");
        pointSources = rhs[k].pointSources;
        planeWaves = rhs[k].planeWaves;
        calculateBoundaryConditionAndSourceTermVector();
        s[k] = sourceTermVector;
        for (const cplx& v : rhs[k].g) anyG = anyG || v != cplx(0, 0);
    }
    pointSources = keepPoints;
    planeWaves = keepPlanes;
    if (nr > 1 && nr <= 8 && P.preconditioner == 0 && P.lockstepRhs) {
        solveLockstep(rhs, s, anyG, t0, pressure, iterations);
        return;
    }
    std::vector<std::unique_ptr<DeviceVector>> b, g;
    for (size_t k = 0; k < nr; k++) {
        b.push_back(std::make_unique<DeviceVector>(ctx_, E.n));
        g.push_back(std::make_unique<DeviceVector>(ctx_, E.n));
        b[k]->upload(s[k]);
        g[k]->upload(rhs[k].g);
    }
    double t1 = t0;
    if (anyG) {   // b_k = s_k + Phi g_k (:578-579)
        assembleRhsOperator();
        stats[7] = phiHMat->stat(1);
        stats[9] = phiHMat->stat(0);
        t1 = now_ms();
        for (size_t k = 0; k < nr; k++) rhsProduct(*b[k], *g[k]);
        phiHMat->clear();
    }
    check(ctx_, bem_synchronize(ctx_), "sync");
    double t2 = now_ms();
    assembleResidualOperator();
    dPhiHMat->assembleBlocksExtra((long)P.acaMaxRank, P.acaRelativeError, P.pivotSeed, P.precision);
    check(ctx_, bem_synchronize(ctx_), "sync");
    double t3 = now_ms();
    stats[8] = dPhiHMat->stat(1);
    stats[10] = dPhiHMat->stat(0);
    stats[15] = dPhiHMat->stat(3);
    const double tol = P.gmresTolerance > 0 ? P.gmresTolerance : 1e-3 * P.acaRelativeError;
    pressure.assign(nr, {});
    iterations.assign(nr, 0);
    DeviceVector x(ctx_, E.n), phi(ctx_, E.n), dphi(ctx_, E.n), p(ctx_, E.n);
    double est = 0;
    bem_prec_t pc = nullptr;
    pc = makePreconditioner();
    for (size_t k = 0; k < nr; k++) {
        iterations[k] = solveSystem(x, *b[k], tol, pc, &est);
        check(ctx_, bem_reconstruct_g(ctx_, x.data(), g[k]->data(), airDensity, frequency, phi.data(), dphi.data(),
                                      p.data()),
              "bem_reconstruct_g");
        pressure[k] = p.download();
    }
    if (pc) {
        double* ps = precondStats;
        bem_precond_stats(pc, ps);
        stats[13] = ps[0];
        stats[14] = ps[2];
        bem_precond_destroy(pc);
    }
    double t4 = now_ms();
    dPhiHMat->clear();
    dPhiResHMat.reset();
    stats[0] = t1 - t0; stats[1] = t2 - t1; stats[2] = t3 - t2; stats[3] = t4 - t3; stats[4] = t4 - t0;
    stats[5] = nr ? (double)iterations[nr - 1] : 0;
    stats[6] = est;
}

// calculateBoundarySolutions for 2..8 right-hand sides without a
// preconditioner: b = s + Phi g for all of them with one pass over Phi, one
// lockstep GMRES (with refinement when set) whose Arnoldi columns share one
// product with DPhi, then the reconstruction per right-hand side
void BoundaryElementSolver::solveLockstep(const std::vector<RightHandSide>& rhs,
                                          const std::vector<std::vector<cplx>>& s, bool anyG, double t0,
                                          std::vector<std::vector<cplx>>& pressure, std::vector<int64_t>& iterations)
{
    auto& E = boundaryElements;
    const auto& P = bemSolverParameters;
    const int R = refineMode(P);
    const int64_t n = E.n, nr = (int64_t)rhs.size();
    std::vector<cplx> Si((size_t)(n * nr)), Gi((size_t)(n * nr));
    for (int64_t k = 0; k < nr; k++)
        for (int64_t i = 0; i < n; i++) {
            Si[i * nr + k] = s[k][i];
            Gi[i * nr + k] = rhs[k].g[i];
        }
    DeviceVector B(ctx_, n * nr), G(ctx_, n * nr), X(ctx_, n * nr);
    B.upload(Si);
    G.upload(Gi);
    double t1 = t0;
    if (anyG) {   // b_k = s_k + Phi g_k (:578-579), Phi read once for all k
        assembleRhsOperator();
        stats[7] = phiHMat->stat(1);
        stats[9] = phiHMat->stat(0);
        t1 = now_ms();
        check(ctx_, bem_hmvm_multi(ctx_, phiHMat->handle(), nr, G.data(), B.data(), 1, R == 1), "b = s + Phi g");
        check(ctx_, bem_synchronize(ctx_), "sync");
        phiHMat->clear();
    }
    double t2 = now_ms();
    assembleResidualOperator();
    dPhiHMat->assembleBlocksExtra((long)P.acaMaxRank, P.acaRelativeError, P.pivotSeed, P.precision);
    check(ctx_, bem_synchronize(ctx_), "sync");
    double t3 = now_ms();
    stats[8] = dPhiHMat->stat(1);
    stats[10] = dPhiHMat->stat(0);
    stats[15] = dPhiHMat->stat(3);
    const double tol = P.gmresTolerance > 0 ? P.gmresTolerance : 1e-3 * P.acaRelativeError;
    std::vector<int64_t> its((size_t)nr, 0);
    std::vector<double> rel((size_t)nr, 0.0);
    int64_t steps = 0;
    bem_hmat_t hres = R == 2 ? dPhiResHMat->handle() : (R == 1 ? dPhiHMat->handle() : nullptr);
    int rc = bem_gmres_multi(ctx_, dPhiHMat->handle(), hres, nr, B.data(), X.data(), 1, tol, P.gmresRestart, 100000,
                             R ? P.refineMaxSteps : 0, its.data(), rel.data(), &steps);
    if (rc != BEM_OK && rc != BEM_ERR_NOCONV) check(ctx_, rc, "bem_gmres_multi");
    if (R) {
        refineStats[0] = (double)steps;
        for (int64_t k = 0; k < nr; k++) refineStats[1] += (double)its[k];
        refineStats[2] = *std::max_element(rel.begin(), rel.end());
    }
    std::vector<cplx> Xi = X.download();
    DeviceVector x(ctx_, n), g(ctx_, n), phi(ctx_, n), dphi(ctx_, n), p(ctx_, n);
    std::vector<cplx> xk((size_t)n);
    pressure.assign((size_t)nr, {});
    iterations.assign((size_t)nr, 0);
    for (int64_t k = 0; k < nr; k++) {
        for (int64_t i = 0; i < n; i++) xk[i] = Xi[i * nr + k];
        x.upload(xk);
        g.upload(rhs[k].g);
        check(ctx_, bem_reconstruct_g(ctx_, x.data(), g.data(), airDensity, frequency, phi.data(), dphi.data(),
                                      p.data()),
              "bem_reconstruct_g");
        pressure[k] = p.download();
        iterations[k] = its[k];
    }
    double t4 = now_ms();
    dPhiHMat->clear();
    dPhiResHMat.reset();
    stats[0] = t1 - t0; stats[1] = t2 - t1; stats[2] = t3 - t2; stats[3] = t4 - t3; stats[4] = t4 - t0;
    stats[5] = (double)its[(size_t)nr - 1];
    stats[6] = rel[(size_t)nr - 1];
}

void BoundaryElementSolver::calculateBoundarySolutionsBatch(const std::vector<double>& freqs,
                                                            const std::vector<RightHandSide>& rhs,
                                                            std::vector<std::vector<std::vector<cplx>>>& pressure,
                                                            std::vector<std::vector<int64_t>>& iterations,
                                                            std::vector<double>& stats16)
{
    auto& E = boundaryElements;
    const auto& P = bemSolverParameters;
    const int64_t nf = (int64_t)freqs.size(), nr = (int64_t)rhs.size(), n = E.n, nq = nf * nr;
    if (nf < 1 || (nr != 1 && nr != 2 && nr != 4) || nq > 8)
        throw Error(BEM_ERR_ARG, "batched solve: 1, 2 or 4 right-hand sides and at most 8 (frequency, right-hand side) pairs");
    if (P.preconditioner != 0) throw Error(BEM_ERR_ARG, "batched solve: no preconditioner");
    const int R = refineMode(P);
    const double t0 = now_ms();
    check(ctx_, bem_set_near_fp64(ctx_, R == 1), "bem_set_near_fp64");
    // wavenumber, coupling and source terms per frequency (the context's
    // boundary-condition arrays are set once: they do not depend on f)
    std::vector<double> kcc(3 * nf);
    std::vector<std::vector<cplx>> src(nq);
    auto keepPoints = pointSources;
    auto keepPlanes = planeWaves;
    bool anyG = false;
    for (int64_t f = nf - 1; f >= 0; f--) {   // ends with the first frequency set on the context
        setFrequency(freqs[f]);
        setWavenumberAndCoupling();
        kcc[3 * f] = wavenumber;
        kcc[3 * f + 1] = couplingParameter.real();
        kcc[3 * f + 2] = couplingParameter.imag();
        for (int64_t k = 0; k < nr; k++) {
            if ((int64_t)rhs[k].g.size() != n) throw Error(BEM_ERR_ARG, "right-hand side of the wrong length");
            pointSources = rhs[k].pointSources;
            planeWaves = rhs[k].planeWaves;
            calculateBoundaryConditionAndSourceTermVector();
            src[f * nr + k] = sourceTermVector;
            for (const cplx& v : rhs[k].g) anyG = anyG || v != cplx(0, 0);
        }
    }
    pointSources = keepPoints;
    planeWaves = keepPlanes;
    auto destroy = [](std::vector<bem_hmat_t>& v) {
        for (bem_hmat_t h : v)
            if (h) bem_hmat_destroy(h);
        v.clear();
    };
    std::vector<bem_hmat_t> phis(nf, nullptr), dphis(nf, nullptr), dres;
    std::vector<cplx> Bi((size_t)(n * nq));
    DeviceVector B(ctx_, n * nq), X(ctx_, n * nq), G(ctx_, n * nr), Bf(ctx_, n * nr);
    std::vector<cplx> Gi((size_t)(n * nr));
    for (int64_t k = 0; k < nr; k++)
        for (int64_t i = 0; i < n; i++) Gi[i * nr + k] = rhs[k].g[i];
    G.upload(Gi);
    double t1 = now_ms(), t2, t3;
    std::vector<double> phiRank(nf, 0.0), phiEvaluated(nf, 0.0);
    try {
        if (anyG) {   // b = s + Phi g, Phi of each frequency read once for its right-hand sides
            check(ctx_, bem_assemble_batch(ctx_, BEM_KIND_PHI, nf, kcc.data(), P.acaRelativeError, (int64_t)P.acaMaxRank,
                                           R == 2 ? BEM_PREC_FP64 : P.precision, P.pivotSeed, phis.data()),
                  "Phi batch");
            t1 = now_ms();
            for (int64_t f = 0; f < nf; f++) {
                phiRank[f] = bem_hmat_stats_value(phis[f], 1);
                phiEvaluated[f] = bem_hmat_stats_value(phis[f], 0);
            }
        }
        std::vector<cplx> sf((size_t)(n * nr));
        for (int64_t f = 0; f < nf; f++) {
            for (int64_t k = 0; k < nr; k++)
                for (int64_t i = 0; i < n; i++) sf[i * nr + k] = src[f * nr + k][i];
            if (anyG) {
                Bf.upload(sf);
                check(ctx_, bem_hmvm_multi(ctx_, phis[f], nr, G.data(), Bf.data(), 1, R == 1), "b = s + Phi g");
                sf = Bf.download();
            }
            for (int64_t k = 0; k < nr; k++)
                for (int64_t i = 0; i < n; i++) Bi[i * nq + f * nr + k] = sf[i * nr + k];
        }
        destroy(phis);
        B.upload(Bi);
        t2 = now_ms();
        if (R == 2) {
            dres.assign(nf, nullptr);
            check(ctx_, bem_assemble_batch(ctx_, BEM_KIND_DPHI, nf, kcc.data(), P.acaRelativeError,
                                           (int64_t)P.acaMaxRank, BEM_PREC_FP64, P.pivotSeed, dres.data()),
                  "DPhi FP64 batch");
        }
        check(ctx_, bem_assemble_batch(ctx_, BEM_KIND_DPHI, nf, kcc.data(), P.acaRelativeError, (int64_t)P.acaMaxRank,
                                       P.precision, P.pivotSeed, dphis.data()),
              "DPhi batch");
        t3 = now_ms();
        const double tol = P.gmresTolerance > 0 ? P.gmresTolerance : 1e-3 * P.acaRelativeError;
        std::vector<int64_t> its(nq, 0);
        std::vector<double> rel(nq, 0.0);
        int64_t steps = 0;
        int rc = bem_gmres_batch(ctx_, nf, dphis.data(), R == 2 ? dres.data() : nullptr, nr, B.data(), X.data(), 1, tol,
                                 P.gmresRestart, 100000, R ? P.refineMaxSteps : 0, its.data(), rel.data(), &steps);
        if (rc != BEM_OK && rc != BEM_ERR_NOCONV) check(ctx_, rc, "bem_gmres_batch");
        std::vector<cplx> Xi = X.download();
        DeviceVector x(ctx_, n), g(ctx_, n), phi(ctx_, n), dphi(ctx_, n), p(ctx_, n);
        std::vector<cplx> xk((size_t)n);
        pressure.assign(nf, std::vector<std::vector<cplx>>(nr));
        iterations.assign(nf, std::vector<int64_t>(nr, 0));
        for (int64_t f = 0; f < nf; f++)
            for (int64_t k = 0; k < nr; k++) {
                for (int64_t i = 0; i < n; i++) xk[i] = Xi[i * nq + f * nr + k];
                x.upload(xk);
                g.upload(rhs[k].g);
                check(ctx_, bem_reconstruct_g(ctx_, x.data(), g.data(), airDensity, freqs[f], phi.data(), dphi.data(),
                                              p.data()),
                      "bem_reconstruct_g");
                pressure[f][k] = p.download();
                iterations[f][k] = its[f * nr + k];
            }
        const double t4 = now_ms();
        stats16.assign(16 * nf, 0.0);
        for (int64_t f = 0; f < nf; f++) {
            double* st = stats16.data() + 16 * f;
            st[0] = (t1 - t0) / nf; st[1] = (t2 - t1) / nf; st[2] = (t3 - t2) / nf; st[3] = (t4 - t3) / nf;
            st[4] = (t4 - t0) / nf;
            st[5] = (double)its[f * nr + nr - 1];
            st[6] = rel[f * nr + nr - 1];
            st[7] = phiRank[f];
            st[9] = phiEvaluated[f];
            st[8] = bem_hmat_stats_value(dphis[f], 1);
            st[10] = bem_hmat_stats_value(dphis[f], 0);
            st[15] = bem_hmat_stats_value(dphis[f], 3);
        }
        if (R) {
            refineStats[0] = (double)steps;
            refineStats[1] = 0;
            for (int64_t q = 0; q < nq; q++) refineStats[1] += (double)its[q];
            refineStats[2] = *std::max_element(rel.begin(), rel.end());
        }
    } catch (...) {
        destroy(phis);
        destroy(dphis);
        destroy(dres);
        throw;
    }
    destroy(dphis);
    destroy(dres);
}

void BoundaryElementSolver::tuneAca(const std::vector<double>& freqs, int64_t choice[3], double* ms)
{
    const auto& P = bemSolverParameters;
    check(ctx_, bem_set_near_fp64(ctx_, refineMode(P) == 1), "bem_set_near_fp64");
    std::vector<double> kcc(3 * freqs.size());
    for (size_t f = 0; f < freqs.size(); f++) {
        setFrequency(freqs[f]);
        setWavenumberAndCoupling();
        kcc[3 * f] = wavenumber;
        kcc[3 * f + 1] = couplingParameter.real();
        kcc[3 * f + 2] = couplingParameter.imag();
    }
    check(ctx_, bem_autotune_aca(ctx_, BEM_KIND_DPHI, P.precision, (int64_t)freqs.size(), kcc.data(),
                                 P.acaRelativeError, (int64_t)P.acaMaxRank, P.pivotSeed, choice, ms),
          "bem_autotune_aca");
}

void BoundaryElementSolver::solveMany(const std::vector<BoundaryElementSolver*>& solvers,
                                      const std::vector<double>& freqs, int threads)
{
    const size_t ns = solvers.size();
    if (ns == 0) return;
    if (freqs.size() != ns) throw Error(BEM_ERR_ARG, "solveMany: one frequency per solver");
    const auto& P0 = solvers[0]->bemSolverParameters;
    const int R = refineMode(P0);
    std::vector<bem_ctx_t> ctxs(ns);
    for (size_t q = 0; q < ns; q++) {
        auto& S = *solvers[q];
        if (!S.boundaryElements.impedancePlanes.empty())
            throw Error(BEM_ERR_ARG, "solveMany: not with an impedance plane");
        S.bemSolverParameters = P0;
        S.setFrequency(freqs[q]);
        S.setWavenumberAndCoupling();
        check(S.ctx_, bem_set_near_fp64(S.ctx_, R == 1), "bem_set_near_fp64");
        ctxs[q] = S.ctx_;
    }
    std::vector<bem_hmat_t> phi(ns, nullptr), dphi(ns, nullptr);
    auto destroy = [](std::vector<bem_hmat_t>& v) {
        for (bem_hmat_t h : v)
            if (h) bem_hmat_destroy(h);
    };
    int rc = bem_assemble_multi(ctxs.data(), (int64_t)ns, BEM_KIND_PHI, P0.acaRelativeError, (int64_t)P0.acaMaxRank,
                                R == 2 ? BEM_PREC_FP64 : P0.precision, P0.pivotSeed, phi.data());
    if (rc == BEM_OK)
        rc = bem_assemble_multi(ctxs.data(), (int64_t)ns, BEM_KIND_DPHI, P0.acaRelativeError, (int64_t)P0.acaMaxRank,
                                P0.precision, P0.pivotSeed, dphi.data());
    if (rc != BEM_OK) {
        destroy(phi);
        destroy(dphi);
        check(ctxs[0], rc, "bem_assemble_multi");
    }
    for (size_t q = 0; q < ns; q++) {
        solvers[q]->presetPhi = phi[q];
        solvers[q]->presetDPhi = dphi[q];
    }
    // the solves: every one takes its preset operators (the frequency is set again, unchanged)
    const int L = std::max(1, std::min<int>(threads, (int)ns));
    std::vector<std::string> errs(L);
    std::vector<int> codes(L, BEM_OK);
    auto work = [&](int t) {
        for (size_t q = t; q < ns; q += L) {
            try {
                solvers[q]->calculateBoundarySolution();
            } catch (const Error& e) {
                codes[t] = e.code;
                errs[t] = e.msg;
                return;
            } catch (const std::exception& e) {
                codes[t] = BEM_ERR_HIP;
                errs[t] = e.what();
                return;
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < L; t++) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
    for (size_t q = 0; q < ns; q++) {   // operators a failed solve did not take
        if (solvers[q]->presetPhi) bem_hmat_destroy(solvers[q]->presetPhi);
        if (solvers[q]->presetDPhi) bem_hmat_destroy(solvers[q]->presetDPhi);
        solvers[q]->presetPhi = solvers[q]->presetDPhi = nullptr;
    }
    for (int t = 0; t < L; t++)
        if (codes[t] != BEM_OK) throw Error(codes[t], errs[t]);
}

void BoundaryElementSolver::regularSolve(double tol, long restart)
{
    setWavenumberAndCoupling();
    calculateBoundaryConditionAndSourceTermVector();
    auto& E = boundaryElements;
    std::memset(stats, 0, sizeof stats);
    double t0 = now_ms();
    bem_ctx_t d = nullptr;
    int rc = bem_create(&d, device_);
    if (rc != BEM_OK) throw Error(rc, "regularSolve: bem_create failed");
    bem_hmat_t h = nullptr;
    try {
        check(d, bem_upload_mesh(d, E.n, E.nodes.data(), E.normal.data(), E.area.data(), E.colloc.data(),
                                 E.averageTriangleDim),
              "bem_upload_mesh");
        const auto& PL = E.impedancePlanes;
        if (!PL.empty())
            check(d, bem_set_plane(d, PL[0].halfSpacePlanePoint, PL[0].halfSpacePlaneNormal), "bem_set_plane");
        std::vector<cplx> a(E.n), bb(E.n), g(E.n);
        for (int64_t e = 0; e < E.n; e++) { a[e] = E.bc[e].a; bb[e] = E.bc[e].b; g[e] = E.bc[e].g; }
        check(d, bem_set_bc(d, (const double*)a.data(), (const double*)bb.data(), (const double*)g.data()), "bem_set_bc");
        const int64_t z = 0, n = E.n;
        const uint8_t dense = 0;
        check(d, bem_upload_partition(d, 1, &z, &n, &z, &n, &dense), "bem_upload_partition");
        check(d, bem_set_frequency(d, wavenumber, couplingParameter.real(), couplingParameter.imag()), "bem_set_frequency");
        DeviceVector b(d, E.n), f(d, E.n), x(d, E.n);
        f.upload(g);
        b.upload(sourceTermVector);
        check(d, bem_assemble(d, BEM_KIND_PHI, 0.0, 0, BEM_PREC_FP64, 1, &h), "dense Phi");
        check(d, bem_hmvm(d, h, f.data(), b.data(), 1), "b = s + Phi f");
        check(d, bem_synchronize(d), "sync");
        bem_hmat_destroy(h);
        h = nullptr;
        check(d, bem_assemble(d, BEM_KIND_DPHI, 0.0, 0, BEM_PREC_FP64, 1, &h), "dense DPhi");
        double t1 = now_ms();
        int64_t its = 0;
        double r = 0;
        x.upload(std::vector<cplx>(E.n, cplx(0, 0)));
        rc = bem_gmres(d, h, b.data(), x.data(), tol, restart, 100000, &its, &r);
        if (rc != BEM_OK && rc != BEM_ERR_NOCONV) check(d, rc, "dense GMRES");
        bem_hmat_destroy(h);
        h = nullptr;
        DeviceVector phi(d, E.n), dphi(d, E.n), p(d, E.n);
        check(d, bem_reconstruct(d, x.data(), airDensity, frequency, phi.data(), dphi.data(), p.data()),
              "bem_reconstruct");
        phiSolution = phi.download();
        dPhiSolution = dphi.download();
        soundPressure = p.download();
        double t2 = now_ms();
        stats[0] = t1 - t0;
        stats[3] = t2 - t1;
        stats[4] = t2 - t0;
        stats[5] = (double)its;
        stats[6] = r;
    } catch (...) {
        if (h) bem_hmat_destroy(h);
        bem_destroy(d);
        throw;
    }
    bem_destroy(d);
}

// BoundaryElementSolver::hMatrixSolve (boundaryelementsolver.cpp:526-744)
void BoundaryElementSolver::hMatrixSolve()
{
    auto& E = boundaryElements;
    const auto& P = bemSolverParameters;
    std::memset(stats, 0, sizeof stats);
    std::memset(precondStats, 0, sizeof precondStats);
    std::memset(refineStats, 0, sizeof refineStats);
    check(ctx_, bem_set_near_fp64(ctx_, refineMode(P) == 1), "bem_set_near_fp64");
    double t0 = now_ms();
    calculateBoundaryConditionAndSourceTermVector();
    assembleRhsOperator();
    check(ctx_, bem_synchronize(ctx_), "sync");
    double t1 = now_ms();
    stats[7] = phiHMat->stat(1);
    stats[9] = phiHMat->stat(0);
    // b = s + Phi f (:578-579)
    DeviceVector b(ctx_, E.n), f(ctx_, E.n), x(ctx_, E.n);
    std::vector<cplx> fv(E.n);
    for (int64_t e = 0; e < E.n; e++) fv[e] = E.bc[e].g;
    f.upload(fv);
    b.upload(sourceTermVector);
    rhsProduct(b, f);
    check(ctx_, bem_synchronize(ctx_), "sync");
    phiHMat->clear();
    double t2 = now_ms();
    assembleResidualOperator();
    if (presetDPhi) {
        dPhiHMat->adopt(presetDPhi);
        presetDPhi = nullptr;
    } else {
        dPhiHMat->assembleBlocksExtra((long)P.acaMaxRank, P.acaRelativeError, P.pivotSeed, P.precision);
    }
    check(ctx_, bem_synchronize(ctx_), "sync");
    double t3 = now_ms();
    stats[8] = dPhiHMat->stat(1);
    stats[10] = dPhiHMat->stat(0);
    stats[15] = dPhiHMat->stat(3);
    // GMRES with tolerance 1e-3 acaRelativeError by default (:670); start
    // vector zero
    double tol = P.gmresTolerance > 0 ? P.gmresTolerance : 1e-3 * P.acaRelativeError;
    double est = 0;
    bem_prec_t pc = nullptr;
    pc = makePreconditioner();
    int64_t its = solveSystem(x, b, tol, pc, &est);
    if (pc) {
        double* ps = precondStats;
        bem_precond_stats(pc, ps);
        stats[13] = ps[0];
        stats[14] = ps[2];
        bem_precond_destroy(pc);
    }
    check(ctx_, bem_synchronize(ctx_), "sync");
    double t4 = now_ms();
    // true residual (:682-684); with refinement the residual of the
    // refinement operator that ended it
    double num = 0, den = 0;
    if (refineMode(P) == 0) {
        DeviceVector r(ctx_, E.n);
        check(ctx_, bem_hmvm(ctx_, dPhiHMat->handle(), x.data(), r.data(), 0), "residual");
        std::vector<cplx> rv = r.download(), bv = b.download();
        for (int64_t e = 0; e < E.n; e++) { num += std::norm(rv[e] - bv[e]); den += std::norm(bv[e]); }
    } else {
        num = est * est;
        den = 1;
    }
    // reconstruction (:729-740) and pressure (:2251-2261)
    DeviceVector phi(ctx_, E.n), dphi(ctx_, E.n), p(ctx_, E.n);
    check(ctx_, bem_reconstruct(ctx_, x.data(), airDensity, frequency, phi.data(), dphi.data(), p.data()),
          "bem_reconstruct");
    phiSolution = phi.download();
    dPhiSolution = dphi.download();
    soundPressure = p.download();
    dPhiHMat->clear();
    dPhiResHMat.reset();
    double t5 = now_ms();
    int64_t nnear = 0, nfar = 0;
    for (auto& B : phiHMat->getMinPartition()) (B.adm ? nfar : nnear)++;
    stats[0] = t1 - t0; stats[1] = t2 - t1; stats[2] = t3 - t2; stats[3] = t4 - t3; stats[4] = t5 - t0;
    stats[5] = (double)its;
    stats[6] = den > 0 ? std::sqrt(num / den) : 0;
    stats[11] = (double)nnear;
    stats[12] = (double)nfar;
}

}  // namespace bem

// ========================================================== solver C ABI

struct bem_solver {
    std::unique_ptr<bem::BoundaryElementSolver> s;
    std::string err;
    // creation arguments, for the extra sweep lanes
    int device = 0;
    int64_t leaf = 0;
    double eta = 0;
    bem::BoundaryElements input;   // input order
    std::vector<std::unique_ptr<bem::BoundaryElementSolver>> lanes;
    int preconditioner = 0;
    long hluRank = 2;
    double hluRelError = 0;
    bool precondLeft = false;   // bem_solver_set_precond_side
    int refine = 0;
    long refineMaxSteps = 10;
    int sweepBatch = 0;   // bem_solver_set_sweep_batch
    // bem_solver_set_autotune: the ACA schedule of each precision, timed on
    // the first batched chunk of the first sweep and then used on every lane
    bool autotune = false;
    bool tuned[2] = {false, false};
    int64_t schedule[2][3] = {};
    double tunedMs[2] = {0, 0};
};

extern "C" {

int bem_solver_create(bem_solver_t* out, int device, int64_t n, const double* nodes9, const double* bc6, int64_t leaf,
                      double eta)
{
    if (!out) return BEM_ERR_ARG;
    *out = nullptr;
    if (n <= 0 || !nodes9 || !bc6 || leaf < 1 || !(eta > 0)) return BEM_ERR_ARG;
    // the cluster tree bisects the elements' centroids: non-finite vertices
    // are refused before it (degenerate elements are refused at the upload)
    for (int64_t q = 0; q < 9 * n; q++)
        if (!std::isfinite(nodes9[q])) return BEM_ERR_ARG;
    auto S = new bem_solver();
    try {
        bem::BoundaryElements E;
        E.n = n;
        E.nodes.assign(nodes9, nodes9 + 9 * n);
        E.bc.resize(n);
        for (int64_t e = 0; e < n; e++) {
            const double* q = bc6 + 6 * e;
            E.bc[e].a = bem::cplx(q[0], q[1]);
            E.bc[e].b = bem::cplx(q[2], q[3]);
            E.bc[e].g = bem::cplx(q[4], q[5]);
        }
        S->device = device;
        S->leaf = leaf;
        S->eta = eta;
        S->input = E;
        S->s = std::make_unique<bem::BoundaryElementSolver>(device, std::move(E), leaf, eta);
    } catch (const bem::Error& e) {
        delete S;
        return e.code;
    } catch (...) {
        delete S;
        return BEM_ERR_HIP;
    }
    *out = S;
    return BEM_OK;
}

int bem_solver_destroy(bem_solver_t s)
{
    if (!s) return BEM_ERR_ARG;
    delete s;
    return BEM_OK;
}

const char* bem_solver_last_error(bem_solver_t s) { return s ? s->err.c_str() : "null solver"; }

bem_ctx_t bem_solver_context(bem_solver_t s) { return s && s->s ? s->s->context() : nullptr; }

int bem_solver_layout(bem_solver_t s, int64_t* perm, int64_t* nblocks, int64_t* blocks5, int64_t cap)
{
    if (!s || !s->s) return BEM_ERR_ARG;
    auto& E = s->s->boundaryElements;
    if (perm) std::memcpy(perm, E.elementIndex.data(), E.n * sizeof(int64_t));
    const auto& P = s->s->partition();
    if (nblocks) *nblocks = (int64_t)P.size();
    if (blocks5) {
        if (cap < (int64_t)P.size()) return BEM_ERR_ARG;
        for (size_t b = 0; b < P.size(); b++) {
            blocks5[5 * b] = P[b].row0; blocks5[5 * b + 1] = P[b].rows;
            blocks5[5 * b + 2] = P[b].col0; blocks5[5 * b + 3] = P[b].cols;
            blocks5[5 * b + 4] = P[b].adm ? 1 : 0;
        }
    }
    return BEM_OK;
}

int bem_solver_set_sources(bem_solver_t s, int64_t npoint, const double* pts, int has_plane, const double* pl)
{
    if (!s || !s->s || npoint < 0 || (npoint && !pts) || (has_plane && !pl)) return BEM_ERR_ARG;
    s->s->pointSources.clear();
    s->s->planeWaves.clear();
    for (int64_t q = 0; q < npoint; q++) {
        const double* v = pts + 6 * q;
        s->s->pointSources.push_back(bem::PointSource{v[0], v[1], v[2], bem::cplx(v[3], v[4])});
    }
    if (has_plane) s->s->planeWaves.push_back(bem::PlaneWave{pl[0], pl[1], pl[2], bem::cplx(pl[3], pl[4])});
    return BEM_OK;
}

int bem_solver_solve(bem_solver_t s, double freq, double c, double rho, int coupling, double aca_tol,
                     int64_t max_rank, double gmres_tol, int64_t restart, uint64_t seed, int precision, double* phi,
                     double* dphi, double* pressure, double* stats16)
{
    if (!s || !s->s || coupling < 0 || coupling > 2 || max_rank < 0 || restart < 1) return BEM_ERR_ARG;
    try {
        auto& S = *s->s;
        S.setFrequency(freq);
        S.setWaveSpeed(c);
        S.setAirDensity(rho);
        S.bemSolverParameters.acaRelativeError = aca_tol;
        S.bemSolverParameters.acaMaxRank = (unsigned long)max_rank;
        S.bemSolverParameters.coupling = (bem::BemCoupling)coupling;
        S.bemSolverParameters.gmresTolerance = gmres_tol;
        S.bemSolverParameters.gmresRestart = (long)restart;
        S.bemSolverParameters.pivotSeed = seed;
        S.bemSolverParameters.precision = precision;
        S.bemSolverParameters.preconditioner = s->preconditioner;
        S.bemSolverParameters.preconditionerRank = s->hluRank;
        S.bemSolverParameters.preconditionerRelError = s->hluRelError;
        S.bemSolverParameters.preconditionLeft = s->precondLeft;
        S.bemSolverParameters.refine = s->refine;
        S.bemSolverParameters.refineMaxSteps = s->refineMaxSteps;
        S.calculateBoundarySolution();
        int64_t n = S.boundaryElements.n;
        if (phi) std::memcpy(phi, S.phiSolution.data(), n * sizeof(bem::cplx));
        if (dphi) std::memcpy(dphi, S.dPhiSolution.data(), n * sizeof(bem::cplx));
        if (pressure) std::memcpy(pressure, S.soundPressure.data(), n * sizeof(bem::cplx));
        if (stats16) std::memcpy(stats16, S.stats, sizeof S.stats);
    } catch (const bem::Error& e) {
        s->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        s->err = e.what();
        return BEM_ERR_HIP;
    }
    return BEM_OK;
}

int bem_solvers_solve_multi(const bem_solver_t* ss, int64_t ns, const double* freqs, double c, double rho, int coupling,
                            double aca_tol, int64_t max_rank, double gmres_tol, int64_t restart, uint64_t seed,
                            int precision, int threads, double* const* pressure, int64_t* iters, double* stats16)
{
    if (!ss || ns < 1 || !freqs || coupling < 0 || coupling > 2 || max_rank < 0 || restart < 1 || threads < 1)
        return BEM_ERR_ARG;
    for (int64_t q = 0; q < ns; q++)
        if (!ss[q] || !ss[q]->s) return BEM_ERR_ARG;
    bem_solver_t s0 = ss[0];
    try {
        std::vector<bem::BoundaryElementSolver*> v(ns);
        for (int64_t q = 0; q < ns; q++) {
            auto& S = *ss[q]->s;
            S.setWaveSpeed(c);
            S.setAirDensity(rho);
            v[q] = &S;
        }
        auto& P = v[0]->bemSolverParameters;
        P.acaRelativeError = aca_tol;
        P.acaMaxRank = (unsigned long)max_rank;
        P.coupling = (bem::BemCoupling)coupling;
        P.gmresTolerance = gmres_tol;
        P.gmresRestart = (long)restart;
        P.pivotSeed = seed;
        P.precision = precision;
        P.preconditioner = s0->preconditioner;
        P.preconditionerRank = s0->hluRank;
        P.preconditionerRelError = s0->hluRelError;
        P.preconditionLeft = s0->precondLeft;
        P.refine = s0->refine;
        P.refineMaxSteps = s0->refineMaxSteps;
        bem::BoundaryElementSolver::solveMany(v, std::vector<double>(freqs, freqs + ns), threads);
        for (int64_t q = 0; q < ns; q++) {
            auto& S = *v[q];
            if (pressure && pressure[q])
                std::memcpy(pressure[q], S.soundPressure.data(), S.boundaryElements.n * sizeof(bem::cplx));
            if (iters) iters[q] = (int64_t)S.stats[5];
            if (stats16) std::memcpy(stats16 + 16 * q, S.stats, sizeof S.stats);
        }
    } catch (const bem::Error& e) {
        s0->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        s0->err = e.what();
        return BEM_ERR_HIP;
    }
    return BEM_OK;
}

int bem_solver_solve_dense(bem_solver_t s, double freq, double c, double rho, int coupling, double gmres_tol,
                           int64_t restart, double* phi, double* dphi, double* pressure, double* stats16)
{
    if (!s || !s->s || coupling < 0 || coupling > 2 || restart < 1 || !(gmres_tol > 0)) return BEM_ERR_ARG;
    try {
        auto& S = *s->s;
        S.setFrequency(freq);
        S.setWaveSpeed(c);
        S.setAirDensity(rho);
        S.bemSolverParameters.coupling = (bem::BemCoupling)coupling;
        S.regularSolve(gmres_tol, (long)restart);
        int64_t n = S.boundaryElements.n;
        if (phi) std::memcpy(phi, S.phiSolution.data(), n * sizeof(bem::cplx));
        if (dphi) std::memcpy(dphi, S.dPhiSolution.data(), n * sizeof(bem::cplx));
        if (pressure) std::memcpy(pressure, S.soundPressure.data(), n * sizeof(bem::cplx));
        if (stats16) std::memcpy(stats16, S.stats, sizeof S.stats);
    } catch (const bem::Error& e) {
        s->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        s->err = e.what();
        return BEM_ERR_HIP;
    }
    return BEM_OK;
}

int bem_solver_set_preconditioner(bem_solver_t s, int kind)
{
    if (!s || !s->s || kind < 0 || kind > 2) return BEM_ERR_ARG;
    s->preconditioner = kind;
    return BEM_OK;
}

int bem_solver_precond_stats(bem_solver_t s, double* out5)
{
    if (!s || !s->s || !out5) return BEM_ERR_ARG;
    std::memcpy(out5, s->s->precondStats, sizeof s->s->precondStats);
    return BEM_OK;
}

int bem_solver_set_refinement(bem_solver_t s, int mode, int64_t max_steps)
{
    if (!s || !s->s || mode < 0 || mode > 2 || max_steps < 0) return BEM_ERR_ARG;
    s->refine = mode;
    s->refineMaxSteps = (long)max_steps;
    return BEM_OK;
}

int bem_solver_set_autotune(bem_solver_t s, int on)
{
    if (!s || !s->s) return BEM_ERR_ARG;
    s->autotune = on != 0;
    if (!on) s->tuned[0] = s->tuned[1] = false;
    return BEM_OK;
}

int bem_solver_set_aca_schedule(bem_solver_t s, int precision, int64_t small_bound, int64_t huge_bound, int queue)
{
    if (!s || !s->s || (precision != BEM_PREC_FP64 && precision != BEM_PREC_FP32) || small_bound < 1 ||
        huge_bound < small_bound || queue < 0 || queue > 1)
        return BEM_ERR_ARG;
    const int pi = precision == BEM_PREC_FP32 ? 1 : 0;
    s->schedule[pi][0] = small_bound;
    s->schedule[pi][1] = huge_bound;
    s->schedule[pi][2] = queue;
    s->tunedMs[pi] = 0;
    s->tuned[pi] = true;
    return BEM_OK;
}

int bem_solver_profile(bem_solver_t s, int enable, double* out14)
{
    if (!s || !s->s) return BEM_ERR_ARG;
    std::vector<bem_ctx_t> cs{s->s->context()};
    for (auto& l : s->lanes) cs.push_back(l->context());
    double ms[7] = {0}, cnt[7] = {0};
    for (bem_ctx_t c : cs) {
        double o[14];
        const int rc = bem_profile(c, enable, o);
        if (rc != BEM_OK) return rc;
        for (int q = 0; q < 7; q++) {
            ms[q] += o[q] * o[7 + q];
            cnt[q] += o[7 + q];
        }
    }
    if (out14)
        for (int q = 0; q < 7; q++) {
            out14[q] = cnt[q] > 0 ? ms[q] / cnt[q] : 0.0;
            out14[7 + q] = cnt[q];
        }
    return BEM_OK;
}

int bem_solver_aca_schedule(bem_solver_t s, int precision, double* out4)
{
    if (!s || !s->s || !out4 || (precision != BEM_PREC_FP64 && precision != BEM_PREC_FP32)) return BEM_ERR_ARG;
    const int pi = precision == BEM_PREC_FP32 ? 1 : 0;
    if (!s->tuned[pi]) {
        out4[0] = out4[1] = out4[2] = out4[3] = -1;
        return BEM_OK;
    }
    for (int q = 0; q < 3; q++) out4[q] = (double)s->schedule[pi][q];
    out4[3] = s->tunedMs[pi];
    return BEM_OK;
}

int bem_solver_set_sweep_batch(bem_solver_t s, int batch)
{
    if (!s || !s->s || batch < 0) return BEM_ERR_ARG;
    s->sweepBatch = batch;
    return BEM_OK;
}

int bem_solver_refine_stats(bem_solver_t s, double* out4)
{
    if (!s || !s->s || !out4) return BEM_ERR_ARG;
    std::memcpy(out4, s->s->refineStats, sizeof s->s->refineStats);
    return BEM_OK;
}

int bem_solver_set_precond_side(bem_solver_t s, int left)
{
    if (!s || !s->s) return BEM_ERR_ARG;
    s->precondLeft = left != 0;
    return BEM_OK;
}

int bem_solver_set_hlu(bem_solver_t s, int64_t rank, double rel_err)
{
    if (!s || !s->s || rank < 0 || !(rel_err >= 0)) return BEM_ERR_ARG;
    s->preconditioner = 2;
    s->hluRank = (long)rank;
    s->hluRelError = rel_err;
    return BEM_OK;
}

int bem_solver_set_impedance_plane(bem_solver_t s, const double* point3, const double* normal3)
{
    if (!s || !s->s) return BEM_ERR_ARG;
    const bool on = normal3 && (normal3[0] != 0 || normal3[1] != 0 || normal3[2] != 0);
    if (on && !point3) return BEM_ERR_ARG;
    std::vector<bem::ImpedancePlane> planes;
    if (on) {
        bem::ImpedancePlane q;
        for (int d = 0; d < 3; d++) {
            q.halfSpacePlanePoint[d] = point3[d];
            q.halfSpacePlaneNormal[d] = normal3[d];
        }
        planes.push_back(q);
    }
    try {
        s->input.impedancePlanes = planes;
        s->s->boundaryElements.impedancePlanes = planes;
        s->s->applyImpedancePlanes();
        for (auto& L : s->lanes) {
            L->boundaryElements.impedancePlanes = planes;
            L->applyImpedancePlanes();
        }
    } catch (const bem::Error& e) {
        s->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        s->err = e.what();
        return BEM_ERR_HIP;
    }
    return BEM_OK;
}

int bem_solver_field(bem_solver_t s, int64_t nobs, const double* obs3, double* pressure2)
{
    if (!s || !s->s || nobs < 0 || (nobs && (!obs3 || !pressure2))) return BEM_ERR_ARG;
    try {
        std::vector<double> obs(obs3, obs3 + 3 * nobs);
        std::vector<bem::cplx> p = s->s->calculateFieldSolution(obs);
        if (nobs) std::memcpy(pressure2, p.data(), nobs * sizeof(bem::cplx));
    } catch (const bem::Error& e) {
        s->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        s->err = e.what();
        return BEM_ERR_HIP;
    }
    return BEM_OK;
}

int bem_solver_field_hmat(bem_solver_t s, int64_t nobs, const double* obs3, double aca_tol, int64_t max_rank,
                          int64_t leaf, double eta, uint64_t seed, double* pressure2, double* stats4)
{
    if (!s || !s->s || nobs < 0 || (nobs && (!obs3 || !pressure2)) || !(aca_tol >= 0) || max_rank < 0)
        return BEM_ERR_ARG;
    try {
        std::vector<double> obs(obs3, obs3 + 3 * nobs);
        std::vector<bem::cplx> p = s->s->calculateFieldSolutionFast(obs, aca_tol, (long)max_rank, leaf, eta, seed, stats4);
        if (nobs) std::memcpy(pressure2, p.data(), nobs * sizeof(bem::cplx));
    } catch (const bem::Error& e) {
        s->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        s->err = e.what();
        return BEM_ERR_HIP;
    }
    return BEM_OK;
}

int bem_solver_sweep(bem_solver_t s, int64_t nf, const double* freqs, double c, double rho, int coupling,
                     double aca_tol, int64_t max_rank, double gmres_tol, int64_t restart, uint64_t seed, int precision,
                     int64_t nrhs, const double* g, const double* planes5, int lanes, double* pressure,
                     int64_t* iters, double* stats16)
{
    if (!s || !s->s || nf < 0 || (nf && !freqs) || nrhs < 1 || lanes < 1 || coupling < 0 || coupling > 2 ||
        max_rank < 0 || restart < 1 || !pressure)
        return BEM_ERR_ARG;
    try {
        const int64_t n = s->s->boundaryElements.n;
        const std::vector<int64_t>& perm = s->s->boundaryElements.elementIndex;   // perm[new] = old
        std::vector<bem::BoundaryElementSolver::RightHandSide> rhs(nrhs);
        for (int64_t k = 0; k < nrhs; k++) {
            rhs[k].g.resize(n);
            for (int64_t e = 0; e < n; e++) {
                const int64_t o = perm[e];
                rhs[k].g[e] = g ? bem::cplx(g[2 * (k * n + o)], g[2 * (k * n + o) + 1]) : s->input.bc[o].g;
            }
            // the solver's point sources belong to every right-hand side; its
            // plane wave too unless planes5 gives each its own
            rhs[k].pointSources = s->s->pointSources;
            if (planes5) {
                const double* q = planes5 + 5 * k;
                if (q[3] != 0 || q[4] != 0) rhs[k].planeWaves.push_back(bem::PlaneWave{q[0], q[1], q[2], bem::cplx(q[3], q[4])});
            } else {
                rhs[k].planeWaves = s->s->planeWaves;
            }
        }
        // batch > 1: frequency chunks of up to `chunk` points, each solved by
        // one batched call; the lanes take chunks instead of single points
        const int64_t chunk = s->sweepBatch > 1 ? std::max<int64_t>(1, std::min<int64_t>(s->sweepBatch, 8 / nrhs)) : 1;
        const int64_t nchunks = (nf + chunk - 1) / chunk;
        const int L = (int)std::min<int64_t>(lanes, std::max<int64_t>(nchunks, 1));
        const int pi = precision == BEM_PREC_FP32 ? 1 : 0;
        if (s->autotune && chunk > 1 && nf > 0 && !s->tuned[pi]) {
            auto& S = *s->s;
            S.setWaveSpeed(c);
            auto& P = S.bemSolverParameters;
            P.acaRelativeError = aca_tol;
            P.acaMaxRank = (unsigned long)max_rank;
            P.coupling = (bem::BemCoupling)coupling;
            P.pivotSeed = seed;
            P.precision = precision;
            P.refine = s->refine;
            std::vector<double> fr(freqs, freqs + std::min(nf, chunk));
            S.tuneAca(fr, s->schedule[pi], &s->tunedMs[pi]);
            s->tuned[pi] = true;
        }
        while ((int)s->lanes.size() < L - 1)
            s->lanes.push_back(std::make_unique<bem::BoundaryElementSolver>(s->device, s->input, s->leaf, s->eta));
        std::vector<bem::BoundaryElementSolver*> solver(L);
        solver[0] = s->s.get();
        for (int q = 1; q < L; q++) solver[q] = s->lanes[q - 1].get();
        if (s->tuned[pi])
            for (int q = 0; q < L; q++)
                bem::check(solver[q]->context(),
                           bem_set_aca_schedule(solver[q]->context(), precision, s->schedule[pi][0],
                                                s->schedule[pi][1], (int)s->schedule[pi][2]),
                           "bem_set_aca_schedule");
        std::vector<std::string> errs(L);
        std::vector<int> codes(L, BEM_OK);
        auto work = [&](int q) {
            try {
                auto& S = *solver[q];
                S.setWaveSpeed(c);
                S.setAirDensity(rho);
                auto& P = S.bemSolverParameters;
                P.acaRelativeError = aca_tol;
                P.acaMaxRank = (unsigned long)max_rank;
                P.coupling = (bem::BemCoupling)coupling;
                P.gmresTolerance = gmres_tol;
                P.gmresRestart = (long)restart;
                P.pivotSeed = seed;
                P.precision = precision;
                P.preconditioner = s->preconditioner;
                P.preconditionerRank = s->hluRank;
                P.preconditionerRelError = s->hluRelError;
                P.preconditionLeft = s->precondLeft;
                P.refine = s->refine;
                P.refineMaxSteps = s->refineMaxSteps;
                if (chunk > 1) {
                    for (int64_t ch = q; ch < nchunks; ch += L) {
                        const int64_t f0 = ch * chunk, f1 = std::min(nf, f0 + chunk);
                        std::vector<double> fr(freqs + f0, freqs + f1);
                        std::vector<std::vector<std::vector<bem::cplx>>> prb;
                        std::vector<std::vector<int64_t>> itb;
                        std::vector<double> st;
                        S.calculateBoundarySolutionsBatch(fr, rhs, prb, itb, st);
                        for (int64_t f = f0; f < f1; f++) {
                            for (int64_t k = 0; k < nrhs; k++) {
                                std::memcpy(pressure + 2 * ((f * nrhs + k) * n), prb[f - f0][k].data(),
                                            n * sizeof(bem::cplx));
                                if (iters) iters[f * nrhs + k] = itb[f - f0][k];
                            }
                            if (stats16) std::memcpy(stats16 + 16 * f, st.data() + 16 * (f - f0), 16 * sizeof(double));
                        }
                    }
                    return;
                }
                std::vector<std::vector<bem::cplx>> pr;
                std::vector<int64_t> its;
                for (int64_t f = q; f < nf; f += L) {
                    S.setFrequency(freqs[f]);
                    S.calculateBoundarySolutions(rhs, pr, its);
                    for (int64_t k = 0; k < nrhs; k++) {
                        std::memcpy(pressure + 2 * ((f * nrhs + k) * n), pr[k].data(), n * sizeof(bem::cplx));
                        if (iters) iters[f * nrhs + k] = its[k];
                    }
                    if (stats16) std::memcpy(stats16 + 16 * f, S.stats, sizeof S.stats);
                }
            } catch (const bem::Error& e) {
                codes[q] = e.code;
                errs[q] = e.msg;
            } catch (const std::exception& e) {
                codes[q] = BEM_ERR_HIP;
                errs[q] = e.what();
            }
        };
        std::vector<std::thread> th;
        for (int q = 1; q < L; q++) th.emplace_back(work, q);
        work(0);
        for (auto& t : th) t.join();
        for (int q = 0; q < L; q++)
            if (codes[q] != BEM_OK) {
                s->err = errs[q];
                return codes[q];
            }
    } catch (const bem::Error& e) {
        s->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        s->err = e.what();
        return BEM_ERR_HIP;
    }
    return BEM_OK;
}

}  // extern "C"
