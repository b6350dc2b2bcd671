// Host H-LU (see bem_hlu.hpp).  Dense kernels are plain loops over
// interleaved complex doubles (std::complex multiplication is avoided in the
// inner loops: without -ffast-math it goes through the NaN-checking
// library routine).
#include "bem_hlu.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <omp.h>

#include "../../include/bem_hip.h"

namespace bemhlu {
namespace {

// blocks at least this large (m n) spawn OpenMP tasks for their independent parts
constexpr double kTaskMin = 96.0 * 96.0;
inline bool big(const Block& B) { return (double)B.m * (double)B.n >= kTaskMin; }

// ------------------------------------------------------------ dense kernels

// C (m x n, ldc) += alpha A (m x k, lda) B (k x n, ldb)
void gemm_nn(int64_t m, int64_t n, int64_t k, cd alpha, const cd* A, int64_t lda, const cd* B, int64_t ldb, cd* C,
             int64_t ldc)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    for (int64_t j = 0; j < n; j++) {
        double* c = reinterpret_cast<double*>(C + j * ldc);
        for (int64_t p = 0; p < k; p++) {
            const cd bv = alpha * B[p + j * ldb];
            const double br = bv.real(), bi = bv.imag();
            if (br == 0.0 && bi == 0.0) continue;
            const double* a = reinterpret_cast<const double*>(A + p * lda);
            for (int64_t i = 0; i < m; i++) {
                const double ar = a[2 * i], ai = a[2 * i + 1];
                c[2 * i] += ar * br - ai * bi;
                c[2 * i + 1] += ar * bi + ai * br;
            }
        }
    }
}

// C (m x n) += alpha A B^T, B n x k (ldb)
void gemm_nt(int64_t m, int64_t n, int64_t k, cd alpha, const cd* A, int64_t lda, const cd* B, int64_t ldb, cd* C,
             int64_t ldc)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    for (int64_t j = 0; j < n; j++) {
        double* c = reinterpret_cast<double*>(C + j * ldc);
        for (int64_t p = 0; p < k; p++) {
            const cd bv = alpha * B[j + p * ldb];
            const double br = bv.real(), bi = bv.imag();
            if (br == 0.0 && bi == 0.0) continue;
            const double* a = reinterpret_cast<const double*>(A + p * lda);
            for (int64_t i = 0; i < m; i++) {
                const double ar = a[2 * i], ai = a[2 * i + 1];
                c[2 * i] += ar * br - ai * bi;
                c[2 * i + 1] += ar * bi + ai * br;
            }
        }
    }
}

// C (m x n) += alpha op(A) B with op(A) = A^T (conj: A^H), A k x m (lda)
void gemm_tn(int64_t m, int64_t n, int64_t k, cd alpha, const cd* A, int64_t lda, const cd* B, int64_t ldb, cd* C,
             int64_t ldc, bool conj)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    const double sg = conj ? -1.0 : 1.0;
    for (int64_t j = 0; j < n; j++) {
        const double* b = reinterpret_cast<const double*>(B + j * ldb);
        for (int64_t i = 0; i < m; i++) {
            const double* a = reinterpret_cast<const double*>(A + i * lda);
            double sr = 0, si = 0;
            for (int64_t p = 0; p < k; p++) {
                const double ar = a[2 * p], ai = sg * a[2 * p + 1], br = b[2 * p], bi = b[2 * p + 1];
                sr += ar * br - ai * bi;
                si += ar * bi + ai * br;
            }
            C[i + j * ldc] += alpha * cd(sr, si);
        }
    }
}

// op(A) as an explicit matrix: 'T' transpose, 'C' conjugate transpose
Mat transposed(const cd* A, int64_t lda, int64_t m, int64_t n, bool conj)
{
    Mat T(n, m);
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = 0; i < m; i++) T(j, i) = conj ? std::conj(A[i + j * lda]) : A[i + j * lda];
    return T;
}
Mat transposed(const Mat& A, bool conj = false) { return transposed(A.a.data(), A.m, A.m, A.n, conj); }

Mat identity(int64_t m)
{
    Mat I(m, m);
    for (int64_t i = 0; i < m; i++) I(i, i) = cd(1, 0);
    return I;
}

// Householder QR of A (m x k) in place: reflectors I - tau v v^H (v_0 = 1
// implicit, v below the diagonal), R on and above it; beta = -sign(Re alpha)
// |(alpha, x)| (LAPACK's zgeqrf convention).  Raw double loops.
struct Householder {
    Mat A;
    std::vector<cd> tau;
    int64_t r = 0;

    explicit Householder(Mat in) : A(std::move(in))
    {
        const int64_t m = A.m, k = A.n;
        r = std::min(m, k);
        tau.assign((size_t)r, cd(0, 0));
        for (int64_t j = 0; j < r; j++) {
            double* x = reinterpret_cast<double*>(A.col(j) + j);
            const int64_t len = m - j;
            double xn2 = 0;
            for (int64_t i = 1; i < len; i++) xn2 += x[2 * i] * x[2 * i] + x[2 * i + 1] * x[2 * i + 1];
            const cd alpha(x[0], x[1]);
            if (xn2 == 0.0 && alpha.imag() == 0.0) continue;
            double beta = std::sqrt(std::norm(alpha) + xn2);
            if (alpha.real() >= 0) beta = -beta;
            tau[(size_t)j] = (cd(beta, 0) - alpha) / beta;
            const cd sc = cd(1, 0) / (alpha - beta);
            const double sr = sc.real(), si = sc.imag();
            for (int64_t i = 1; i < len; i++) {
                const double a = x[2 * i], b = x[2 * i + 1];
                x[2 * i] = a * sr - b * si;
                x[2 * i + 1] = a * si + b * sr;
            }
            x[0] = beta;
            x[1] = 0.0;
            // trailing columns: (I - conj(tau) v v^H)
            const cd ct = std::conj(tau[(size_t)j]);
            for (int64_t c = j + 1; c < k; c++) reflect(x, len, ct, reinterpret_cast<double*>(A.col(c) + j));
        }
    }

    // y <- (I - t v v^H) y over len entries, v_0 = 1
    static void reflect(const double* v, int64_t len, cd t, double* y)
    {
        double wr = y[0], wi = y[1];
        for (int64_t i = 1; i < len; i++) {   // w = v^H y
            wr += v[2 * i] * y[2 * i] + v[2 * i + 1] * y[2 * i + 1];
            wi += v[2 * i] * y[2 * i + 1] - v[2 * i + 1] * y[2 * i];
        }
        const cd w = t * cd(wr, wi);
        const double ar = w.real(), ai = w.imag();
        y[0] -= ar;
        y[1] -= ai;
        for (int64_t i = 1; i < len; i++) {
            y[2 * i] -= v[2 * i] * ar - v[2 * i + 1] * ai;
            y[2 * i + 1] -= v[2 * i] * ai + v[2 * i + 1] * ar;
        }
    }

    Mat R() const
    {
        Mat out(r, A.n);
        for (int64_t c = 0; c < A.n; c++)
            for (int64_t i = 0; i <= std::min(c, r - 1); i++) out(i, c) = A(i, c);
        return out;
    }

    // Q [T; 0] for T (r x c): m x c
    Mat apply_q(const Mat& T) const
    {
        Mat M(A.m, T.n);
        for (int64_t c = 0; c < T.n; c++)
            for (int64_t i = 0; i < T.m; i++) M(i, c) = T(i, c);
        for (int64_t j = r - 1; j >= 0; j--) {
            if (tau[(size_t)j] == cd(0, 0)) continue;
            const double* v = reinterpret_cast<const double*>(A.col(j) + j);
            for (int64_t c = 0; c < M.n; c++)
                reflect(v, A.m - j, tau[(size_t)j], reinterpret_cast<double*>(M.col(c) + j));
        }
        return M;
    }
};

void qr(const Mat& A, Mat& Q, Mat& R)
{
    Householder H(A);
    R = H.R();
    Q = H.apply_q(identity(H.r));
}

// One-sided (Hestenes) Jacobi SVD of A (m x n, n <= m): A = X diag(s) Y^H
// with s descending; X m x n, Y n x n.  Column norms are carried through
// the rotations (a_pp - t g, a_qq + t g) and refreshed every sweep.
void svd_tall(const Mat& Ain, Mat& X, std::vector<double>& s, Mat& Y)
{
    Mat A = Ain;
    const int64_t m = A.m, n = A.n;
    Y = identity(n);
    std::vector<double> d((size_t)n);
    for (int sweep = 0; sweep < 60; sweep++) {
        for (int64_t j = 0; j < n; j++) {
            const double* a = reinterpret_cast<const double*>(A.col(j));
            double t = 0;
            for (int64_t i = 0; i < 2 * m; i++) t += a[i] * a[i];
            d[(size_t)j] = t;
        }
        bool rotated = false;
        for (int64_t p = 0; p < n - 1; p++)
            for (int64_t q = p + 1; q < n; q++) {
                const double al = d[(size_t)p], be = d[(size_t)q];
                if (al == 0.0 || be == 0.0) continue;
                const double* ap = reinterpret_cast<const double*>(A.col(p));
                const double* aq = reinterpret_cast<const double*>(A.col(q));
                double gr = 0, gi = 0;   // a_p^H a_q
                for (int64_t i = 0; i < m; i++) {
                    gr += ap[2 * i] * aq[2 * i] + ap[2 * i + 1] * aq[2 * i + 1];
                    gi += ap[2 * i] * aq[2 * i + 1] - ap[2 * i + 1] * aq[2 * i];
                }
                const double g = std::sqrt(gr * gr + gi * gi);
                if (g == 0.0 || g <= 1e-15 * std::sqrt(al * be)) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * g);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
                const double er = gr / g, ei = -gi / g;   // e^{-i phi}
                auto rot = [&](double* u, double* v, int64_t len) {
                    for (int64_t i = 0; i < len; i++) {
                        const double ur = u[2 * i], ui = u[2 * i + 1];
                        const double vr = er * v[2 * i] - ei * v[2 * i + 1], vi = er * v[2 * i + 1] + ei * v[2 * i];
                        u[2 * i] = c * ur - sn * vr;
                        u[2 * i + 1] = c * ui - sn * vi;
                        v[2 * i] = sn * ur + c * vr;
                        v[2 * i + 1] = sn * ui + c * vi;
                    }
                };
                rot(reinterpret_cast<double*>(A.col(p)), reinterpret_cast<double*>(A.col(q)), m);
                rot(reinterpret_cast<double*>(Y.col(p)), reinterpret_cast<double*>(Y.col(q)), n);
                d[(size_t)p] = al - t * g;
                d[(size_t)q] = be + t * g;
            }
        if (!rotated) break;
    }
    std::vector<double> nrm((size_t)n);
    for (int64_t j = 0; j < n; j++) {
        const double* a = reinterpret_cast<const double*>(A.col(j));
        double t = 0;
        for (int64_t i = 0; i < 2 * m; i++) t += a[i] * a[i];
        nrm[(size_t)j] = std::sqrt(t);
    }
    std::vector<int64_t> ord((size_t)n);
    for (int64_t j = 0; j < n; j++) ord[(size_t)j] = j;
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return nrm[(size_t)a] > nrm[(size_t)b]; });
    X = Mat(m, n);
    Mat Ys(n, n);
    s.assign((size_t)n, 0.0);
    for (int64_t jj = 0; jj < n; jj++) {
        const int64_t j = ord[(size_t)jj];
        s[(size_t)jj] = nrm[(size_t)j];
        const double inv = nrm[(size_t)j] > 0 ? 1.0 / nrm[(size_t)j] : 0.0;
        for (int64_t i = 0; i < m; i++) X(i, jj) = A(i, j) * inv;
        for (int64_t i = 0; i < n; i++) Ys(i, jj) = Y(i, j);
    }
    Y = std::move(Ys);
}

// A = X diag(s) Y^H for any shape (X m x r, Y n x r, r = min(m, n)): the
// Jacobi sweeps run over the shorter side
void svd(const Mat& A, Mat& X, std::vector<double>& s, Mat& Y)
{
    if (A.n <= A.m) {
        svd_tall(A, X, s, Y);
        return;
    }
    // A^T = X' S Y'^H  =>  A = conj(Y') S X'^T
    Mat Xt, Yt;
    svd_tall(transposed(A), Xt, s, Yt);
    X = Mat(Yt.m, Yt.n);
    Y = Mat(Xt.m, Xt.n);
    for (size_t q = 0; q < X.a.size(); q++) X.a[q] = std::conj(Yt.a[q]);
    for (size_t q = 0; q < Y.a.size(); q++) Y.a[q] = std::conj(Xt.a[q]);
}

// GoloubReinschSvd::minRankforError (golubReinschSVD.cpp:551-585); a zero
// matrix gets rank 0
int64_t min_rank(const std::vector<double>& s, int64_t maxRank, double relErr)
{
    const int64_t cnt = (int64_t)s.size();
    double total = 0;
    for (double v : s) total += v * v;
    if (total == 0.0) return 0;
    if (relErr <= 0) return maxRank <= 0 ? cnt : std::min(cnt, maxRank);
    const int64_t maxIndex = maxRank > 0 ? std::min(maxRank, cnt) : cnt;
    double residue = total;
    const double err2 = relErr * relErr;
    for (int64_t r = 1; r <= maxIndex; r++) {
        residue -= s[(size_t)(r - 1)] * s[(size_t)(r - 1)];
        if (residue / total < err2) return r;
    }
    return maxIndex;
}

// U W^T truncated in place: QR of both factors, SVD of the core
// (HArithm::RkMatRankReduction, harithm.cpp:158-227)
void truncate(Mat& U, Mat& W, const Params& p)
{
    const int64_t k = U.n;
    if (k == 0) return;
    const bool full_rank_ok = k <= std::min(U.m, W.m);   // no more columns than the block's rank bound
    if (full_rank_ok && ((p.relErr == 0 && p.maxRank > 0 && k <= p.maxRank) || (p.maxRank <= 0 && p.relErr <= 0)))
        return;
    Householder Hu(std::move(U)), Hw(std::move(W));
    const Mat Ru = Hu.R(), Rw = Hw.R();
    Mat C(Ru.m, Rw.m);   // Ru Rw^T
    gemm_nt(Ru.m, Rw.m, k, cd(1, 0), Ru.a.data(), Ru.m, Rw.a.data(), Rw.m, C.a.data(), C.m);
    Mat X, Y;
    std::vector<double> s;
    svd(C, X, s, Y);
    const int64_t r = std::min<int64_t>(min_rank(s, p.maxRank, p.relErr), (int64_t)s.size());
    Mat Xs(X.m, r), Yc(Y.m, r);
    for (int64_t j = 0; j < r; j++) {
        for (int64_t i = 0; i < X.m; i++) Xs(i, j) = X(i, j) * s[(size_t)j];
        for (int64_t i = 0; i < Y.m; i++) Yc(i, j) = std::conj(Y(i, j));
    }
    U = Hu.apply_q(Xs);
    W = Hw.apply_q(Yc);
}

// [U1 | alpha U2], [W1 | W2] (rows u0 / w0 of U2 / W2 onward), truncated
void concat_truncate(Mat& U, Mat& W, const cd* U2, int64_t ldu, const cd* W2, int64_t ldw, int64_t k2, cd alpha,
                     const Params& p)
{
    const int64_t m = U.m, n = W.m, k1 = U.n;
    Mat Un(m, k1 + k2), Wn(n, k1 + k2);
    if (k1) {
        std::memcpy(Un.a.data(), U.a.data(), sizeof(cd) * (size_t)(m * k1));
        std::memcpy(Wn.a.data(), W.a.data(), sizeof(cd) * (size_t)(n * k1));
    }
    for (int64_t l = 0; l < k2; l++) {
        for (int64_t i = 0; i < m; i++) Un(i, k1 + l) = alpha * U2[i + l * ldu];
        for (int64_t i = 0; i < n; i++) Wn(i, k1 + l) = W2[i + l * ldw];
    }
    U = std::move(Un);
    W = std::move(Wn);
    truncate(U, W, p);
}

// ------------------------------------------------------- H-matrix algebra

// Y (rows of B) += alpha B X (X: rows of B's columns), ncol columns
void hmul(const Block& B, cd alpha, const cd* X, int64_t ldx, int64_t ncol, cd* Y, int64_t ldy)
{
    switch (B.type) {
    case Block::Full:
        gemm_nn(B.m, ncol, B.n, alpha, B.F.a.data(), B.F.m, X, ldx, Y, ldy);
        break;
    case Block::LowRank: {
        const int64_t k = B.U.n;
        if (!k) break;
        Mat T(k, ncol);
        gemm_tn(k, ncol, B.n, cd(1, 0), B.W.a.data(), B.W.m, X, ldx, T.a.data(), T.m, false);
        gemm_nn(B.m, ncol, k, alpha, B.U.a.data(), B.U.m, T.a.data(), T.m, Y, ldy);
        break;
    }
    case Block::Sub: {
        // sons 0, 1 write the first row half, 2, 3 the second: one task each
        const bool par = big(B);
        for (int h = 0; h < 2; h++)
#pragma omp task if (par) default(shared) firstprivate(h)
            for (int q = 2 * h; q < 2 * h + 2; q++) {
                const Block& S = *B.s[q];
                hmul(S, alpha, X + (S.c0 - B.c0), ldx, ncol, Y + (S.r0 - B.r0), ldy);
            }
#pragma omp taskwait
        break;
    }
    }
}

// Y (rows of B's columns) += alpha B^T X (X: rows of B)
void hmulT(const Block& B, cd alpha, const cd* X, int64_t ldx, int64_t ncol, cd* Y, int64_t ldy)
{
    switch (B.type) {
    case Block::Full:
        gemm_tn(B.n, ncol, B.m, alpha, B.F.a.data(), B.F.m, X, ldx, Y, ldy, false);
        break;
    case Block::LowRank: {
        const int64_t k = B.U.n;
        if (!k) break;
        Mat T(k, ncol);
        gemm_tn(k, ncol, B.m, cd(1, 0), B.U.a.data(), B.U.m, X, ldx, T.a.data(), T.m, false);
        gemm_nn(B.n, ncol, k, alpha, B.W.a.data(), B.W.m, T.a.data(), T.m, Y, ldy);
        break;
    }
    case Block::Sub: {
        // sons 0, 2 write the first column half, 1, 3 the second
        const bool par = big(B);
        for (int h = 0; h < 2; h++)
#pragma omp task if (par) default(shared) firstprivate(h)
            for (int q = h; q < 4; q += 2) {
                const Block& S = *B.s[q];
                hmulT(S, alpha, X + (S.r0 - B.r0), ldx, ncol, Y + (S.c0 - B.c0), ldy);
            }
#pragma omp taskwait
        break;
    }
    }
}

void to_dense(const Block& B, cd* D, int64_t ldd)
{
    switch (B.type) {
    case Block::Full:
        for (int64_t j = 0; j < B.n; j++)
            for (int64_t i = 0; i < B.m; i++) D[i + j * ldd] = B.F(i, j);
        break;
    case Block::LowRank: {
        for (int64_t j = 0; j < B.n; j++)
            for (int64_t i = 0; i < B.m; i++) D[i + j * ldd] = cd(0, 0);
        gemm_nt(B.m, B.n, B.U.n, cd(1, 0), B.U.a.data(), B.U.m, B.W.a.data(), B.W.m, D, ldd);
        break;
    }
    case Block::Sub:
        for (int q = 0; q < 4; q++) {
            const Block& S = *B.s[q];
            to_dense(S, D + (S.r0 - B.r0) + (S.c0 - B.c0) * ldd, ldd);
        }
        break;
    }
}

Mat dense_of(const Block& B)
{
    Mat D(B.m, B.n);
    to_dense(B, D.a.data(), D.m);
    return D;
}

// C += alpha U W^T (U rows from C's first row, W rows from C's first column)
void add_lr(Block& C, const cd* U, int64_t ldu, const cd* W, int64_t ldw, int64_t k, cd alpha, const Params& p)
{
    if (k == 0) return;
    switch (C.type) {
    case Block::Full:
        gemm_nt(C.m, C.n, k, alpha, U, ldu, W, ldw, C.F.a.data(), C.F.m);
        break;
    case Block::LowRank:
        concat_truncate(C.U, C.W, U, ldu, W, ldw, k, alpha, p);
        break;
    case Block::Sub: {
        const bool par = big(C);
        for (int q = 0; q < 4; q++) {
            Block& S = *C.s[q];
#pragma omp task if (par) default(shared) firstprivate(q)
            add_lr(*C.s[q], U + (S.r0 - C.r0), ldu, W + (S.c0 - C.c0), ldw, k, alpha, p);
        }
#pragma omp taskwait
        break;
    }
    }
}

// A B as U W^T (A: tau x rho, B: rho x sigma), sums truncated with p
void lr_product(const Block& A, const Block& B, Mat& U, Mat& W, const Params& p)
{
    if (A.type == Block::LowRank) {
        const int64_t k = A.U.n;
        U = A.U;
        W = Mat(B.n, k);
        if (k) hmulT(B, cd(1, 0), A.W.a.data(), A.W.m, k, W.a.data(), W.m);
        return;
    }
    if (B.type == Block::LowRank) {
        const int64_t k = B.U.n;
        U = Mat(A.m, k);
        W = B.W;
        if (k) hmul(A, cd(1, 0), B.U.a.data(), B.U.m, k, U.a.data(), U.m);
        return;
    }
    if (A.type == Block::Full && B.type == Block::Full) {   // rank |rho|
        U = A.F;
        W = transposed(B.F);
        truncate(U, W, p);
        return;
    }
    if (A.type == Block::Full) {   // tau is a leaf: (A B)^T = B^T A^T, |tau| columns
        Mat At = transposed(A.F);
        W = Mat(B.n, A.m);
        hmulT(B, cd(1, 0), At.a.data(), At.m, A.m, W.a.data(), W.m);
        U = identity(A.m);
        truncate(U, W, p);
        return;
    }
    if (B.type == Block::Full) {   // sigma is a leaf: A B, |sigma| columns
        U = Mat(A.m, B.n);
        hmul(A, cd(1, 0), B.F.a.data(), B.F.m, B.n, U.a.data(), U.m);
        W = identity(B.n);
        truncate(U, W, p);
        return;
    }
    // both subdivided: the four son products, each a truncated sum, joined
    // and truncated once
    std::vector<Mat> Uq(4), Wq(4);
    const bool par = (double)A.m * (double)B.n >= kTaskMin;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
#pragma omp task if (par) default(shared) firstprivate(i, j)
        {
            Mat Us, Ws;
            for (int k = 0; k < 2; k++) {
                Mat Ut, Wt;
                lr_product(*A.s[2 * i + k], *B.s[2 * k + j], Ut, Wt, p);
                if (Us.m == 0 && Us.n == 0) {
                    Us = std::move(Ut);
                    Ws = std::move(Wt);
                } else {
                    concat_truncate(Us, Ws, Ut.a.data(), Ut.m, Wt.a.data(), Wt.m, Ut.n, cd(1, 0), p);
                }
            }
            const Block& Ai = *A.s[2 * i];
            const Block& Bj = *B.s[j];
            // pad to tau x sigma
            Mat Up(A.m, Us.n), Wp(B.n, Ws.n);
            for (int64_t l = 0; l < Us.n; l++) {
                for (int64_t r = 0; r < Us.m; r++) Up(Ai.r0 - A.r0 + r, l) = Us(r, l);
                for (int64_t r = 0; r < Ws.m; r++) Wp(Bj.c0 - B.c0 + r, l) = Ws(r, l);
            }
            Uq[(size_t)(2 * i + j)] = std::move(Up);
            Wq[(size_t)(2 * i + j)] = std::move(Wp);
        }
#pragma omp taskwait
    int64_t ktot = 0;
    for (const Mat& M : Uq) ktot += M.n;
    U = Mat(A.m, ktot);
    W = Mat(B.n, ktot);
    int64_t off = 0;
    for (size_t q = 0; q < Uq.size(); q++) {
        if (Uq[q].n) {
            std::memcpy(U.col(off), Uq[q].a.data(), sizeof(cd) * Uq[q].a.size());
            std::memcpy(W.col(off), Wq[q].a.data(), sizeof(cd) * Wq[q].a.size());
        }
        off += Uq[q].n;
    }
    truncate(U, W, p);
}

// C -= A B
void mulsub(Block& C, const Block& A, const Block& B, const Params& p)
{
    if (C.type == Block::Sub && A.type == Block::Sub && B.type == Block::Sub) {
        const bool par = big(C);
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) {
#pragma omp task if (par) default(shared) firstprivate(i, j)
                for (int k = 0; k < 2; k++) mulsub(*C.s[2 * i + j], *A.s[2 * i + k], *B.s[2 * k + j], p);
            }
#pragma omp taskwait
        return;
    }
    if (C.type == Block::Full && A.type != Block::LowRank && B.type != Block::LowRank) {
        if (C.n <= C.m) {   // C -= A dense(B)
            Mat Bd = dense_of(B);
            hmul(A, cd(-1, 0), Bd.a.data(), Bd.m, Bd.n, C.F.a.data(), C.F.m);
        } else {            // C^T -= B^T dense(A)^T
            Mat At = transposed(dense_of(A));
            Mat Pt(C.n, C.m);
            hmulT(B, cd(1, 0), At.a.data(), At.m, At.n, Pt.a.data(), Pt.m);
            for (int64_t j = 0; j < C.n; j++)
                for (int64_t i = 0; i < C.m; i++) C.F(i, j) -= Pt(j, i);
        }
        return;
    }
    Mat U, W;
    lr_product(A, B, U, W, p);
    add_lr(C, U.a.data(), U.m, W.a.data(), W.m, U.n, cd(-1, 0), p);
}

// ---------------------------------------------------- triangular pieces

// X := Q^H X for the rows of a factored leaf (X m x ncol, ldx)
void apply_qh(const Mat& Q, cd* X, int64_t ldx, int64_t ncol)
{
    Mat T(Q.n, ncol);
    gemm_tn(Q.n, ncol, Q.m, cd(1, 0), Q.a.data(), Q.m, X, ldx, T.a.data(), T.m, true);
    for (int64_t j = 0; j < ncol; j++)
        for (int64_t i = 0; i < Q.n; i++) X[i + j * ldx] = T(i, j);
}

// L X = X for the dense columns X (rows of L)
void solveL_dense(const Block& L, cd* X, int64_t ldx, int64_t ncol)
{
    if (L.type == Block::Full) {
        apply_qh(L.Q, X, ldx, ncol);
        return;
    }
    const Block &L11 = *L.s[0], &L21 = *L.s[2], &L22 = *L.s[3];
    solveL_dense(L11, X, ldx, ncol);
    hmul(L21, cd(-1, 0), X, ldx, ncol, X + L11.m, ldx);
    solveL_dense(L22, X + L11.m, ldx, ncol);
}

// U X = X for the dense columns X (rows of U)
void solveU_dense(const Block& U, cd* X, int64_t ldx, int64_t ncol)
{
    if (U.type == Block::Full) {
        const Mat& R = U.F;
        const int64_t m = R.m;
        for (int64_t j = 0; j < ncol; j++) {
            cd* x = X + j * ldx;
            for (int64_t i = m - 1; i >= 0; i--) {
                cd t = x[i];
                for (int64_t c = i + 1; c < m; c++) t -= R(i, c) * x[c];
                x[i] = t / R(i, i);
            }
        }
        return;
    }
    const Block &U11 = *U.s[0], &U12 = *U.s[1], &U22 = *U.s[3];
    solveU_dense(U22, X + U11.m, ldx, ncol);
    hmul(U12, cd(-1, 0), X + U11.m, ldx, ncol, X, ldx);
    solveU_dense(U11, X, ldx, ncol);
}

// U^T Y = Y for the dense columns Y (rows of U's columns)
void solveUT_dense(const Block& U, cd* Y, int64_t ldy, int64_t ncol)
{
    if (U.type == Block::Full) {
        const Mat& R = U.F;   // R^T is lower triangular
        const int64_t m = R.m;
        for (int64_t j = 0; j < ncol; j++) {
            cd* y = Y + j * ldy;
            for (int64_t i = 0; i < m; i++) {
                cd t = y[i];
                for (int64_t c = 0; c < i; c++) t -= R(c, i) * y[c];
                y[i] = t / R(i, i);
            }
        }
        return;
    }
    const Block &U11 = *U.s[0], &U12 = *U.s[1], &U22 = *U.s[3];
    solveUT_dense(U11, Y, ldy, ncol);
    hmulT(U12, cd(-1, 0), Y, ldy, ncol, Y + U11.m, ldy);
    solveUT_dense(U22, Y + U11.m, ldy, ncol);
}

// L X = Z with Z an H-matrix block (rows of L), in place
// (HArithm::forwSubsMatVal, harithm.cpp:1450-1517)
void solveL_block(const Block& L, Block& Z, const Params& p)
{
    if (Z.type == Block::Sub && L.type == Block::Sub) {
        const bool par = big(Z);
        for (int j = 0; j < 2; j++)
#pragma omp task if (par) default(shared) firstprivate(j)
        {
            Block &Z1 = *Z.s[j], &Z2 = *Z.s[2 + j];
            solveL_block(*L.s[0], Z1, p);
            mulsub(Z2, *L.s[2], Z1, p);
            solveL_block(*L.s[3], Z2, p);
        }
#pragma omp taskwait
        return;
    }
    if (Z.type == Block::Sub) throw Error{"H-LU: subdivided block right of a leaf diagonal block"};
    Mat& X = Z.type == Block::Full ? Z.F : Z.U;
    if (X.n) solveL_dense(L, X.a.data(), X.m, X.n);
}

// X U = Z with Z an H-matrix block (columns of U), in place
// (HArithm::forwSubsMatValTransposed, harithm.cpp:1528-1595)
void solveUr_block(const Block& U, Block& Z, const Params& p)
{
    if (Z.type == Block::Sub && U.type == Block::Sub) {
        const bool par = big(Z);
        for (int i = 0; i < 2; i++)
#pragma omp task if (par) default(shared) firstprivate(i)
        {
            Block &Z1 = *Z.s[2 * i], &Z2 = *Z.s[2 * i + 1];
            solveUr_block(*U.s[0], Z1, p);
            mulsub(Z2, Z1, *U.s[1], p);
            solveUr_block(*U.s[3], Z2, p);
        }
#pragma omp taskwait
        return;
    }
    if (Z.type == Block::Sub) throw Error{"H-LU: subdivided block below a leaf diagonal block"};
    if (Z.type == Block::LowRank) {   // (U W^T) U^-1 = U (U^-T W)^T
        if (Z.W.n) solveUT_dense(U, Z.W.a.data(), Z.W.m, Z.W.n);
        return;
    }
    Mat Ft = transposed(Z.F);   // X U = F  <=>  U^T X^T = F^T
    solveUT_dense(U, Ft.a.data(), Ft.m, Ft.n);
    Z.F = transposed(Ft);
}

// recursiveLUDecomposition (harithm.cpp:1204-1296) in place
void lu(Block& A, const Params& p)
{
    if (A.type == Block::Full) {
        if (A.m != A.n) throw Error{"H-LU: non-square diagonal leaf"};
        Mat Q, R;
        qr(A.F, Q, R);
        for (int64_t i = 0; i < R.m; i++)
            if (R(i, i) == cd(0, 0)) throw Error{"H-LU: singular diagonal leaf"};
        A.Q = std::move(Q);
        A.F = std::move(R);
        return;
    }
    if (A.type != Block::Sub) throw Error{"H-LU: low-rank diagonal block"};
    lu(*A.s[0], p);
    const bool par = big(A);
#pragma omp task if (par) default(shared)
    solveL_block(*A.s[0], *A.s[1], p);    // U12
#pragma omp task if (par) default(shared)
    solveUr_block(*A.s[0], *A.s[2], p);   // L21
#pragma omp taskwait
    mulsub(*A.s[3], *A.s[2], *A.s[1], p);
    lu(*A.s[3], p);
}

// dense L (lower = true) or U of a factored diagonal block
void factor_dense(const Block& A, bool lower, cd* D, int64_t ldd)
{
    if (A.type == Block::Full) {
        const Mat& S = lower ? A.Q : A.F;
        for (int64_t j = 0; j < A.n; j++)
            for (int64_t i = 0; i < A.m; i++) D[i + j * ldd] = S(i, j);
        return;
    }
    const Block &A11 = *A.s[0], &A12 = *A.s[1], &A21 = *A.s[2], &A22 = *A.s[3];
    factor_dense(A11, lower, D, ldd);
    factor_dense(A22, lower, D + A11.m + A11.n * ldd, ldd);
    if (lower) to_dense(A21, D + A11.m, ldd);
    else to_dense(A12, D + A11.n * ldd, ldd);
}

void count(const Block& B, double& entries, int64_t& maxk)
{
    switch (B.type) {
    case Block::Full: entries += (double)B.F.a.size() + (double)B.Q.a.size(); break;
    case Block::LowRank:
        entries += (double)(B.U.a.size() + B.W.a.size());
        maxk = std::max(maxk, B.U.n);
        break;
    case Block::Sub:
        for (int q = 0; q < 4; q++) count(*B.s[q], entries, maxk);
        break;
    }
}

struct Builder {
    std::map<std::pair<int64_t, int64_t>, int64_t> sons;             // (start, size) -> size of son 1
    std::map<std::tuple<int64_t, int64_t, int64_t, int64_t>, int64_t> part;
    const int64_t* rank;
    const double* const* data;
    const int64_t* blocks5;
    Params p;

    std::unique_ptr<Block> make(int64_t r0, int64_t m, int64_t c0, int64_t n)
    {
        auto B = std::make_unique<Block>();
        B->r0 = r0; B->m = m; B->c0 = c0; B->n = n;
        auto it = part.find(std::make_tuple(r0, m, c0, n));
        if (it != part.end()) {
            const int64_t b = it->second;
            const cd* src = reinterpret_cast<const cd*>(data[b]);
            if (rank[b] < 0) {
                B->type = Block::Full;
                B->F = Mat(m, n);
                std::memcpy(B->F.a.data(), src, sizeof(cd) * (size_t)(m * n));
            } else {
                const int64_t k = rank[b];
                B->type = Block::LowRank;
                B->U = Mat(m, k);
                B->W = Mat(n, k);
                const cd* Us = src;
                const cd* al = src + m * k;
                const cd* Vs = al + k;   // k x n column-major
                for (int64_t l = 0; l < k; l++) {
                    for (int64_t i = 0; i < m; i++) B->U(i, l) = Us[i + l * m] * al[l];
                    for (int64_t j = 0; j < n; j++) B->W(j, l) = Vs[l + j * k];
                }
                truncate(B->U, B->W, p);   // BlockCluster::compressAdmissibleBlock (hmatrix.cpp:101-140)
            }
            return B;
        }
        auto rs = sons.find({r0, m}), cs = sons.find({c0, n});
        if (rs == sons.end() || cs == sons.end())
            throw Error{"H-LU: block (" + std::to_string(r0) + "," + std::to_string(m) + ")x(" + std::to_string(c0) +
                        "," + std::to_string(n) + ") is neither in the partition nor splittable"};
        const int64_t m1 = rs->second, n1 = cs->second;
        B->type = Block::Sub;
        const int64_t rr[4] = {r0, r0, r0 + m1, r0 + m1}, mm[4] = {m1, m1, m - m1, m - m1};
        const int64_t cc[4] = {c0, c0 + n1, c0, c0 + n1}, nn[4] = {n1, n - n1, n1, n - n1};
        const bool par = (double)m * (double)n >= kTaskMin;
        std::string err[4];
        for (int q = 0; q < 4; q++)
#pragma omp task if (par) default(shared) firstprivate(q)
        {
            try {
                B->s[q] = make(rr[q], mm[q], cc[q], nn[q]);
            } catch (const Error& e) {
                err[q] = e.msg;
            }
        }
#pragma omp taskwait
        for (int q = 0; q < 4; q++)
            if (!err[q].empty()) throw Error{err[q]};
        return B;
    }
};

// the structure the factorisation relies on (checked before any task runs):
// diagonal blocks are square and full or subdivided, and no subdivided block
// shares its rows or columns with a full diagonal leaf
void leaf_diagonals(const Block& B, std::vector<std::pair<int64_t, int64_t>>& out)
{
    if (B.type == Block::Sub) {
        leaf_diagonals(*B.s[0], out);
        leaf_diagonals(*B.s[3], out);
        return;
    }
    if (B.type != Block::Full || B.m != B.n) throw Error{"H-LU: diagonal block neither square full nor subdivided"};
    out.push_back({B.r0, B.m});
}
void check_sub(const Block& B, const std::map<std::pair<int64_t, int64_t>, bool>& leafs)
{
    if (B.type != Block::Sub) return;
    if (leafs.count({B.r0, B.m}) || leafs.count({B.c0, B.n}))
        throw Error{"H-LU: subdivided block in the rows or columns of a diagonal leaf"};
    for (int q = 0; q < 4; q++) check_sub(*B.s[q], leafs);
}

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

HLU::HLU(int64_t n, int64_t nclusters, const int64_t* clusters2, int64_t nb, const int64_t* blocks5,
         const int64_t* rank, const double* const* data, Params p)
    : n_(n), p_(p)
{
    if (n <= 0 || nclusters <= 0 || nb <= 0) throw Error{"H-LU: empty matrix"};
    Builder bld;
    bld.rank = rank;
    bld.data = data;
    bld.blocks5 = blocks5;
    bld.p = p;
    // sons of each cluster: the largest cluster with the same start, and the rest
    std::map<int64_t, std::vector<int64_t>> by_start;
    for (int64_t c = 0; c < nclusters; c++) by_start[clusters2[2 * c]].push_back(clusters2[2 * c + 1]);
    std::map<std::pair<int64_t, int64_t>, bool> exists;
    for (int64_t c = 0; c < nclusters; c++) exists[{clusters2[2 * c], clusters2[2 * c + 1]}] = true;
    for (auto& kv : by_start) {
        auto& sz = kv.second;
        std::sort(sz.begin(), sz.end());
        sz.erase(std::unique(sz.begin(), sz.end()), sz.end());
        for (size_t q = 1; q < sz.size(); q++) {
            const int64_t s1 = sz[q - 1], size = sz[q];
            if (exists.count({kv.first + s1, size - s1})) bld.sons[{kv.first, size}] = s1;
        }
    }
    for (int64_t b = 0; b < nb; b++) {
        const int64_t* q = blocks5 + 5 * b;
        if ((q[4] != 0) != (rank[b] >= 0)) throw Error{"H-LU: block type and rank disagree"};
        bld.part[std::make_tuple(q[0], q[1], q[2], q[3])] = b;
    }
    std::string err;
#pragma omp parallel
#pragma omp single
    {
        try {
            root_ = bld.make(0, n, 0, n);
        } catch (const Error& e) {
            err = e.msg;
        }
    }
    if (!err.empty()) throw Error{err};
    std::vector<std::pair<int64_t, int64_t>> diag;
    leaf_diagonals(*root_, diag);
    std::map<std::pair<int64_t, int64_t>, bool> leafs;
    for (auto& d : diag) leafs[d] = true;
    check_sub(*root_, leafs);
}

void HLU::factor()
{
    auto t0 = std::chrono::steady_clock::now();
    std::string err;
#pragma omp parallel
#pragma omp single
    {
        try {
            lu(*root_, p_);
        } catch (const Error& e) {
            err = e.msg;
        }
    }
    if (!err.empty()) throw Error{err};
    factored_ = true;
    t_factor_ = seconds_since(t0);
}

void HLU::solve(const cd* b, cd* x) const
{
    if (!factored_) throw Error{"H-LU: not factored"};
    if (x != b) std::memcpy(x, b, sizeof(cd) * (size_t)n_);
#pragma omp parallel
#pragma omp single
    {
        solveL_dense(*root_, x, n_, 1);   // forwardSubstitution (harithm.cpp:1314-1340)
        solveU_dense(*root_, x, n_, 1);   // backwardSubstitution (harithm.cpp:1386-1412)
    }
}

Mat HLU::product() const
{
    Mat L(n_, n_), U(n_, n_), P(n_, n_);
    factor_dense(*root_, true, L.a.data(), n_);
    factor_dense(*root_, false, U.a.data(), n_);
    gemm_nn(n_, n_, n_, cd(1, 0), L.a.data(), n_, U.a.data(), n_, P.a.data(), n_);
    return P;
}

void HLU::stats(double* out3) const
{
    double entries = 0;
    int64_t maxk = 0;
    count(*root_, entries, maxk);
    out3[0] = entries;
    out3[1] = (double)maxk;
    out3[2] = t_factor_;
}

namespace {
// factored diagonal leaves and off-diagonal leaf blocks of the tree
void collect(const Block& B, std::vector<const Block*>& diag, std::vector<const Block*>& off)
{
    if (B.type == Block::Sub) {
        for (int q = 0; q < 4; q++) collect(*B.s[q], diag, off);
        return;
    }
    if (B.r0 == B.c0) {
        if (B.type != Block::Full || B.m != B.n || B.Q.m != B.m) throw Error{"H-LU: unfactored diagonal leaf"};
        diag.push_back(&B);
    } else {
        off.push_back(&B);
    }
}
}  // namespace

void HLU::flatten(Flat& F) const
{
    if (!factored_) throw Error{"H-LU: not factored"};
    std::vector<const Block*> diag, off;
    collect(*root_, diag, off);
    std::sort(diag.begin(), diag.end(), [](const Block* a, const Block* b) { return a->r0 < b->r0; });
    F = Flat();
    F.n = n_;
    const int64_t K = (int64_t)diag.size();
    std::vector<int64_t> leaf_of((size_t)n_, -1);
    int64_t s = 0;
    for (int64_t k = 0; k < K; k++) {
        const Block& D = *diag[k];
        if (D.r0 != s) throw Error{"H-LU: diagonal leaves do not tile the matrix"};
        const int64_t m = D.m;
        for (int64_t i = 0; i < m; i++) leaf_of[(size_t)(s + i)] = k;
        FlatLeaf L{s, m, (int64_t)F.data.size(), 0};
        for (int64_t j = 0; j < m; j++)   // Q^H
            for (int64_t i = 0; i < m; i++) F.data.push_back(std::conj(D.Q(j, i)));
        L.rinv = (int64_t)F.data.size();
        F.data.resize(F.data.size() + (size_t)(m * m), cd(0, 0));
        cd* X = F.data.data() + L.rinv;   // R X = I by back substitution, column by column
        for (int64_t j = 0; j < m; j++)
            for (int64_t i = j; i >= 0; i--) {
                cd t = i == j ? cd(1, 0) : cd(0, 0);
                for (int64_t c = i + 1; c <= j; c++) t -= D.F(i, c) * X[c + j * m];
                X[i + j * m] = t / D.F(i, i);
            }
        F.leaf.push_back(L);
        F.maxm = std::max(F.maxm, m);
        s += m;
    }
    if (s != n_) throw Error{"H-LU: diagonal leaves do not tile the matrix"};
    std::vector<std::vector<FlatItem>> tat[2], cat[2];
    for (int d = 0; d < 2; d++) {
        tat[d].resize((size_t)K);
        cat[d].resize((size_t)K);
    }
    for (const Block* pb : off) {
        const Block& B = *pb;
        const bool lower = B.r0 > B.c0;
        const int d = lower ? 0 : 1;
        if (lower ? B.c0 + B.n > B.r0 : B.r0 + B.m > B.c0) throw Error{"H-LU: off-diagonal block crosses the diagonal"};
        const int64_t k0 = leaf_of[(size_t)B.r0], k1 = leaf_of[(size_t)(B.r0 + B.m - 1)];
        if (F.leaf[k0].s != B.r0 || F.leaf[k1].s + F.leaf[k1].m != B.r0 + B.m)
            throw Error{"H-LU: block rows are not whole diagonal leaves"};
        if (B.type == Block::Full) {
            const int64_t a = (int64_t)F.data.size();
            F.data.insert(F.data.end(), B.F.a.begin(), B.F.a.end());
            for (int64_t k = k0; k <= k1; k++)
                cat[d][k].push_back(FlatItem{a + F.leaf[k].s - B.r0, B.m, B.c0, 0, (int32_t)B.n, -1});
            continue;
        }
        const int64_t r = B.U.n;
        if (r == 0) continue;   // a zero block
        const int64_t au = (int64_t)F.data.size();
        F.data.insert(F.data.end(), B.U.a.begin(), B.U.a.end());
        const int64_t aw = (int64_t)F.data.size();
        F.data.insert(F.data.end(), B.W.a.begin(), B.W.a.end());
        const int64_t t = F.tsize;
        F.tsize += r;
        tat[d][lower ? k0 : k1].push_back(FlatItem{aw, B.n, B.c0, t, (int32_t)B.n, (int32_t)r});
        for (int64_t k = k0; k <= k1; k++)
            cat[d][k].push_back(FlatItem{au + F.leaf[k].s - B.r0, B.m, B.c0, t, (int32_t)B.n, (int32_t)r});
    }
    for (int d = 0; d < 2; d++) {
        F.tptr[d].assign(1, 0);
        F.cptr[d].assign(1, 0);
        for (int64_t k = 0; k < K; k++) {
            F.titem[d].insert(F.titem[d].end(), tat[d][k].begin(), tat[d][k].end());
            F.citem[d].insert(F.citem[d].end(), cat[d][k].begin(), cat[d][k].end());
            F.tptr[d].push_back((int64_t)F.titem[d].size());
            F.cptr[d].push_back((int64_t)F.citem[d].size());
        }
    }
}

void Flat::solve(const cd* b, cd* x) const
{
    if (x != b) std::memcpy(x, b, sizeof(cd) * (size_t)n);
    std::vector<cd> t((size_t)tsize), acc((size_t)maxm);
    const int64_t K = (int64_t)leaf.size();
    for (int d = 0; d < 2; d++)
        for (int64_t s = 0; s < K; s++) {
            const int64_t k = d == 0 ? s : K - 1 - s;
            const FlatLeaf& L = leaf[k];
            for (int64_t q = tptr[d][k]; q < tptr[d][k + 1]; q++) {
                const FlatItem& T = titem[d][q];
                for (int l = 0; l < T.k; l++) {
                    cd v(0, 0);
                    for (int64_t j = 0; j < T.n; j++) v += data[T.a + j + l * T.ld] * x[T.c0 + j];
                    t[T.t + l] = v;
                }
            }
            for (int64_t i = 0; i < L.m; i++) acc[i] = x[L.s + i];
            for (int64_t q = cptr[d][k]; q < cptr[d][k + 1]; q++) {
                const FlatItem& C = citem[d][q];
                if (C.k < 0) {
                    for (int64_t j = 0; j < C.n; j++)
                        for (int64_t i = 0; i < L.m; i++) acc[i] -= data[C.a + i + j * C.ld] * x[C.c0 + j];
                } else {
                    for (int l = 0; l < C.k; l++)
                        for (int64_t i = 0; i < L.m; i++) acc[i] -= data[C.a + i + l * C.ld] * t[C.t + l];
                }
            }
            const cd* M = data.data() + (d == 0 ? L.qh : L.rinv);
            for (int64_t i = 0; i < L.m; i++) {
                cd v(0, 0);
                for (int64_t j = 0; j < L.m; j++) v += M[i + j * L.m] * acc[j];
                x[L.s + i] = v;
            }
        }
}

}  // namespace bemhlu

// ------------------------------------------------------------------ C ABI

struct bem_hlu {
    std::unique_ptr<bemhlu::HLU> h;
    std::unique_ptr<bemhlu::Flat> flat;   // formed on first use
};

void bemhlu::flatten(bem_hlu* h, Flat& F) { h->h->flatten(F); }

namespace {
thread_local std::string g_hlu_err;

template <class F>
int guarded(F&& f)
{
    try {
        return f();
    } catch (const bemhlu::Error& e) {
        g_hlu_err = e.msg;
    } catch (const std::exception& e) {
        g_hlu_err = e.what();
    }
    return BEM_ERR_ARG;
}
}  // namespace

extern "C" {

int bem_hlu_create(int64_t n, int64_t nclusters, const int64_t* clusters2, int64_t nb, const int64_t* blocks5,
                   const int64_t* rank, const double* const* data, int64_t max_rank, double rel_err, bem_hlu_t* out)
{
    if (!out || !clusters2 || !blocks5 || !rank || !data || n <= 0 || nclusters <= 0 || nb <= 0 || rel_err < 0) {
        g_hlu_err = "bem_hlu_create: bad arguments";
        return BEM_ERR_ARG;
    }
    *out = nullptr;
    return guarded([&]() {
        bemhlu::Params p;
        p.maxRank = max_rank;
        p.relErr = rel_err;
        auto H = std::make_unique<bem_hlu>();
        H->h = std::make_unique<bemhlu::HLU>(n, nclusters, clusters2, nb, blocks5, rank, data, p);
        H->h->factor();
        *out = H.release();
        return BEM_OK;
    });
}

int bem_hlu_solve(bem_hlu_t h, const double* b2, double* x2)
{
    if (!h || !b2 || !x2) return BEM_ERR_ARG;
    return guarded([&]() {
        h->h->solve(reinterpret_cast<const bemhlu::cd*>(b2), reinterpret_cast<bemhlu::cd*>(x2));
        return BEM_OK;
    });
}

int bem_hlu_solve_flat(bem_hlu_t h, const double* b2, double* x2)
{
    if (!h || !b2 || !x2) return BEM_ERR_ARG;
    return guarded([&]() {
        if (!h->flat) {
            auto F = std::make_unique<bemhlu::Flat>();
            h->h->flatten(*F);
            h->flat = std::move(F);
        }
        h->flat->solve(reinterpret_cast<const bemhlu::cd*>(b2), reinterpret_cast<bemhlu::cd*>(x2));
        return BEM_OK;
    });
}

int bem_hlu_dense(bem_hlu_t h, double* out2)
{
    if (!h || !out2) return BEM_ERR_ARG;
    return guarded([&]() {
        bemhlu::Mat P = h->h->product();
        std::memcpy(out2, P.a.data(), sizeof(bemhlu::cd) * P.a.size());
        return BEM_OK;
    });
}

int bem_hlu_stats(bem_hlu_t h, double* out3)
{
    if (!h || !out3) return BEM_ERR_ARG;
    h->h->stats(out3);
    return BEM_OK;
}

int bem_hlu_destroy(bem_hlu_t h)
{
    if (!h) return BEM_ERR_ARG;
    delete h;
    return BEM_OK;
}

const char* bem_hlu_error(void) { return g_hlu_err.c_str(); }

}  // extern "C"
