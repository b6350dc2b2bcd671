// Host H-LU of an assembled H-matrix, used as the GMRES preconditioner
// (HArithm::LUDecomposition / recursiveLUDecomposition,
// HMatrix/harithm.cpp:1192-1296, on the truncated copy
// HMatrix(const HMatrix&, rank, relError), hmatrix.cpp:770-776, 230-304).
//
// The factors keep the block partition of the matrix: diagonal leaves are
// factored A = Q R (the reference's qrInLUDecomp, harithm.h:13: L leaf = Q,
// U leaf = R), off-diagonal blocks of L and U come from triangular solves
// with H-matrix right-hand sides, and the Schur complements are formed with
// truncated H-matrix products.  Every truncation is QR of both factors + SVD
// of the small core, keeping the smallest rank whose discarded singular
// values hold less than relErr of the Frobenius norm, capped at maxRank
// (GoloubReinschSvd::minRankforError, golubReinschSVD.cpp:551-585).
#pragma once
#include <complex>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

struct bem_hlu;

namespace bemhlu {

using cd = std::complex<double>;

// column-major m x n
struct Mat {
    int64_t m = 0, n = 0;
    std::vector<cd> a;
    Mat() = default;
    Mat(int64_t rows, int64_t cols) : m(rows), n(cols), a((size_t)(rows * cols), cd(0, 0)) {}
    cd& operator()(int64_t i, int64_t j) { return a[(size_t)(i + j * m)]; }
    const cd& operator()(int64_t i, int64_t j) const { return a[(size_t)(i + j * m)]; }
    cd* col(int64_t j) { return a.data() + j * m; }
    const cd* col(int64_t j) const { return a.data() + j * m; }
};

struct Block {
    enum Type { Full, LowRank, Sub };
    Type type = Full;
    int64_t r0 = 0, m = 0, c0 = 0, n = 0;
    Mat F;      // Full: the m x n block; factored diagonal leaf: R
    Mat Q;      // factored diagonal leaf: the unitary factor
    Mat U, W;   // LowRank: U W^T (U m x k, W n x k)
    std::unique_ptr<Block> s[4];   // Sub: 11, 12, 21, 22
};

struct Params {
    int64_t maxRank = 2;   // bemparameters.h:27 (preconditionerRank)
    double relErr = 0;     // bemparameters.h:28 (preconditionerRelError)
};

struct Error {
    std::string msg;
};

// The factors flattened for a leaf-by-leaf substitution (the form the device
// preconditioner runs; forwardSubstitution / backwardSubstitution,
// harithm.cpp:1314-1412, in the same arithmetic up to the order of the
// additions).  Direction 0 solves L y = b over the diagonal leaves in
// increasing order, direction 1 solves U x = y in decreasing order; leaf k
// of direction d is
//   z_k := M_k (z_k - sum of the contributions of leaf k),
// M_k = Q_k^H (direction 0) or R_k^-1 (direction 1), after the low-rank
// blocks whose rows start (in that direction) at leaf k have formed their
// vectors t_B = W_B^T z_{cols B}.  A contribution is the leaf's row slice of
// an off-diagonal block times the solved part of z: full F (rows, columns
// c0 ..) or low-rank U (rows) t_B.  Every matrix is column-major in data.
struct FlatLeaf {
    int64_t s, m;        // rows s .. s + m - 1
    int64_t qh, rinv;    // offsets in data of Q^H and R^-1 (m x m)
};
struct FlatItem {
    int64_t a;     // offset in data: F or U at the leaf's first row, or W
    int64_t ld;    // leading dimension of that matrix
    int64_t c0;    // first column of the block
    int64_t t;     // offset of t_B in the t buffer (low-rank contribution, t item)
    int32_t n;     // columns of the block (of F for a full contribution)
    int32_t k;     // rank (low-rank contribution, t item); -1 for a full contribution
};
struct Flat {
    int64_t n = 0, tsize = 0, maxm = 0;
    std::vector<FlatLeaf> leaf;
    std::vector<int64_t> tptr[2], cptr[2];   // per direction: items of leaf k in [ptr[k], ptr[k + 1])
    std::vector<FlatItem> titem[2], citem[2];
    std::vector<cd> data;
    // x = U^-1 L^-1 b in the schedule's order (x may equal b)
    void solve(const cd* b, cd* x) const;
};

class HLU {
public:
    // clusters2: (start, size) of every cluster of the (row = column) cluster
    // tree, any order; blocks5: (row0, rows, col0, cols, admissible) of the
    // partition; rank[b] = -1: data[b] is the m x n block column-major,
    // else U (m x k column-major), alpha (k), V (k x n column-major) packed, the
    // block being U diag(alpha) V (interleaved complex doubles).
    HLU(int64_t n, int64_t nclusters, const int64_t* clusters2, int64_t nb, const int64_t* blocks5,
        const int64_t* rank, const double* const* data, Params p);
    void factor();
    // x = U^-1 L^-1 b (n complex each; x may equal b)
    void solve(const cd* b, cd* x) const;
    // the product L U as a dense n x n matrix (tests)
    Mat product() const;
    int64_t n() const { return n_; }
    // [0] stored complex entries, [1] largest rank, [2] factor seconds
    void stats(double* out3) const;
    // the factors in the leaf-by-leaf form (after factor())
    void flatten(Flat& F) const;

private:
    int64_t n_;
    Params p_;
    std::unique_ptr<Block> root_;
    double t_factor_ = 0;
    bool factored_ = false;
};

// the leaf-by-leaf form of a factorisation made by bem_hlu_create
void flatten(bem_hlu* h, Flat& F);

}  // namespace bemhlu
