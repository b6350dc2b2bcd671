// gfx950 kernels and the C ABI (include/bem_hip.h) of the H-matrix BEM path.
//
// Layout in HBM (all cluster order):
//   mesh      SoA doubles (collocation point, normal, area, 16 quadrature
//             points per element as q * n + j), near-singular CSR with its
//             polar sub-triangle geometry, self-integral radii
//   dense     every non-admissible block column-major, packed back to back,
//             complex FP64; processed as 64 x 64 tiles
//   low rank  U (m x k col-major), alpha (k), V (k x n row-major) per block,
//             packed back to back after compaction to the final ranks
//   MVM       per tile / per low-rank block partial vectors, summed per row
//             in a fixed order (no atomics: deterministic results)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/bem_hip.h"
#include "bem_hlu.hpp"
#include "bem_kernels.hpp"
#include "bem_setup.hpp"

using namespace bemdev;

namespace {

constexpr int kWG = 256;        // workgroup size of the block kernels (4 wavefronts)
constexpr int kTile = 64;       // dense tile edge
constexpr int kCapFirst = 48;   // first-pass ACA rank capacity per block
constexpr int kCapMax = 1016;   // second-pass capacity limit (FP64 LDS: (4 cap + 2) x 16 B <= 64 KiB)
// ACA workgroup size classes by m + n, per storage precision [FP64, FP32]:
// up to kSmallBlock one wavefront, up to kHugeBlock kMidThreads, above kHugeThreads
constexpr int64_t kSmallBlock[2] = {384, 384};
constexpr int64_t kHugeBlock[2] = {(int64_t)1 << 40, 2048};
constexpr double kTiny = 10.0 * 2.220446049250313e-16;  // global.h:53

struct Tile {
    int64_t dest;      // element offset of the tile origin in dense storage
    int64_t part;      // offset of this tile's partial vector (MVM)
    int32_t ld;        // rows of the owning block
    int32_t row0, col0, rows, cols;   // absolute first row/col, extent
    int32_t pad;
};

struct Fix {
    int64_t dest;      // element offset in dense storage
    int32_t i, j;
    int64_t t;         // near-singular pair index, -1 for the self entry
};

// ------------------------------------------------------------------ kernels

__global__ void k_quadpoints(int64_t n, const double* __restrict__ nodes, double* qx, double* qy, double* qz)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double* v = nodes + 9 * j;
    for (int q = 0; q < 16; q++) {
        double x1 = kTriX[q][0], x2 = kTriX[q][1];
        double x3 = (1.0 - x2) - x1;
        qx[q * n + j] = (x1 * v[0] + x2 * v[3]) + x3 * v[6];
        qy[q * n + j] = (x1 * v[1] + x2 * v[4]) + x3 * v[7];
        qz[q * n + j] = (x1 * v[2] + x2 * v[5]) + x3 * v[8];
    }
}

// The (row, column) pairs of a dense tile in column-major order, pair p on
// thread p mod kWG, so that every wavefront's lanes are busy but in the
// tile's last round.  (With one row per lane, blocks of fewer than 64 rows
// left lanes idle in every round and the few-row tiles of blocks just over
// 64 rows left most of them idle: three quarters of the lanes worked on the
// 20480-triangle sphere's tiles, against 99 % for the pair order.)
struct TilePairs {
    int r, c, dr, dc;
    __device__ TilePairs(int tid, int rows)
    {
        c = tid / rows;
        r = tid - c * rows;
        dc = kWG / rows;
        dr = kWG - dc * rows;
    }
    __device__ void next(int rows)   // pair p + kWG
    {
        r += dr;
        c += dc;
        if (r >= rows) {
            r -= rows;
            c++;
        }
    }
};

// Dense near-field tiles (K1): regular quadrature of every (row, col) pair of
// a 64 x 64 tile with the row and column data staged in LDS.  Self and
// near-singular pairs are skipped here and written by k_fixup.  The
// plane-free kernel is budgeted for four wavefronts per SIMD (no spills);
// the plane's would spill at that budget and keeps the compiler's choice.
template <bool PLANE>
__global__ __launch_bounds__(kWG, PLANE ? 1 : 4) void k_dense(DevMesh M, DevFreq F, const Tile* __restrict__ tiles,
                                               bem_c64* outPhi, bem_c64* outDPhi)
{
    __shared__ double sq[3][16][kTile];
    __shared__ double sn[3][kTile], sp[3][kTile], sa[kTile];
    __shared__ double rv[PLANE ? 9 : 1][kTile], rn[PLANE ? 3 : 1][kTile];   // mirror images
    __shared__ double srow[6][kTile];   // the rows' collocation points and normals
    const Tile T = tiles[blockIdx.x];
    const int tid = threadIdx.x;
    const int64_t n = M.n;
    for (int idx = tid; idx < 16 * kTile; idx += kWG) {
        int q = idx / kTile, c = idx % kTile;
        if (c < T.cols) {
            int64_t j = T.col0 + c;
            sq[0][q][c] = M.qx[q * n + j];
            sq[1][q][c] = M.qy[q * n + j];
            sq[2][q][c] = M.qz[q * n + j];
        }
    }
    for (int c = tid; c < T.cols; c += kWG) {
        int64_t j = T.col0 + c;
        sn[0][c] = M.nx[j]; sn[1][c] = M.ny[j]; sn[2][c] = M.nz[j];
        sp[0][c] = M.px[j]; sp[1][c] = M.py[j]; sp[2][c] = M.pz[j];
        sa[c] = M.area[j];
        if (PLANE) {
            for (int d = 0; d < 9; d++) rv[d][c] = M.rvtx[d * n + j];
            rn[0][c] = M.rnx[j]; rn[1][c] = M.rny[j]; rn[2][c] = M.rnz[j];
        }
    }
    for (int r = tid; r < T.rows; r += kWG) {
        const int64_t i = T.row0 + r;
        srow[0][r] = M.px[i]; srow[1][r] = M.py[i]; srow[2][r] = M.pz[i];
        srow[3][r] = M.nx[i]; srow[4][r] = M.ny[i]; srow[5][r] = M.nz[i];
    }
    __syncthreads();
    TilePairs w(tid, T.rows);
    for (int p = tid; p < T.rows * T.cols; p += kWG, w.next(T.rows)) {
        const int r = w.r, c = w.c;
        const int64_t i = T.row0 + r, j = T.col0 + c;
        const double px = srow[0][r], py = srow[1][r], pz = srow[2][r];
        const double npx = srow[3][r], npy = srow[4][r], npz = srow[5][r];
        double dx = px - sp[0][c], dy = py - sp[1][c], dz = pz - sp[2][c];
        if (i == j || sqrt(fma(dz, dz, fma(dy, dy, dx * dx))) < M.near_lim) continue;
        bem_ops4 o = regular_ops(F.k, px, py, pz, npx, npy, npz, sn[0][c], sn[1][c], sn[2][c], sa[c], &sq[0][0][c],
                                 &sq[1][0][c], &sq[2][0][c], kTile);
        bem_c64 X, Y;
        bem_xy(&o, F.cc, 0, &X, &Y);
        int64_t d = T.dest + r + (int64_t)c * T.ld;
        bem_c64 ep = bem_phi_entry(X, Y, F.subst[j], F.cphi[j]);
        bem_c64 ed = bem_dphi_entry(X, Y, F.subst[j], F.cd[j]);
        if (PLANE) {   // impedance plane: + the mirror image's entry (refl_xy on the staged image)
            const ElemVtx e{rv[0][c], rv[1][c], rv[2][c], rv[3][c], rv[4][c], rv[5][c], rv[6][c], rv[7][c], rv[8][c]};
            bem_ops4 ro = regular_ops_vtx<1>(F.k, px, py, pz, npx, npy, npz, rn[0][c], rn[1][c], rn[2][c], sa[c], e);
            bem_xy(&ro, F.cc, 0, &X, &Y);
            ep = c64_add(ep, bem_phi_entry(X, Y, F.subst[j], F.cphi[j]));
            ed = c64_add(ed, bem_dphi_entry(X, Y, F.subst[j], F.cd[j]));
        }
        if (outPhi) outPhi[d] = ep;
        if (outDPhi) outDPhi[d] = ed;
    }
}

// Self and near-singular entries (K2): every near-singular pair of the CSR
// (items 0..nnz-1, in the cost order `order`) and every self entry (items
// nnz..nnz+n-1), one per thread, into the per-pair table `sing` of the
// operator `kind`.  Dense blocks scatter from the table (k_scatter), the ACA
// reads it for the rare singular entries of admissible blocks.
__global__ __launch_bounds__(kWG) void k_singular(DevMesh M, DevFreq F, int kind, const int32_t* __restrict__ order,
                                                  int64_t nitems, int64_t nnz, bem_c64* sing, bem_c64* sing_dphi)
{
    int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nitems) return;
    int64_t t = order[u];
    bem_ops4 o;
    int64_t i, j;
    if (t >= nnz) {
        i = j = t - nnz;
        o = self_ops(F.k, M.self_geo + kSelf * i);
    } else {
        i = M.nr_row[t];
        j = M.nr_col[t];
        o = near_ops(F.k, M.px[i], M.py[i], M.pz[i], M.nx[i], M.ny[i], M.nz[i], M.nx[j], M.ny[j], M.nz[j],
                     M.nr_geo + M.nr_off[t], M.nr_nsub[t]);
    }
    // sing_dphi (optional): the DPhi entry of the same pair from the same
    // operators (the Phi assembly computing both near fields)
    bem_c64 e = bc_entry(F, kind, o, i, j);
    bem_c64 e2 = sing_dphi ? bc_entry(F, BEM_KIND_DPHI, o, i, j) : e;
    if (M.rvtx) {
        bem_c64 X, Y;
        refl_xy(M, F, i, j, &X, &Y);
        e = c64_add(e, kind == 0 ? bem_phi_entry(X, Y, F.subst[j], F.cphi[j]) : bem_dphi_entry(X, Y, F.subst[j], F.cd[j]));
        if (sing_dphi) e2 = c64_add(e2, bem_dphi_entry(X, Y, F.subst[j], F.cd[j]));
    }
    sing[t] = e;
    if (sing_dphi) sing_dphi[t] = e2;
}

// FP32 dense near-field tiles: as k_dense with the regular quadrature of
// regular_ops_f32, written as complex FP32
template <bool PLANE>
__global__ __launch_bounds__(kWG) void k_dense32(DevMesh M, DevFreq32 F, const int32_t* __restrict__ subst, int kind,
                                                 const Tile* __restrict__ tiles, bem_c32* out, bem_c32* out_dphi)
{
    __shared__ double sq[3][16][kTile];
    __shared__ double sp[3][kTile];
    __shared__ float sn[3][kTile], sa[kTile];
    __shared__ double spr[3][kTile];   // the rows' collocation points and normals
    __shared__ float snr[3][kTile];
    const Tile T = tiles[blockIdx.x];
    const int tid = threadIdx.x;
    const int64_t n = M.n;
    for (int idx = tid; idx < 16 * kTile; idx += kWG) {
        int q = idx / kTile, c = idx % kTile;
        if (c < T.cols) {
            int64_t j = T.col0 + c;
            sq[0][q][c] = M.qx[q * n + j];
            sq[1][q][c] = M.qy[q * n + j];
            sq[2][q][c] = M.qz[q * n + j];
        }
    }
    for (int c = tid; c < T.cols; c += kWG) {
        int64_t j = T.col0 + c;
        sn[0][c] = (float)M.nx[j]; sn[1][c] = (float)M.ny[j]; sn[2][c] = (float)M.nz[j];
        sp[0][c] = M.px[j]; sp[1][c] = M.py[j]; sp[2][c] = M.pz[j];
        sa[c] = (float)M.area[j];
    }
    for (int r = tid; r < T.rows; r += kWG) {
        const int64_t i = T.row0 + r;
        spr[0][r] = M.px[i]; spr[1][r] = M.py[i]; spr[2][r] = M.pz[i];
        snr[0][r] = (float)M.nx[i]; snr[1][r] = (float)M.ny[i]; snr[2][r] = (float)M.nz[i];
    }
    __syncthreads();
    TilePairs w(tid, T.rows);
    for (int p = tid; p < T.rows * T.cols; p += kWG, w.next(T.rows)) {
        const int r = w.r, c = w.c;
        const int64_t i = T.row0 + r, j = T.col0 + c;
        const double px = spr[0][r], py = spr[1][r], pz = spr[2][r];
        const float npx = snr[0][r], npy = snr[1][r], npz = snr[2][r];
        double dx = px - sp[0][c], dy = py - sp[1][c], dz = pz - sp[2][c];
        if (i == j || sqrt(fma(dz, dz, fma(dy, dy, dx * dx))) < M.near_lim) continue;
        bem_c32f4 o = regular_ops_f32(F, px, py, pz, npx, npy, npz, sn[0][c], sn[1][c], sn[2][c], sa[c],
                                      &sq[0][0][c], &sq[1][0][c], &sq[2][0][c], kTile);
        bem_c32 e = bc_entry_f32(F, subst, kind, o, j);
        bem_c32 e2 = out_dphi ? bc_entry_f32(F, subst, BEM_KIND_DPHI, o, j) : e;   // same operators
        if (PLANE) {   // impedance plane: + the mirror image's entry
            const bem_c32f4 ro = refl_ops_f32(F, M, i, j);
            bem_c32 m = bc_entry_f32(F, subst, kind, ro, j);
            e = c32(e.re + m.re, e.im + m.im);
            if (out_dphi) {
                m = bc_entry_f32(F, subst, BEM_KIND_DPHI, ro, j);
                e2 = c32(e2.re + m.re, e2.im + m.im);
            }
        }
        out[T.dest + r + (int64_t)c * T.ld] = e;
        if (out_dphi) out_dphi[T.dest + r + (int64_t)c * T.ld] = e2;
    }
}

// singular entries of the dense blocks from the (FP64) table
template <class C>
__global__ void k_scatter(const Fix* __restrict__ fx, int64_t nfix, int64_t nnz, const bem_c64* __restrict__ sing,
                          C* out)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfix) return;
    Fix f = fx[t];
    bem_c64 v = sing[f.t >= 0 ? f.t : nnz + f.i];
    C o;
    o.re = v.re;
    o.im = v.im;
    out[f.dest] = o;
}

// FP32 copy of an FP64 near field (the operator GMRES iterates with when the
// FP64 one is kept for the refinement residual)
__global__ void k_round32(int64_t len, const bem_c64* __restrict__ src, bem_c32* __restrict__ dst)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= len) return;
    const bem_c64 v = src[t];
    dst[t] = c32((float)v.re, (float)v.im);
}

__global__ void k_entries(DevMesh M, DevFreq F, int mode, int64_t np, const int64_t* ij, bem_c64* out)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    int64_t i = ij[2 * t], j = ij[2 * t + 1];
    if (mode == 2) {   // the free-space operators
        bem_ops4 o = ops_general(M, F.k, i, j);
        out[4 * t] = o.L; out[4 * t + 1] = o.M; out[4 * t + 2] = o.Mt; out[4 * t + 3] = o.N;
    } else {
        out[t] = entry_general(M, F, mode, i, j);
    }
}

// ------------------------------------------------------ field evaluation

__constant__ static const double kTri7[7][3] = BEM_TRI7;
constexpr int64_t kFieldChunk = 1024;   // boundary elements per partial sum of the field

// L and M of one element for the observation point p (BemOperatorField,
// boundaryelementsolver.cpp:2319-2353; vertices and normal from LDS column c)
__device__ __forceinline__ void field_ops(double k, double px, double py, double pz, const double (*sv)[kTile],
                                          const double (*sn)[kTile], int c, double area, bem_c64* L, bem_c64* Mk)
{
    *L = c64(0, 0);
    *Mk = c64(0, 0);
#pragma unroll
    for (int q = 0; q < 7; q++) {
        double x1 = kTri7[q][0], x2 = kTri7[q][1];
        double x3 = (1.0 - x2) - x1;
        double qx = (x1 * sv[0][c] + x2 * sv[3][c]) + x3 * sv[6][c];
        double qy = (x1 * sv[1][c] + x2 * sv[4][c]) + x3 * sv[7][c];
        double qz = (x1 * sv[2][c] + x2 * sv[5][c]) + x3 * sv[8][c];
        bem_field_point_accumulate(L, Mk, k, px, py, pz, sn[0][c], sn[1][c], sn[2][c], qx, qy, qz, kTri7[q][2]);
    }
    *L = c64_scale(*L, area);
    *Mk = c64_scale(*Mk, area);
}

// M phi - L dphi of one element
__device__ __forceinline__ bem_c64 field_term(double k, double px, double py, double pz, const double (*sv)[kTile],
                                              const double (*sn)[kTile], int c, double area, bem_c64 phi,
                                              bem_c64 dphi)
{
    bem_c64 L, Mk;
    field_ops(k, px, py, pz, sv, sn, c, area, &L, &Mk);
    return bem_field_term(L, Mk, phi, dphi);
}

// Field potential of the boundary solution at observation points
// (calculateFieldSolutionRegular / BemOperatorField,
// boundaryelementsolver.cpp:2262-2353), one 64-point tile x one chunk of
// kFieldChunk elements per workgroup.  With an impedance plane each element's
// term is followed by its mirror image's (:2305-2311) and the two are added
// first.  Canonical order: inside a chunk four partial sums over element index
// mod 4 (ascending), combined (s0 + s1) + (s2 + s3); k_field_reduce adds the
// chunks in order.
__global__ __launch_bounds__(kWG) void k_field(DevMesh M, double k, int64_t nobs, const double* __restrict__ obs,
                                               const bem_c64* __restrict__ phi, const bem_c64* __restrict__ dphi,
                                               int nchunk, bem_c64* __restrict__ part)
{
    __shared__ double sv[9][kTile], rv[9][kTile];
    __shared__ double sn[3][kTile], rn[3][kTile], sa[kTile];
    __shared__ bem_c64 sphi[kTile], sdphi[kTile];
    __shared__ bem_c64 red[4][kTile];
    const int tid = threadIdx.x, r = tid & (kTile - 1), cg = tid >> 6;
    const int64_t n = M.n;
    const int64_t i = (int64_t)blockIdx.x * kTile + r;
    const bool live = i < nobs;
    const double px = live ? obs[i] : 0.0, py = live ? obs[nobs + i] : 0.0, pz = live ? obs[2 * nobs + i] : 0.0;
    const int64_t j0 = (int64_t)blockIdx.y * kFieldChunk;
    const int64_t j1 = j0 + kFieldChunk < n ? j0 + kFieldChunk : n;
    bem_c64 acc = c64(0, 0);
    for (int64_t t0 = j0; t0 < j1; t0 += kTile) {
        const int cnt = (int)(j1 - t0 < kTile ? j1 - t0 : kTile);
        __syncthreads();
        for (int idx = tid; idx < 9 * kTile; idx += kWG) {
            int d = idx / kTile, c = idx % kTile;
            if (c < cnt) {
                sv[d][c] = M.vtx[(int64_t)d * n + t0 + c];
                if (M.rvtx) rv[d][c] = M.rvtx[(int64_t)d * n + t0 + c];
            }
        }
        for (int c = tid; c < cnt; c += kWG) {
            sn[0][c] = M.nx[t0 + c]; sn[1][c] = M.ny[t0 + c]; sn[2][c] = M.nz[t0 + c];
            if (M.rvtx) { rn[0][c] = M.rnx[t0 + c]; rn[1][c] = M.rny[t0 + c]; rn[2][c] = M.rnz[t0 + c]; }
            sa[c] = M.area[t0 + c];
            sphi[c] = phi[t0 + c];
            sdphi[c] = dphi[t0 + c];
        }
        __syncthreads();
        if (!live) continue;
        for (int c = cg; c < cnt; c += 4) {
            bem_c64 v = field_term(k, px, py, pz, sv, sn, c, sa[c], sphi[c], sdphi[c]);
            if (M.rvtx) v = c64_add(v, field_term(k, px, py, pz, rv, rn, c, sa[c], sphi[c], sdphi[c]));
            acc = c64_add(acc, v);
        }
    }
    red[cg][r] = acc;
    __syncthreads();
    if (cg == 0 && live)
        part[i * nchunk + blockIdx.y] = c64_add(c64_add(red[0][r], red[1][r]), c64_add(red[2][r], red[3][r]));
}

// ---- field through the observation x boundary H-matrix
// (calculateFieldSolutionFast, boundaryelementsolver.cpp:2397-2476).
// A dense block's 64 x 64 tile: partial[t][r] = sum_c M phi_c - sum_c L dphi_c
// over the tile's columns (column c taken by thread group c mod 4; the four
// sums of each operator combined (s0 + s1) + (s2 + s3) before the difference).
struct FieldTile {
    int64_t row0, col0, part;
    int32_t rows, cols;
};

__global__ __launch_bounds__(kWG) void k_field_dense(DevMesh M, double k, int64_t nobs, const double* __restrict__ obs,
                                                     const FieldTile* __restrict__ tiles,
                                                     const bem_c64* __restrict__ phi, const bem_c64* __restrict__ dphi,
                                                     bem_c64* __restrict__ part)
{
    __shared__ double sv[9][kTile];
    __shared__ double sn[3][kTile], sa[kTile];
    __shared__ bem_c64 sphi[kTile], sdphi[kTile];
    __shared__ bem_c64 redM[4][kTile], redL[4][kTile];
    const FieldTile T = tiles[blockIdx.x];
    const int tid = threadIdx.x, r = tid & (kTile - 1), cg = tid >> 6;
    const int64_t n = M.n;
    for (int idx = tid; idx < 9 * kTile; idx += kWG) {
        int d = idx / kTile, c = idx % kTile;
        if (c < T.cols) sv[d][c] = M.vtx[(int64_t)d * n + T.col0 + c];
    }
    for (int c = tid; c < T.cols; c += kWG) {
        sn[0][c] = M.nx[T.col0 + c]; sn[1][c] = M.ny[T.col0 + c]; sn[2][c] = M.nz[T.col0 + c];
        sa[c] = M.area[T.col0 + c];
        sphi[c] = phi[T.col0 + c];
        sdphi[c] = dphi[T.col0 + c];
    }
    __syncthreads();
    bem_c64 sM = c64(0, 0), sL = c64(0, 0);
    if (r < T.rows) {
        const int64_t i = T.row0 + r;
        const double px = obs[i], py = obs[nobs + i], pz = obs[2 * nobs + i];
        for (int c = cg; c < T.cols; c += 4) {
            bem_c64 L, Mk;
            field_ops(k, px, py, pz, sv, sn, c, sa[c], &L, &Mk);
            sM = c64_fma(Mk, sphi[c], sM);
            sL = c64_fma(L, sdphi[c], sL);
        }
    }
    redM[cg][r] = sM;
    redL[cg][r] = sL;
    __syncthreads();
    if (cg == 0 && r < T.rows)
        part[T.part + r] = c64_sub(c64_add(c64_add(redM[0][r], redM[1][r]), c64_add(redM[2][r], redM[3][r])),
                                   c64_add(c64_add(redL[0][r], redL[1][r]), c64_add(redL[2][r], redL[3][r])));
}

__global__ void k_field_reduce(int64_t nobs, int nchunk, const bem_c64* __restrict__ part, bem_c64* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nobs) return;
    bem_c64 acc = part[i * nchunk];
    for (int q = 1; q < nchunk; q++) acc = c64_add(acc, part[i * nchunk + q]);
    out[i] = acc;
}

// Fixed device scratch of bem_field_noalloc (the field's per-chunk partial
// sums), one copy per GPU for the life of the process.  Host code never
// passes `scratch` itself to a kernel: its device address on each GPU comes
// from hipGetSymbolAddress (scratch_address below).
constexpr size_t kFieldScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) char scratch[kFieldScratchBytes];

// -------------------------------------------------- workgroup reductions

__device__ __forceinline__ double wave_sum(double x)
{
    for (int s = 32; s > 0; s >>= 1) x = x + __shfl_down(x, s, 64);
    return x;
}
__device__ __forceinline__ bem_c64 wave_sum(bem_c64 x)
{
    for (int s = 32; s > 0; s >>= 1) {
        x.re = x.re + __shfl_down(x.re, s, 64);
        x.im = x.im + __shfl_down(x.im, s, 64);
    }
    return x;
}

// An admissible block's two low-rank operators (the M kernel's factors and
// the L kernel's, from the ACA scratch): partial[i] = (U0 diag(A0) V0 phi_J)_i
// - (U1 diag(A1) V1 dphi_J)_i, one workgroup per block; t_l = A_l (V_l . x_J)
// by wavefront l mod 4 (lane-strided sums and the wavefront tree), then the
// rows in order l = 0, 1, ...
struct FieldLR {
    const bem_c64 *U0, *V0, *A0, *U1, *V1, *A1;
    int32_t r0, r1;
    int64_t rows, col0, cols, part;
};

__global__ __launch_bounds__(kWG) void k_field_lr(const FieldLR* __restrict__ blocks, const bem_c64* __restrict__ phi,
                                                  const bem_c64* __restrict__ dphi, bem_c64* __restrict__ part)
{
    extern __shared__ bem_c64 st[];   // t of the M kernel (r0), then of the L kernel (r1)
    const FieldLR B = blocks[blockIdx.x];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int l = w; l < B.r0 + B.r1; l += kWG / 64) {
        const bool second = l >= B.r0;
        const int ll = second ? l - B.r0 : l;
        const bem_c64* v = (second ? B.V1 : B.V0) + (int64_t)ll * B.cols;
        const bem_c64* x = (second ? dphi : phi) + B.col0;
        bem_c64 p = c64(0, 0);
        for (int64_t j = lane; j < B.cols; j += 64) p = c64_fma(v[j], x[j], p);
        p = wave_sum(p);
        if (lane == 0) st[l] = c64_mul((second ? B.A1 : B.A0)[ll], p);
    }
    __syncthreads();
    for (int64_t i = threadIdx.x; i < B.rows; i += kWG) {
        bem_c64 a0 = c64(0, 0), a1 = c64(0, 0);
        for (int l = 0; l < B.r0; l++) a0 = c64_fma(B.U0[(int64_t)l * B.rows + i], st[l], a0);
        for (int l = 0; l < B.r1; l++) a1 = c64_fma(B.U1[(int64_t)l * B.rows + i], st[B.r0 + l], a1);
        part[B.part + i] = c64_sub(a0, a1);
    }
}


// ---- complex arithmetic over the ACA storage type: bem_c64 uses the
// bit-exact FP64 contract of bem_detmath.h, bem_c32 the same formulas in FP32
__device__ __forceinline__ bem_c64 cx_mul(bem_c64 a, bem_c64 b) { return c64_mul(a, b); }
__device__ __forceinline__ bem_c32 cx_mul(bem_c32 a, bem_c32 b) { return c32_mul(a, b); }
__device__ __forceinline__ bem_c64 cx_fma(bem_c64 a, bem_c64 b, bem_c64 acc) { return c64_fma(a, b, acc); }
__device__ __forceinline__ bem_c32 cx_fma(bem_c32 a, bem_c32 b, bem_c32 acc) { return c32_fma(a, b, acc); }
__device__ __forceinline__ bem_c64 cx_add(bem_c64 a, bem_c64 b) { return c64_add(a, b); }
__device__ __forceinline__ bem_c32 cx_add(bem_c32 a, bem_c32 b) { return c32(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ bem_c64 cx_sub(bem_c64 a, bem_c64 b) { return c64_sub(a, b); }
__device__ __forceinline__ bem_c32 cx_sub(bem_c32 a, bem_c32 b) { return c32(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ double cx_abs2(bem_c64 a) { return c64_abs2(a); }
__device__ __forceinline__ double cx_abs2(bem_c32 a) { return (double)fmaf(a.re, a.re, a.im * a.im); }
__device__ __forceinline__ bem_c64 cx_inv(bem_c64 a) { return c64_div(c64(1.0, 0.0), a); }
__device__ __forceinline__ bem_c32 cx_inv(bem_c32 a)
{
    float d = fmaf(a.re, a.re, a.im * a.im);
    return c32(a.re / d, -a.im / d);
}
template <class C> __device__ __forceinline__ C cx_zero();
template <> __device__ __forceinline__ bem_c64 cx_zero<bem_c64>() { return c64(0, 0); }
template <> __device__ __forceinline__ bem_c32 cx_zero<bem_c32>() { return c32(0, 0); }
__device__ __forceinline__ bem_c64 ld64(const bem_c64* p) { return *p; }
__device__ __forceinline__ bem_c64 ld64(const bem_c32* p)
{
    bem_c32 v = *p;
    return c64((double)v.re, (double)v.im);
}
// loads from device global memory through a pointer the compiler cannot
// place (factor storage reached through argument arrays): the explicit global
// address space gives global_load with a scalar base instead of flat loads,
// which also wait on the LDS counter
typedef float bem_v2f __attribute__((ext_vector_type(2)));
typedef double bem_v2d __attribute__((ext_vector_type(2)));
__device__ __forceinline__ bem_c32 ldg(const bem_c32* p)
{
    const bem_v2f x = *(const __attribute__((address_space(1))) bem_v2f*)p;
    return c32(x.x, x.y);
}
__device__ __forceinline__ bem_c64 ldg(const bem_c64* p)
{
    const bem_v2d x = *(const __attribute__((address_space(1))) bem_v2d*)p;
    return c64(x.x, x.y);
}
__device__ __forceinline__ bem_c64 ldg64(const bem_c64* p) { return ldg(p); }
__device__ __forceinline__ bem_c64 ldg64(const bem_c32* p)
{
    bem_c32 v = ldg(p);
    return c64((double)v.re, (double)v.im);
}
template <class C> __device__ __forceinline__ C from64(bem_c64 z);
template <> __device__ __forceinline__ bem_c64 from64<bem_c64>(bem_c64 z) { return z; }
template <> __device__ __forceinline__ bem_c32 from64<bem_c32>(bem_c64 z) { return c32((float)z.re, (float)z.im); }

struct ArgMax { double v; int64_t i; };

__device__ __forceinline__ ArgMax better(ArgMax a, ArgMax b)
{
    if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
    return a;
}

// first index of the largest |v_i|^2 over [0, len); index `excl` counts as 0.
// NT threads (NT / 64 wavefronts); every thread gets the result.  GLOBAL: v
// lies in device global memory (the ACA's factors) and is read through the
// global address space; otherwise (e.g. LDS) through generic loads.
template <bool GLOBAL, class C> __device__ __forceinline__ C ld_any(const C* p)
{
    if constexpr (GLOBAL) return ldg(p);
    else return *p;
}
template <int NT, class C, bool GLOBAL = false>
__device__ ArgMax block_argmax(const C* __restrict__ v, int64_t len, int64_t excl, ArgMax* sred)
{
    const int tid = threadIdx.x;
    ArgMax a{-1.0, (int64_t)1 << 62};
    int64_t t = tid;
    // four elements per thread loaded together, then compared in index order
    for (; t + 3 * NT < len; t += 4 * NT) {
        C e[4];
#pragma unroll
        for (int u = 0; u < 4; u++) e[u] = ld_any<GLOBAL>(v + t + u * NT);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t tu = t + u * NT;
            double x = tu == excl ? 0.0 : cx_abs2(e[u]);
            if (x > a.v) { a.v = x; a.i = tu; }
        }
    }
    for (; t < len; t += NT) {
        double x = t == excl ? 0.0 : cx_abs2(ld_any<GLOBAL>(v + t));
        if (x > a.v) { a.v = x; a.i = t; }
    }
    for (int s = 32; s > 0; s >>= 1) {
        ArgMax o;
        o.v = __shfl_xor(a.v, s, 64);
        o.i = __shfl_xor(a.i, s, 64);
        a = better(a, o);
    }
    if (NT == 64) return a;
    __syncthreads();
    if ((tid & 63) == 0) sred[tid >> 6] = a;
    __syncthreads();
    ArgMax r = sred[0];
    for (int w = 1; w < NT / 64; w++) r = better(r, sred[w]);
    return r;
}

struct MeshJob;

struct AcaArgs {
    DevMesh M;
    DevFreq F;
    DevFreq32 F32;
    int kind;
    const int32_t* blist;
    const int64_t *row0, *rows, *col0, *cols;
    const int64_t *uoff, *voff, *aoff;
    const int32_t* cap;
    void *U, *V, *A;      // storage of the ACA type C
    int32_t* rank_out;
    int32_t* flag_out;    // 1: rank capacity reached, 2: non-finite
    int64_t* evals_out;
    double tol;
    int64_t maxrank;
    uint64_t seed;
    const int64_t* piv_off;
    const int32_t* piv;
    const bem_c64* sing;  // self / near-singular entries (k_singular), FP64
    int64_t nnz;
    // several frequencies in one launch (nf > 0): list entries are items
    // f * nbk + b; frequency f has wavenumber kf[f], coupling ccf[f], its
    // singular table singf[f], factor storage Uf/Vf/Af[f] (the same per-block
    // offsets for every f) and its results at rank_out / flag_out /
    // evals_out + f * nbk
    int nf;
    int64_t nbk;
    const double* kf;
    const bem_c64* ccf;
    const bem_c64* const* singf;
    void* const* Uf;
    void* const* Vf;
    void* const* Af;
    // persistent launch: workgroups take list positions from this counter
    // (zero at launch) until nitems are taken; null: one workgroup per item
    int* queue;
    int nitems;
    // field operators (FIELD instantiations): rows are observation points
    // obs[d * nobs + i] (cluster order of the observation tree), columns
    // boundary elements; kind 0 the M kernel (multiplies phi), 1 the L
    // kernel (multiplies dphi)
    const double* obs = nullptr;
    int64_t nobs = 0;
    // several contexts (meshes) in one launch: list entries are
    // (job << kJobShift) | block, job j's mesh, frequency, partition, ACA
    // scratch and results in jobs[j]
    const MeshJob* jobs = nullptr;
    // per block: 0 when no (row, column) pair of the block can be a self or
    // near-singular pair (its collocation boxes lie apart by more than the
    // near limit), so that its rows and columns skip the near-pair checks;
    // null: every block is checked
    const int8_t* bnear = nullptr;
};

// One context's part of a multi-mesh ACA launch: the AcaArgs fields that
// differ between contexts
struct MeshJob {
    DevMesh M;
    DevFreq F;
    DevFreq32 F32;
    const int64_t *row0, *rows, *col0, *cols;
    const int64_t *uoff, *voff, *aoff;
    const int32_t* cap;
    void *U, *V, *A;
    int32_t* rank_out;
    int32_t* flag_out;
    int64_t* evals_out;
    const int64_t* piv_off;
    const int32_t* piv;
    const bem_c64* sing;
    int64_t nnz;
    const int8_t* bnear;
};
constexpr int kJobShift = 20;                       // blocks per context < 2^20
constexpr int32_t kJobMask = (1 << kJobShift) - 1;

// The dot products are processed kDotGroup at a time per wavefront: per
// element the new vector and the group's basis vectors are loaded together
// and the group's totals are reduced together (wave_sum_spread); each task
// keeps the canonical lane-strided / pairwise order.  Accumulation is FP64
// for both types.
constexpr int kDotGroup = 4;

// Totals over the 64 lanes of NV values per lane (NV a power of two up to
// 64), each by the pairwise tree of the device reduction order (partial p
// plus partial p + s for s = 32, 16, ..., 1), with the values spread over
// the lanes as the tree narrows: while a lane holds more than one value it
// keeps half of them at each step and trades the other half with its partner
// (lane ^ s), so that each addition pairs the same two partials as the tree
// (addition commutes, so the sums are the same bits) with far fewer lane
// exchanges (34 instead of 192 for 16 values).  Lane (64 / NV) v returns the
// total of value v.
template <int NV>
__device__ __forceinline__ double wave_sum_spread(const double (&x)[NV], int lane)
{
    static_assert(NV >= 1 && NV <= 64 && (NV & (NV - 1)) == 0, "NV: a power of two up to 64");
    double cur[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) cur[v] = x[v];
    int nv = NV;
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) {
        if (nv > 1) {
            const bool hi = lane & s;
            const int h = nv / 2;
#pragma unroll
            for (int j = 0; j < h; j++) {
                const double send = hi ? cur[j] : cur[h + j], keep = hi ? cur[h + j] : cur[j];
                cur[j] = keep + __shfl_xor(send, s, 64);
            }
            nv = h;
        } else {
            cur[0] = cur[0] + __shfl_xor(cur[0], s, 64);
        }
    }
    return cur[0];
}

// tasks over vector v (length len) against basis W (stride len): k = 0 is
// |v|^2 into out[base], k >= 1 is v^H W_{k-1} into out[base + 2k].  The
// group's loads of one element are issued together, unconditionally: a task
// without a basis vector (k = 0, or k >= ntask) reads v itself and discards
// the value, so that no load sits behind a branch and the group pays one
// memory latency per element instead of one per task.
template <int NT, class C>
__device__ __forceinline__ void dot_group_pass(const C* __restrict__ v, const C* __restrict__ W, int64_t len,
                                               int ntask, bem_c64* out, int base)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int kb = w * kDotGroup; kb < ntask; kb += (NT / 64) * kDotGroup) {
        const C* src[kDotGroup];
        bem_c64 acc[kDotGroup];
#pragma unroll
        for (int g = 0; g < kDotGroup; g++) {
            const int k = kb + g;
            src[g] = (k >= 1 && k < ntask) ? W + (int64_t)(k - 1) * len : v;
            acc[g] = c64(0, 0);
        }
        // branch-free: every task accumulates (the products of tasks without
        // a basis vector are never written out) and |v|^2 is accumulated
        // beside them, taking task 0's place in the first group
        double nrm = 0.0;
        for (int64_t t = lane; t < len; t += 64) {
            const bem_c64 cv = ldg64(v + t);
            C wv[kDotGroup];
#pragma unroll
            for (int g = 0; g < kDotGroup; g++) wv[g] = ldg(src[g] + t);
            nrm = nrm + c64_abs2(cv);
#pragma unroll
            for (int g = 0; g < kDotGroup; g++) acc[g] = c64_fma(c64_conj(cv), ld64(&wv[g]), acc[g]);
        }
        if (kb == 0) acc[0] = c64(nrm, 0.0);
        double x[2 * kDotGroup];
#pragma unroll
        for (int g = 0; g < kDotGroup; g++) {
            x[2 * g] = acc[g].re;
            x[2 * g + 1] = acc[g].im;
        }
        const double tot = wave_sum_spread<2 * kDotGroup>(x, lane);
        constexpr int kStride = 64 / (2 * kDotGroup);
        if ((lane & (kStride - 1)) == 0) {
            const int v = lane / kStride, k = kb + (v >> 1);
            if (k < ntask) {
                double* o = (double*)(out + (k == 0 ? base : base + 2 * k));
                o[v & 1] = tot;
            }
        }
    }
}

// sdots[0] |row|^2, [1] |col|^2, [2 + 2l] row^H V_l, [3 + 2l] col^H U_l (l < r)
template <int NT, class C>
__device__ __forceinline__ void aca_dots(const C* col, const C* row, const C* U, const C* V, int64_t m, int64_t n,
                                         int r, bem_c64* sdots)
{
    dot_group_pass<NT>(col, U, m, r + 1, sdots, 1);
    dot_group_pass<NT>(row, V, n, r + 1, sdots, 0);
}

// Frobenius-norm estimate update of partialPivotACA (hmatrix.cpp:568-572);
// every thread computes it identically from the LDS dot products
template <class C>
__device__ __forceinline__ double aca_update_norm(C alpha, const C* A, int r, const bem_c64* sdots, bem_c64& S)
{
    double gain = (cx_abs2(alpha) * sdots[0].re) * sdots[1].re;
    bem_c64 s = S;
    s.re = s.re + gain;
    bem_c64 ca2 = c64_scale(c64_conj(ld64(&alpha)), 2.0);
    for (int l = 0; l < r; l++) {
        bem_c64 t = c64_mul(c64_mul(c64_mul(ca2, ld64(A + l)), sdots[3 + 2 * l]), sdots[2 + 2 * l]);
        s = c64_add(s, t);
    }
    S = s;
    return gain;
}

// Cross-approximation correction of one residual entry,
//   s = sum_{l < r} w_l * B[l * ld + idx]   (W_FIRST)   or
//   s = sum_{l < r} B[l * ld + idx] * w_l   (!W_FIRST),
// accumulated strictly in the order l = 0, 1, ..., r - 1.  The loads of B are
// issued kCrossUnroll at a time ahead of their products so that the L2 / HBM
// latency is paid once per kCrossUnroll terms rather than once per term.
constexpr int kCrossUnroll = 8;
// Offsets are 32-bit from the block's (wavefront-uniform) base: a block's
// factors hold at most kCapMax x (m + n) < 2^31 entries.
template <bool W_FIRST, class C>
__device__ __forceinline__ C cross_sum(const C* w, const C* B, int64_t ld64, int64_t idx64, int r)
{
    const uint32_t ld = (uint32_t)ld64, idx = (uint32_t)idx64;
    C s = cx_zero<C>();
    int l = 0;
    for (; l + kCrossUnroll <= r; l += kCrossUnroll) {
        C v[kCrossUnroll];
#pragma unroll
        for (int u = 0; u < kCrossUnroll; u++) v[u] = ldg(B + ((uint32_t)(l + u) * ld + idx));
#pragma unroll
        for (int u = 0; u < kCrossUnroll; u++) s = W_FIRST ? cx_fma(w[l + u], v[u], s) : cx_fma(v[u], w[l + u], s);
    }
    for (; l < r; l++) {
        C v = ldg(B + ((uint32_t)l * ld + idx));
        s = W_FIRST ? cx_fma(w[l], v, s) : cx_fma(v, w[l], s);
    }
    return s;
}

__device__ __forceinline__ bool near_or_self(const DevMesh& M, int64_t i, int64_t j)
{
    if (i == j) return true;
    double dx = M.px[i] - M.px[j], dy = M.py[i] - M.py[j], dz = M.pz[i] - M.pz[j];
    return sqrt(fma(dz, dz, fma(dy, dy, dx * dx))) < M.near_lim;
}

// regular entry of (i, j), or false for a self / near-singular pair; PLANE:
// plus the mirror image's entry (a template argument so that the plain
// kernels keep their register budget)
template <bool PLANE>
__device__ __forceinline__ bool regular_entry(const AcaArgs& a, int64_t i, int64_t j, bem_c64* e, bool nearb = true)
{
    const DevMesh& M = a.M;
    if (nearb && near_or_self(M, i, j)) return false;
    bem_ops4 o = regular_ops_vtx<1>(a.F.k, M.px[i], M.py[i], M.pz[i], M.nx[i], M.ny[i], M.nz[i], M.nx[j], M.ny[j],
                                    M.nz[j], M.area[j], load_vtx(M, j));
    *e = bc_entry(a.F, a.kind, o, i, j);
    if (PLANE) *e = c64_add(*e, refl_entry(M, a.F, a.kind, i, j));
    return true;
}
template <bool PLANE>
__device__ __forceinline__ bool regular_entry(const AcaArgs& a, int64_t i, int64_t j, bem_c32* e, bool nearb = true)
{
    const DevMesh& M = a.M;
    if (nearb && near_or_self(M, i, j)) return false;
    // only the operators the entry uses (the others stay zero, and the
    // entry formula then gives the same value)
    const int need = bc_needs_f32(a.F32, a.F.subst, a.kind, j);
    bem_c32f4 o = regular_ops_f32_edge<1>(a.F32, M, i, j, need & 1, need & 2);
    *e = bc_entry_f32(a.F32, a.F.subst, a.kind, o, j);
    if (PLANE) {
        bem_c32 r = bc_entry_f32(a.F32, a.F.subst, a.kind, refl_ops_f32(a.F32, M, i, j), j);
        *e = c32(e->re + r.re, e->im + r.im);
    }
    return true;
}

// field operator entry of observation point i and element j: the 7-point
// rule of BemOperatorField (boundaryelementsolver.cpp:2319-2353) with the
// points formed as k_field forms them; kind 0 implicitPhiMatrixField (M,
// :2647-2683), kind 1 implicitDPhiMatrixField (L, :2685-2719)
__device__ __forceinline__ bem_c64 field_entry(const AcaArgs& a, int64_t i, int64_t j)
{
    const DevMesh& M = a.M;
    const int64_t n = M.n;
    const double px = a.obs[i], py = a.obs[a.nobs + i], pz = a.obs[2 * a.nobs + i];
    const double nqx = M.nx[j], nqy = M.ny[j], nqz = M.nz[j];
    bem_c64 L = c64(0, 0), Mk = c64(0, 0);
    for (int q = 0; q < 7; q++) {
        double x1 = kTri7[q][0], x2 = kTri7[q][1];
        double x3 = (1.0 - x2) - x1;
        double qx = (x1 * M.vtx[j] + x2 * M.vtx[3 * n + j]) + x3 * M.vtx[6 * n + j];
        double qy = (x1 * M.vtx[n + j] + x2 * M.vtx[4 * n + j]) + x3 * M.vtx[7 * n + j];
        double qz = (x1 * M.vtx[2 * n + j] + x2 * M.vtx[5 * n + j]) + x3 * M.vtx[8 * n + j];
        bem_field_point_accumulate(&L, &Mk, a.F.k, px, py, pz, nqx, nqy, nqz, qx, qy, qz, kTri7[q][2]);
    }
    return c64_scale(a.kind == 0 ? Mk : L, M.area[j]);
}

// self / near-singular entry of (i, j) from the table; t = pair index or
// -1 for the self entry
template <class C>
__device__ __forceinline__ C singular_entry(const AcaArgs& a, int64_t i, int64_t t)
{
    return from64<C>(a.sing[t >= 0 ? t : a.nnz + i]);
}

// any entry (the pivot probe of the extra variant)
template <bool PLANE, class C, bool FIELD = false>
__device__ __forceinline__ C entry_any(const AcaArgs& a, int64_t i, int64_t j, bool nearb = true)
{
    if constexpr (FIELD) return from64<C>(field_entry(a, i, j));
    C e;
    if (regular_entry<PLANE>(a, i, j, &e, nearb)) return e;
    return singular_entry<C>(a, i, i == j ? -1 : near_find(a.M, i, (int32_t)j));
}

// Row i of the block minus the current cross approximation, into dst.
// Regular entries first; the few self / near-singular entries of the row
// (CSR row r0 + i) in a second pass, so that the main loop has no divergent
// branch.
template <int NT, class C, bool PLANE, bool FIELD = false>
__device__ void aca_row(const AcaArgs& a, int64_t r0, int64_t c0, int64_t m, int64_t n, int r, int64_t i,
                        const C* U, const C* V, const C* A, C* dst, C* sw, bool nearb)
{
    for (int l = threadIdx.x; l < r; l += NT) sw[l] = cx_mul(ldg(U + (int64_t)l * m + i), ldg(A + l));
    __syncthreads();
    const int64_t gi = r0 + i;
    for (int64_t j = threadIdx.x; j < n; j += NT) {
        C e;
        if constexpr (FIELD) e = from64<C>(field_entry(a, gi, c0 + j));
        else if (!regular_entry<PLANE>(a, gi, c0 + j, &e, nearb)) continue;
        dst[j] = cx_sub(e, cross_sum<true>(sw, V, n, j, r));
    }
    // no self or near-singular entries between observation points and
    // elements, nor in a block whose collocation boxes lie apart
    if (FIELD || !nearb) {
        __syncthreads();
        return;
    }
    const int64_t p0 = a.M.nr_ptr[gi], p1 = a.M.nr_ptr[gi + 1];
    for (int64_t t = p0 + threadIdx.x; t <= p1; t += NT) {
        int64_t gj = t < p1 ? a.M.nr_col[t] : gi;
        if (gj < c0 || gj >= c0 + n) continue;
        int64_t j = gj - c0;
        C e = singular_entry<C>(a, gi, t < p1 ? t : -1);
        dst[j] = cx_sub(e, cross_sum<true>(sw, V, n, j, r));
    }
    __syncthreads();
}

// column j of the block minus the current cross approximation, into dst;
// the near-singular rows of column c0 + j are the near list of row c0 + j
// (the distance criterion is symmetric)
template <int NT, class C, bool PLANE, bool FIELD = false>
__device__ void aca_col(const AcaArgs& a, int64_t r0, int64_t c0, int64_t m, int64_t n, int r, int64_t j,
                        const C* U, const C* V, const C* A, C* dst, C* sw, bool nearb)
{
    for (int l = threadIdx.x; l < r; l += NT) sw[l] = cx_mul(ldg(A + l), ldg(V + (int64_t)l * n + j));
    __syncthreads();
    const int64_t gj = c0 + j;
    for (int64_t i = threadIdx.x; i < m; i += NT) {
        C e;
        if constexpr (FIELD) e = from64<C>(field_entry(a, r0 + i, gj));
        else if (!regular_entry<PLANE>(a, r0 + i, gj, &e, nearb)) continue;
        dst[i] = cx_sub(e, cross_sum<false>(sw, U, m, i, r));
    }
    if (FIELD || !nearb) {
        __syncthreads();
        return;
    }
    const int64_t p0 = a.M.nr_ptr[gj], p1 = a.M.nr_ptr[gj + 1];
    for (int64_t t = p0 + threadIdx.x; t <= p1; t += NT) {
        int64_t gi = t < p1 ? a.M.nr_col[t] : gj;
        if (gi < r0 || gi >= r0 + m) continue;
        int64_t i = gi - r0;
        int64_t pair = t < p1 ? near_find(a.M, gi, (int32_t)gj) : -1;
        C e = singular_entry<C>(a, gi, pair);
        dst[i] = cx_sub(e, cross_sum<false>(sw, U, m, i, r));
    }
    __syncthreads();
}

// The entry counter of an ACA is the same in every lane; it is kept in scalar
// registers (readfirstlane after each update).  In one compiled form of the
// batched one-workgroup-per-item kernel its vector-register copy was spilled,
// its register reused and the count stored wrong while every factor was
// right; the scalar form does not depend on that register allocation.
__device__ __forceinline__ int64_t uni64(int64_t x)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// ACA of one admissible block b by a workgroup of NT threads.  Plain
// variant = BlockCluster::partialPivotACA (hmatrix.cpp:471-593); extra
// variant = BlockCluster::partialPivotACAextra (hmatrix.cpp:595-745).
// Dynamic LDS (aca_shm_bytes): sw[cap] and sA[cap] (type C: cross-sum
// weights / probe terms, and a copy of alpha for the loops every thread
// runs in rank order), then sdots[2 cap + 2] (FP64).
template <bool EXTRA, int NT, class C, bool PLANE, bool FIELD = false>
__device__ void aca_block(const AcaArgs& a, const int b, char* dyn, ArgMax* sred)
{
    const int64_t m = a.rows[b], n = a.cols[b], r0 = a.row0[b], c0 = a.col0[b];
    const int cap = a.cap[b];
    const bool nearb = !FIELD && (!a.bnear || a.bnear[b] != 0);
    int64_t maxRank = m < n ? m : n;
    if (a.maxrank > 0 && a.maxrank < maxRank) maxRank = a.maxrank;
    C* U = (C*)a.U + a.uoff[b];
    C* V = (C*)a.V + a.voff[b];
    C* A = (C*)a.A + a.aoff[b];
    const size_t cbytes = ((size_t)cap * sizeof(C) + 15) & ~(size_t)15;
    C* sw = (C*)dyn;
    C* sA = (C*)(dyn + cbytes);
    bem_c64* sdots = (bem_c64*)(dyn + 2 * cbytes);
    const double tiny2 = kTiny * kTiny, tol2 = a.tol * a.tol;
    int r = 0, flag = 0;
    int64_t ev = 0;
    bem_c64 S = c64(0, 0);
    if (!EXTRA) {
        int64_t j = bem_rand_bounded(a.seed, 0, (uint64_t)b, 0, n);
        aca_col<NT, C, PLANE, FIELD>(a, r0, c0, m, n, 0, j, U, V, A, U, sw, nearb);
        ev = uni64(ev + m);
        int64_t i = -1;
        while (r < maxRank) {
            if (r >= cap) { flag = 1; break; }
            const C* wcol = r >= 1 ? U + (int64_t)(r - 1) * m : U;
            ArgMax am = block_argmax<NT, C, true>(wcol, m, r >= 1 ? i : -1, sred);
            i = am.i;
            if (am.v <= tiny2 && r >= 1) break;
            C* row = V + (int64_t)r * n;
            aca_row<NT, C, PLANE, FIELD>(a, r0, c0, m, n, r, i, U, V, A, row, sw, nearb);
            ev = uni64(ev + n);
            ArgMax ar = block_argmax<NT, C, true>(row, n, -1, sred);
            if (ar.v <= tiny2 && r >= 1) break;
            C* col = U + (int64_t)r * m;
            aca_col<NT, C, PLANE, FIELD>(a, r0, c0, m, n, r, ar.i, U, V, A, col, sw, nearb);
            ev = uni64(ev + m);
            C alpha = cx_inv(row[ar.i]);
            if (threadIdx.x == 0) A[r] = sA[r] = alpha;
            aca_dots<NT>(col, row, U, V, m, n, r, sdots);
            __syncthreads();
            double gain = aca_update_norm(alpha, sA, r, sdots, S);
            r++;
            __syncthreads();
            if (!(isfinite(gain))) { flag = 2; break; }
            if (a.tol > 0 && gain < tol2 * S.re) break;
        }
    } else {
        const int64_t p0 = a.piv_off[b], p1 = a.piv_off[b + 1];
        bool done = false;
        for (int64_t p = p0; p < p1 && r < maxRank && !done; p++) {
            int64_t i = a.piv[2 * p], j = a.piv[2 * p + 1];
            C pe = entry_any<PLANE, C, FIELD>(a, r0 + i, c0 + j, nearb);
            ev = uni64(ev + 1);
            if (cx_abs2(pe) < tiny2) continue;
            // the approximation at (i, j): its rank terms formed in parallel,
            // then summed in rank order by every thread
            for (int l = threadIdx.x; l < r; l += NT)
                sw[l] = cx_mul(cx_mul(ldg(U + (int64_t)l * m + i), sA[l]), ldg(V + (int64_t)l * n + j));
            __syncthreads();
            C approx = cx_zero<C>();
            for (int l = 0; l < r; l++) approx = cx_add(approx, sw[l]);
            __syncthreads();
            if (cx_abs2(cx_sub(pe, approx)) < tol2 * cx_abs2(pe)) continue;
            if (r >= cap) { flag = 1; break; }
            aca_col<NT, C, PLANE, FIELD>(a, r0, c0, m, n, r, j, U, V, A, U + (int64_t)r * m, sw, nearb);
            ev = uni64(ev + m);
            bool first = true;
            while (r < maxRank) {
                if (r >= cap) { flag = 1; done = true; break; }
                const C* wcol = first ? U + (int64_t)r * m : U + (int64_t)(r - 1) * m;
                ArgMax am = block_argmax<NT, C, true>(wcol, m, first ? -1 : i, sred);
                i = am.i;
                first = false;
                if (cx_abs2(wcol[i]) <= tiny2 && r >= 1) break;
                C* row = V + (int64_t)r * n;
                aca_row<NT, C, PLANE, FIELD>(a, r0, c0, m, n, r, i, U, V, A, row, sw, nearb);
                ev = uni64(ev + n);
                ArgMax ar = block_argmax<NT, C, true>(row, n, -1, sred);
                if (ar.v <= tiny2 && r >= 1) break;
                C* col = U + (int64_t)r * m;
                aca_col<NT, C, PLANE, FIELD>(a, r0, c0, m, n, r, ar.i, U, V, A, col, sw, nearb);
                ev = uni64(ev + m);
                C alpha = cx_inv(row[ar.i]);
                if (threadIdx.x == 0) A[r] = sA[r] = alpha;
                aca_dots<NT>(col, row, U, V, m, n, r, sdots);
                __syncthreads();
                bem_c64 oldS = S;
                double gain = aca_update_norm(alpha, sA, r, sdots, S);
                r++;
                __syncthreads();
                if (!(isfinite(gain))) { flag = 2; done = true; break; }
                double target = tol2 * S.re;
                if (a.tol > 0 && gain < target) {
                    if (gain < 0.1 * target) { r--; S = oldS; }
                    break;
                }
            }
        }
    }
    if (threadIdx.x == 0) {
        a.rank_out[b] = r;
        a.flag_out[b] = flag;
        a.evals_out[b] = ev;
    }
}

// ACA (K3): one workgroup of NT threads per listed admissible block
// FP32 ACA limited to 128 VGPRs: four wavefronts per SIMD at every size; with
// the impedance plane (its mirror-image quadrature in every entry) three, so
// that the one-wavefront kernels spill less
template <int NT, class C> constexpr int kAcaWaves = 1;
template <> constexpr int kAcaWaves<64, bem_c32> = 4;
template <> constexpr int kAcaWaves<128, bem_c32> = 4;

// one listed item of a batched launch (nf > 0; the single operator's
// launches take aca_block directly): block b of frequency f with that
// frequency's wavenumber, coupling, table and storage
template <bool EXTRA, int NT, class C, bool PLANE>
__device__ __forceinline__ void aca_item(const AcaArgs& a, int item, char* dyn, ArgMax* sred)
{
    const int f = (int)(item / a.nbk);
    const int b = (int)(item - (int64_t)f * a.nbk);
    AcaArgs ai = a;
    const double k = a.kf[f];
    const bem_c64 cc = a.ccf[f];
    ai.F.k = k;
    ai.F.cc = cc;
    ai.F32.k = (float)k;
    ai.F32.krev = k / (2.0 * BEM_PI);
    ai.F32.cc = c32((float)cc.re, (float)cc.im);
    ai.sing = a.singf[f];
    ai.U = a.Uf[f];
    ai.V = a.Vf[f];
    ai.A = a.Af[f];
    ai.rank_out = a.rank_out + (int64_t)f * a.nbk;
    ai.flag_out = a.flag_out + (int64_t)f * a.nbk;
    ai.evals_out = a.evals_out + (int64_t)f * a.nbk;
    aca_block<EXTRA, NT, C, PLANE>(ai, b, dyn, sred);
}

// one listed (context, block) item of a multi-mesh launch
template <bool EXTRA, int NT, class C, bool PLANE>
__device__ __forceinline__ void aca_job(const AcaArgs& a, int item, char* dyn, ArgMax* sred)
{
    const MeshJob& J = a.jobs[item >> kJobShift];
    AcaArgs ai = a;
    ai.M = J.M;
    ai.F = J.F;
    ai.F32 = J.F32;
    ai.row0 = J.row0; ai.rows = J.rows; ai.col0 = J.col0; ai.cols = J.cols;
    ai.uoff = J.uoff; ai.voff = J.voff; ai.aoff = J.aoff; ai.cap = J.cap;
    ai.U = J.U; ai.V = J.V; ai.A = J.A;
    ai.rank_out = J.rank_out; ai.flag_out = J.flag_out; ai.evals_out = J.evals_out;
    ai.piv_off = J.piv_off; ai.piv = J.piv;
    ai.sing = J.sing; ai.nnz = J.nnz;
    ai.bnear = J.bnear;
    ai.nf = 0;
    aca_block<EXTRA, NT, C, PLANE>(ai, item & kJobMask, dyn, sred);
}

// MODE 0: one workgroup per listed block of the single operator; 1: one per
// listed (frequency, block) item; 3: one per listed (context, block) item;
// 2 / 4: persistent workgroups taking the items of 1 / 3 from the queue
// counter (separate instantiations, so that the plain kernel keeps its
// register budget)
template <bool EXTRA, int NT, class C, bool PLANE, int MODE, bool FIELD = false>
__global__ __launch_bounds__(NT, (PLANE && kAcaWaves<NT, C> > 3 ? 3 : kAcaWaves<NT, C>)) void k_aca(AcaArgs a)
{
    extern __shared__ __attribute__((aligned(16))) char dyn[];
    __shared__ ArgMax sred[NT / 64];
    if constexpr (MODE == 0) {
        aca_block<EXTRA, NT, C, PLANE, FIELD>(a, a.blist[blockIdx.x], dyn, sred);
        return;
    }
    if constexpr (MODE == 1) {
        aca_item<EXTRA, NT, C, PLANE>(a, a.blist[blockIdx.x], dyn, sred);
        return;
    }
    if constexpr (MODE == 3) {
        aca_job<EXTRA, NT, C, PLANE>(a, a.blist[blockIdx.x], dyn, sred);
        return;
    }
    __shared__ int s_pos;
    while (true) {   // every workgroup leaves once the counter passes nitems
        if (threadIdx.x == 0) s_pos = atomicAdd(a.queue, 1);
        __syncthreads();
        const int pos = s_pos;
        __syncthreads();
        if (pos >= a.nitems) break;
        if constexpr (MODE == 4) aca_job<EXTRA, NT, C, PLANE>(a, a.blist[pos], dyn, sred);
        else aca_item<EXTRA, NT, C, PLANE>(a, a.blist[pos], dyn, sred);
        __syncthreads();
    }
}

// copy the first rank columns/rows of the ACA scratch into the packed
// storage (rank 0 becomes a zero rank-1 block, hmatrix.cpp:587-592)
template <class C>
__global__ __launch_bounds__(kWG) void k_compact(const int32_t* blist, const int64_t* rows, const int64_t* cols,
                                                 const int64_t* suoff, const int64_t* svoff, const int64_t* saoff,
                                                 const C* sU, const C* sV, const C* sA, const int32_t* srank,
                                                 const int64_t* duoff, const int64_t* dvoff, const int64_t* daoff,
                                                 C* dU, C* dV, C* dA)
{
    const int b = blist[blockIdx.x];
    const int64_t m = rows[b], n = cols[b];
    const int r = srank[b];
    const int64_t nu = (int64_t)(r > 0 ? r : 1) * m, nv = (int64_t)(r > 0 ? r : 1) * n;
    const C z = cx_zero<C>();
    for (int64_t t = threadIdx.x; t < nu; t += kWG) dU[duoff[b] + t] = r > 0 ? sU[suoff[b] + t] : z;
    for (int64_t t = threadIdx.x; t < nv; t += kWG) dV[dvoff[b] + t] = r > 0 ? sV[svoff[b] + t] : z;
    for (int t = threadIdx.x; t < (r > 0 ? r : 1); t += kWG) dA[daoff[b] + t] = r > 0 ? sA[saoff[b] + t] : z;
}

// ------------------------------------------------------------- H-MVM
// Block storage is complex FP64 (bem_c64) or FP32 (bem_c32); products
// accumulate in FP64 and vectors are FP64 in both cases.

// near-field tiles (K4): partial[tile][row] = sum_c F[row, c] x[c]; the
// body takes x and the partial vectors with element stride XS (one column
// of XS interleaved ones)
template <class S, int XS>
__device__ __forceinline__ void mvm_dense_tile(const Tile& T, const S* __restrict__ F, const bem_c64* __restrict__ x,
                                               bem_c64* __restrict__ part)
{
    __shared__ bem_c64 red[4][kTile];
    const int tid = threadIdx.x, r = tid & (kTile - 1), cg = tid >> 6;
    bem_c64 acc = c64(0, 0);
    if (r < T.rows) {
        const S* f = F + T.dest + r;
        for (int c = cg; c < T.cols; c += 4)
            acc = c64_fma(ld64(f + (int64_t)c * T.ld), x[(int64_t)(T.col0 + c) * XS], acc);
    }
    red[cg][r] = acc;
    __syncthreads();
    if (cg == 0 && r < T.rows)
        part[(T.part + r) * XS] = c64_add(c64_add(c64_add(red[0][r], red[1][r]), red[2][r]), red[3][r]);
}

template <class S>
__global__ __launch_bounds__(kWG) void k_mvm_dense(const Tile* __restrict__ tiles, const S* __restrict__ F,
                                                   const bem_c64* __restrict__ x, bem_c64* __restrict__ part)
{
    mvm_dense_tile<S, 1>(tiles[blockIdx.x], F, x, part);
}

// low-rank blocks, phase 1 (K5a): t_b[l] = alpha_l (V_b[l, :] . x_J), one
// wavefront per (block, l0) task covering rows l0 .. l0 + kLrRows - 1 of V_b.
// FP64 storage: every row over all 64 lanes (lane-strided sums and the
// wavefront tree, the order the restatement follows), each x element loaded
// once for the group.  FP32 storage: 16 lanes per row with four loads in
// flight per lane and a 16-lane tree (a sixth of the shuffles).
constexpr int kLrRows = 4;
static_assert(kLrRows * 16 == 64, "FP32 k_lr_t: 16 lanes per row");
// (body with x of element stride XS)
template <class S, int XS>
__device__ __forceinline__ void lr_t_task(int64_t task, const int32_t* __restrict__ tasks,
                                          const int64_t* __restrict__ col0, const int64_t* __restrict__ cols,
                                          const int32_t* __restrict__ rank, const int64_t* __restrict__ voff,
                                          const int64_t* __restrict__ aoff, const int64_t* __restrict__ toff,
                                          const S* __restrict__ V, const S* __restrict__ A,
                                          const bem_c64* __restrict__ x, bem_c64* __restrict__ t)
{
    const int lane = threadIdx.x & 63;
    const int b = tasks[2 * task], l0 = tasks[2 * task + 1];
    const int64_t n = cols[b];
    const int nl = min(kLrRows, rank[b] - l0);
    const S* v = V + voff[b] + (int64_t)l0 * n;
    const bem_c64* xs = x + col0[b] * XS;
    if constexpr (sizeof(S) == sizeof(bem_c32)) {
        const int u = lane >> 4, q = lane & 15;   // row l0 + u, columns q, q + 16, ...
        bem_c64 p = c64(0, 0);
        if (u < nl) {
            // four independent accumulators (columns q + 16 e + 64 i in pe[e]),
            // so that consecutive multiply-adds do not wait on each other
            const S* vr = v + (int64_t)u * n;
            bem_c64 pe[4];
#pragma unroll
            for (int e = 0; e < 4; e++) pe[e] = c64(0, 0);
            int64_t j = q;
            for (; j + 48 < n; j += 64) {
                S vv[4];
                bem_c64 xv[4];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    vv[e] = vr[j + 16 * e];
                    xv[e] = xs[(j + 16 * e) * XS];
                }
#pragma unroll
                for (int e = 0; e < 4; e++) pe[e] = c64_fma(ld64(&vv[e]), xv[e], pe[e]);
            }
            // at most three columns are left (j + 48 >= n)
            if (j < n) pe[0] = c64_fma(ld64(vr + j), xs[j * XS], pe[0]);
            if (j + 16 < n) pe[1] = c64_fma(ld64(vr + j + 16), xs[(j + 16) * XS], pe[1]);
            if (j + 32 < n) pe[2] = c64_fma(ld64(vr + j + 32), xs[(j + 32) * XS], pe[2]);
            p = c64_add(c64_add(pe[0], pe[1]), c64_add(pe[2], pe[3]));
        }
        for (int s = 8; s > 0; s >>= 1) {
            p.re = p.re + __shfl_xor(p.re, s, 16);
            p.im = p.im + __shfl_xor(p.im, s, 16);
        }
        if (q == 0 && u < nl) t[toff[b] + l0 + u] = c64_mul(ld64(A + aoff[b] + l0 + u), p);
        return;
    }
    bem_c64 p[kLrRows];
#pragma unroll
    for (int u = 0; u < kLrRows; u++) p[u] = c64(0, 0);
    for (int64_t j = lane; j < n; j += 64) {
        const bem_c64 xv = xs[j * XS];
#pragma unroll
        for (int u = 0; u < kLrRows; u++)
            if (u < nl) p[u] = c64_fma(ld64(v + (int64_t)u * n + j), xv, p[u]);
    }
#pragma unroll
    for (int u = 0; u < kLrRows; u++) p[u] = wave_sum(p[u]);
    if (lane == 0) {   // wave_sum leaves the full sums in lane 0
#pragma unroll
        for (int u = 0; u < kLrRows; u++)
            if (u < nl) t[toff[b] + l0 + u] = c64_mul(ld64(A + aoff[b] + l0 + u), p[u]);
    }
}

template <class S>
__global__ __launch_bounds__(kWG) void k_lr_t(const int32_t* __restrict__ tasks, int64_t ntask,
                                              const int64_t* __restrict__ col0, const int64_t* __restrict__ cols,
                                              const int32_t* __restrict__ rank, const int64_t* __restrict__ voff,
                                              const int64_t* __restrict__ aoff, const int64_t* __restrict__ toff,
                                              const S* __restrict__ V, const S* __restrict__ A,
                                              const bem_c64* __restrict__ x, bem_c64* __restrict__ t)
{
    const int64_t task = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
    if (task >= ntask) return;
    lr_t_task<S, 1>(task, tasks, col0, cols, rank, voff, aoff, toff, V, A, x, t);
}

// low-rank blocks, phase 2 (K5b): partial_b[i] = sum_l U_b[i, l] t_b[l] over a
// chunk of kWG rows of one block
// (body: task `task`, t staged in st, partial vectors of element stride XS)
template <class S, int XS>
__device__ __forceinline__ void lr_u_task(int64_t task, const int32_t* __restrict__ tasks,
                                          const int64_t* __restrict__ rows, const int32_t* __restrict__ rank,
                                          const int64_t* __restrict__ uoff, const int64_t* __restrict__ toff,
                                          const int64_t* __restrict__ poff, const S* __restrict__ U,
                                          const bem_c64* __restrict__ t, bem_c64* st, bem_c64* __restrict__ part)
{
    const int b = tasks[2 * task], chunk = tasks[2 * task + 1];
    const int64_t m = rows[b];
    const int k = rank[b];
    for (int l = threadIdx.x; l < k; l += kWG) st[l] = t[toff[b] + l];
    __syncthreads();
    const int64_t i = (int64_t)chunk * kWG + threadIdx.x;
    if (i >= m) return;
    const S* u = U + uoff[b] + i;
    bem_c64 acc = c64(0, 0);
    for (int l = 0; l < k; l++) acc = c64_fma(ld64(u + (int64_t)l * m), st[l], acc);
    part[(poff[b] + i) * XS] = acc;
}

template <class S>
__global__ __launch_bounds__(kWG) void k_lr_u(const int32_t* __restrict__ tasks, const int64_t* __restrict__ rows,
                                              const int32_t* __restrict__ rank, const int64_t* __restrict__ uoff,
                                              const int64_t* __restrict__ toff, const int64_t* __restrict__ poff,
                                              const S* __restrict__ U, const bem_c64* __restrict__ t,
                                              bem_c64* __restrict__ part)
{
    extern __shared__ bem_c64 st[];
    lr_u_task<S, 1>(blockIdx.x, tasks, rows, rank, uoff, toff, poff, U, t, st, part);
}

// y_i (+)= sum of the partial vectors covering row i in partition order:
// 16 lanes per row take every 16th partial, then a fixed shuffle tree
__global__ void k_mvm_gather(int64_t n, const int64_t* __restrict__ gptr, const int64_t* __restrict__ gidx,
                             const bem_c64* __restrict__ part, bem_c64* y, int accumulate)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = g >> 4;
    const int q = (int)(g & 15);
    bem_c64 acc = c64(0, 0);
    if (i < n)
        for (int64_t p = gptr[i] + q; p < gptr[i + 1]; p += 16) acc = c64_add(acc, part[gidx[p]]);
    for (int s = 8; s > 0; s >>= 1) {
        acc.re = acc.re + __shfl_down(acc.re, s, 16);
        acc.im = acc.im + __shfl_down(acc.im, s, 16);
    }
    if (i < n && q == 0) y[i] = accumulate ? c64_add(y[i], acc) : acc;
}

// ------------------------------------------------------------- GMRES

constexpr int64_t kDotChunk = 128;   // rows per wavefront of the Arnoldi dot products
constexpr int kDotVecs = 8;          // basis vectors per wavefront (w loaded once for them)

// part[l * nchunk + chunk] = sum over the chunk's rows of conj(V_l) w; one
// wavefront per (chunk, group of kDotVecs basis vectors)
__global__ __launch_bounds__(kWG) void k_cgs_dots(int64_t n, int cnt, const bem_c64* __restrict__ V,
                                                  const bem_c64* __restrict__ w, bem_c64* __restrict__ part)
{
    const int l0 = (blockIdx.y * (kWG / 64) + (threadIdx.x >> 6)) * kDotVecs;
    if (l0 >= cnt) return;
    const int lane = threadIdx.x & 63;
    const int nv = min(kDotVecs, cnt - l0);
    const int64_t i0 = (int64_t)blockIdx.x * kDotChunk, i1 = i0 + kDotChunk < n ? i0 + kDotChunk : n;
    bem_c64 p[kDotVecs];
#pragma unroll
    for (int u = 0; u < kDotVecs; u++) p[u] = c64(0, 0);
    for (int64_t i = i0 + lane; i < i1; i += 64) {
        const bem_c64 wi = w[i];
#pragma unroll
        for (int u = 0; u < kDotVecs; u++)
            if (u < nv) p[u] = c64_fma(c64_conj(V[(int64_t)(l0 + u) * n + i]), wi, p[u]);
    }
#pragma unroll
    for (int u = 0; u < kDotVecs; u++) p[u] = wave_sum(p[u]);
    if (lane == 0) {
#pragma unroll
        for (int u = 0; u < kDotVecs; u++)
            if (u < nv) part[(int64_t)(l0 + u) * gridDim.x + blockIdx.x] = p[u];
    }
}

// h[l] = sum over chunks (lane-strided, then the wavefront tree: a fixed
// order); hacc[l] = h[l] (first pass) or += h[l].  One wavefront per l.
__global__ __launch_bounds__(64) void k_cgs_reduce(int cnt, int nchunk, const bem_c64* __restrict__ part, bem_c64* h,
                                                   bem_c64* hacc, int first)
{
    const int l = blockIdx.x;
    if (l >= cnt) return;
    bem_c64 s = c64(0, 0);
    for (int q = threadIdx.x; q < nchunk; q += 64) s = c64_add(s, part[(int64_t)l * nchunk + q]);
    s = wave_sum(s);
    if (threadIdx.x == 0) {
        h[l] = s;
        hacc[l] = first ? s : c64_add(hacc[l], s);
    }
}

// One Arnoldi column on the device: h[0..j] = hacc, h[j+1] = |w|; apply the
// previous Givens rotations, make the new one, update g and store the
// column of the triangular factor; resid = |g[j+1]| / |b|.
__global__ void k_arnoldi(int j, int m, const bem_c64* hacc, const double* hn, bem_c64* H, double* cs_c,
                          bem_c64* cs_s, bem_c64* g, double bn, double* resid)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    bem_c64* h = H + (int64_t)j * (m + 1);
    for (int l = 0; l <= j; l++) h[l] = hacc[l];
    h[j + 1] = c64(*hn, 0.0);
    for (int l = 0; l < j; l++) {
        bem_c64 a0 = h[l], b0 = h[l + 1];
        h[l] = c64_add(c64_scale(a0, cs_c[l]), c64_mul(cs_s[l], b0));
        h[l + 1] = c64_sub(c64_scale(b0, cs_c[l]), c64_mul(c64_conj(cs_s[l]), a0));
    }
    double na = sqrt(c64_abs2(h[j])), nb2 = sqrt(c64_abs2(h[j + 1]));
    double t = sqrt(na * na + nb2 * nb2);
    double cc;
    bem_c64 ss;
    if (na == 0) { cc = 0; ss = c64(1, 0); }
    else {
        cc = na / t;
        ss = c64_scale(c64_mul(c64_scale(h[j], 1.0 / na), c64_conj(h[j + 1])), 1.0 / t);
    }
    cs_c[j] = cc;
    cs_s[j] = ss;
    h[j] = c64_add(c64_scale(h[j], cc), c64_mul(ss, h[j + 1]));
    h[j + 1] = c64(0, 0);
    bem_c64 gj = g[j];
    g[j] = c64_scale(gj, cc);
    g[j + 1] = c64_neg(c64_mul(c64_conj(ss), gj));
    *resid = sqrt(c64_abs2(g[j + 1])) / bn;
}

// w -= sum_l V_l h_l
__global__ void k_orth_update(int64_t n, int cnt, const bem_c64* V, const bem_c64* h, bem_c64* w)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bem_c64 acc = w[i];
    int l = 0;
    for (; l + 8 <= cnt; l += 8) {   // loads first, then the products in order
        bem_c64 v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = V[(int64_t)(l + u) * n + i];
#pragma unroll
        for (int u = 0; u < 8; u++) acc = c64_fms(h[l + u], v[u], acc);
    }
    for (; l < cnt; l++) acc = c64_fms(h[l], V[(int64_t)l * n + i], acc);
    w[i] = acc;
}

// partial sums of |w|^2 per workgroup
__global__ __launch_bounds__(kWG) void k_norm2_part(int64_t n, const bem_c64* w, double* part)
{
    __shared__ double red[kWG / 64];
    double p = 0;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG)
        p = p + c64_abs2(w[i]);
    p = wave_sum(p);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = p;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void k_norm2_final(int nparts, const double* part, double* out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0;
        for (int q = 0; q < nparts; q++) s += part[q];
        *out = sqrt(s);
    }
}

// dst = src / *nrm (0 if the norm is 0)
__global__ void k_scale_by_norm(int64_t n, const bem_c64* src, const double* nrm, bem_c64* dst)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = *nrm;
    dst[i] = d > 0 ? c64_scale(src[i], 1.0 / d) : c64(0, 0);
}

// x += sum_l V_l y_l
__global__ void k_combine(int64_t n, int cnt, const bem_c64* V, const bem_c64* y, bem_c64* x, int accumulate)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bem_c64 acc = accumulate ? x[i] : c64(0, 0);
    for (int l = 0; l < cnt; l++) acc = c64_fma(V[(int64_t)l * n + i], y[l], acc);
    x[i] = acc;
}

// ------------------------------------------- leaf-block preconditioner
// M = block diagonal of the operator's diagonal leaf blocks (the leaf level
// of the reference's H-LU, HArithm::LUDecomposition harithm.cpp:1192-1296),
// each factored by LU with partial pivoting in FP64.

constexpr int kPrecMax = 64;   // largest diagonal block

struct PrecBlock { int64_t row0, lu; int32_t m, pad; };

// one workgroup per diagonal block: copy to LDS (column-major), factor,
// write L\U and the pivot rows back; flag[0] = 1 on a zero pivot
template <class S>
__global__ __launch_bounds__(kWG) void k_blu_factor(const PrecBlock* __restrict__ blk, const int64_t* __restrict__ src,
                                                    const S* __restrict__ dense, bem_c64* __restrict__ lu,
                                                    int32_t* __restrict__ piv, int32_t* flag)
{
    __shared__ bem_c64 a[kPrecMax * kPrecMax];
    __shared__ ArgMax sred[kWG / 64];
    __shared__ int prow;
    const PrecBlock B = blk[blockIdx.x];
    const int m = B.m, tid = threadIdx.x;
    const S* d = dense + src[blockIdx.x];
    for (int t = tid; t < m * m; t += kWG) a[t] = ld64(d + t);
    __syncthreads();
    for (int k = 0; k < m; k++) {
        ArgMax am = block_argmax<kWG>(a + k * m + k, m - k, -1, sred);
        if (tid == 0) {
            prow = k + (int)am.i;
            piv[B.row0 + k] = prow;
            if (!(am.v > 0)) flag[0] = 1;
        }
        __syncthreads();
        const int p = prow;
        if (p != k)
            for (int j = tid; j < m; j += kWG) {
                bem_c64 t = a[j * m + k];
                a[j * m + k] = a[j * m + p];
                a[j * m + p] = t;
            }
        __syncthreads();
        const bem_c64 inv = c64_div(c64(1.0, 0.0), a[k * m + k]);
        for (int i = k + 1 + tid; i < m; i += kWG) a[k * m + i] = c64_mul(a[k * m + i], inv);
        __syncthreads();
        const int w = m - k - 1;
        for (int t = tid; t < w * w; t += kWG) {
            const int i = k + 1 + t % w, j = k + 1 + t / w;
            a[j * m + i] = c64_fms(a[k * m + i], a[j * m + k], a[j * m + i]);
        }
        __syncthreads();
    }
    for (int t = tid; t < m * m; t += kWG) lu[B.lu + t] = a[t];
}

// z = M^-1 v on every diagonal block: one wavefront per block, z in LDS
__global__ __launch_bounds__(64) void k_blu_apply(const PrecBlock* __restrict__ blk, const bem_c64* __restrict__ lu,
                                                  const int32_t* __restrict__ piv, const bem_c64* __restrict__ v,
                                                  bem_c64* __restrict__ z)
{
    __shared__ bem_c64 y[kPrecMax];
    const PrecBlock B = blk[blockIdx.x];
    const int m = B.m, i = threadIdx.x;
    const bem_c64* a = lu + B.lu;
    if (i < m) y[i] = v[B.row0 + i];
    __syncthreads();
    if (i == 0)
        for (int k = 0; k < m; k++) {
            const int p = piv[B.row0 + k];
            if (p != k) { bem_c64 t = y[k]; y[k] = y[p]; y[p] = t; }
        }
    __syncthreads();
    for (int k = 0; k < m; k++) {          // unit lower triangle
        const bem_c64 yk = y[k];
        if (i > k && i < m) y[i] = c64_fms(a[k * m + i], yk, y[i]);
        __syncthreads();
    }
    for (int k = m - 1; k >= 0; k--) {     // upper triangle
        if (i == k) y[k] = c64_div(y[k], a[k * m + k]);
        __syncthreads();
        const bem_c64 yk = y[k];
        if (i < k) y[i] = c64_fms(a[k * m + i], yk, y[i]);
        __syncthreads();
    }
    if (i < m) z[B.row0 + i] = y[i];
}

// H-LU substitution on the device (the flattened factors of bemhlu::Flat;
// forwardSubstitution / backwardSubstitution, harithm.cpp:1314-1412): one
// workgroup of kHluThreads walks the diagonal leaves of one direction in
// order (dir 0: L y = z, increasing; dir 1: U x = y, decreasing), z in place.
// Per leaf, with the item headers staged in LDS (kHluHdr per round): the
// low-rank blocks that start there form t_B = W_B^T z_J from tasks of
// kHluTChunk columns (every wavefront takes tasks; the chunk sums are added
// in chunk order); the leaf's row slices of the blocks left of it (L) or
// right of it (U) are applied to the solved part of z in tasks of kHluCols
// columns, task tau going to wavefront tau mod 16 with all its loads issued
// together, lane i holding rows i and i + 64; the partial sums are
// subtracted from z_k in wavefront order; z_k := M z_k with M = Q^H or R^-1
// of the leaf (wavefront w taking columns w, w + 16, ...).  Every offset was
// checked against the buffers on the host.
constexpr int kHluThreads = 1024;
constexpr int kHluWaves = kHluThreads / 64;
constexpr int kHluMaxLeaf = 128;
constexpr int kHluCols = 6;
struct HluLeafD { int64_t s, m, qh, rinv; };
struct HluItemD { int64_t a, ld, c0, t; int32_t n, k; };
static_assert(sizeof(HluLeafD) == sizeof(bemhlu::FlatLeaf), "leaf layout");
static_assert(sizeof(HluItemD) == sizeof(bemhlu::FlatItem), "item layout");

// a0 (+ a1) += columns j0 .. j0 + nj - 1 (nj <= NC) of a row slice times x,
// all loads of the group issued together; TWO: rows lane + 64 as well
template <int NC, bool TWO>
__device__ __forceinline__ void hlu_cols(const bem_c64* col, int64_t ld, const bem_c64* x, int j0, int nj, int lane,
                                         int m, bem_c64& a0, bem_c64& a1)
{
    bem_c64 xv[NC], f0[NC], f1[NC];
#pragma unroll
    for (int u = 0; u < NC; u++) {
        const bool in = u < nj;
        xv[u] = in ? x[j0 + u] : c64(0, 0);
        f0[u] = in && lane < m ? col[lane + (int64_t)(j0 + u) * ld] : c64(0, 0);
        if (TWO) f1[u] = in && lane + 64 < m ? col[lane + 64 + (int64_t)(j0 + u) * ld] : c64(0, 0);
    }
#pragma unroll
    for (int u = 0; u < NC; u++) {
        a0 = c64_fma(f0[u], xv[u], a0);
        if (TWO) a1 = c64_fma(f1[u], xv[u], a1);
    }
}

// item headers staged in LDS per round, and the columns of one t task
constexpr int kHluHdr = 256;
constexpr int kHluTChunk = 512;

// tb[0 .. nq]: first task of each staged item (nt(item) tasks each)
template <class NT>
__device__ __forceinline__ void hlu_task_bases(const HluItemD* hdr, int nq, int32_t* tb, NT nt)
{
    if (threadIdx.x == 0) {
        int32_t s = 0;
        for (int i = 0; i < nq; i++) {
            tb[i] = s;
            s += nt(hdr[i]);
        }
        tb[nq] = s;
    }
    __syncthreads();
}

// the staged item holding task: tb[i] <= task < tb[i + 1]
__device__ __forceinline__ int hlu_item_of(const int32_t* tb, int nq, int task)
{
    int lo = 0, hi = nq;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (tb[mid] <= task) lo = mid;
        else hi = mid;
    }
    return lo;
}

// t vectors of the items [qb, qe): (item, rank index, chunk of kHluTChunk
// columns) tasks over all wavefronts into gpart, then each t_B[l] sums its
// chunks in chunk order
__device__ void hlu_t_phase(const HluItemD* __restrict__ titem, int64_t qb, int64_t qe,
                            const bem_c64* __restrict__ data, const bem_c64* z, bem_c64* tbuf, bem_c64* gpart,
                            HluItemD* hdr, int32_t* tb)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    auto t_tasks = [](const HluItemD& T) { return T.k * ((T.n + kHluTChunk - 1) / kHluTChunk); };
    for (int64_t q0 = qb; q0 < qe; q0 += kHluHdr) {
        const int nq = (int)min((int64_t)kHluHdr, qe - q0);
        if (threadIdx.x < nq) hdr[threadIdx.x] = titem[q0 + threadIdx.x];
        __syncthreads();
        hlu_task_bases(hdr, nq, tb, t_tasks);
        const int ntask = tb[nq];
        for (int task = w; task < ntask; task += kHluWaves) {
            const int i = hlu_item_of(tb, nq, task);
            const HluItemD T = hdr[i];
            const int nch = (T.n + kHluTChunk - 1) / kHluTChunk;
            const int rel = task - tb[i], l = rel / nch, j0 = (rel - l * nch) * kHluTChunk;
            const int len = min(kHluTChunk, T.n - j0);
            const bem_c64* col = data + T.a + (int64_t)l * T.ld + j0;
            const bem_c64* zc = z + T.c0 + j0;
            bem_c64 wv[kHluTChunk / 64], xv[kHluTChunk / 64];
#pragma unroll
            for (int u = 0; u < kHluTChunk / 64; u++) {
                const int j = lane + 64 * u;
                wv[u] = j < len ? col[j] : c64(0, 0);
                xv[u] = j < len ? zc[j] : c64(0, 0);
            }
            bem_c64 v = c64(0, 0);
#pragma unroll
            for (int u = 0; u < kHluTChunk / 64; u++) v = c64_fma(wv[u], xv[u], v);
            v = wave_sum(v);
            if (lane == 0) gpart[task] = v;
        }
        __syncthreads();
        if (threadIdx.x < nq) {
            const HluItemD T = hdr[threadIdx.x];
            const int nch = (T.n + kHluTChunk - 1) / kHluTChunk;
            for (int l = 0; l < T.k; l++) {
                const bem_c64* g = gpart + tb[threadIdx.x] + l * nch;
                bem_c64 v = g[0];
                for (int ch = 1; ch < nch; ch++) v = c64_add(v, g[ch]);
                tbuf[T.t + l] = v;
            }
        }
        __syncthreads();
    }
}

// a0 (rows lane) and a1 (rows lane + 64) of this wavefront += the
// contributions of the items [qb, qe) to a leaf of m rows: (item, kHluCols
// columns) tasks, task tau on wavefront tau mod 16
__device__ void hlu_c_phase(const HluItemD* __restrict__ citem, int64_t qb, int64_t qe, int m,
                            const bem_c64* __restrict__ data, const bem_c64* z, const bem_c64* tbuf, HluItemD* hdr,
                            int32_t* tb, bem_c64& a0, bem_c64& a1)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    auto c_tasks = [](const HluItemD& C) { return ((C.k < 0 ? C.n : C.k) + kHluCols - 1) / kHluCols; };
    for (int64_t q0 = qb; q0 < qe; q0 += kHluHdr) {
        const int nq = (int)min((int64_t)kHluHdr, qe - q0);
        if (threadIdx.x < nq) hdr[threadIdx.x] = citem[q0 + threadIdx.x];
        __syncthreads();
        hlu_task_bases(hdr, nq, tb, c_tasks);
        const int ntask = tb[nq];
        for (int task = w; task < ntask; task += kHluWaves) {
            const int i = hlu_item_of(tb, nq, task);
            const HluItemD C = hdr[i];
            const int nc = C.k < 0 ? C.n : C.k;
            const int j0 = (task - tb[i]) * kHluCols, nj = min(kHluCols, nc - j0);
            const bem_c64* x = C.k < 0 ? z + C.c0 : tbuf + C.t;
            const bem_c64* col = data + C.a;
            if (m <= 64)
                hlu_cols<kHluCols, false>(col, C.ld, x, j0, nj, lane, m, a0, a1);
            else
                for (int h = 0; h < kHluCols; h += kHluCols / 2)
                    hlu_cols<kHluCols / 2, true>(col, C.ld, x, j0 + h, min(kHluCols / 2, nj - h), lane, m, a0, a1);
        }
        __syncthreads();
    }
}

// z_k := z_k - (sum of the wavefronts' a0 / a1); MUL: then z_k := M z_k
template <bool MUL>
__device__ void hlu_finish(const HluLeafD& L, const bem_c64* __restrict__ M, bem_c64 a0, bem_c64 a1, bem_c64* z,
                           bem_c64 (*part)[kHluMaxLeaf], bem_c64* acc)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, m = (int)L.m;
    part[w][lane] = a0;
    part[w][lane + 64] = a1;
    __syncthreads();
    if (threadIdx.x < m) {
        bem_c64 v = z[L.s + threadIdx.x];
        for (int g = 0; g < kHluWaves; g++) v = c64_sub(v, part[g][threadIdx.x]);
        if (MUL) acc[threadIdx.x] = v;
        else z[L.s + threadIdx.x] = v;
    }
    __syncthreads();
    if (!MUL) return;
    bem_c64 y0 = c64(0, 0), y1 = c64(0, 0);
    for (int j = w; j < m; j += kHluWaves) {
        const bem_c64 av = acc[j];
        if (lane < m) y0 = c64_fma(M[lane + j * m], av, y0);
        if (lane + 64 < m) y1 = c64_fma(M[lane + 64 + j * m], av, y1);
    }
    part[w][lane] = y0;
    part[w][lane + 64] = y1;
    __syncthreads();
    if (threadIdx.x < m) {
        bem_c64 v = part[0][threadIdx.x];
        for (int g = 1; g < kHluWaves; g++) v = c64_add(v, part[g][threadIdx.x]);
        z[L.s + threadIdx.x] = v;
    }
    __syncthreads();
}

// MODE 0: one workgroup walks the leaves kb .. ke - 1 of the panel in the
// direction's order (t vectors, contributions, z_k := M (z_k - sum)) with
// the panel's inner item lists; MODE 1: workgroup b forms the t vectors of
// the outer items of leaf kb + b (gpart + b gstride); MODE 2: workgroup b
// subtracts the outer contributions of leaf kb + b.  Outer items only read
// z outside the panel, solved before the panel starts.
template <int MODE>
__global__ __launch_bounds__(kHluThreads) void k_hlu(int dir, int64_t kb, int64_t ke,
                                                     const HluLeafD* __restrict__ leaf,
                                                     const int64_t* __restrict__ tptr,
                                                     const HluItemD* __restrict__ titem,
                                                     const int64_t* __restrict__ cptr,
                                                     const HluItemD* __restrict__ citem,
                                                     const bem_c64* __restrict__ data, bem_c64* tbuf, bem_c64* gpart,
                                                     int64_t gstride, bem_c64* z)
{
    __shared__ bem_c64 part[kHluWaves][kHluMaxLeaf];
    __shared__ bem_c64 acc[kHluMaxLeaf];
    __shared__ HluItemD hdr[kHluHdr];
    __shared__ int32_t tb[kHluHdr + 1];
    if (MODE == 1) {
        const int64_t k = kb + blockIdx.x;
        hlu_t_phase(titem, tptr[k], tptr[k + 1], data, z, tbuf, gpart + (int64_t)blockIdx.x * gstride, hdr, tb);
        return;
    }
    if (MODE == 2) {
        const int64_t k = kb + blockIdx.x;
        const HluLeafD L = leaf[k];
        bem_c64 a0 = c64(0, 0), a1 = c64(0, 0);
        hlu_c_phase(citem, cptr[k], cptr[k + 1], (int)L.m, data, z, tbuf, hdr, tb, a0, a1);
        hlu_finish<false>(L, nullptr, a0, a1, z, part, acc);
        return;
    }
    // the next leaf's header and item ranges are loaded one step ahead
    int64_t k = dir == 0 ? kb : ke - 1;
    HluLeafD L = leaf[k];
    int64_t t0 = tptr[k], t1 = tptr[k + 1], c0 = cptr[k], c1 = cptr[k + 1];
    for (int64_t step = kb; step < ke; step++) {
        const int64_t kn = step + 1 < ke ? (dir == 0 ? step + 1 : kb + ke - 2 - step) : k;
        const HluLeafD Ln = leaf[kn];
        const int64_t t0n = tptr[kn], t1n = tptr[kn + 1], c0n = cptr[kn], c1n = cptr[kn + 1];
        const int m = (int)L.m;
        const bem_c64* M = data + (dir == 0 ? L.qh : L.rinv);
        if (m <= 64) {   // Q^H / R^-1 columns w, w + 16, .. and z_k ahead of the contributions
            const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
            bem_c64 mv[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int j = w + kHluWaves * u;
                mv[u] = lane < m && j < m ? M[lane + j * m] : c64(0, 0);
            }
            const bem_c64 zk = threadIdx.x < m ? z[L.s + threadIdx.x] : c64(0, 0);
            hlu_t_phase(titem, t0, t1, data, z, tbuf, gpart, hdr, tb);
            bem_c64 a0 = c64(0, 0), a1 = c64(0, 0);
            hlu_c_phase(citem, c0, c1, m, data, z, tbuf, hdr, tb, a0, a1);
            part[w][lane] = a0;
            __syncthreads();
            if (threadIdx.x < m) {
                bem_c64 v = zk;
                for (int g = 0; g < kHluWaves; g++) v = c64_sub(v, part[g][threadIdx.x]);
                acc[threadIdx.x] = v;
            }
            __syncthreads();
            bem_c64 y0 = c64(0, 0);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int j = w + kHluWaves * u;
                if (j < m) y0 = c64_fma(mv[u], acc[j], y0);
            }
            part[w][lane] = y0;
            __syncthreads();
            if (threadIdx.x < m) {
                bem_c64 v = part[0][threadIdx.x];
                for (int g = 1; g < kHluWaves; g++) v = c64_add(v, part[g][threadIdx.x]);
                z[L.s + threadIdx.x] = v;
            }
            __syncthreads();
        } else {
            hlu_t_phase(titem, t0, t1, data, z, tbuf, gpart, hdr, tb);
            bem_c64 a0 = c64(0, 0), a1 = c64(0, 0);
            hlu_c_phase(citem, c0, c1, m, data, z, tbuf, hdr, tb, a0, a1);
            hlu_finish<true>(L, M, a0, a1, z, part, acc);
        }
        k = kn;
        L = Ln;
        t0 = t0n; t1 = t1n; c0 = c0n; c1 = c1n;
    }
}

// The flattened factors split by panels of consecutive diagonal leaves
// (at least kHluPanelRows rows each): per direction, the items of every
// leaf whose block columns lie wholly outside its panel on the solved side
// (outer: before the panel for L, after it for U) and the rest (inner).
constexpr int64_t kHluPanelRows = 512;
struct HluSplit {
    std::vector<std::pair<int64_t, int64_t>> panel;   // leaves [k0, k1)
    std::vector<int64_t> tptr[2][2], cptr[2][2];      // [direction][outer]
    std::vector<bemhlu::FlatItem> titem[2][2], citem[2][2];
    std::vector<char> has_t[2], has_c[2];             // [direction][panel]: outer work present
    int64_t maxleaves = 1, gstride = 1;
};

void hlu_split(const bemhlu::Flat& F, HluSplit& S)
{
    const int64_t K = (int64_t)F.leaf.size();
    std::vector<int64_t> pan((size_t)K);
    for (int64_t k0 = 0; k0 < K;) {
        int64_t k1 = k0, rows = 0;
        while (k1 < K && rows < kHluPanelRows) rows += F.leaf[k1++].m;
        for (int64_t k = k0; k < k1; k++) pan[k] = (int64_t)S.panel.size();
        S.panel.push_back({k0, k1});
        S.maxleaves = std::max(S.maxleaves, k1 - k0);
        k0 = k1;
    }
    auto rounds = [&](const std::vector<bemhlu::FlatItem>& it, int64_t q0, int64_t q1) {
        for (int64_t r = q0; r < q1; r += kHluHdr) {
            int64_t s = 0;
            for (int64_t q = r; q < std::min(r + kHluHdr, q1); q++)
                s += (int64_t)it[q].k * ((it[q].n + kHluTChunk - 1) / kHluTChunk);
            S.gstride = std::max(S.gstride, s);
        }
    };
    for (int d = 0; d < 2; d++) {
        S.has_t[d].assign(S.panel.size(), 0);
        S.has_c[d].assign(S.panel.size(), 0);
        for (int o = 0; o < 2; o++) {
            S.tptr[d][o].assign(1, 0);
            S.cptr[d][o].assign(1, 0);
        }
        for (int64_t k = 0; k < K; k++) {
            const auto& P = S.panel[pan[k]];
            const int64_t lo = F.leaf[P.first].s, hi = F.leaf[P.second - 1].s + F.leaf[P.second - 1].m;
            auto outer = [&](const bemhlu::FlatItem& I) { return d == 0 ? I.c0 + I.n <= lo : I.c0 >= hi; };
            for (int64_t q = F.tptr[d][k]; q < F.tptr[d][k + 1]; q++) {
                const int o = outer(F.titem[d][q]) ? 1 : 0;
                S.titem[d][o].push_back(F.titem[d][q]);
                if (o) S.has_t[d][pan[k]] = 1;
            }
            for (int64_t q = F.cptr[d][k]; q < F.cptr[d][k + 1]; q++) {
                const int o = outer(F.citem[d][q]) ? 1 : 0;
                S.citem[d][o].push_back(F.citem[d][q]);
                if (o) S.has_c[d][pan[k]] = 1;
            }
            for (int o = 0; o < 2; o++) {
                S.tptr[d][o].push_back((int64_t)S.titem[d][o].size());
                S.cptr[d][o].push_back((int64_t)S.citem[d][o].size());
                rounds(S.titem[d][o], S.tptr[d][o][k], S.tptr[d][o][k + 1]);
            }
        }
    }
}

// every access k_hlu makes inside its buffers (checked before upload)
bool hlu_split_ok(const bemhlu::Flat& F, const HluSplit& S, std::string& why)
{
    const int64_t N = F.n, D = (int64_t)F.data.size(), T = F.tsize, K = (int64_t)F.leaf.size();
    auto bad = [&](const std::string& w) {
        why = w;
        return false;
    };
    if (K < 1) return bad("no diagonal leaves");
    for (const auto& L : F.leaf)
        if (L.m < 1 || L.m > kHluMaxLeaf || L.s < 0 || L.s + L.m > N || L.qh < 0 || L.qh + L.m * L.m > D ||
            L.rinv < 0 || L.rinv + L.m * L.m > D)
            return bad("diagonal leaf of " + std::to_string(L.m) + " rows (at most " + std::to_string(kHluMaxLeaf) +
                       ") or outside the factors");
    // the last element a matrix slice touches: rows 0 .. rows - 1 of columns 0 .. cols - 1
    auto inside = [&](int64_t a, int64_t ld, int64_t rows, int64_t cols) {
        if (rows <= 0 || cols <= 0) return a >= 0;
        return a >= 0 && ld >= rows && a + (rows - 1) + (cols - 1) * ld < D;
    };
    for (const auto& P : S.panel)
        if (P.first < 0 || P.second > K || P.first >= P.second) return bad("panels");
    for (int d = 0; d < 2; d++)
        for (int o = 0; o < 2; o++) {
            const auto &tp = S.tptr[d][o], &cp = S.cptr[d][o];
            if ((int64_t)tp.size() != K + 1 || (int64_t)cp.size() != K + 1 || tp[0] != 0 || cp[0] != 0 ||
                tp[K] != (int64_t)S.titem[d][o].size() || cp[K] != (int64_t)S.citem[d][o].size())
                return bad("item lists");
            for (int64_t k = 0; k < K; k++) {
                if (tp[k + 1] < tp[k] || cp[k + 1] < cp[k]) return bad("item lists");
                const int64_t m = F.leaf[k].m;
                int64_t round = 0;
                for (int64_t q = tp[k]; q < tp[k + 1]; q++) {
                    const auto& I = S.titem[d][o][q];
                    if (I.k < 0 || I.n < 0 || I.c0 < 0 || I.c0 + I.n > N || I.t < 0 || I.t + I.k > T ||
                        !inside(I.a, I.ld, I.n, I.k))
                        return bad("t item outside the factors");
                    if ((q - tp[k]) % kHluHdr == 0) round = 0;
                    round += (int64_t)I.k * ((I.n + kHluTChunk - 1) / kHluTChunk);
                    if (round > S.gstride) return bad("t tasks of a round");
                }
                for (int64_t q = cp[k]; q < cp[k + 1]; q++) {
                    const auto& I = S.citem[d][o][q];
                    const bool ok = I.k < 0 ? I.n >= 0 && I.c0 >= 0 && I.c0 + I.n <= N && inside(I.a, I.ld, m, I.n)
                                            : I.t >= 0 && I.t + I.k <= T && inside(I.a, I.ld, m, I.k);
                    if (!ok) return bad("contribution outside the factors");
                }
            }
        }
    return true;
}

// v = 0 (a kernel rather than a memset: also safe under graph capture)
__global__ void k_fill0(int64_t n, bem_c64* v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = c64(0.0, 0.0);
}

// x += z
__global__ void k_axpy1(int64_t n, const bem_c64* z, bem_c64* x)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = c64_add(x[i], z[i]);
}

// r = b - r
__global__ void k_residual(int64_t n, const bem_c64* b, bem_c64* r)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = c64_sub(b[i], r[i]);
}

__global__ void k_reconstruct(int64_t n, const bem_c64* x, const int32_t* subst, const bem_c64* al,
                              const bem_c64* be, const bem_c64* f, double s, bem_c64* phi, bem_c64* dphi,
                              bem_c64* pr)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bem_c64 ph = x[i], dp = x[i];
    if (subst[i]) dp = c64_div(c64_sub(f[i], c64_mul(al[i], ph)), be[i]);
    else ph = c64_div(c64_sub(f[i], c64_mul(be[i], dp)), al[i]);
    if (phi) phi[i] = ph;
    if (dphi) dphi[i] = dp;
    if (pr) pr[i] = c64_mul(c64(0.0, s), ph);
}

template <class C>
__global__ void k_has_nonfinite(int64_t len, const C* v, int32_t* flag)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len && !(isfinite(v[i].re) && isfinite(v[i].im))) *flag = 1;
}

// ------------------------------------------- several right-hand sides
// Vectors of Q right-hand sides are interleaved, v[i * Q + q], so that one
// pass over every block of the operator serves all Q products (Q = 1, 2, 4,
// 8).  Partial vectors are interleaved the same way.  The products
// accumulate in FP64 for either block storage.  For Q >= 4 the near-field
// tiles and the U t phase of the low-rank blocks run on the FP64 matrix
// cores (V_MFMA_F64_16X16X4F64, a complex product as four real ones); the
// VALU kernels stay for Q < 4 and as their measured alternative.

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma_f64(double a, double b, v4d c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// C (16 x 16 complex) += A (16 x 4) B (4 x 16) for one K step of four: this
// lane holds A[lane & 15][lane >> 4] in a and B[lane >> 4][lane & 15] in b;
// C[r][c] is held by lane c + 16 (r % 4), element r / 4 (the layout measured
// on gfx950 with a one-wavefront probe)
__device__ __forceinline__ void mfma_cstep(bem_c64 a, bem_c64 b, v4d& cr, v4d& ci)
{
    cr = mfma_f64(a.re, b.re, cr);
    cr = mfma_f64(-a.im, b.im, cr);
    ci = mfma_f64(a.re, b.im, ci);
    ci = mfma_f64(a.im, b.re, ci);
}

// near-field tiles, Q right-hand sides on the VALU (body: R columns of
// vectors with XS interleaved ones)
template <class S, int R, int XS>
__device__ __forceinline__ void mvm_dense_tile_q(const Tile& T, const S* __restrict__ F,
                                                 const bem_c64* __restrict__ x, bem_c64* __restrict__ part)
{
    __shared__ bem_c64 red[4][kTile * R];
    const int tid = threadIdx.x, r = tid & (kTile - 1), cg = tid >> 6;
    bem_c64 acc[R];
#pragma unroll
    for (int q = 0; q < R; q++) acc[q] = c64(0, 0);
    if (r < T.rows) {
        const S* f = F + T.dest + r;
        for (int c = cg; c < T.cols; c += 4) {
            const bem_c64 fv = ld64(f + (int64_t)c * T.ld);
            const bem_c64* xc = x + (int64_t)(T.col0 + c) * XS;
#pragma unroll
            for (int q = 0; q < R; q++) acc[q] = c64_fma(fv, xc[q], acc[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < R; q++) red[cg][r * R + q] = acc[q];
    __syncthreads();
    for (int o = tid; o < T.rows * R; o += kWG)
        part[(T.part + o / R) * XS + o % R] = c64_add(c64_add(c64_add(red[0][o], red[1][o]), red[2][o]), red[3][o]);
}

template <class S, int Q>
__global__ __launch_bounds__(kWG) void k_mvm_dense_q(const Tile* __restrict__ tiles, const S* __restrict__ F,
                                                     const bem_c64* __restrict__ x, bem_c64* __restrict__ part)
{
    mvm_dense_tile_q<S, Q, Q>(tiles[blockIdx.x], F, x, part);
}

// near-field tiles on the matrix cores: wavefront w takes rows 16 w ..
// 16 w + 15 of the tile over its columns in K steps of four
template <class S, int Q>
__global__ __launch_bounds__(kWG) void k_mvm_dense_mfma(const Tile* __restrict__ tiles, const S* __restrict__ F,
                                                        const bem_c64* __restrict__ x, bem_c64* __restrict__ part)
{
    static_assert(Q <= 16, "one 16-column tile of right-hand sides");
    const Tile T = tiles[blockIdx.x];
    const int lane = threadIdx.x & 63, rb = 16 * (threadIdx.x >> 6), lo = lane & 15, kq = lane >> 4;
    if (rb >= T.rows) return;
    const bool rowok = rb + lo < T.rows;
    const S* f = F + T.dest + rb + lo;
    v4d cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
    for (int kk = 0; kk < T.cols; kk += 4) {
        const int c = kk + kq;
        const bool cok = c < T.cols;
        const bem_c64 a = rowok && cok ? ld64(f + (int64_t)c * T.ld) : c64(0, 0);
        const bem_c64 b = lo < Q && cok ? x[(int64_t)(T.col0 + c) * Q + lo] : c64(0, 0);
        mfma_cstep(a, b, cr, ci);
    }
    if (lo < Q) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = rb + kq + 4 * v;
            if (row < T.rows) part[(T.part + row) * Q + lo] = c64(cr[v], ci[v]);
        }
    }
}

// low-rank phase 1, Q right-hand sides: t_b[l][q] = alpha_l (V_b[l, :] . x_J[:, q]),
// one wavefront per (block, l0) task of kLrRows rows of V (rows l0 .. l0 + NR - 1
// per pass, NR = kLrRows / Q rounded up to one).  Body: R columns of x with
// XS interleaved ones, t with R.
template <class S, int R, int XS>
__device__ __forceinline__ void lr_t_task_q(int64_t task, const int32_t* __restrict__ tasks,
                                            const int64_t* __restrict__ col0, const int64_t* __restrict__ cols,
                                            const int32_t* __restrict__ rank, const int64_t* __restrict__ voff,
                                            const int64_t* __restrict__ aoff, const int64_t* __restrict__ toff,
                                            const S* __restrict__ V, const S* __restrict__ A,
                                            const bem_c64* __restrict__ x, bem_c64* __restrict__ t)
{
    constexpr int NR = kLrRows / R > 0 ? kLrRows / R : 1;
    const int lane = threadIdx.x & 63;
    const int b = tasks[2 * task], lt = tasks[2 * task + 1];
    const int64_t n = cols[b];
    const bem_c64* xs = x + col0[b] * XS;
    for (int l0 = lt; l0 < lt + kLrRows && l0 < rank[b]; l0 += NR) {
        const int nl = min(NR, rank[b] - l0);
        const S* v = V + voff[b] + (int64_t)l0 * n;
        bem_c64 p[NR][R];
#pragma unroll
        for (int u = 0; u < NR; u++)
#pragma unroll
            for (int q = 0; q < R; q++) p[u][q] = c64(0, 0);
        for (int64_t j = lane; j < n; j += 64) {
            bem_c64 xv[R];
#pragma unroll
            for (int q = 0; q < R; q++) xv[q] = xs[j * XS + q];
#pragma unroll
            for (int u = 0; u < NR; u++)
                if (u < nl) {
                    const bem_c64 vv = ld64(v + (int64_t)u * n + j);
#pragma unroll
                    for (int q = 0; q < R; q++) p[u][q] = c64_fma(vv, xv[q], p[u][q]);
                }
        }
#pragma unroll
        for (int u = 0; u < NR; u++)
#pragma unroll
            for (int q = 0; q < R; q++) p[u][q] = wave_sum(p[u][q]);
        if (lane == 0) {
#pragma unroll
            for (int u = 0; u < NR; u++)
                if (u < nl) {
                    const bem_c64 al = ld64(A + aoff[b] + l0 + u);
#pragma unroll
                    for (int q = 0; q < R; q++) t[(toff[b] + l0 + u) * R + q] = c64_mul(al, p[u][q]);
                }
        }
    }
}

template <class S, int Q>
__global__ __launch_bounds__(kWG) void k_lr_t_q(const int32_t* __restrict__ tasks, int64_t ntask,
                                                const int64_t* __restrict__ col0, const int64_t* __restrict__ cols,
                                                const int32_t* __restrict__ rank, const int64_t* __restrict__ voff,
                                                const int64_t* __restrict__ aoff, const int64_t* __restrict__ toff,
                                                const S* __restrict__ V, const S* __restrict__ A,
                                                const bem_c64* __restrict__ x, bem_c64* __restrict__ t)
{
    const int64_t task = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
    if (task >= ntask) return;
    lr_t_task_q<S, Q, Q>(task, tasks, col0, cols, rank, voff, aoff, toff, V, A, x, t);
}

// low-rank phase 2, Q right-hand sides on the VALU: partial_b[i][q] =
// sum_l U_b[i, l] t_b[l][q] over a chunk of kWG rows (body: t of R columns,
// partial vectors with XS interleaved ones)
template <class S, int R, int XS>
__device__ __forceinline__ void lr_u_task_q(int64_t task, const int32_t* __restrict__ tasks,
                                            const int64_t* __restrict__ rows, const int32_t* __restrict__ rank,
                                            const int64_t* __restrict__ uoff, const int64_t* __restrict__ toff,
                                            const int64_t* __restrict__ poff, const S* __restrict__ U,
                                            const bem_c64* __restrict__ t, bem_c64* __restrict__ part)
{
    const int b = tasks[2 * task], chunk = tasks[2 * task + 1];
    const int64_t m = rows[b];
    const int k = rank[b];
    const int64_t i = (int64_t)chunk * kWG + threadIdx.x;
    if (i >= m) return;
    const S* u = U + uoff[b] + i;
    const bem_c64* tb = t + toff[b] * R;   // the same k R values for every thread (cached)
    bem_c64 acc[R];
#pragma unroll
    for (int q = 0; q < R; q++) acc[q] = c64(0, 0);
    for (int l = 0; l < k; l++) {
        const bem_c64 uv = ld64(u + (int64_t)l * m);
#pragma unroll
        for (int q = 0; q < R; q++) acc[q] = c64_fma(uv, tb[l * R + q], acc[q]);
    }
#pragma unroll
    for (int q = 0; q < R; q++) part[(poff[b] + i) * XS + q] = acc[q];
}

template <class S, int Q>
__global__ __launch_bounds__(kWG) void k_lr_u_q(const int32_t* __restrict__ tasks, const int64_t* __restrict__ rows,
                                                const int32_t* __restrict__ rank, const int64_t* __restrict__ uoff,
                                                const int64_t* __restrict__ toff, const int64_t* __restrict__ poff,
                                                const S* __restrict__ U, const bem_c64* __restrict__ t,
                                                bem_c64* __restrict__ part)
{
    lr_u_task_q<S, Q, Q>(blockIdx.x, tasks, rows, rank, uoff, toff, poff, U, t, part);
}

// low-rank phase 2 on the matrix cores: the chunk's kWG rows as 16 row tiles,
// wavefront w taking tiles w, w + 4, ...; K = the block's rank in steps of four
template <class S, int Q>
__global__ __launch_bounds__(kWG) void k_lr_u_mfma(const int32_t* __restrict__ tasks, const int64_t* __restrict__ rows,
                                                   const int32_t* __restrict__ rank, const int64_t* __restrict__ uoff,
                                                   const int64_t* __restrict__ toff, const int64_t* __restrict__ poff,
                                                   const S* __restrict__ U, const bem_c64* __restrict__ t,
                                                   bem_c64* __restrict__ part)
{
    const int b = tasks[2 * blockIdx.x], chunk = tasks[2 * blockIdx.x + 1];
    const int64_t m = rows[b];
    const int k = rank[b];
    const int lane = threadIdx.x & 63, lo = lane & 15, kq = lane >> 4;
    const S* u = U + uoff[b];
    const bem_c64* tb = t + toff[b] * Q;
    for (int tile = threadIdx.x >> 6; tile < kWG / 16; tile += kWG / 64) {
        const int64_t r0 = (int64_t)chunk * kWG + 16 * tile;
        if (r0 >= m) break;
        const bool rowok = r0 + lo < m;
        v4d cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
        for (int kk = 0; kk < k; kk += 4) {
            const int l = kk + kq;
            const bool lok = l < k;
            const bem_c64 a = rowok && lok ? ld64(u + (r0 + lo) + (int64_t)l * m) : c64(0, 0);
            const bem_c64 bb = lo < Q && lok ? tb[l * Q + lo] : c64(0, 0);
            mfma_cstep(a, bb, cr, ci);
        }
        if (lo < Q) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int64_t row = r0 + kq + 4 * v;
                if (row < m) part[(poff[b] + row) * Q + lo] = c64(cr[v], ci[v]);
            }
        }
    }
}

// y[i][q] (+)= sum of the partial vectors covering row i, in partition order
// per lane group as k_mvm_gather
template <int Q>
__global__ void k_mvm_gather_q(int64_t n, const int64_t* __restrict__ gptr, const int64_t* __restrict__ gidx,
                               const bem_c64* __restrict__ part, bem_c64* y, int accumulate)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = g >> 4;
    const int lq = (int)(g & 15);
    bem_c64 acc[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) acc[q] = c64(0, 0);
    if (i < n)
        for (int64_t p = gptr[i] + lq; p < gptr[i + 1]; p += 16) {
            const bem_c64* src = part + gidx[p] * Q;
#pragma unroll
            for (int q = 0; q < Q; q++) acc[q] = c64_add(acc[q], src[q]);
        }
#pragma unroll
    for (int q = 0; q < Q; q++)
        for (int s = 8; s > 0; s >>= 1) {
            acc[q].re = acc[q].re + __shfl_down(acc[q].re, s, 16);
            acc[q].im = acc[q].im + __shfl_down(acc[q].im, s, 16);
        }
    if (i < n && lq == 0) {
#pragma unroll
        for (int q = 0; q < Q; q++) y[i * Q + q] = accumulate ? c64_add(y[i * Q + q], acc[q]) : acc[q];
    }
}

// ---- products of several operators in one pass (the frequencies of a
// batch): operator g = blockIdx.y owns columns g R .. g R + R - 1 of the Q
// interleaved vectors.  The operators share the context's partition (tiles,
// block extents, partial-vector offsets, the row-chunk tasks of phase 2);
// each has its own stored values, ranks and phase-1 tasks.  The input is
// first regrouped by operator (xg: operator g's R columns, R-interleaved, at
// g n R) and the partial vectors are kept the same way (operator g's at
// g P R, P partials per column), so that every operator's kernels read and
// write contiguous vectors exactly as its single-operator kernels do; with
// R = 1 the arithmetic is theirs.  A padding group (pad[g]: columns beyond
// the solve's right-hand sides) writes zero partial vectors and reads
// nothing.
constexpr int kMaxOps = 8;
struct OpsArg {
    int32_t pad[kMaxOps];
    const void* F[kMaxOps];                  // near-field storage read
    const int32_t* tt[kMaxOps];              // phase-1 (block, l0) tasks
    const int32_t* ut[kMaxOps];              // phase-2 (block, row chunk) tasks
    int64_t ntt[kMaxOps], nut[kMaxOps];
    const int32_t* rank[kMaxOps];
    const int64_t *voff[kMaxOps], *aoff[kMaxOps], *toff[kMaxOps], *uoff[kMaxOps];
    const void *U[kMaxOps], *V[kMaxOps], *A[kMaxOps];
    bem_c64* t[kMaxOps];                     // R columns per operator
    int64_t n, P;                            // rows; partial vector entries per column
};

// xg[(g n + i) R + r] = x[i Q + g R + r]
template <int R, int Q>
__global__ void k_group_cols(int64_t n, const bem_c64* __restrict__ x, bem_c64* __restrict__ xg)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * Q) return;
    const int64_t i = e / Q;
    const int q = (int)(e - i * Q), g = q / R, r = q - g * R;
    xg[((int64_t)g * n + i) * R + r] = x[e];
}

template <class S, int R>
__global__ __launch_bounds__(kWG) void k_mvm_dense_ops(const Tile* __restrict__ tiles, OpsArg O,
                                                       const bem_c64* __restrict__ xg, bem_c64* __restrict__ part)
{
    const int g = blockIdx.y;
    bem_c64* pg = part + (int64_t)g * O.P * R;
    if (O.pad[g]) {
        const Tile T = tiles[blockIdx.x];
        for (int o = threadIdx.x; o < T.rows * R; o += kWG) pg[T.part * R + o] = c64(0, 0);
        return;
    }
    const S* F = (const S*)O.F[g];
    const bem_c64* x = xg + (int64_t)g * O.n * R;
    if constexpr (R == 1) mvm_dense_tile<S, 1>(tiles[blockIdx.x], F, x, pg);
    else mvm_dense_tile_q<S, R, R>(tiles[blockIdx.x], F, x, pg);
}

template <class S, int R>
__global__ __launch_bounds__(kWG) void k_lr_t_ops(OpsArg O, const int64_t* __restrict__ col0,
                                                  const int64_t* __restrict__ cols, const bem_c64* __restrict__ xg)
{
    const int g = blockIdx.y;
    const int64_t task = (int64_t)blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
    if (task >= O.ntt[g]) return;
    const bem_c64* x = xg + (int64_t)g * O.n * R;
    if constexpr (R == 1)
        lr_t_task<S, 1>(task, O.tt[g], col0, cols, O.rank[g], O.voff[g], O.aoff[g], O.toff[g], (const S*)O.V[g],
                        (const S*)O.A[g], x, O.t[g]);
    else
        lr_t_task_q<S, R, R>(task, O.tt[g], col0, cols, O.rank[g], O.voff[g], O.aoff[g], O.toff[g],
                             (const S*)O.V[g], (const S*)O.A[g], x, O.t[g]);
}

template <class S, int R>
__global__ __launch_bounds__(kWG) void k_lr_u_ops(OpsArg O, const int64_t* __restrict__ rows,
                                                  const int64_t* __restrict__ poff, bem_c64* __restrict__ part)
{
    extern __shared__ bem_c64 st[];
    const int g = blockIdx.y;
    if ((int64_t)blockIdx.x >= O.nut[g]) return;   // uniform per workgroup
    bem_c64* pg = part + (int64_t)g * O.P * R;
    if (O.pad[g]) {
        const int b = O.ut[g][2 * blockIdx.x], chunk = O.ut[g][2 * blockIdx.x + 1];
        const int64_t i = (int64_t)chunk * kWG + threadIdx.x;
        if (i < rows[b])
            for (int q = 0; q < R; q++) pg[(poff[b] + i) * R + q] = c64(0, 0);
        return;
    }
    if constexpr (R == 1)
        lr_u_task<S, 1>(blockIdx.x, O.ut[g], rows, O.rank[g], O.uoff[g], O.toff[g], poff, (const S*)O.U[g], O.t[g],
                        st, pg);
    else
        lr_u_task_q<S, R, R>(blockIdx.x, O.ut[g], rows, O.rank[g], O.uoff[g], O.toff[g], poff, (const S*)O.U[g],
                             O.t[g], pg);
}

// y[i Q + g R + r] = sum of operator g's partial vectors covering row i (in
// the order and with the shuffle tree of k_mvm_gather / k_mvm_gather_q)
template <int R, int Q>
__global__ void k_mvm_gather_ops(int64_t n, int64_t P, const int64_t* __restrict__ gptr,
                                 const int64_t* __restrict__ gidx, const bem_c64* __restrict__ part, bem_c64* y)
{
    const int64_t gl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = gl >> 4;
    const int lq = (int)(gl & 15);
    bem_c64 acc[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) acc[q] = c64(0, 0);
    if (i < n)
        for (int64_t p = gptr[i] + lq; p < gptr[i + 1]; p += 16) {
            const int64_t o = gidx[p] * R;
#pragma unroll
            for (int q = 0; q < Q; q++) acc[q] = c64_add(acc[q], part[(int64_t)(q / R) * P * R + o + q % R]);
        }
#pragma unroll
    for (int q = 0; q < Q; q++)
        for (int s = 8; s > 0; s >>= 1) {
            acc[q].re = acc[q].re + __shfl_down(acc[q].re, s, 16);
            acc[q].im = acc[q].im + __shfl_down(acc[q].im, s, 16);
        }
    if (i < n && lq == 0) {
#pragma unroll
        for (int q = 0; q < Q; q++) y[i * Q + q] = acc[q];
    }
}

// ---- lockstep GMRES over Q right-hand sides.  The Krylov bases are
// interleaved like the vectors (V[(l n + i) Q + q]); the current column j is
// read from device memory (*jp), so that one iteration's launches form a
// graph that is captured once and replayed for every j.

constexpr int kArnMax = 8;   // right-hand sides of one lockstep solve

// X = column *jp of the bases (the product's input)
template <int Q>
__global__ void k_basis_col(int64_t n, const int* __restrict__ jp, const bem_c64* __restrict__ V, bem_c64* __restrict__ X)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n * Q) X[t] = V[(int64_t)(*jp) * n * Q + t];
}

// part[(l nchunk + chunk) Q + q] = sum over the chunk's rows of conj(V_l) w
// for l <= *jp; one wavefront per (chunk, group of vectors)
template <int Q>
__global__ __launch_bounds__(kWG) void k_cgs_dots_q(int64_t n, const int* __restrict__ jp,
                                                    const bem_c64* __restrict__ V, const bem_c64* __restrict__ w,
                                                    bem_c64* __restrict__ part)
{
    constexpr int G = kDotVecs / Q > 0 ? kDotVecs / Q : 1;
    const int cnt = *jp + 1;
    const int l0 = (blockIdx.y * (kWG / 64) + (threadIdx.x >> 6)) * G;
    if (l0 >= cnt) return;
    const int lane = threadIdx.x & 63;
    const int nv = min(G, cnt - l0);
    const int64_t i0 = (int64_t)blockIdx.x * kDotChunk, i1 = i0 + kDotChunk < n ? i0 + kDotChunk : n;
    bem_c64 p[G][Q];
#pragma unroll
    for (int u = 0; u < G; u++)
#pragma unroll
        for (int q = 0; q < Q; q++) p[u][q] = c64(0, 0);
    for (int64_t i = i0 + lane; i < i1; i += 64) {
        bem_c64 wi[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) wi[q] = w[i * Q + q];
#pragma unroll
        for (int u = 0; u < G; u++)
            if (u < nv) {
                const bem_c64* vr = V + ((int64_t)(l0 + u) * n + i) * Q;
#pragma unroll
                for (int q = 0; q < Q; q++) p[u][q] = c64_fma(c64_conj(vr[q]), wi[q], p[u][q]);
            }
    }
#pragma unroll
    for (int u = 0; u < G; u++)
#pragma unroll
        for (int q = 0; q < Q; q++) p[u][q] = wave_sum(p[u][q]);
    if (lane == 0) {
#pragma unroll
        for (int u = 0; u < G; u++)
            if (u < nv)
#pragma unroll
                for (int q = 0; q < Q; q++) part[((int64_t)(l0 + u) * gridDim.x + blockIdx.x) * Q + q] = p[u][q];
    }
}

// h[l Q + q] = sum over chunks; hacc = h (first pass) or += h.  One
// wavefront per l <= *jp.
template <int Q>
__global__ __launch_bounds__(64) void k_cgs_reduce_q(const int* __restrict__ jp, int nchunk,
                                                     const bem_c64* __restrict__ part, bem_c64* h, bem_c64* hacc,
                                                     int first)
{
    const int l = blockIdx.x;
    if (l > *jp) return;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        bem_c64 s = c64(0, 0);
        for (int c = threadIdx.x; c < nchunk; c += 64) s = c64_add(s, part[((int64_t)l * nchunk + c) * Q + q]);
        s = wave_sum(s);
        if (threadIdx.x == 0) {
            h[l * Q + q] = s;
            hacc[l * Q + q] = first ? s : c64_add(hacc[l * Q + q], s);
        }
    }
}

// w -= sum_{l <= *jp} V_l h_l
template <int Q>
__global__ void k_orth_update_q(int64_t n, const int* __restrict__ jp, const bem_c64* __restrict__ V,
                                const bem_c64* __restrict__ h, bem_c64* __restrict__ w)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * Q) return;
    const int q = (int)(t % Q);
    const int cnt = *jp + 1;
    bem_c64 acc = w[t];
    for (int l = 0; l < cnt; l++) acc = c64_fms(h[l * Q + q], V[(int64_t)l * n * Q + t], acc);
    w[t] = acc;
}

// partial sums of |w_q|^2 per workgroup: part[block Q + q]
template <int Q>
__global__ __launch_bounds__(kWG) void k_norm2_part_q(int64_t n, const bem_c64* __restrict__ w, double* part)
{
    __shared__ double red[kWG / 64][Q];
    double p[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) p[q] = 0;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG)
#pragma unroll
        for (int q = 0; q < Q; q++) p[q] = p[q] + c64_abs2(w[i * Q + q]);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        p[q] = wave_sum(p[q]);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][q] = p[q];
    }
    __syncthreads();
    if (threadIdx.x < Q) {
        const int q = threadIdx.x;
        part[blockIdx.x * Q + q] = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
    }
}

template <int Q>
__global__ void k_norm2_final_q(int nparts, const double* __restrict__ part, double* out)
{
    const int q = threadIdx.x;
    if (blockIdx.x != 0 || q >= Q) return;
    double s = 0;
    for (int p = 0; p < nparts; p++) s += part[p * Q + q];
    out[q] = sqrt(s);
}

// V column *jp + 1 = w / |w| (0 for a zero norm)
template <int Q>
__global__ void k_scale_next_q(int64_t n, const int* __restrict__ jp, const bem_c64* __restrict__ w,
                               const double* __restrict__ nrm, bem_c64* __restrict__ V)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * Q) return;
    const double d = nrm[t % Q];
    V[(int64_t)(*jp + 1) * n * Q + t] = d > 0 ? c64_scale(w[t], 1.0 / d) : c64(0, 0);
}

// Arnoldi column *jp for every right-hand side (thread q), as k_arnoldi:
// H_q column j, the Givens rotations, g_q and resid[q] = |g_q[j + 1]| / |b_q|;
// then *jp += 1.  H is Q x (m + 1) x m, cs_c / cs_s / g are Q x (m + 2).
template <int Q>
__global__ void k_arnoldi_q(int* jp, int m, const bem_c64* __restrict__ hacc, const double* __restrict__ hn,
                            bem_c64* H, double* cs_c, bem_c64* cs_s, bem_c64* g, const double* __restrict__ bn,
                            double* resid)
{
    const int q = threadIdx.x;
    const int j = *jp;
    if (q < Q) {
        bem_c64* h = H + ((int64_t)q * m + j) * (m + 1);
        double* cc_q = cs_c + (int64_t)q * (m + 2);
        bem_c64* ss_q = cs_s + (int64_t)q * (m + 2);
        bem_c64* g_q = g + (int64_t)q * (m + 2);
        for (int l = 0; l <= j; l++) h[l] = hacc[l * Q + q];
        h[j + 1] = c64(hn[q], 0.0);
        for (int l = 0; l < j; l++) {
            bem_c64 a0 = h[l], b0 = h[l + 1];
            h[l] = c64_add(c64_scale(a0, cc_q[l]), c64_mul(ss_q[l], b0));
            h[l + 1] = c64_sub(c64_scale(b0, cc_q[l]), c64_mul(c64_conj(ss_q[l]), a0));
        }
        double na = sqrt(c64_abs2(h[j])), nb2 = sqrt(c64_abs2(h[j + 1]));
        double tt = sqrt(na * na + nb2 * nb2);
        double cc;
        bem_c64 ss;
        if (na == 0) { cc = 0; ss = c64(1, 0); }
        else {
            cc = na / tt;
            ss = c64_scale(c64_mul(c64_scale(h[j], 1.0 / na), c64_conj(h[j + 1])), 1.0 / tt);
        }
        cc_q[j] = cc;
        ss_q[j] = ss;
        h[j] = c64_add(c64_scale(h[j], cc), c64_mul(ss, h[j + 1]));
        h[j + 1] = c64(0, 0);
        bem_c64 gj = g_q[j];
        g_q[j] = c64_scale(gj, cc);
        g_q[j + 1] = c64_neg(c64_mul(c64_conj(ss), gj));
        resid[q] = bn[q] > 0 ? sqrt(c64_abs2(g_q[j + 1])) / bn[q] : 0.0;
    }
    __syncthreads();
    if (q == 0) *jp = j + 1;
}

// x_q += sum_{l < cnt} V_l y[l Q + q]
template <int Q>
__global__ void k_combine_q(int64_t n, int cnt, const bem_c64* __restrict__ V, const bem_c64* __restrict__ y,
                            bem_c64* __restrict__ x)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * Q) return;
    const int q = (int)(t % Q);
    bem_c64 acc = x[t];
    for (int l = 0; l < cnt; l++) acc = c64_fma(V[(int64_t)l * n * Q + t], y[l * Q + q], acc);
    x[t] = acc;
}

// V column 0 = r / beta (per right-hand side; 0 where beta = 0)
template <int Q>
__global__ void k_scale_first_q(int64_t n, const bem_c64* __restrict__ r, const double* __restrict__ beta,
                                bem_c64* __restrict__ V)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * Q) return;
    const double d = beta[t % Q];
    V[t] = d > 0 ? c64_scale(r[t], 1.0 / d) : c64(0, 0);
}

// r = b - r over n Q entries
__global__ void k_residual_n(int64_t len, const bem_c64* __restrict__ b, bem_c64* __restrict__ r)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < len) r[t] = c64_sub(b[t], r[t]);
}

// w columns between two interleaved layouts: dst[i ds + doff + q] =
// src[i ss + soff + q] for q < w
__global__ void k_cols(int64_t n, int w, const bem_c64* __restrict__ src, int ss, int soff, bem_c64* __restrict__ dst,
                       int ds, int doff)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * w) return;
    const int64_t i = t / w;
    const int q = (int)(t % w);
    dst[i * ds + doff + q] = src[i * ss + soff + q];
}

// dst[i Q + q] = src[i ld + q] (q < nq), 0 for nq <= q < Q
__global__ void k_repack(int64_t n, int nq, int64_t ld, int Q, const bem_c64* __restrict__ src, bem_c64* __restrict__ dst)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * Q) return;
    const int64_t i = t / Q;
    const int q = (int)(t % Q);
    dst[t] = q < nq ? src[i * ld + q] : c64(0, 0);
}

// ------------------------------------------------------------ host side

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
};

// Recycled host vectors: the per-assembly offset, rank and task arrays are
// taken from here and given back, so repeated assemblies touch memory that is
// already mapped instead of faulting in fresh pages every time.
struct VecPool {
    std::vector<std::vector<int64_t>> v64;
    std::vector<std::vector<int32_t>> v32;
    template <class T>
    std::vector<std::vector<T>>& list();
    template <class T>
    std::vector<T> take(size_t n, T fill)
    {
        auto& L = list<T>();
        std::vector<T> v;
        if (!L.empty()) {
            v = std::move(L.back());
            L.pop_back();
        }
        v.assign(n, fill);
        return v;
    }
    template <class T>
    void give(std::vector<T>& v)
    {
        auto& L = list<T>();
        if (L.size() < 64 && v.capacity()) L.push_back(std::move(v));
        v = std::vector<T>();
    }
};
template <>
std::vector<std::vector<int64_t>>& VecPool::list<int64_t>() { return v64; }
template <>
std::vector<std::vector<int32_t>>& VecPool::list<int32_t>() { return v32; }

}  // namespace

struct bem_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side[2] = {nullptr, nullptr};   // concurrent ACA launches (default priority)
    hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
    std::string err;
    // mesh
    int64_t n = 0;
    double hbar = 0;
    std::vector<double> h_nodes, h_normal, h_area, h_colloc;
    std::vector<void*> allocs_mesh;
    DevMesh mesh{};
    // impedance plane (point, normal) and the mirror images of the mesh
    bool have_plane = false;
    double plane_pt[3] = {0, 0, 0}, plane_nrm[3] = {0, 0, 0};
    std::vector<void*> allocs_plane;
    bemsetup::NearList near;
    // bc
    std::vector<bem_c64> h_alpha, h_beta, h_f;
    bem_c64 *d_alpha = nullptr, *d_beta = nullptr, *d_f = nullptr;
    bool have_bc = false;
    // frequency
    double k = 0;
    bem_c64 cc{0, 0};
    int32_t* d_subst = nullptr;
    bem_c64 *d_cphi = nullptr, *d_cd = nullptr;
    bem_c32 *d_cphi32 = nullptr, *d_cd32 = nullptr;   // FP32 copies for the FP32 assembly
    int64_t freq_n = 0;                                // length of the arrays above
    std::vector<int32_t> h_subst;
    bool have_freq = false;
    // partition
    std::vector<bemsetup::Block> blocks;
    int64_t *d_row0 = nullptr, *d_rows = nullptr, *d_col0 = nullptr, *d_cols = nullptr;
    int8_t* d_bnear = nullptr;   // per block: may hold self / near-singular pairs (aca_block)
    std::vector<int64_t> dense_off;          // per block (-1 if admissible)
    int64_t dense_total = 0;
    std::vector<Tile> tiles;
    Tile* d_tiles = nullptr;
    std::vector<Fix> fixes;
    Fix* d_fixes = nullptr;
    bool piv_ok = false;                 // cached DPhi pivot lists
    uint64_t piv_seed = 0;
    int64_t* d_pivoff = nullptr;
    int32_t* d_piv = nullptr;
    std::vector<void*> allocs_piv;
    int32_t* d_sing_order = nullptr;     // evaluation order of the singular items
    bem_c64* d_sing = nullptr;           // per near pair, then per element (self)
    // near-field pairing (bem_set_near_pairing): the Phi assembly also forms
    // the DPhi near field (dense blocks and singular table) from the same
    // operators and keeps it for the next DPhi assembly of the same state
    bool pair_near = false;
    bool prec_left = false;   // bem_set_precond_side: left preconditioning (the reference's form)
    uint64_t gen = 0;                    // bumped by every change of mesh, partition, BC, frequency, plane
    bem_c64* d_sing2 = nullptr;          // the kept DPhi singular table
    void* st_dense = nullptr;            // the kept DPhi dense blocks
    void* st_dense64 = nullptr;          // and their FP64 form (FP32 operators with near64)
    int st_prec = -1;
    bool st_near64 = false;
    // the same for a batched Phi assembly: the DPhi near field and singular
    // table of each of its frequencies, kept for the next batched DPhi
    // assembly of the same state and frequency list
    struct KeptBatch {
        std::vector<double> kcc;
        std::vector<void*> dense, dense64, sing;
        uint64_t gen = 0;
        int prec = -1;
        bool near64 = false;
    } kb;
    // FP32 assemblies form the near field in FP64 and keep it beside the FP32
    // copy (bem_set_near_fp64): the refinement residual's operator
    bool near64 = false;
    // products of several operators (the frequencies of a batched solve) in
    // one pass per kernel, not one operator after another (bem_set_fuse_ops)
    bool fuse_ops = true;
    uint64_t st_gen = 0;
    const bem_c64* sing_cur = nullptr;   // table the running assembly's ACA reads
    int64_t nnz = 0;
    std::vector<int32_t> adm_list;           // admissible block ids, large first
    int32_t* d_adm_list = nullptr;
    std::vector<int64_t> lr_part;            // per block partial offset (low rank)
    int64_t* d_lr_part = nullptr;
    int64_t part_total = 0;
    bem_c64* d_part = nullptr;
    int64_t *d_gptr = nullptr, *d_gidx = nullptr;
    bool have_part = false;
    // GMRES workspace
    bem_c64* d_gm = nullptr;
    int64_t gm_cap = 0;
    // profiling
    bool prof = false;
    double prof_ms[7] = {0};
    int64_t prof_cnt[7] = {0};
    struct Pending { int slot; hipEvent_t e0, e1; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    std::vector<void*> allocs_part;
    // device buffer pool: assemblies reuse their multi-GB scratch instead of
    // calling hipMalloc/hipFree every time
    std::multimap<size_t, void*> pool;
    std::unordered_map<void*, size_t> live;
    // pinned host staging for the per-assembly metadata uploads and the
    // rank / flag downloads (one copy each instead of one per array)
    char* stage = nullptr;
    size_t stage_cap = 0;
    hipEvent_t stage_done = nullptr;
    bool stage_busy = false;
    VecPool vp;
    // device address of `scratch` on this context's GPU (bem_create)
    void* scratch_dev = nullptr;
    // GMRES: residual estimates of the last two Arnoldi columns
    double* gm_res = nullptr;
    // several right-hand sides: Q-interleaved partial vectors (part_total x
    // partq_cap), the lockstep GMRES's pinned residual ring (2 x kArnMax),
    // matrix cores for Q >= 4 products, graph-captured Arnoldi iterations
    bem_c64* d_partq = nullptr;
    int partq_cap = 0;
    bem_c64* d_xq = nullptr;             // n x partq_cap: vectors regrouped by operator (fused products)
    double* gm_resq = nullptr;
    bool use_mfma = true;
    bool use_graphs = true;
    // ACA size classes by m + n per storage precision [FP64, FP32]: up to
    // cls_small one wavefront, up to cls_huge the middle class, above the
    // largest; batched assemblies on persistent workgroups (queue counter)
    int64_t cls_small[2] = {kSmallBlock[0], kSmallBlock[1]};
    int64_t cls_huge[2] = {kHugeBlock[0], kHugeBlock[1]};
    bool aca_queue = false;
    bool capturing = false;   // Timer records no events while a graph is captured
    hipEvent_t gm_ev[2] = {nullptr, nullptr};
};

struct bem_hmat {
    bem_ctx* ctx = nullptr;
    int kind = 0, prec = 0;
    size_t esz = sizeof(bem_c64);   // bytes per stored entry (FP64 or FP32 complex)
    void* dense = nullptr;
    void* dense64 = nullptr;   // FP64 near field of an FP32 operator (bem_set_near_fp64), else null
    void *U = nullptr, *V = nullptr, *A = nullptr;
    void* meta = nullptr;    // one allocation holding the offset, rank and task arrays below
    int64_t *d_uoff = nullptr, *d_voff = nullptr, *d_aoff = nullptr;
    int32_t* d_rank = nullptr;
    // two-phase low-rank MVM: (block, l) and (block, row chunk) work lists
    int32_t *d_ttask = nullptr, *d_utask = nullptr;
    int64_t nttask = 0, nutask = 0;
    int64_t* d_toff = nullptr;
    bem_c64* d_t = nullptr;
    int maxrank = 1;
    int64_t t_total = 0;      // sum of the ranks (entries of d_t)
    bem_c64* d_tq = nullptr;  // t vectors of Q right-hand sides (t_total x tq_cap)
    int tq_cap = 0;
    std::vector<int64_t> uoff, voff, aoff;
    std::vector<int32_t> rank;
    double stats[9] = {0};
};

struct bem_prec {
    bem_ctx* ctx = nullptr;
    int64_t nblk = 0;
    PrecBlock* d_blk = nullptr;
    bem_c64* d_lu = nullptr;
    int32_t* d_piv = nullptr;
    // H-LU (bem_precond_create_hlu): the flattened factors on the device
    int64_t hlu_nleaf = 0;
    HluLeafD* hlu_leaf = nullptr;
    int64_t* hlu_tptr[2][2] = {};      // [direction][outer]
    int64_t* hlu_cptr[2][2] = {};
    HluItemD* hlu_titem[2][2] = {};
    HluItemD* hlu_citem[2][2] = {};
    std::vector<std::pair<int64_t, int64_t>> hlu_panel;
    std::vector<char> hlu_has_t[2], hlu_has_c[2];
    int64_t hlu_gstride = 1;
    bem_c64* hlu_data = nullptr;
    bem_c64* hlu_t = nullptr;
    bem_c64* hlu_gpart = nullptr;   // chunk sums of the t vectors of one round
    std::vector<void*> hlu_owned;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> hlu_ev;   // per application, resolved by bem_precond_stats
    double setup_ms = 0, apply_ms = 0, entries = 0, factor_ms = 0;
    int64_t applies = 0;
};

namespace {

// Device time per kernel family: a pair of events is recorded around each
// launch on the context stream and resolved only when bem_profile() reads
// the totals, so profiling adds no synchronisation to the timed work.
struct Timer {
    bem_ctx* c;
    int slot;
    hipStream_t st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    Timer(bem_ctx* c_, int s, hipStream_t stream = nullptr);
    ~Timer();
};

hipEvent_t pool_get(bem_ctx* c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void resolve_pending(bem_ctx* c)
{
    if (c->pending.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (hipStream_t t : c->side) (void)hipStreamSynchronize(t);
    for (auto& p : c->pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, p.e0, p.e1) == hipSuccess) {
            c->prof_ms[p.slot] += ms;
            c->prof_cnt[p.slot]++;
        }
        c->event_pool.push_back(p.e0);
        c->event_pool.push_back(p.e1);
    }
    c->pending.clear();
}

Timer::Timer(bem_ctx* c_, int s, hipStream_t stream) : c(c_), slot(s), st(stream ? stream : c_->stream)
{
    if (c->prof && !c->capturing) {
        e0 = pool_get(c);
        e1 = pool_get(c);
        (void)hipEventRecord(e0, st);
    }
}

Timer::~Timer()
{
    if (e1) {
        (void)hipEventRecord(e1, st);
        c->pending.push_back(bem_ctx::Pending{slot, e0, e1});
        if (c->pending.size() > 4096) resolve_pending(c);
    }
}

int set_err(bem_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPC(ctx, call)                                                                                     \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess)                                                                               \
            return set_err(ctx, BEM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));            \
    } while (0)

void pool_release_all(bem_ctx* c)
{
    for (auto& kv : c->pool) (void)hipFree(kv.second);
    c->pool.clear();
}

int dalloc_bytes(bem_ctx* c, void** p, size_t bytes)
{
    bytes = ((bytes ? bytes : 1) + 255) & ~(size_t)255;
    auto it = c->pool.lower_bound(bytes);
    if (it != c->pool.end() && it->first <= 2 * bytes + ((size_t)1 << 20)) {
        *p = it->second;
        c->live[*p] = it->first;
        c->pool.erase(it);
        return BEM_OK;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {   // give the pooled memory back and try once more
        (void)hipGetLastError();
        pool_release_all(c);
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) return set_err(c, BEM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->live[*p] = bytes;
    return BEM_OK;
}

void pfree(bem_ctx* c, void* p)
{
    if (!p) return;
    auto it = c->live.find(p);
    if (it == c->live.end()) { (void)hipFree(p); return; }
    c->pool.insert({it->second, p});
    c->live.erase(it);
}

template <class T>
int dalloc(bem_ctx* c, T** p, size_t count, std::vector<void*>* owner)
{
    void* q = nullptr;
    int rc = dalloc_bytes(c, &q, count * sizeof(T));
    *p = (T*)q;
    if (rc) return rc;
    if (owner) owner->push_back(q);
    return BEM_OK;
}

template <class T>
int dupload(bem_ctx* c, T** p, const T* src, size_t count, std::vector<void*>* owner)
{
    int rc = dalloc(c, p, count, owner);
    if (rc) return rc;
    if (count) HIPC(c, hipMemcpyAsync(*p, src, count * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return BEM_OK;
}

void free_list(bem_ctx* c, std::vector<void*>& v)
{
    for (void* p : v) pfree(c, p);
    v.clear();
}

// pinned staging buffer of at least `bytes`, free for writing (the previous
// transfer through it has completed)
int stage_get(bem_ctx* c, size_t bytes, char** out)
{
    if (c->stage_busy) {
        HIPC(c, hipEventSynchronize(c->stage_done));
        c->stage_busy = false;
    }
    if (c->stage_cap < bytes) {
        if (c->stage) (void)hipHostFree(c->stage);
        c->stage = nullptr;
        c->stage_cap = 0;
        size_t cap = std::max(bytes, (size_t)1 << 20);
        HIPC(c, hipHostMalloc((void**)&c->stage, cap, hipHostMallocDefault));
        c->stage_cap = cap;
    }
    *out = c->stage;
    return BEM_OK;
}

// Several host arrays uploaded into one device allocation with one copy
// through the pinned staging buffer.  Pieces are 256-byte aligned.
struct Packer {
    std::vector<std::pair<const void*, size_t>> parts;
    std::vector<size_t> offs;
    size_t total = 0;
    char* dbase = nullptr;
    template <class T>
    size_t add(const std::vector<T>& v)
    {
        size_t o = total;
        parts.push_back({v.data(), v.size() * sizeof(T)});
        offs.push_back(o);
        total = (o + v.size() * sizeof(T) + 255) & ~(size_t)255;
        return parts.size() - 1;
    }
    template <class T>
    T* ptr(size_t k) const { return (T*)(dbase + offs[k]); }
    int upload(bem_ctx* c, std::vector<void*>* owner)
    {
        int rc;
        void* d = nullptr;
        if ((rc = dalloc_bytes(c, &d, total ? total : 1))) return rc;
        dbase = (char*)d;
        if (owner) owner->push_back(d);
        if (!total) return BEM_OK;
        char* h = nullptr;
        if ((rc = stage_get(c, total, &h))) return rc;
        for (size_t k = 0; k < parts.size(); k++)
            if (parts[k].second) std::memcpy(h + offs[k], parts[k].first, parts[k].second);
        HIPC(c, hipMemcpyAsync(dbase, h, total, hipMemcpyHostToDevice, c->stream));
        HIPC(c, hipEventRecord(c->stage_done, c->stream));
        c->stage_busy = true;
        return BEM_OK;
    }
};

DevFreq dev_freq(const bem_ctx* c)
{
    DevFreq F;
    F.k = c->k;
    F.cc = c->cc;
    F.subst = c->d_subst;
    F.cphi = c->d_cphi;
    F.cd = c->d_cd;
    return F;
}

DevFreq32 dev_freq32(const bem_ctx* c)
{
    DevFreq32 F;
    F.k = (float)c->k;
    F.krev = c->k / (2.0 * BEM_PI);
    F.cc.re = (float)c->cc.re;
    F.cc.im = (float)c->cc.im;
    F.cphi = c->d_cphi32;
    F.cd = c->d_cd32;
    return F;
}

// a change of the operators: the kept DPhi near field no longer applies
// release a kept DPhi near field
void drop_kept(bem_ctx* c)
{
    for (auto* v : {&c->kb.dense, &c->kb.dense64, &c->kb.sing}) {
        for (void* q : *v)
            if (q) pfree(c, q);
        v->clear();
    }
    c->kb.kcc.clear();
    if (c->st_dense) {
        pfree(c, c->st_dense);
        c->st_dense = nullptr;
    }
    if (c->st_dense64) {
        pfree(c, c->st_dense64);
        c->st_dense64 = nullptr;
    }
}

void bump(bem_ctx* c)
{
    c->gen++;
    drop_kept(c);
}

// vertices a, b, c as 9 SoA rows, and the FP32 path's edge form c, a - c, b - c
void vertex_soa(int64_t n, const double* nodes9, std::vector<double>& vtx, std::vector<double>& edg)
{
    vtx.resize(9 * n);
    edg.resize(9 * n);
    for (int64_t e = 0; e < n; e++) {
        for (int d = 0; d < 9; d++) vtx[d * n + e] = nodes9[9 * e + d];
        for (int d = 0; d < 3; d++) {
            const double cv = nodes9[9 * e + 6 + d];
            edg[d * n + e] = cv;
            edg[(3 + d) * n + e] = nodes9[9 * e + d] - cv;
            edg[(6 + d) * n + e] = nodes9[9 * e + 3 + d] - cv;
        }
    }
}

// mirror images of the current mesh in the impedance plane (or none)
int build_plane(bem_ctx* c)
{
    free_list(c, c->allocs_plane);
    DevMesh& M = c->mesh;
    M.rvtx = M.redg = M.rnx = M.rny = M.rnz = nullptr;
    if (!c->have_plane || c->n <= 0) return BEM_OK;
    const int64_t n = c->n;
    std::vector<double> rnodes, vtx, edg, nrm(3 * n);
    std::vector<bemsetup::Vec3> rnormal;
    bemsetup::reflect_elements(n, c->h_nodes.data(), c->plane_pt, c->plane_nrm, rnodes, rnormal);
    vertex_soa(n, rnodes.data(), vtx, edg);
    for (int64_t e = 0; e < n; e++) {
        nrm[e] = rnormal[e].x; nrm[n + e] = rnormal[e].y; nrm[2 * n + e] = rnormal[e].z;
    }
    double *d_vtx, *d_edg, *d_nrm;
    int rc;
    if ((rc = dupload(c, &d_vtx, vtx.data(), vtx.size(), &c->allocs_plane))) return rc;
    if ((rc = dupload(c, &d_edg, edg.data(), edg.size(), &c->allocs_plane))) return rc;
    if ((rc = dupload(c, &d_nrm, nrm.data(), nrm.size(), &c->allocs_plane))) return rc;
    M.rvtx = d_vtx;
    M.redg = d_edg;
    M.rnx = d_nrm; M.rny = d_nrm + n; M.rnz = d_nrm + 2 * n;
    return BEM_OK;
}

int launch_check(bem_ctx* c, const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(c, BEM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return BEM_OK;
}

unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

// The address of `scratch` on `device` (the current device), looked up once
// per device and kept: every GPU holds its own copy at its own address.
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_addr;
hipError_t scratch_address(int device, void** out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[device] = p;
    *out = p;
    return hipSuccess;
}

// =================================================================== C ABI

extern "C" {

int bem_create(bem_ctx_t* out, int device)
{
    if (!out) return BEM_ERR_ARG;
    *out = nullptr;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) return BEM_ERR_HIP;
    if (device < 0 || device >= nd) return BEM_ERR_ARG;
    auto c = new bem_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side[0], hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side[1], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->stage_done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->gm_ev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->gm_ev[1], hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void**)&c->gm_res, 2 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&c->gm_resq, 2 * kArnMax * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        scratch_address(device, &c->scratch_dev) != hipSuccess) {
        delete c;
        return BEM_ERR_HIP;
    }
    *out = c;
    return BEM_OK;
}

int bem_destroy(bem_ctx_t c)
{
    if (!c) return BEM_ERR_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    bump(c);
    free_list(c, c->allocs_mesh);
    free_list(c, c->allocs_plane);
    free_list(c, c->allocs_part);
    free_list(c, c->allocs_piv);
    for (void* p : {(void*)c->d_alpha, (void*)c->d_beta, (void*)c->d_f, (void*)c->d_subst, (void*)c->d_cphi,
                    (void*)c->d_cd, (void*)c->d_cphi32, (void*)c->d_cd32, (void*)c->d_gm, (void*)c->d_partq,
                    (void*)c->d_xq})
        pfree(c, p);
    for (auto& kv : c->live) (void)hipFree(kv.first);
    c->live.clear();
    pool_release_all(c);
    resolve_pending(c);
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    for (hipStream_t t : c->side) { (void)hipStreamSynchronize(t); (void)hipStreamDestroy(t); }
    (void)hipEventDestroy(c->ev_fork);
    for (hipEvent_t e : c->ev_join) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->gm_ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(c->stage_done);
    if (c->stage) (void)hipHostFree(c->stage);
    if (c->gm_res) (void)hipHostFree(c->gm_res);
    if (c->gm_resq) (void)hipHostFree(c->gm_resq);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return BEM_OK;
}

const char* bem_last_error(bem_ctx_t c) { return c ? c->err.c_str() : "null context"; }

int bem_upload_mesh(bem_ctx_t c, int64_t n, const double* nodes9, const double* normal3, const double* area,
                    const double* colloc3, double hbar)
{
    if (!c || n <= 0 || !nodes9 || !normal3 || !area || !colloc3 || n > (int64_t)INT32_MAX)
        return set_err(c, BEM_ERR_ARG, "bem_upload_mesh: bad arguments");
    if (!(std::isfinite(hbar) && hbar > 0)) return set_err(c, BEM_ERR_ARG, "bem_upload_mesh: hbar must be > 0");
    // a degenerate or non-finite element would give its whole column
    // non-finite entries; the kernels are never given one
    for (int64_t e = 0; e < n; e++) {
        bool ok = std::isfinite(area[e]) && area[e] > 0;
        for (int d = 0; d < 9 && ok; d++) ok = std::isfinite(nodes9[9 * e + d]);
        for (int d = 0; d < 3 && ok; d++) ok = std::isfinite(normal3[3 * e + d]) && std::isfinite(colloc3[3 * e + d]);
        if (!ok)
            return set_err(c, BEM_ERR_ARG, "bem_upload_mesh: element " + std::to_string(e) +
                                               " is degenerate (area not > 0) or has a non-finite coordinate");
    }
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipStreamSynchronize(c->stream));
    free_list(c, c->allocs_mesh);
    bump(c);
    c->have_part = false;   // partition, boundary conditions and frequency belong to the old mesh
    c->have_bc = false;
    c->have_freq = false;
    c->n = n;
    c->hbar = hbar;
    c->h_nodes.assign(nodes9, nodes9 + 9 * n);
    c->h_normal.assign(normal3, normal3 + 3 * n);
    c->h_area.assign(area, area + n);
    c->h_colloc.assign(colloc3, colloc3 + 3 * n);
    // frequency-independent geometry of the singular and near-singular rules
    bemsetup::Elements E;
    E.n = n;
    E.nodes = c->h_nodes;
    E.normal.resize(n);
    E.colloc.resize(n);
    E.area = c->h_area;
    E.hbar = hbar;
    for (int64_t e = 0; e < n; e++) {
        E.normal[e] = bemsetup::Vec3{normal3[3 * e], normal3[3 * e + 1], normal3[3 * e + 2]};
        E.colloc[e] = bemsetup::Vec3{colloc3[3 * e], colloc3[3 * e + 1], colloc3[3 * e + 2]};
    }
    c->near = bemsetup::build_near(E);
    std::vector<double> self = bemsetup::build_self(E);
    std::vector<double> soa(7 * n);
    for (int64_t e = 0; e < n; e++) {
        soa[e] = colloc3[3 * e]; soa[n + e] = colloc3[3 * e + 1]; soa[2 * n + e] = colloc3[3 * e + 2];
        soa[3 * n + e] = normal3[3 * e]; soa[4 * n + e] = normal3[3 * e + 1]; soa[5 * n + e] = normal3[3 * e + 2];
        soa[6 * n + e] = area[e];
    }
    std::vector<double> vtx, edg;
    vertex_soa(n, nodes9, vtx, edg);
    double *d_soa, *d_nodes, *d_q, *d_vtx, *d_edg;
    int rc;
    if ((rc = dupload(c, &d_soa, soa.data(), soa.size(), &c->allocs_mesh))) return rc;
    if ((rc = dupload(c, &d_vtx, vtx.data(), vtx.size(), &c->allocs_mesh))) return rc;
    if ((rc = dupload(c, &d_edg, edg.data(), edg.size(), &c->allocs_mesh))) return rc;
    if ((rc = dupload(c, &d_nodes, c->h_nodes.data(), c->h_nodes.size(), &c->allocs_mesh))) return rc;
    if ((rc = dalloc(c, &d_q, (size_t)48 * n, &c->allocs_mesh))) return rc;
    DevMesh& M = c->mesh;
    M.n = n;
    M.px = d_soa; M.py = d_soa + n; M.pz = d_soa + 2 * n;
    M.nx = d_soa + 3 * n; M.ny = d_soa + 4 * n; M.nz = d_soa + 5 * n;
    M.area = d_soa + 6 * n;
    M.qx = d_q; M.qy = d_q + 16 * n; M.qz = d_q + 32 * n;
    M.vtx = d_vtx;
    M.edg = d_edg;
    hipLaunchKernelGGL(k_quadpoints, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, d_nodes, d_q, d_q + 16 * n,
                       d_q + 32 * n);
    if ((rc = launch_check(c, "k_quadpoints"))) return rc;
    int64_t *d_ptr, *d_off;
    int32_t *d_col, *d_nsub;
    double *d_geo, *d_self;
    if ((rc = dupload(c, &d_ptr, c->near.rowptr.data(), c->near.rowptr.size(), &c->allocs_mesh))) return rc;
    if ((rc = dupload(c, &d_col, c->near.col.data(), c->near.col.size(), &c->allocs_mesh))) return rc;
    if ((rc = dupload(c, &d_nsub, c->near.nsub.data(), c->near.nsub.size(), &c->allocs_mesh))) return rc;
    if ((rc = dupload(c, &d_off, c->near.geo_off.data(), c->near.geo_off.size(), &c->allocs_mesh))) return rc;
    if ((rc = dupload(c, &d_geo, c->near.geo.data(), c->near.geo.size(), &c->allocs_mesh))) return rc;
    if ((rc = dupload(c, &d_self, self.data(), self.size(), &c->allocs_mesh))) return rc;
    // row of every pair, and the evaluation order of the singular items
    // (self entries first, then near pairs by decreasing sub-triangle count)
    const int64_t nnz = c->near.rowptr[n];
    std::vector<int32_t> prow(nnz), order(nnz + n);
    for (int64_t i = 0; i < n; i++)
        for (int64_t t = c->near.rowptr[i]; t < c->near.rowptr[i + 1]; t++) prow[t] = (int32_t)i;
    for (int64_t i = 0; i < n; i++) order[i] = (int32_t)(nnz + i);
    for (int64_t t = 0; t < nnz; t++) order[n + t] = (int32_t)t;
    std::stable_sort(order.begin() + n, order.end(),
                     [&](int32_t a, int32_t b) { return c->near.nsub[a] > c->near.nsub[b]; });
    int32_t* d_prow;
    if ((rc = dupload(c, &d_prow, prow.data(), prow.size(), &c->allocs_mesh))) return rc;
    if ((rc = dupload(c, &c->d_sing_order, order.data(), order.size(), &c->allocs_mesh))) return rc;
    if ((rc = dalloc(c, &c->d_sing, (size_t)(nnz + n), &c->allocs_mesh))) return rc;
    if ((rc = dalloc(c, &c->d_sing2, (size_t)(nnz + n), &c->allocs_mesh))) return rc;
    c->nnz = nnz;
    M.nr_ptr = d_ptr; M.nr_row = d_prow; M.nr_col = d_col; M.nr_nsub = d_nsub; M.nr_off = d_off; M.nr_geo = d_geo;
    M.self_geo = d_self;
    M.near_lim = 4.0 * hbar;
    if ((rc = build_plane(c))) return rc;
    HIPC(c, hipStreamSynchronize(c->stream));
    return BEM_OK;
}

int bem_set_near_pairing(bem_ctx_t c, int on)
{
    if (!c) return BEM_ERR_ARG;
    c->pair_near = on != 0;
    if (!c->pair_near) bump(c);
    return BEM_OK;
}

int bem_set_near_fp64(bem_ctx_t c, int on)
{
    if (!c) return BEM_ERR_ARG;
    if (c->near64 != (on != 0)) bump(c);   // a kept near field of the other form no longer applies
    c->near64 = on != 0;
    return BEM_OK;
}

int bem_set_plane(bem_ctx_t c, const double* point3, const double* normal3)
{
    if (!c) return BEM_ERR_ARG;
    const bool on = normal3 && (normal3[0] != 0 || normal3[1] != 0 || normal3[2] != 0);
    if (on && !point3) return set_err(c, BEM_ERR_ARG, "bem_set_plane: no plane point");
    if (on && !(std::isfinite(normal3[0]) && std::isfinite(normal3[1]) && std::isfinite(normal3[2]) &&
                std::isfinite(point3[0]) && std::isfinite(point3[1]) && std::isfinite(point3[2])))
        return set_err(c, BEM_ERR_ARG, "bem_set_plane: non-finite plane");
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipStreamSynchronize(c->stream));
    bump(c);
    c->have_plane = on;
    for (int d = 0; d < 3; d++) {
        c->plane_pt[d] = on ? point3[d] : 0.0;
        c->plane_nrm[d] = on ? normal3[d] : 0.0;
    }
    int rc = build_plane(c);
    if (rc) return rc;
    HIPC(c, hipStreamSynchronize(c->stream));
    return BEM_OK;
}

int bem_set_bc(bem_ctx_t c, const double* a, const double* b, const double* f)
{
    if (!c || c->n <= 0 || !a || !b || !f) return set_err(c, BEM_ERR_ARG, "bem_set_bc: bad arguments / no mesh");
    HIPC(c, hipSetDevice(c->device));
    int64_t n = c->n;
    c->h_alpha.assign((const bem_c64*)a, (const bem_c64*)a + n);
    c->h_beta.assign((const bem_c64*)b, (const bem_c64*)b + n);
    c->h_f.assign((const bem_c64*)f, (const bem_c64*)f + n);
    for (bem_c64** p : {&c->d_alpha, &c->d_beta, &c->d_f}) {
        pfree(c, *p);
        *p = nullptr;
    }
    int rc;
    if ((rc = dupload(c, &c->d_alpha, c->h_alpha.data(), n, nullptr))) return rc;
    if ((rc = dupload(c, &c->d_beta, c->h_beta.data(), n, nullptr))) return rc;
    if ((rc = dupload(c, &c->d_f, c->h_f.data(), n, nullptr))) return rc;
    c->have_bc = true;
    bump(c);
    c->have_freq = false;
    HIPC(c, hipStreamSynchronize(c->stream));
    return BEM_OK;
}

int bem_set_frequency(bem_ctx_t c, double k, double cc_re, double cc_im)
{
    if (!c || !c->have_bc) return set_err(c, BEM_ERR_STATE, "bem_set_frequency: mesh/boundary conditions missing");
    if (!(k > 0) || !std::isfinite(k)) return set_err(c, BEM_ERR_ARG, "bem_set_frequency: k must be > 0");
    HIPC(c, hipSetDevice(c->device));
    int64_t n = c->n;
    // substitution rule and column coefficients (bem_detmath.h bem_xy notes)
    std::vector<int32_t> subst(n);
    std::vector<bem_c64> cphi(n), cd(n);
    for (int64_t j = 0; j < n; j++) {
        bem_c64 al = c->h_alpha[j], be = c->h_beta[j];
        subst[j] = std::hypot(al.re, al.im) < std::hypot(be.re, be.im);
        if (subst[j]) {
            cphi[j] = c64_div(c64(-1.0, 0.0), be);
            cd[j] = c64_div(c64_neg(al), be);
        } else {
            cphi[j] = c64_div(c64(1.0, 0.0), al);
            cd[j] = c64_div(be, al);
        }
        // a = b = 0 (or a non-finite a, b) leaves the column without finite
        // coefficients; the kernels are never given such a column
        if (!(std::isfinite(cphi[j].re) && std::isfinite(cphi[j].im) && std::isfinite(cd[j].re) &&
              std::isfinite(cd[j].im))) {
            c->have_freq = false;
            return set_err(c, BEM_ERR_ARG, "bem_set_frequency: boundary condition of element " + std::to_string(j) +
                                               " (cluster order) has a = b = 0 or a non-finite coefficient");
        }
    }
    c->k = k;
    c->cc = c64(cc_re, cc_im);
    c->h_subst = subst;
    if (c->freq_n != n) {   // (re)size for the current mesh
        for (void** p : {(void**)&c->d_subst, (void**)&c->d_cphi, (void**)&c->d_cd, (void**)&c->d_cphi32,
                         (void**)&c->d_cd32}) {
            pfree(c, *p);
            *p = nullptr;
        }
        c->freq_n = 0;
        int rc;
        if ((rc = dalloc(c, &c->d_subst, n, nullptr))) return rc;
        if ((rc = dalloc(c, &c->d_cphi, n, nullptr))) return rc;
        if ((rc = dalloc(c, &c->d_cd, n, nullptr))) return rc;
        if ((rc = dalloc(c, &c->d_cphi32, n, nullptr))) return rc;
        if ((rc = dalloc(c, &c->d_cd32, n, nullptr))) return rc;
        c->freq_n = n;
    }
    std::vector<bem_c32> cphi32(n), cd32(n);
    for (int64_t j = 0; j < n; j++) {
        cphi32[j].re = (float)cphi[j].re; cphi32[j].im = (float)cphi[j].im;
        cd32[j].re = (float)cd[j].re; cd32[j].im = (float)cd[j].im;
    }
    HIPC(c, hipMemcpyAsync(c->d_cphi32, cphi32.data(), n * sizeof(bem_c32), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(c->d_cd32, cd32.data(), n * sizeof(bem_c32), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(c->d_subst, subst.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(c->d_cphi, cphi.data(), n * sizeof(bem_c64), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(c->d_cd, cd.data(), n * sizeof(bem_c64), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    c->have_freq = true;
    bump(c);
    return BEM_OK;
}

int bem_upload_partition(bem_ctx_t c, int64_t nb, const int64_t* row0, const int64_t* rows, const int64_t* col0,
                         const int64_t* cols, const uint8_t* adm)
{
    if (!c || c->n <= 0 || nb <= 0 || !row0 || !rows || !col0 || !cols || !adm)
        return set_err(c, BEM_ERR_ARG, "bem_upload_partition: bad arguments / no mesh");
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipStreamSynchronize(c->stream));
    int64_t n = c->n;
    // every entry covered exactly once and every block inside the matrix
    std::vector<int64_t> cover(n, 0);
    for (int64_t b = 0; b < nb; b++) {
        if (rows[b] <= 0 || cols[b] <= 0 || row0[b] < 0 || col0[b] < 0 || row0[b] + rows[b] > n ||
            col0[b] + cols[b] > n || rows[b] > INT32_MAX || cols[b] > INT32_MAX)
            return set_err(c, BEM_ERR_ARG, "bem_upload_partition: block outside the matrix");
        for (int64_t i = row0[b]; i < row0[b] + rows[b]; i++) cover[i] += cols[b];
    }
    for (int64_t i = 0; i < n; i++)
        if (cover[i] != n) return set_err(c, BEM_ERR_ARG, "bem_upload_partition: blocks do not tile the matrix");
    free_list(c, c->allocs_part);
    free_list(c, c->allocs_piv);
    c->piv_ok = false;
    c->blocks.resize(nb);
    for (int64_t b = 0; b < nb; b++) c->blocks[b] = bemsetup::Block{row0[b], rows[b], col0[b], cols[b], adm[b] != 0};
    // dense storage and tiles
    c->dense_off.assign(nb, -1);
    c->tiles.clear();
    c->adm_list.clear();
    c->lr_part.assign(nb, -1);
    int64_t doff = 0, poff = 0;
    std::vector<std::vector<int64_t>> contrib(n);  // partial offsets per row, partition order
    for (int64_t b = 0; b < nb; b++) {
        const auto& B = c->blocks[b];
        if (!B.adm) {
            c->dense_off[b] = doff;
            for (int64_t tc = 0; tc < B.cols; tc += kTile)
                for (int64_t tr = 0; tr < B.rows; tr += kTile) {
                    Tile T;
                    T.dest = doff + tr + tc * B.rows;
                    T.ld = (int32_t)B.rows;
                    T.row0 = (int32_t)(B.row0 + tr);
                    T.col0 = (int32_t)(B.col0 + tc);
                    T.rows = (int32_t)std::min<int64_t>(kTile, B.rows - tr);
                    T.cols = (int32_t)std::min<int64_t>(kTile, B.cols - tc);
                    T.part = poff;
                    T.pad = 0;
                    for (int r = 0; r < T.rows; r++) contrib[T.row0 + r].push_back(poff + r);
                    poff += T.rows;
                    c->tiles.push_back(T);
                }
            doff += B.rows * B.cols;
        } else {
            c->adm_list.push_back((int32_t)b);
            c->lr_part[b] = poff;
            for (int64_t r = 0; r < B.rows; r++) contrib[B.row0 + r].push_back(poff + r);
            poff += B.rows;
        }
    }
    c->dense_total = doff;
    c->part_total = poff;
    // largest admissible blocks first (load balance of the one-block-per-
    // workgroup kernels)
    std::stable_sort(c->adm_list.begin(), c->adm_list.end(), [&](int32_t a, int32_t b) {
        return c->blocks[a].rows + c->blocks[a].cols > c->blocks[b].rows + c->blocks[b].cols;
    });
    // self / near-singular entries inside dense blocks
    c->fixes.clear();
    const auto& NL = c->near;
    for (int64_t b = 0; b < nb; b++) {
        const auto& B = c->blocks[b];
        if (B.adm) continue;
        for (int64_t i = B.row0; i < B.row0 + B.rows; i++) {
            if (i >= B.col0 && i < B.col0 + B.cols)
                c->fixes.push_back(Fix{c->dense_off[b] + (i - B.row0) + (i - B.col0) * B.rows, (int32_t)i, (int32_t)i, -1});
            for (int64_t t = NL.rowptr[i]; t < NL.rowptr[i + 1]; t++) {
                int64_t j = NL.col[t];
                if (j >= B.col0 && j < B.col0 + B.cols)
                    c->fixes.push_back(Fix{c->dense_off[b] + (i - B.row0) + (j - B.col0) * B.rows, (int32_t)i, (int32_t)j, t});
            }
        }
    }
    // admissible blocks whose row and column collocation boxes lie further
    // apart than the near limit (with a margin far above any rounding of the
    // device's distance) hold no self or near-singular pair
    std::vector<int8_t> bnear(nb, 1);
    {
        auto box = [&](int64_t s0, int64_t cnt, double lo[3], double hi[3]) {
            for (int d = 0; d < 3; d++) lo[d] = hi[d] = c->h_colloc[3 * s0 + d];
            for (int64_t e = s0; e < s0 + cnt; e++)
                for (int d = 0; d < 3; d++) {
                    lo[d] = std::min(lo[d], c->h_colloc[3 * e + d]);
                    hi[d] = std::max(hi[d], c->h_colloc[3 * e + d]);
                }
        };
        const double lim = c->mesh.near_lim * (1.0 + 1e-9);
        for (int32_t b : c->adm_list) {
            const auto& B = c->blocks[b];
            double rl[3], rh[3], cl[3], ch[3];
            box(B.row0, B.rows, rl, rh);
            box(B.col0, B.cols, cl, ch);
            double d2 = 0;
            for (int d = 0; d < 3; d++) {
                const double g = std::max({0.0, cl[d] - rh[d], rl[d] - ch[d]});
                d2 += g * g;
            }
            if (std::sqrt(d2) > lim) bnear[b] = 0;
        }
    }
    std::vector<int64_t> gptr(n + 1, 0), gidx;
    for (int64_t i = 0; i < n; i++) gptr[i + 1] = gptr[i] + (int64_t)contrib[i].size();
    gidx.reserve(gptr[n]);
    for (int64_t i = 0; i < n; i++) gidx.insert(gidx.end(), contrib[i].begin(), contrib[i].end());
    std::vector<int64_t> r0(nb), rs(nb), c0(nb), cs(nb);
    for (int64_t b = 0; b < nb; b++) { r0[b] = row0[b]; rs[b] = rows[b]; c0[b] = col0[b]; cs[b] = cols[b]; }
    int rc;
    auto* ap = &c->allocs_part;
    if ((rc = dupload(c, &c->d_row0, r0.data(), nb, ap))) return rc;
    if ((rc = dupload(c, &c->d_rows, rs.data(), nb, ap))) return rc;
    if ((rc = dupload(c, &c->d_col0, c0.data(), nb, ap))) return rc;
    if ((rc = dupload(c, &c->d_cols, cs.data(), nb, ap))) return rc;
    if ((rc = dupload(c, &c->d_bnear, bnear.data(), nb, ap))) return rc;
    if ((rc = dupload(c, &c->d_tiles, c->tiles.data(), c->tiles.size(), ap))) return rc;
    if ((rc = dupload(c, &c->d_fixes, c->fixes.data(), c->fixes.size(), ap))) return rc;
    if ((rc = dupload(c, &c->d_adm_list, c->adm_list.data(), c->adm_list.size(), ap))) return rc;
    if ((rc = dupload(c, &c->d_lr_part, c->lr_part.data(), nb, ap))) return rc;
    if ((rc = dalloc(c, &c->d_part, c->part_total, ap))) return rc;
    if ((rc = dupload(c, &c->d_gptr, gptr.data(), gptr.size(), ap))) return rc;
    if ((rc = dupload(c, &c->d_gidx, gidx.data(), gidx.size(), ap))) return rc;
    HIPC(c, hipStreamSynchronize(c->stream));
    pfree(c, c->d_partq);   // sized for the previous partition
    pfree(c, c->d_xq);
    c->d_partq = nullptr;
    c->d_xq = nullptr;
    c->partq_cap = 0;
    c->have_part = true;
    bump(c);
    return BEM_OK;
}

int bem_hmat_destroy(bem_hmat_t h)
{
    if (!h) return BEM_ERR_ARG;
    (void)hipSetDevice(h->ctx->device);
    (void)hipStreamSynchronize(h->ctx->stream);
    for (hipStream_t t : h->ctx->side) (void)hipStreamSynchronize(t);
    for (void* p : {(void*)h->dense, h->dense64, (void*)h->U, (void*)h->V, (void*)h->A, h->meta, (void*)h->d_t,
                    (void*)h->d_tq})
        pfree(h->ctx, p);
    for (auto* v : {&h->uoff, &h->voff, &h->aoff}) h->ctx->vp.give(*v);
    h->ctx->vp.give(h->rank);
    delete h;
    return BEM_OK;
}

}  // extern "C"

namespace {

// ACA over the blocks in `list` with per-block capacity `capf(b)` into freshly
// allocated scratch; returns ranks/flags/evals per partition block
struct AcaScratch {
    void *U = nullptr, *V = nullptr, *A = nullptr;
    int64_t *d_uoff = nullptr, *d_voff = nullptr, *d_aoff = nullptr;
    int32_t *d_cap = nullptr, *d_rank = nullptr, *d_flag = nullptr, *d_list = nullptr;
    int64_t* d_ev = nullptr;
    std::vector<int64_t> uoff, voff, aoff;
    int64_t tu = 0, tv = 0, ta = 0;   // allocated entries of U, V, alpha
    std::vector<void*> owned;
    bem_ctx* ctx = nullptr;
    ~AcaScratch()
    {
        if (!ctx) return;
        free_list(ctx, owned);
        ctx->vp.give(uoff); ctx->vp.give(voff); ctx->vp.give(aoff);
    }
};

// workgroup of the largest class (FP64 never forms it: kHugeBlock[0])
template <class C> constexpr int kHugeThreads = 256;
// workgroup of the middle class: FP32 two wavefronts (several lanes share the
// chip), FP64 four (its points run one after the other, so one block's pivot
// chain matters more)
template <class C> constexpr int kMidThreads = 128;
template <> constexpr int kMidThreads<bem_c64> = 256;

template <int NT, class C>
void launch_aca_nt(bool extra, bool plane, dim3 g, size_t shm, hipStream_t st, const AcaArgs& aq)
{
    // batched items (mode 1) take the impedance plane; the work-queue and
    // multi-mesh modes run without one (the callers check)
    const int mode = aq.jobs ? (aq.queue ? 4 : 3) : aq.nf == 0 ? 0 : (aq.queue ? 2 : 1);
    if (mode == 3) {
        if (extra) hipLaunchKernelGGL((k_aca<true, NT, C, false, 3>), g, dim3(NT), shm, st, aq);
        else hipLaunchKernelGGL((k_aca<false, NT, C, false, 3>), g, dim3(NT), shm, st, aq);
        return;
    }
    if (mode == 4) {
        if (extra) hipLaunchKernelGGL((k_aca<true, NT, C, false, 4>), g, dim3(NT), shm, st, aq);
        else hipLaunchKernelGGL((k_aca<false, NT, C, false, 4>), g, dim3(NT), shm, st, aq);
        return;
    }
    if (mode == 1) {
        if (extra) {
            if (plane) hipLaunchKernelGGL((k_aca<true, NT, C, true, 1>), g, dim3(NT), shm, st, aq);
            else hipLaunchKernelGGL((k_aca<true, NT, C, false, 1>), g, dim3(NT), shm, st, aq);
        } else {
            if (plane) hipLaunchKernelGGL((k_aca<false, NT, C, true, 1>), g, dim3(NT), shm, st, aq);
            else hipLaunchKernelGGL((k_aca<false, NT, C, false, 1>), g, dim3(NT), shm, st, aq);
        }
        return;
    }
    if (mode == 2) {
        if (extra) hipLaunchKernelGGL((k_aca<true, NT, C, false, 2>), g, dim3(NT), shm, st, aq);
        else hipLaunchKernelGGL((k_aca<false, NT, C, false, 2>), g, dim3(NT), shm, st, aq);
        return;
    }
    if (extra) {
        if (plane) hipLaunchKernelGGL((k_aca<true, NT, C, true, 0>), g, dim3(NT), shm, st, aq);
        else hipLaunchKernelGGL((k_aca<true, NT, C, false, 0>), g, dim3(NT), shm, st, aq);
    } else {
        if (plane) hipLaunchKernelGGL((k_aca<false, NT, C, true, 0>), g, dim3(NT), shm, st, aq);
        else hipLaunchKernelGGL((k_aca<false, NT, C, false, 0>), g, dim3(NT), shm, st, aq);
    }
}

// one size class of blocks on stream st, storage type C
template <class C>
void launch_aca(int q, bool extra, unsigned nblocks, size_t shm, hipStream_t st, const AcaArgs& aq)
{
    const dim3 g(nblocks);
    if constexpr (sizeof(C) == sizeof(bem_c64)) {
        if (aq.obs) {   // field operators (FP64, one workgroup per block)
            const int nt = q == 0 ? kHugeThreads<C> : q == 1 ? kMidThreads<C> : 64;
            if (nt == 64) {
                if (extra) hipLaunchKernelGGL((k_aca<true, 64, C, false, 0, true>), g, dim3(64), shm, st, aq);
                else hipLaunchKernelGGL((k_aca<false, 64, C, false, 0, true>), g, dim3(64), shm, st, aq);
            } else {
                if (extra) hipLaunchKernelGGL((k_aca<true, 256, C, false, 0, true>), g, dim3(256), shm, st, aq);
                else hipLaunchKernelGGL((k_aca<false, 256, C, false, 0, true>), g, dim3(256), shm, st, aq);
            }
            return;
        }
    }
    const bool plane = aq.M.rvtx != nullptr;
    if (q == 0) launch_aca_nt<kHugeThreads<C>, C>(extra, plane, g, shm, st, aq);
    else if (q == 1) launch_aca_nt<kMidThreads<C>, C>(extra, plane, g, shm, st, aq);
    else launch_aca_nt<64, C>(extra, plane, g, shm, st, aq);
}

// A block partition other than the context's (the field's observation x
// boundary partition): host extents, device extent arrays, observation points
struct PartView {
    const std::vector<bemsetup::Block>* blocks;
    const int64_t *row0, *rows, *col0, *cols;
    const double* obs;
    int64_t nobs;
};

// dynamic LDS of aca_block for rank capacity cap and storage size esz
size_t aca_shm_bytes(int cap, size_t esz)
{
    const size_t cbytes = ((size_t)cap * esz + 15) & ~(size_t)15;
    return 2 * cbytes + (size_t)(2 * cap + 2) * sizeof(bem_c64);
}

// The set-up half of run_aca: scratch at the given capacities, offsets,
// the listed blocks ordered by size class (cls, largest first within each),
// the results array, and the kernel arguments (blist = the ordered list)
int aca_prepare(bem_ctx* c, int kind, int prec, double tol, int64_t maxrank, uint64_t seed,
                const std::vector<int32_t>& list, const std::vector<int32_t>& cap, const int64_t* d_pivoff,
                const int32_t* d_piv, AcaScratch& S, AcaArgs& a, std::vector<int32_t> (&cls)[3], size_t& shm,
                const PartView* pv = nullptr)
{
    const std::vector<bemsetup::Block>& blocks = pv ? *pv->blocks : c->blocks;
    int64_t nb = (int64_t)blocks.size();
    S.ctx = c;
    S.uoff = c->vp.take<int64_t>(nb, 0); S.voff = c->vp.take<int64_t>(nb, 0); S.aoff = c->vp.take<int64_t>(nb, 0);
    int64_t tu = 0, tv = 0, ta = 0;
    int maxcap = 1;
    for (int32_t b : list) {
        const auto& B = blocks[b];
        S.uoff[b] = tu; S.voff[b] = tv; S.aoff[b] = ta;
        tu += B.rows * cap[b]; tv += B.cols * cap[b]; ta += cap[b];
        maxcap = std::max(maxcap, cap[b]);
    }
    int rc;
    const size_t esz = prec == BEM_PREC_FP32 ? sizeof(bem_c32) : sizeof(bem_c64);
    if ((rc = dalloc_bytes(c, &S.U, tu * esz))) return rc;
    S.owned.push_back(S.U);
    if ((rc = dalloc_bytes(c, &S.V, tv * esz))) return rc;
    S.owned.push_back(S.V);
    if ((rc = dalloc_bytes(c, &S.A, ta * esz))) return rc;
    S.owned.push_back(S.A);
    S.tu = tu; S.tv = tv; S.ta = ta;
    // size classes: m + n <= kSmallBlock one wavefront, <= kHugeBlock
    // kMidThreads, larger kHugeThreads (measured: with several lanes sharing
    // the chip, smaller workgroups than one block's pivot chain alone would
    // want give the higher rate).  The huge and large classes run on their
    // own streams so
    // that their long tails overlap the small blocks (and the caller's dense
    // tiles); within each class the largest blocks start first.
    for (auto& v : cls) v.clear();
    for (int32_t b : list) {
        int64_t s2 = blocks[b].rows + blocks[b].cols;
        const int pi = prec == BEM_PREC_FP32 ? 1 : 0;
        cls[s2 > c->cls_huge[pi] ? 0 : (s2 > c->cls_small[pi] ? 1 : 2)].push_back(b);
    }
    std::vector<int32_t> ordered;
    for (auto& v : cls) ordered.insert(ordered.end(), v.begin(), v.end());
    // offsets, capacities and the block list in one upload; ranks, flags and
    // evaluation counts in one allocation written by the kernel for every
    // listed block (the host reads listed blocks only)
    Packer pk;
    size_t k_uoff = pk.add(S.uoff), k_voff = pk.add(S.voff), k_aoff = pk.add(S.aoff);
    size_t k_cap = pk.add(cap), k_list = pk.add(ordered);
    if ((rc = pk.upload(c, &S.owned))) return rc;
    S.d_uoff = pk.ptr<int64_t>(k_uoff); S.d_voff = pk.ptr<int64_t>(k_voff); S.d_aoff = pk.ptr<int64_t>(k_aoff);
    S.d_cap = pk.ptr<int32_t>(k_cap); S.d_list = pk.ptr<int32_t>(k_list);
    void* res = nullptr;
    if ((rc = dalloc_bytes(c, &res, (size_t)nb * 16))) return rc;
    S.owned.push_back(res);
    S.d_ev = (int64_t*)res;
    S.d_rank = (int32_t*)((char*)res + (size_t)nb * 8);
    S.d_flag = S.d_rank + nb;
    a = AcaArgs();
    a.M = c->mesh;
    a.F = dev_freq(c);
    a.F32 = dev_freq32(c);
    a.kind = kind;
    a.blist = S.d_list;
    if (pv) {
        a.row0 = pv->row0; a.rows = pv->rows; a.col0 = pv->col0; a.cols = pv->cols;
        a.obs = pv->obs; a.nobs = pv->nobs;
    } else {
        a.row0 = c->d_row0; a.rows = c->d_rows; a.col0 = c->d_col0; a.cols = c->d_cols;
    }
    a.uoff = S.d_uoff; a.voff = S.d_voff; a.aoff = S.d_aoff; a.cap = S.d_cap;
    a.U = S.U; a.V = S.V; a.A = S.A;
    a.rank_out = S.d_rank; a.flag_out = S.d_flag; a.evals_out = S.d_ev;
    a.tol = tol; a.maxrank = maxrank; a.seed = seed;
    a.piv_off = d_pivoff; a.piv = d_piv;
    a.sing = c->sing_cur ? c->sing_cur : c->d_sing; a.nnz = c->nnz;
    a.bnear = pv ? nullptr : c->d_bnear;
    a.nf = 0; a.nbk = nb; a.kf = nullptr; a.ccf = nullptr; a.singf = nullptr;
    a.Uf = a.Vf = a.Af = nullptr;
    a.queue = nullptr; a.nitems = 0;
    shm = aca_shm_bytes(maxcap, esz);
    return BEM_OK;
}

// The result half of run_aca: ranks, flags and evaluation counts of the
// nb blocks of S's partition (after the launches, on c's stream)
int aca_collect(bem_ctx* c, const AcaScratch& S, int64_t nb, std::vector<int32_t>& rank, std::vector<int32_t>& flag,
                std::vector<int64_t>& ev)
{
    int rc;
    rank = c->vp.take<int32_t>(nb, 0); flag = c->vp.take<int32_t>(nb, 0); ev = c->vp.take<int64_t>(nb, 0);
    char* h = nullptr;
    if ((rc = stage_get(c, (size_t)nb * 16, &h))) return rc;
    HIPC(c, hipMemcpyAsync(h, S.d_ev, (size_t)nb * 16, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    std::memcpy(ev.data(), h, (size_t)nb * 8);
    std::memcpy(rank.data(), h + (size_t)nb * 8, (size_t)nb * 4);
    std::memcpy(flag.data(), h + (size_t)nb * 12, (size_t)nb * 4);
    return BEM_OK;
}

int run_aca(bem_ctx* c, int kind, int prec, double tol, int64_t maxrank, uint64_t seed, const std::vector<int32_t>& list,
            const std::vector<int32_t>& cap, const int64_t* d_pivoff, const int32_t* d_piv,
            const std::function<int()>& main_work, AcaScratch& S,
            std::vector<int32_t>& rank, std::vector<int32_t>& flag, std::vector<int64_t>& ev,
            const PartView* pv = nullptr)
{
    const int64_t nb = (int64_t)(pv ? pv->blocks->size() : c->blocks.size());
    AcaArgs a;
    std::vector<int32_t> cls[3];
    size_t shm = 0;
    int rc;
    if ((rc = aca_prepare(c, kind, prec, tol, maxrank, seed, list, cap, d_pivoff, d_piv, S, a, cls, shm, pv)))
        return rc;
    HIPC(c, hipEventRecord(c->ev_fork, c->stream));
    size_t off = 0;
    for (int q = 0; q < 3; q++) {
        AcaArgs aq = a;
        aq.blist = S.d_list + off;
        off += cls[q].size();
        if (q == 2 && (rc = main_work())) return rc;   // e.g. the dense tiles, overlapping the large classes
        if (cls[q].empty()) continue;
        hipStream_t st = q < 2 ? c->side[q] : c->stream;
        if (q < 2) HIPC(c, hipStreamWaitEvent(st, c->ev_fork, 0));
        Timer t(c, 2, st);
        bool x = kind == BEM_KIND_DPHI;
        if (prec == BEM_PREC_FP32) launch_aca<bem_c32>(q, x, (unsigned)cls[q].size(), shm, st, aq);
        else launch_aca<bem_c64>(q, x, (unsigned)cls[q].size(), shm, st, aq);
        if ((rc = launch_check(c, "k_aca"))) return rc;
    }
    for (int q = 0; q < 2; q++) {
        HIPC(c, hipEventRecord(c->ev_join[q], c->side[q]));
        HIPC(c, hipStreamWaitEvent(c->stream, c->ev_join[q], 0));
    }
    return aca_collect(c, S, nb, rank, flag, ev);
}

// Packs the ACA results of h's assembly (first pass S1, re-run of the
// overflowing blocks S2, final ranks `rank`) into h's storage -- or keeps the
// first pass's scratch as the storage when nothing overflowed and no block
// has rank 0 -- builds the MVM task lists, checks the near field for
// non-finite entries and fills h->stats except the timings ([4], [6], [7]).
int finish_hmat(bem_ctx* c, bem_hmat* h, bool f32, AcaScratch& S1, AcaScratch& S2, std::vector<int32_t>& rank,
                std::vector<int64_t>& evs, std::vector<int32_t>& over, std::vector<int32_t>& keep, hipEvent_t ev_done)
{
    const int64_t nb = (int64_t)c->blocks.size();
    int rc;
    // packed storage at the final ranks
    h->uoff = c->vp.take<int64_t>(nb, 0); h->voff = c->vp.take<int64_t>(nb, 0);
    h->aoff = c->vp.take<int64_t>(nb, 0); h->rank = c->vp.take<int32_t>(nb, -1);
    int64_t tu = 0, tv = 0, ta = 0;
    double maxr = 0, sumr = 0;
    int64_t evaluated = 0;
    for (int64_t b = 0; b < nb; b++) {
        const auto& B = c->blocks[b];
        if (!B.adm) { evaluated += B.rows * B.cols; continue; }
        int32_t r = rank[b] > 0 ? rank[b] : 1;
        h->rank[b] = r;
        h->uoff[b] = tu; h->voff[b] = tv; h->aoff[b] = ta;
        tu += B.rows * r; tv += B.cols * r; ta += r;
        maxr = std::max(maxr, (double)r);
        sumr += r;
        evaluated += evs[b];
    }
    // When no block overflowed its first capacity and every block has rank >= 1,
    // the first pass's scratch is the storage: its layout per block is the
    // packed one (rank rows of U, then of V, then alpha), with the unused
    // capacity left between blocks.  Otherwise the rank rows are copied into
    // packed storage (k_compact; a rank-0 block becomes a zero rank-1 block).
    bool direct = over.empty();
    for (int32_t b : keep)
        if (rank[b] < 1) direct = false;
    if (direct) {
        for (int32_t b : c->adm_list) {
            h->uoff[b] = S1.uoff[b];
            h->voff[b] = S1.voff[b];
            h->aoff[b] = S1.aoff[b];
        }
        h->U = S1.U; h->V = S1.V; h->A = S1.A;
        auto& o = S1.owned;
        o.erase(std::remove_if(o.begin(), o.end(), [&](void* q) { return q == h->U || q == h->V || q == h->A; }),
                o.end());
    } else {
        if ((rc = dalloc_bytes(c, &h->U, (size_t)tu * h->esz))) return rc;
        if ((rc = dalloc_bytes(c, &h->V, (size_t)tv * h->esz))) return rc;
        if ((rc = dalloc_bytes(c, &h->A, (size_t)ta * h->esz))) return rc;
    }
    // offsets, ranks, MVM task lists and the packing lists: one upload
    std::vector<int32_t>& rk = rank;   // first-pass ranks with the re-run ranks merged in
    std::vector<void*> own2;
    // (block, l) tasks of the first MVM phase, (block, row chunk) tasks of
    // the second; sized first, then filled by index
    std::vector<int64_t> toff = c->vp.take<int64_t>(nb, 0);
    int64_t tot = 0, ntt = 0, nchunks = 0;
    for (int32_t b : c->adm_list) {
        tot += h->rank[b];
        ntt += (h->rank[b] + kLrRows - 1) / kLrRows;
        nchunks += (c->blocks[b].rows + kWG - 1) / kWG;
    }
    std::vector<int32_t> tt = c->vp.take<int32_t>((size_t)2 * ntt, 0), ut = c->vp.take<int32_t>((size_t)2 * nchunks, 0);
    {
        int64_t pt = 0, pu = 0, acc = 0;
        for (int32_t b : c->adm_list) {
            const int r = h->rank[b];
            toff[b] = acc;
            acc += r;
            h->maxrank = std::max(h->maxrank, r);
            for (int l = 0; l < r; l += kLrRows, pt++) { tt[2 * pt] = b; tt[2 * pt + 1] = l; }
            const int64_t nq = (c->blocks[b].rows + kWG - 1) / kWG;
            for (int64_t q = 0; q < nq; q++, pu++) { ut[2 * pu] = b; ut[2 * pu + 1] = (int32_t)q; }
        }
    }
    h->nttask = ntt;
    h->t_total = tot;
    h->nutask = nchunks;
    Packer pk;
    size_t k_uoff = pk.add(h->uoff), k_voff = pk.add(h->voff), k_aoff = pk.add(h->aoff);
    size_t k_rank = pk.add(h->rank), k_srank = pk.add(rk), k_tt = pk.add(tt), k_ut = pk.add(ut);
    size_t k_toff = pk.add(toff), k_keep = pk.add(keep), k_over = pk.add(over);
    rc = pk.upload(c, nullptr);
    h->meta = pk.dbase;   // owned by h from here on, also on failure
    if (rc) return rc;
    h->d_uoff = pk.ptr<int64_t>(k_uoff); h->d_voff = pk.ptr<int64_t>(k_voff); h->d_aoff = pk.ptr<int64_t>(k_aoff);
    h->d_rank = pk.ptr<int32_t>(k_rank);
    int32_t* d_srank = pk.ptr<int32_t>(k_srank);
    h->d_ttask = pk.ptr<int32_t>(k_tt); h->d_utask = pk.ptr<int32_t>(k_ut); h->d_toff = pk.ptr<int64_t>(k_toff);
    if ((rc = dalloc(c, &h->d_t, tot, nullptr))) return rc;
    auto compact = [&](const std::vector<int32_t>& lst, const int32_t* d_lst, const AcaScratch& Sx) -> int {
        if (lst.empty()) return BEM_OK;
        dim3 g((unsigned)lst.size());
        if (f32)
            hipLaunchKernelGGL(k_compact<bem_c32>, g, dim3(kWG), 0, c->stream, d_lst, c->d_rows, c->d_cols,
                               Sx.d_uoff, Sx.d_voff, Sx.d_aoff, (const bem_c32*)Sx.U, (const bem_c32*)Sx.V,
                               (const bem_c32*)Sx.A, d_srank, h->d_uoff, h->d_voff, h->d_aoff, (bem_c32*)h->U,
                               (bem_c32*)h->V, (bem_c32*)h->A);
        else
            hipLaunchKernelGGL(k_compact<bem_c64>, g, dim3(kWG), 0, c->stream, d_lst, c->d_rows, c->d_cols,
                               Sx.d_uoff, Sx.d_voff, Sx.d_aoff, (const bem_c64*)Sx.U, (const bem_c64*)Sx.V,
                               (const bem_c64*)Sx.A, d_srank, h->d_uoff, h->d_voff, h->d_aoff, (bem_c64*)h->U,
                               (bem_c64*)h->V, (bem_c64*)h->A);
        return launch_check(c, "k_compact");
    };
    if (!direct && (rc = compact(keep, pk.ptr<int32_t>(k_keep), S1))) return rc;
    if (!direct && (rc = compact(over, pk.ptr<int32_t>(k_over), S2))) return rc;
    if (ev_done) HIPC(c, hipEventRecord(ev_done, c->stream));
    // NaN check of the dense part (the reference reports NaN blocks,
    // hmatrix.cpp:812-820)
    int32_t* d_flag;
    if ((rc = dalloc(c, &d_flag, 1, &own2))) return rc;
    HIPC(c, hipMemsetAsync(d_flag, 0, sizeof(int32_t), c->stream));
    if (c->dense_total) {
        if (f32)
            hipLaunchKernelGGL(k_has_nonfinite<bem_c32>, dim3(nblk(c->dense_total, 256)), dim3(256), 0, c->stream,
                               c->dense_total, (const bem_c32*)h->dense, d_flag);
        else
            hipLaunchKernelGGL(k_has_nonfinite<bem_c64>, dim3(nblk(c->dense_total, 256)), dim3(256), 0, c->stream,
                               c->dense_total, (const bem_c64*)h->dense, d_flag);
    }
    int32_t hflag = 0;
    HIPC(c, hipMemcpyAsync(&hflag, d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    free_list(c, own2);
    if (hflag) return set_err(c, BEM_ERR_NAN, "bem_assemble: non-finite entry in the near field");
    size_t nadm = c->adm_list.size();
    h->stats[0] = (double)evaluated;
    h->stats[1] = maxr;
    h->stats[2] = nadm ? sumr / (double)nadm : 0;
    h->stats[3] = (double)((c->dense_total + tu + tv + ta) * (int64_t)h->esz);
    h->stats[5] = (double)over.size();
    h->stats[8] = (double)((c->dense_total + (direct ? S1.tu + S1.tv + S1.ta : tu + tv + ta)) * (int64_t)h->esz +
                           (h->dense64 ? c->dense_total * (int64_t)sizeof(bem_c64) : 0));
    for (auto* v : {&tt, &ut}) c->vp.give(*v);
    c->vp.give(toff);
    return BEM_OK;
}

// one frequency of a batched assembly
struct FreqJob {
    double k;
    bem_c64 cc;
    const bem_c64* sing;
};

// run_aca for several frequencies at once: one launch per size class over
// the flattened (frequency, block) items, same-size blocks of all
// frequencies next to each other, largest first; persistent workgroups on a
// queue counter when c->aca_queue.  Frequency f's factors go to S[f]
// (identical per-block offsets) and its ranks / flags / evaluation counts
// to rank[f] / flag[f] / ev[f].  The offset arrays live in S[0].
int run_aca_batch(bem_ctx* c, int kind, int prec, double tol, int64_t maxrank, uint64_t seed,
                  const std::vector<FreqJob>& jobs, const std::vector<int32_t>& list, const std::vector<int32_t>& cap,
                  const int64_t* d_pivoff, const int32_t* d_piv, const std::function<int()>& main_work,
                  std::vector<AcaScratch>& S, std::vector<std::vector<int32_t>>& rank,
                  std::vector<std::vector<int32_t>>& flag, std::vector<std::vector<int64_t>>& ev)
{
    const int64_t nb = (int64_t)c->blocks.size();
    const int nf = (int)jobs.size();
    const size_t esz = prec == BEM_PREC_FP32 ? sizeof(bem_c32) : sizeof(bem_c64);
    int rc;
    std::vector<int64_t> uoff(nb, 0), voff(nb, 0), aoff(nb, 0);
    int64_t tu = 0, tv = 0, ta = 0;
    int maxcap = 1;
    for (int32_t b : list) {
        const auto& B = c->blocks[b];
        uoff[b] = tu; voff[b] = tv; aoff[b] = ta;
        tu += B.rows * cap[b]; tv += B.cols * cap[b]; ta += cap[b];
        maxcap = std::max(maxcap, cap[b]);
    }
    std::vector<uint64_t> Up(nf), Vp(nf), Ap(nf), Sp(nf);
    std::vector<double> kf(nf);
    std::vector<bem_c64> ccf(nf);
    for (int f = 0; f < nf; f++) {
        AcaScratch& s = S[f];
        s.ctx = c;
        s.uoff = c->vp.take<int64_t>(nb, 0); s.voff = c->vp.take<int64_t>(nb, 0); s.aoff = c->vp.take<int64_t>(nb, 0);
        std::copy(uoff.begin(), uoff.end(), s.uoff.begin());
        std::copy(voff.begin(), voff.end(), s.voff.begin());
        std::copy(aoff.begin(), aoff.end(), s.aoff.begin());
        if ((rc = dalloc_bytes(c, &s.U, tu * esz))) return rc;
        s.owned.push_back(s.U);
        if ((rc = dalloc_bytes(c, &s.V, tv * esz))) return rc;
        s.owned.push_back(s.V);
        if ((rc = dalloc_bytes(c, &s.A, ta * esz))) return rc;
        s.owned.push_back(s.A);
        s.tu = tu; s.tv = tv; s.ta = ta;
        Up[f] = (uint64_t)s.U; Vp[f] = (uint64_t)s.V; Ap[f] = (uint64_t)s.A; Sp[f] = (uint64_t)jobs[f].sing;
        kf[f] = jobs[f].k;
        ccf[f] = jobs[f].cc;
    }
    const int pi = prec == BEM_PREC_FP32 ? 1 : 0;
    std::vector<int32_t> cls[3];
    for (int32_t b : list) {
        int64_t s2 = c->blocks[b].rows + c->blocks[b].cols;
        cls[s2 > c->cls_huge[pi] ? 0 : (s2 > c->cls_small[pi] ? 1 : 2)].push_back(b);
    }
    std::vector<int32_t> items;
    size_t ncls[3];
    for (int q = 0; q < 3; q++) {
        for (int32_t b : cls[q])
            for (int f = 0; f < nf; f++) items.push_back((int32_t)(f * nb + b));
        ncls[q] = cls[q].size() * nf;
    }
    if ((int64_t)nf * nb >= ((int64_t)1 << 31)) return set_err(c, BEM_ERR_ARG, "bem_assemble_batch: too many items");
    Packer pk;
    size_t k_uoff = pk.add(uoff), k_voff = pk.add(voff), k_aoff = pk.add(aoff), k_cap = pk.add(cap);
    size_t k_items = pk.add(items), k_kf = pk.add(kf), k_ccf = pk.add(ccf), k_sp = pk.add(Sp);
    size_t k_up = pk.add(Up), k_vp = pk.add(Vp), k_ap = pk.add(Ap);
    if ((rc = pk.upload(c, &S[0].owned))) return rc;
    for (int f = 0; f < nf; f++) {   // the shared offsets (owned by S[0])
        S[f].d_uoff = pk.ptr<int64_t>(k_uoff); S[f].d_voff = pk.ptr<int64_t>(k_voff);
        S[f].d_aoff = pk.ptr<int64_t>(k_aoff); S[f].d_cap = pk.ptr<int32_t>(k_cap);
    }
    void* res = nullptr;
    if ((rc = dalloc_bytes(c, &res, (size_t)nf * nb * 16 + 256))) return rc;
    S[0].owned.push_back(res);
    int64_t* d_ev = (int64_t*)res;
    int32_t* d_rank = (int32_t*)((char*)res + (size_t)nf * nb * 8);
    int32_t* d_flag = d_rank + (size_t)nf * nb;
    int* d_queue = (int*)((char*)res + (size_t)nf * nb * 16);
    // the work queue's persistent form is not built with the impedance plane
    const bool useq = c->aca_queue && !c->mesh.rvtx;
    if (useq) HIPC(c, hipMemsetAsync(d_queue, 0, 3 * sizeof(int), c->stream));
    AcaArgs a;
    a.M = c->mesh;
    a.F = dev_freq(c);
    a.F32 = dev_freq32(c);
    a.kind = kind;
    a.row0 = c->d_row0; a.rows = c->d_rows; a.col0 = c->d_col0; a.cols = c->d_cols;
    a.uoff = S[0].d_uoff; a.voff = S[0].d_voff; a.aoff = S[0].d_aoff; a.cap = S[0].d_cap;
    a.U = a.V = a.A = nullptr;
    a.rank_out = d_rank; a.flag_out = d_flag; a.evals_out = d_ev;
    a.tol = tol; a.maxrank = maxrank; a.seed = seed;
    a.piv_off = d_pivoff; a.piv = d_piv;
    a.sing = nullptr; a.nnz = c->nnz;
    a.bnear = c->d_bnear;
    a.nf = nf; a.nbk = nb;
    a.kf = pk.ptr<double>(k_kf); a.ccf = pk.ptr<bem_c64>(k_ccf);
    a.singf = pk.ptr<const bem_c64*>(k_sp);
    a.Uf = pk.ptr<void*>(k_up); a.Vf = pk.ptr<void*>(k_vp); a.Af = pk.ptr<void*>(k_ap);
    a.queue = nullptr; a.nitems = 0;
    size_t shm = aca_shm_bytes(maxcap, esz);
    HIPC(c, hipEventRecord(c->ev_fork, c->stream));
    size_t off = 0;
    // resident workgroups of each class on the chip (256 CUs x 4 SIMDs x
    // 4 waves per SIMD, divided by the class's wavefronts per workgroup)
    const int nt[3] = {pi ? kHugeThreads<bem_c32> : kHugeThreads<bem_c64>,
                       pi ? kMidThreads<bem_c32> : kMidThreads<bem_c64>, 64};
    for (int q = 0; q < 3; q++) {
        AcaArgs aq = a;
        aq.blist = pk.ptr<int32_t>(k_items) + off;
        off += ncls[q];
        if (q == 2 && (rc = main_work())) return rc;
        if (!ncls[q]) continue;
        hipStream_t st = q < 2 ? c->side[q] : c->stream;
        if (q < 2) HIPC(c, hipStreamWaitEvent(st, c->ev_fork, 0));
        unsigned grid = (unsigned)ncls[q];
        if (useq) {
            aq.queue = d_queue + q;
            aq.nitems = (int)ncls[q];
            grid = (unsigned)std::min<size_t>(ncls[q], (size_t)(256 * 4 * 4 * 64 / nt[q]));
        }
        Timer t(c, 2, st);
        bool x = kind == BEM_KIND_DPHI;
        if (prec == BEM_PREC_FP32) launch_aca<bem_c32>(q, x, grid, shm, st, aq);
        else launch_aca<bem_c64>(q, x, grid, shm, st, aq);
        if ((rc = launch_check(c, "k_aca (batch)"))) return rc;
    }
    for (int q = 0; q < 2; q++) {
        HIPC(c, hipEventRecord(c->ev_join[q], c->side[q]));
        HIPC(c, hipStreamWaitEvent(c->stream, c->ev_join[q], 0));
    }
    std::vector<char> h((size_t)nf * nb * 16);
    HIPC(c, hipMemcpyAsync(h.data(), res, h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    rank.assign(nf, {}); flag.assign(nf, {}); ev.assign(nf, {});
    for (int f = 0; f < nf; f++) {
        rank[f] = c->vp.take<int32_t>(nb, 0); flag[f] = c->vp.take<int32_t>(nb, 0); ev[f] = c->vp.take<int64_t>(nb, 0);
        std::memcpy(ev[f].data(), h.data() + (size_t)f * nb * 8, (size_t)nb * 8);
        std::memcpy(rank[f].data(), h.data() + (size_t)nf * nb * 8 + (size_t)f * nb * 4, (size_t)nb * 4);
        std::memcpy(flag[f].data(), h.data() + (size_t)nf * nb * 12 + (size_t)f * nb * 4, (size_t)nb * 4);
    }
    return BEM_OK;
}

// the DPhi variant's pivot lists of the partition (getNormalFilteredPivotIndices
// for every admissible block), built on the host once per mesh, partition
// and seed and kept on the device
int ensure_pivot_lists(bem_ctx* c, uint64_t seed)
{
    if (c->piv_ok && c->piv_seed == seed) return BEM_OK;
    int rc;
    free_list(c, c->allocs_piv);
    c->piv_ok = false;
    bemsetup::Elements E;
    E.n = c->n;
    E.normal.resize(c->n);
    for (int64_t e = 0; e < c->n; e++)
        E.normal[e] = bemsetup::Vec3{c->h_normal[3 * e], c->h_normal[3 * e + 1], c->h_normal[3 * e + 2]};
    std::vector<int64_t> pivoff;
    std::vector<int32_t> piv;
    bemsetup::build_pivot_lists(E, c->blocks, seed, pivoff, piv);
    if ((rc = dupload(c, &c->d_pivoff, pivoff.data(), pivoff.size(), &c->allocs_piv))) return rc;
    if ((rc = dupload(c, &c->d_piv, piv.data(), piv.size(), &c->allocs_piv))) return rc;
    c->piv_ok = true;
    c->piv_seed = seed;
    return BEM_OK;
}

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

int bem_assemble(bem_ctx_t c, int kind, double tol, int64_t maxrank, int precision, uint64_t seed, bem_hmat_t* out)
{
    if (!c || !out) return set_err(c, BEM_ERR_ARG, "bem_assemble: null argument");
    *out = nullptr;
    if (!c->have_part || !c->have_freq) return set_err(c, BEM_ERR_STATE, "bem_assemble: partition or frequency missing");
    if (kind != BEM_KIND_PHI && kind != BEM_KIND_DPHI) return set_err(c, BEM_ERR_ARG, "bem_assemble: bad kind");
    if (precision != BEM_PREC_FP64 && precision != BEM_PREC_FP32)
        return set_err(c, BEM_ERR_ARG, "bem_assemble: precision must be BEM_PREC_FP64 or BEM_PREC_FP32");
    const bool f32 = precision == BEM_PREC_FP32;
    if (!(tol >= 0)) return set_err(c, BEM_ERR_ARG, "bem_assemble: negative tolerance");
    HIPC(c, hipSetDevice(c->device));
    double t0 = now_ms();
    auto h = new bem_hmat();
    h->ctx = c;
    h->kind = kind;
    h->prec = precision;
    h->esz = f32 ? sizeof(bem_c32) : sizeof(bem_c64);
    int64_t nb = (int64_t)c->blocks.size();
    int rc;
    // near-field pairing: a DPhi assembly takes the near field the Phi
    // assembly of the same state formed; a Phi assembly forms both
    // FP32 operator with its near field formed in FP64 (kept for the
    // refinement residual, rounded for the FP32 copy)
    const bool n64 = f32 && c->near64;
    const bool adopt = kind == BEM_KIND_DPHI && c->st_dense && c->st_gen == c->gen && c->st_prec == precision &&
                       c->st_near64 == n64;
    const bool pairing = c->pair_near && kind == BEM_KIND_PHI;
    void *dense2 = nullptr, *dense2_64 = nullptr;
    auto fail = [&](int code) {
        c->sing_cur = nullptr;
        if (dense2) pfree(c, dense2);
        if (dense2_64) pfree(c, dense2_64);
        bem_hmat_destroy(h);
        return code;
    };
    // dense near field
    if (adopt) {
        h->dense = c->st_dense;
        h->dense64 = c->st_dense64;
        c->st_dense = nullptr;
        c->st_dense64 = nullptr;
    } else {
        drop_kept(c);   // kept for another state or precision
        if ((rc = dalloc_bytes(c, &h->dense, (size_t)c->dense_total * h->esz))) return fail(rc);
        if (n64 && (rc = dalloc_bytes(c, &h->dense64, (size_t)c->dense_total * sizeof(bem_c64)))) return fail(rc);
    }
    if (pairing && (rc = dalloc_bytes(c, &dense2, (size_t)c->dense_total * h->esz))) return fail(rc);
    if (pairing && n64 && (rc = dalloc_bytes(c, &dense2_64, (size_t)c->dense_total * sizeof(bem_c64))))
        return fail(rc);
    c->sing_cur = adopt ? c->d_sing2 : c->d_sing;
    // FP64 near-field targets: the operator's own storage (FP64 path) or its
    // kept FP64 form (FP32 with near64)
    bem_c64* own64 = (bem_c64*)(n64 ? h->dense64 : h->dense);
    bem_c64* pair64 = (bem_c64*)(n64 ? dense2_64 : dense2);
    bem_c64* outPhi = kind == BEM_KIND_PHI ? own64 : nullptr;
    bem_c64* outDPhi = kind == BEM_KIND_DPHI ? own64 : pair64;
    DevFreq F = dev_freq(c);
    DevFreq32 F32 = dev_freq32(c);
    hipEvent_t ev0, ev1, ev2;
    HIPC(c, hipEventCreate(&ev0)); HIPC(c, hipEventCreate(&ev1)); HIPC(c, hipEventCreate(&ev2));
    HIPC(c, hipEventRecord(ev0, c->stream));
    if (!adopt) {
        Timer t(c, 1);
        int64_t nitems = c->nnz + c->n;
        hipLaunchKernelGGL(k_singular, dim3(nblk(nitems, kWG)), dim3(kWG), 0, c->stream, c->mesh, F, kind,
                           c->d_sing_order, nitems, c->nnz, c->d_sing, pairing ? c->d_sing2 : nullptr);
        if ((rc = launch_check(c, "k_singular"))) return fail(rc);
    }
    // dense near-field tiles + scatter of their singular entries: enqueued on
    // the main stream while the large ACA classes run on the side streams
    auto dense_work = [&]() -> int {
        int rc2;
        if (adopt) return BEM_OK;
        if (!c->tiles.empty()) {
            Timer t(c, 0);
            const dim3 g((unsigned)c->tiles.size());
            const bool plane = c->mesh.rvtx != nullptr;
            if (f32 && !n64)
                hipLaunchKernelGGL(plane ? k_dense32<true> : k_dense32<false>, g, dim3(kWG), 0, c->stream, c->mesh,
                                   F32, c->d_subst, kind, c->d_tiles, (bem_c32*)h->dense, (bem_c32*)dense2);
            else
                hipLaunchKernelGGL(plane ? k_dense<true> : k_dense<false>, g, dim3(kWG), 0, c->stream, c->mesh, F,
                                   c->d_tiles, outPhi, outDPhi);
            if ((rc2 = launch_check(c, "k_dense"))) return rc2;
        }
        if (!c->fixes.empty()) {
            dim3 g(nblk((int64_t)c->fixes.size(), kWG));
            for (int q = 0; q < (dense2 ? 2 : 1); q++) {
                const bem_c64* tab = q ? c->d_sing2 : c->d_sing;
                if (f32 && !n64)
                    hipLaunchKernelGGL(k_scatter<bem_c32>, g, dim3(kWG), 0, c->stream, c->d_fixes,
                                       (int64_t)c->fixes.size(), c->nnz, tab, (bem_c32*)(q ? dense2 : h->dense));
                else
                    hipLaunchKernelGGL(k_scatter<bem_c64>, g, dim3(kWG), 0, c->stream, c->d_fixes,
                                       (int64_t)c->fixes.size(), c->nnz, tab, q ? pair64 : own64);
                if ((rc2 = launch_check(c, "k_scatter"))) return rc2;
            }
        }
        if (n64 && c->dense_total) {   // the FP32 copies GMRES iterates with
            const dim3 g(nblk(c->dense_total, 256));
            hipLaunchKernelGGL(k_round32, g, dim3(256), 0, c->stream, c->dense_total, (const bem_c64*)h->dense64,
                               (bem_c32*)h->dense);
            if (dense2)
                hipLaunchKernelGGL(k_round32, g, dim3(256), 0, c->stream, c->dense_total, (const bem_c64*)dense2_64,
                                   (bem_c32*)dense2);
            if ((rc2 = launch_check(c, "k_round32"))) return rc2;
        }
        return BEM_OK;
    };
    // pivot lists of the DPhi variant (per mesh, partition and seed; cached)
    std::vector<void*> tmp_owned;
    int64_t* d_pivoff = nullptr;
    int32_t* d_piv = nullptr;
    if (kind == BEM_KIND_DPHI) {
        if ((rc = ensure_pivot_lists(c, seed))) return fail(rc);
        d_pivoff = c->d_pivoff;
        d_piv = c->d_piv;
    }
    // ACA, first pass with a modest rank capacity, then the overflowing
    // blocks again with their full capacity
    std::vector<int32_t> cap = c->vp.take<int32_t>(nb, 0);
    auto full_cap = [&](int32_t b) {
        int64_t m = std::min(c->blocks[b].rows, c->blocks[b].cols);
        if (maxrank > 0) m = std::min(m, maxrank);
        return (int32_t)m;
    };
    for (int32_t b : c->adm_list) cap[b] = std::min<int32_t>(full_cap(b), kCapFirst);
    AcaScratch S1, S2;
    std::vector<int32_t> rank, flag, rank2, flag2;
    std::vector<int64_t> evs, evs2;
    if ((rc = run_aca(c, kind, precision, tol, maxrank, seed, c->adm_list, cap, d_pivoff, d_piv, dense_work, S1, rank, flag,
                      evs)))
        return fail(rc);
    HIPC(c, hipEventRecord(ev1, c->stream));
    std::vector<int32_t> over, keep;
    for (int32_t b : c->adm_list) {
        if (flag[b] == 2) {
            free_list(c, tmp_owned);
            return fail(set_err(c, BEM_ERR_NAN, "bem_assemble: non-finite value in ACA block " + std::to_string(b)));
        }
        (flag[b] == 1 ? over : keep).push_back(b);
    }
    if (!over.empty()) {
        std::vector<int32_t> cap2 = c->vp.take<int32_t>(nb, 0);
        for (int32_t b : over) cap2[b] = std::min<int32_t>(full_cap(b), kCapMax);
        if ((rc = run_aca(c, kind, precision, tol, maxrank, seed, over, cap2, d_pivoff, d_piv, [] { return BEM_OK; }, S2, rank2,
                          flag2, evs2)))
            return fail(rc);
        for (int32_t b : over) {
            rank[b] = rank2[b];
            evs[b] = evs2[b];
            if (flag2[b]) {
                free_list(c, tmp_owned);
                return fail(set_err(c, flag2[b] == 2 ? BEM_ERR_NAN : BEM_ERR_ARG,
                                    "bem_assemble: ACA block " + std::to_string(b) +
                                        (flag2[b] == 2 ? " non-finite"
                                                       : " needs rank > " + std::to_string(kCapMax) + "; lower the tolerance or set max_rank")));
            }
        }
        c->vp.give(cap2);
    }
    free_list(c, tmp_owned);
    if ((rc = finish_hmat(c, h, f32, S1, S2, rank, evs, over, keep, ev2))) return fail(rc);
    float ms_dense = 0, ms_aca = 0;
    (void)hipEventElapsedTime(&ms_dense, ev0, ev1);
    (void)hipEventElapsedTime(&ms_aca, ev1, ev2);
    (void)hipEventDestroy(ev0); (void)hipEventDestroy(ev1); (void)hipEventDestroy(ev2);
    h->stats[4] = now_ms() - t0;
    h->stats[6] = ms_dense;
    h->stats[7] = ms_aca;
    for (auto* v : {&cap, &rank, &flag, &rank2, &flag2}) c->vp.give(*v);
    for (auto* v : {&evs, &evs2}) c->vp.give(*v);
    if (pairing) {   // the DPhi near field waits for the next DPhi assembly of this state
        c->st_dense = dense2;
        c->st_dense64 = dense2_64;
        c->st_gen = c->gen;
        c->st_prec = precision;
        c->st_near64 = n64;
        dense2 = nullptr;
        dense2_64 = nullptr;
    }
    c->sing_cur = nullptr;
    *out = h;
    return BEM_OK;
}

int bem_assemble_batch(bem_ctx_t c, int kind, int64_t nf, const double* kcc3, double tol, int64_t maxrank,
                       int precision, uint64_t seed, bem_hmat_t* outs)
{
    if (!c || !outs || !kcc3 || nf < 1 || nf > 64) return set_err(c, BEM_ERR_ARG, "bem_assemble_batch: bad arguments");
    if (!c->have_part || !c->have_freq)
        return set_err(c, BEM_ERR_STATE, "bem_assemble_batch: partition or boundary conditions missing");
    if (kind != BEM_KIND_PHI && kind != BEM_KIND_DPHI) return set_err(c, BEM_ERR_ARG, "bem_assemble_batch: bad kind");
    if (precision != BEM_PREC_FP64 && precision != BEM_PREC_FP32)
        return set_err(c, BEM_ERR_ARG, "bem_assemble_batch: bad precision");
    if (!(tol >= 0)) return set_err(c, BEM_ERR_ARG, "bem_assemble_batch: negative tolerance");
    for (int64_t f = 0; f < nf; f++)
        if (!(kcc3[3 * f] > 0) || !std::isfinite(kcc3[3 * f]))
            return set_err(c, BEM_ERR_ARG, "bem_assemble_batch: k must be > 0");
    HIPC(c, hipSetDevice(c->device));
    const double t0 = now_ms();
    const double k0 = c->k;
    const bem_c64 cc0 = c->cc;
    const int64_t nb = (int64_t)c->blocks.size();
    const bool f32 = precision == BEM_PREC_FP32;
    const size_t esz = f32 ? sizeof(bem_c32) : sizeof(bem_c64);
    std::vector<bem_hmat*> H((size_t)nf, nullptr);
    std::vector<bem_c64*> sing((size_t)nf, nullptr);
    std::vector<void*> own;
    auto use = [&](int64_t f) {
        c->k = kcc3[3 * f];
        c->cc = c64(kcc3[3 * f + 1], kcc3[3 * f + 2]);
    };
    auto fail = [&](int code) {
        c->k = k0;
        c->cc = cc0;
        c->sing_cur = nullptr;
        for (auto* h : H)
            if (h) bem_hmat_destroy(h);
        free_list(c, own);
        return code;
    };
    int rc;
    const int64_t nitems = c->nnz + c->n;
    const bool n64 = f32 && c->near64;   // FP32 operators with their FP64 near field (bem_set_near_fp64)
    // near-field pairing: a batched DPhi assembly takes the near fields the
    // batched Phi assembly of the same state and frequencies formed; a
    // batched Phi assembly forms both
    const std::vector<double> kccv(kcc3, kcc3 + 3 * nf);
    const bool adopt = kind == BEM_KIND_DPHI && c->kb.gen == c->gen && c->kb.prec == precision &&
                       c->kb.near64 == n64 && c->kb.kcc == kccv;
    const bool pairing = c->pair_near && kind == BEM_KIND_PHI;
    std::vector<void*> pd((size_t)nf, nullptr), pd64((size_t)nf, nullptr), ps((size_t)nf, nullptr);
    auto fail_pair = [&](int code) {
        for (auto* v : {&pd, &pd64, &ps})
            for (void* q : *v)
                if (q) pfree(c, q);
        return fail(code);
    };
    for (int64_t f = 0; f < nf; f++) {
        H[f] = new bem_hmat();
        H[f]->ctx = c;
        H[f]->kind = kind;
        H[f]->prec = precision;
        H[f]->esz = esz;
        if (adopt) {   // the kept blocks become the operators' own; the tables are freed with the batch
            H[f]->dense = c->kb.dense[f];
            H[f]->dense64 = c->kb.dense64[f];
            sing[f] = (bem_c64*)c->kb.sing[f];
            own.push_back(sing[f]);
            c->kb.dense[f] = c->kb.dense64[f] = c->kb.sing[f] = nullptr;
        }
    }
    drop_kept(c);   // what is kept (or left of it) does not belong to these
    for (int64_t f = 0; f < nf && !adopt; f++) {
        if ((rc = dalloc_bytes(c, &H[f]->dense, (size_t)c->dense_total * esz))) return fail_pair(rc);
        if (n64 && (rc = dalloc_bytes(c, &H[f]->dense64, (size_t)c->dense_total * sizeof(bem_c64))))
            return fail_pair(rc);
        if ((rc = dalloc(c, &sing[f], (size_t)nitems, &own))) return fail_pair(rc);
        if (pairing) {
            bem_c64* t = nullptr;
            if ((rc = dalloc_bytes(c, &pd[f], (size_t)c->dense_total * esz))) return fail_pair(rc);
            if (n64 && (rc = dalloc_bytes(c, &pd64[f], (size_t)c->dense_total * sizeof(bem_c64))))
                return fail_pair(rc);
            if ((rc = dalloc(c, &t, (size_t)nitems, nullptr))) return fail_pair(rc);
            ps[f] = t;
        }
    }
    hipEvent_t ev0, ev1, ev2;
    HIPC(c, hipEventCreate(&ev0)); HIPC(c, hipEventCreate(&ev1)); HIPC(c, hipEventCreate(&ev2));
    HIPC(c, hipEventRecord(ev0, c->stream));
    for (int64_t f = 0; f < nf && !adopt; f++) {   // self and near-singular entries per frequency
        use(f);
        Timer t(c, 1);
        hipLaunchKernelGGL(k_singular, dim3(nblk(nitems, kWG)), dim3(kWG), 0, c->stream, c->mesh, dev_freq(c), kind,
                           c->d_sing_order, nitems, c->nnz, sing[f], (bem_c64*)ps[f]);
        if ((rc = launch_check(c, "k_singular"))) return fail_pair(rc);
    }
    // the dense tiles of every frequency, on the main stream while the large
    // ACA classes run on the side streams
    auto dense_work = [&]() -> int {
        int rc2;
        for (int64_t f = 0; f < nf && !adopt; f++) {
            use(f);
            bem_c64* d64 = (bem_c64*)(n64 ? H[f]->dense64 : H[f]->dense);   // the FP64 near field, if any
            bem_c64* p64 = (bem_c64*)(n64 ? pd64[f] : pd[f]);                // and the paired DPhi one
            if (!c->tiles.empty()) {
                Timer t(c, 0);
                const dim3 g((unsigned)c->tiles.size());
                const bool plane = c->mesh.rvtx != nullptr;
                if (f32 && !n64)
                    hipLaunchKernelGGL(plane ? k_dense32<true> : k_dense32<false>, g, dim3(kWG), 0, c->stream, c->mesh,
                                       dev_freq32(c), c->d_subst, kind, c->d_tiles, (bem_c32*)H[f]->dense,
                                       (bem_c32*)pd[f]);
                else
                    hipLaunchKernelGGL(plane ? k_dense<true> : k_dense<false>, g, dim3(kWG), 0, c->stream, c->mesh,
                                       dev_freq(c), c->d_tiles, kind == BEM_KIND_PHI ? d64 : nullptr,
                                       kind == BEM_KIND_DPHI ? d64 : p64);
                if ((rc2 = launch_check(c, "k_dense"))) return rc2;
            }
            if (!c->fixes.empty()) {
                const dim3 g(nblk((int64_t)c->fixes.size(), kWG));
                for (int q = 0; q < (pairing ? 2 : 1); q++) {
                    const bem_c64* tab = q ? (const bem_c64*)ps[f] : sing[f];
                    if (f32 && !n64)
                        hipLaunchKernelGGL(k_scatter<bem_c32>, g, dim3(kWG), 0, c->stream, c->d_fixes,
                                           (int64_t)c->fixes.size(), c->nnz, tab, (bem_c32*)(q ? pd[f] : H[f]->dense));
                    else
                        hipLaunchKernelGGL(k_scatter<bem_c64>, g, dim3(kWG), 0, c->stream, c->d_fixes,
                                           (int64_t)c->fixes.size(), c->nnz, tab, q ? p64 : d64);
                    if ((rc2 = launch_check(c, "k_scatter"))) return rc2;
                }
            }
            if (n64 && c->dense_total) {
                const dim3 g(nblk(c->dense_total, 256));
                hipLaunchKernelGGL(k_round32, g, dim3(256), 0, c->stream, c->dense_total, (const bem_c64*)d64,
                                   (bem_c32*)H[f]->dense);
                if (pairing)
                    hipLaunchKernelGGL(k_round32, g, dim3(256), 0, c->stream, c->dense_total, (const bem_c64*)p64,
                                       (bem_c32*)pd[f]);
                if ((rc2 = launch_check(c, "k_round32"))) return rc2;
            }
        }
        return BEM_OK;
    };
    int64_t* d_pivoff = nullptr;
    int32_t* d_piv = nullptr;
    if (kind == BEM_KIND_DPHI) {
        if ((rc = ensure_pivot_lists(c, seed))) return fail_pair(rc);
        d_pivoff = c->d_pivoff;
        d_piv = c->d_piv;
    }
    auto full_cap = [&](int32_t b) {
        int64_t m = std::min(c->blocks[b].rows, c->blocks[b].cols);
        if (maxrank > 0) m = std::min(m, maxrank);
        return (int32_t)m;
    };
    std::vector<int32_t> cap(nb, 0);
    for (int32_t b : c->adm_list) cap[b] = std::min<int32_t>(full_cap(b), kCapFirst);
    std::vector<FreqJob> jobs((size_t)nf);
    for (int64_t f = 0; f < nf; f++) jobs[f] = FreqJob{kcc3[3 * f], c64(kcc3[3 * f + 1], kcc3[3 * f + 2]), sing[f]};
    std::vector<AcaScratch> S1((size_t)nf);
    std::vector<std::vector<int32_t>> rank, flag;
    std::vector<std::vector<int64_t>> evs;
    if ((rc = run_aca_batch(c, kind, precision, tol, maxrank, seed, jobs, c->adm_list, cap, d_pivoff, d_piv, dense_work,
                            S1, rank, flag, evs)))
        return fail_pair(rc);
    HIPC(c, hipEventRecord(ev1, c->stream));
    for (int64_t f = 0; f < nf; f++) {
        std::vector<int32_t> over, keep;
        for (int32_t b : c->adm_list) {
            if (flag[f][b] == 2)
                return fail_pair(set_err(c, BEM_ERR_NAN, "bem_assemble_batch: non-finite value in ACA block " +
                                                        std::to_string(b)));
            (flag[f][b] == 1 ? over : keep).push_back(b);
        }
        AcaScratch S2;
        if (!over.empty()) {   // the overflowing blocks of this frequency with their full capacity
            std::vector<int32_t> cap2(nb, 0), rank2, flag2;
            std::vector<int64_t> evs2;
            for (int32_t b : over) cap2[b] = std::min<int32_t>(full_cap(b), kCapMax);
            use(f);
            c->sing_cur = sing[f];
            rc = run_aca(c, kind, precision, tol, maxrank, seed, over, cap2, d_pivoff, d_piv, [] { return BEM_OK; }, S2,
                         rank2, flag2, evs2);
            c->sing_cur = nullptr;
            if (rc) return fail_pair(rc);
            for (int32_t b : over) {
                rank[f][b] = rank2[b];
                evs[f][b] = evs2[b];
                if (flag2[b])
                    return fail_pair(set_err(c, flag2[b] == 2 ? BEM_ERR_NAN : BEM_ERR_ARG,
                                        "bem_assemble_batch: ACA block " + std::to_string(b) +
                                            (flag2[b] == 2 ? " non-finite" : " needs rank > " + std::to_string(kCapMax))));
            }
            c->vp.give(rank2); c->vp.give(flag2); c->vp.give(evs2);
        }
        if ((rc = finish_hmat(c, H[f], f32, S1[f], S2, rank[f], evs[f], over, keep, nullptr))) return fail_pair(rc);
        c->vp.give(rank[f]); c->vp.give(flag[f]); c->vp.give(evs[f]);
    }
    HIPC(c, hipEventRecord(ev2, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    float ms_dense = 0, ms_aca = 0;
    (void)hipEventElapsedTime(&ms_dense, ev0, ev1);
    (void)hipEventElapsedTime(&ms_aca, ev1, ev2);
    (void)hipEventDestroy(ev0); (void)hipEventDestroy(ev1); (void)hipEventDestroy(ev2);
    const double wall = now_ms() - t0;
    for (int64_t f = 0; f < nf; f++) {   // the batch's times, shared out evenly
        H[f]->stats[4] = wall / nf;
        H[f]->stats[6] = ms_dense / nf;
        H[f]->stats[7] = ms_aca / nf;
        outs[f] = H[f];
        H[f] = nullptr;
    }
    if (pairing) {   // the DPhi near fields wait for the next batched DPhi assembly of this state
        c->kb.kcc = kccv;
        c->kb.dense = pd;
        c->kb.dense64 = pd64;
        c->kb.sing = ps;
        c->kb.gen = c->gen;
        c->kb.prec = precision;
        c->kb.near64 = n64;
    }
    c->k = k0;
    c->cc = cc0;
    free_list(c, own);
    return BEM_OK;
}

int bem_assemble_multi(const bem_ctx_t* ctxs, int64_t nctx, int kind, double tol, int64_t maxrank, int precision,
                       uint64_t seed, bem_hmat_t* outs)
{
    if (!ctxs || !outs || nctx < 1 || nctx > 2048 || !ctxs[0]) return BEM_ERR_ARG;
    bem_ctx* lead = ctxs[0];
    if (kind != BEM_KIND_PHI && kind != BEM_KIND_DPHI) return set_err(lead, BEM_ERR_ARG, "bem_assemble_multi: bad kind");
    if (precision != BEM_PREC_FP64 && precision != BEM_PREC_FP32)
        return set_err(lead, BEM_ERR_ARG, "bem_assemble_multi: bad precision");
    if (!(tol >= 0)) return set_err(lead, BEM_ERR_ARG, "bem_assemble_multi: negative tolerance");
    for (int64_t q = 0; q < nctx; q++) {
        bem_ctx* c = ctxs[q];
        if (!c) return set_err(lead, BEM_ERR_ARG, "bem_assemble_multi: null context");
        if (c->device != lead->device) return set_err(lead, BEM_ERR_ARG, "bem_assemble_multi: contexts on other devices");
        if (!c->have_part || !c->have_freq)
            return set_err(lead, BEM_ERR_STATE, "bem_assemble_multi: partition or frequency missing");
        if (c->mesh.rvtx) return set_err(lead, BEM_ERR_ARG, "bem_assemble_multi: not with an impedance plane");
        if ((int64_t)c->blocks.size() > kJobMask)
            return set_err(lead, BEM_ERR_ARG, "bem_assemble_multi: more than 2^20 blocks in a context");
        for (int64_t e = 0; e < q; e++)
            if (ctxs[e] == c) return set_err(lead, BEM_ERR_ARG, "bem_assemble_multi: a context listed twice");
    }
    HIPC(lead, hipSetDevice(lead->device));
    const double t0 = now_ms();
    const bool f32 = precision == BEM_PREC_FP32;
    const int pi = f32 ? 1 : 0;
    const size_t esz = f32 ? sizeof(bem_c32) : sizeof(bem_c64);
    std::vector<bem_hmat*> H((size_t)nctx, nullptr);
    std::vector<AcaScratch> S1((size_t)nctx);
    std::vector<AcaArgs> A((size_t)nctx);
    std::vector<std::vector<int32_t>> caps((size_t)nctx);
    std::vector<std::array<std::vector<int32_t>, 3>> cls((size_t)nctx);
    std::vector<void*> lead_own;
    auto fail = [&](int code) {
        for (auto* h : H)
            if (h) bem_hmat_destroy(h);
        free_list(lead, lead_own);
        return code;
    };
    int rc;
    size_t shm = 0;
    for (int64_t q = 0; q < nctx; q++) {   // per context: near field, singular table, ACA scratch
        bem_ctx* c = ctxs[q];
        drop_kept(c);
        const bool n64 = f32 && c->near64;
        H[q] = new bem_hmat();
        H[q]->ctx = c;
        H[q]->kind = kind;
        H[q]->prec = precision;
        H[q]->esz = esz;
        if ((rc = dalloc_bytes(c, &H[q]->dense, (size_t)c->dense_total * esz))) return fail(rc);
        if (n64 && (rc = dalloc_bytes(c, &H[q]->dense64, (size_t)c->dense_total * sizeof(bem_c64)))) return fail(rc);
        const int64_t nitems = c->nnz + c->n;
        {
            Timer t(c, 1);
            hipLaunchKernelGGL(k_singular, dim3(nblk(nitems, kWG)), dim3(kWG), 0, c->stream, c->mesh, dev_freq(c), kind,
                               c->d_sing_order, nitems, c->nnz, c->d_sing, nullptr);
            if ((rc = launch_check(c, "k_singular"))) return fail(rc);
        }
        bem_c64* d64 = (bem_c64*)(n64 ? H[q]->dense64 : H[q]->dense);
        if (!c->tiles.empty()) {
            Timer t(c, 0);
            const dim3 g((unsigned)c->tiles.size());
            if (f32 && !n64)
                hipLaunchKernelGGL(k_dense32<false>, g, dim3(kWG), 0, c->stream, c->mesh, dev_freq32(c), c->d_subst, kind,
                                   c->d_tiles, (bem_c32*)H[q]->dense, (bem_c32*)nullptr);
            else
                hipLaunchKernelGGL(k_dense<false>, g, dim3(kWG), 0, c->stream, c->mesh, dev_freq(c), c->d_tiles,
                                   kind == BEM_KIND_PHI ? d64 : nullptr, kind == BEM_KIND_DPHI ? d64 : nullptr);
            if ((rc = launch_check(c, "k_dense"))) return fail(rc);
        }
        if (!c->fixes.empty()) {
            const dim3 g(nblk((int64_t)c->fixes.size(), kWG));
            if (f32 && !n64)
                hipLaunchKernelGGL(k_scatter<bem_c32>, g, dim3(kWG), 0, c->stream, c->d_fixes, (int64_t)c->fixes.size(),
                                   c->nnz, c->d_sing, (bem_c32*)H[q]->dense);
            else
                hipLaunchKernelGGL(k_scatter<bem_c64>, g, dim3(kWG), 0, c->stream, c->d_fixes, (int64_t)c->fixes.size(),
                                   c->nnz, c->d_sing, d64);
            if ((rc = launch_check(c, "k_scatter"))) return fail(rc);
        }
        if (n64 && c->dense_total) {
            hipLaunchKernelGGL(k_round32, dim3(nblk(c->dense_total, 256)), dim3(256), 0, c->stream, c->dense_total,
                               (const bem_c64*)d64, (bem_c32*)H[q]->dense);
            if ((rc = launch_check(c, "k_round32"))) return fail(rc);
        }
        int64_t* d_pivoff = nullptr;
        int32_t* d_piv = nullptr;
        if (kind == BEM_KIND_DPHI) {
            if ((rc = ensure_pivot_lists(c, seed))) return fail(rc);
            d_pivoff = c->d_pivoff;
            d_piv = c->d_piv;
        }
        const int64_t nb = (int64_t)c->blocks.size();
        caps[q].assign(nb, 0);
        for (int32_t b : c->adm_list) {
            int64_t m = std::min(c->blocks[b].rows, c->blocks[b].cols);
            if (maxrank > 0) m = std::min(m, maxrank);
            caps[q][b] = (int32_t)std::min<int64_t>(m, kCapFirst);
        }
        // every context classifies its blocks by the lead's schedule
        const int64_t keep_s = c->cls_small[pi], keep_h = c->cls_huge[pi];
        c->cls_small[pi] = lead->cls_small[pi];
        c->cls_huge[pi] = lead->cls_huge[pi];
        size_t shm_q = 0;
        std::vector<int32_t> cq[3];
        c->sing_cur = c->d_sing;
        rc = aca_prepare(c, kind, precision, tol, maxrank, seed, c->adm_list, caps[q], d_pivoff, d_piv, S1[q], A[q], cq,
                         shm_q);
        c->sing_cur = nullptr;
        c->cls_small[pi] = keep_s;
        c->cls_huge[pi] = keep_h;
        if (rc) return fail(rc);
        for (int k = 0; k < 3; k++) cls[q][k] = std::move(cq[k]);
        shm = std::max(shm, shm_q);
        // the lead's launches wait for this context's near field, table and uploads
        HIPC(c, hipEventRecord(c->ev_fork, c->stream));
        HIPC(lead, hipStreamWaitEvent(lead->stream, c->ev_fork, 0));
    }
    // the jobs, and the items of every size class: all contexts' blocks,
    // the largest first
    std::vector<MeshJob> jobs((size_t)nctx);
    for (int64_t q = 0; q < nctx; q++) {
        const AcaArgs& a = A[q];
        MeshJob& J = jobs[q];
        J.M = a.M; J.F = a.F; J.F32 = a.F32;
        J.row0 = a.row0; J.rows = a.rows; J.col0 = a.col0; J.cols = a.cols;
        J.uoff = a.uoff; J.voff = a.voff; J.aoff = a.aoff; J.cap = a.cap;
        J.U = a.U; J.V = a.V; J.A = a.A;
        J.rank_out = a.rank_out; J.flag_out = a.flag_out; J.evals_out = a.evals_out;
        J.piv_off = a.piv_off; J.piv = a.piv;
        J.sing = a.sing; J.nnz = a.nnz;
        J.bnear = a.bnear;
    }
    std::vector<int32_t> items;
    size_t ncls[3];
    for (int k = 0; k < 3; k++) {
        std::vector<std::pair<int64_t, int32_t>> v;
        for (int64_t q = 0; q < nctx; q++)
            for (int32_t b : cls[q][k])
                v.push_back({ctxs[q]->blocks[b].rows + ctxs[q]->blocks[b].cols, (int32_t)((q << kJobShift) | b)});
        std::stable_sort(v.begin(), v.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
        for (auto& e : v) items.push_back(e.second);
        ncls[k] = v.size();
    }
    MeshJob* d_jobs = nullptr;
    int32_t* d_items = nullptr;
    int* d_queue = nullptr;
    if ((rc = dupload(lead, &d_jobs, jobs.data(), jobs.size(), &lead_own)) ||
        (!items.empty() && (rc = dupload(lead, &d_items, items.data(), items.size(), &lead_own))) ||
        (rc = dalloc(lead, &d_queue, (size_t)3, &lead_own)))
        return fail(rc);
    if (lead->aca_queue) HIPC(lead, hipMemsetAsync(d_queue, 0, 3 * sizeof(int), lead->stream));
    AcaArgs a = A[0];
    a.jobs = d_jobs;
    HIPC(lead, hipEventRecord(lead->ev_fork, lead->stream));
    const int nt[3] = {pi ? kHugeThreads<bem_c32> : kHugeThreads<bem_c64>,
                       pi ? kMidThreads<bem_c32> : kMidThreads<bem_c64>, 64};
    size_t off = 0;
    for (int k = 0; k < 3; k++) {
        AcaArgs aq = a;
        aq.blist = d_items + off;
        off += ncls[k];
        if (!ncls[k]) continue;
        hipStream_t st = k < 2 ? lead->side[k] : lead->stream;
        if (k < 2) HIPC(lead, hipStreamWaitEvent(st, lead->ev_fork, 0));
        unsigned grid = (unsigned)ncls[k];
        if (lead->aca_queue) {
            aq.queue = d_queue + k;
            aq.nitems = (int)ncls[k];
            grid = (unsigned)std::min<size_t>(ncls[k], (size_t)(256 * 4 * 4 * 64 / nt[k]));
        }
        Timer t(lead, 2, st);
        const bool x = kind == BEM_KIND_DPHI;
        if (f32) launch_aca<bem_c32>(k, x, grid, shm, st, aq);
        else launch_aca<bem_c64>(k, x, grid, shm, st, aq);
        if ((rc = launch_check(lead, "k_aca (multi)"))) return fail(rc);
    }
    for (int k = 0; k < 2; k++) {
        HIPC(lead, hipEventRecord(lead->ev_join[k], lead->side[k]));
        HIPC(lead, hipStreamWaitEvent(lead->stream, lead->ev_join[k], 0));
    }
    HIPC(lead, hipStreamSynchronize(lead->stream));
    const double t1 = now_ms();
    // per context: results, re-run of the overflowing blocks, packing
    for (int64_t q = 0; q < nctx; q++) {
        bem_ctx* c = ctxs[q];
        const int64_t nb = (int64_t)c->blocks.size();
        std::vector<int32_t> rank, flag;
        std::vector<int64_t> evs;
        if ((rc = aca_collect(c, S1[q], nb, rank, flag, evs))) return fail(rc);
        std::vector<int32_t> over, keep;
        for (int32_t b : c->adm_list) {
            if (flag[b] == 2)
                return fail(set_err(lead, BEM_ERR_NAN, "bem_assemble_multi: non-finite value in ACA block " +
                                                          std::to_string(b) + " of context " + std::to_string(q)));
            (flag[b] == 1 ? over : keep).push_back(b);
        }
        AcaScratch S2;
        if (!over.empty()) {
            std::vector<int32_t> cap2(nb, 0), rank2, flag2;
            std::vector<int64_t> evs2;
            for (int32_t b : over) {
                int64_t m = std::min(c->blocks[b].rows, c->blocks[b].cols);
                if (maxrank > 0) m = std::min(m, maxrank);
                cap2[b] = (int32_t)std::min<int64_t>(m, kCapMax);
            }
            c->sing_cur = c->d_sing;
            rc = run_aca(c, kind, precision, tol, maxrank, seed, over, cap2, kind == BEM_KIND_DPHI ? c->d_pivoff : nullptr,
                         kind == BEM_KIND_DPHI ? c->d_piv : nullptr, [] { return BEM_OK; }, S2, rank2, flag2, evs2);
            c->sing_cur = nullptr;
            if (rc) return fail(rc);
            for (int32_t b : over) {
                rank[b] = rank2[b];
                evs[b] = evs2[b];
                if (flag2[b])
                    return fail(set_err(lead, flag2[b] == 2 ? BEM_ERR_NAN : BEM_ERR_ARG,
                                        "bem_assemble_multi: ACA block " + std::to_string(b) + " of context " +
                                            std::to_string(q) +
                                            (flag2[b] == 2 ? " non-finite"
                                                           : " needs rank > " + std::to_string(kCapMax))));
            }
        }
        if ((rc = finish_hmat(c, H[q], f32, S1[q], S2, rank, evs, over, keep, nullptr))) return fail(rc);
        HIPC(c, hipStreamSynchronize(c->stream));
    }
    const double wall = now_ms() - t0;
    for (int64_t q = 0; q < nctx; q++) {   // the launch's times, shared out evenly
        H[q]->stats[4] = wall / nctx;
        H[q]->stats[7] = (t1 - t0) / nctx;
        outs[q] = H[q];
        H[q] = nullptr;
    }
    free_list(lead, lead_own);
    return BEM_OK;
}

int bem_autotune_aca(bem_ctx_t c, int kind, int precision, int64_t nf, const double* kcc3, double tol,
                     int64_t maxrank, uint64_t seed, int64_t* choice3, double* best_ms)
{
    if (!c || !kcc3 || nf < 1 || nf > 64 || (precision != BEM_PREC_FP64 && precision != BEM_PREC_FP32))
        return set_err(c, BEM_ERR_ARG, "bem_autotune_aca: bad arguments");
    const int pi = precision == BEM_PREC_FP32 ? 1 : 0;
    const int64_t keep_small = c->cls_small[pi], keep_huge = c->cls_huge[pi];
    const bool keep_queue = c->aca_queue;
    // one timed assembly (the fastest of two runs) under the schedule s
    auto timed = [&](int64_t small, int64_t huge, bool queue, double* ms) -> int {
        c->cls_small[pi] = small;
        c->cls_huge[pi] = huge;
        c->aca_queue = queue;
        double best = 1e300;
        for (int rep = 0; rep < 2; rep++) {
            std::vector<bem_hmat_t> hs((size_t)nf, nullptr);
            const double t0 = now_ms();
            int rc;
            if (nf == 1) {
                const double k0 = c->k;
                const bem_c64 cc0 = c->cc;
                c->k = kcc3[0];
                c->cc = c64(kcc3[1], kcc3[2]);
                rc = bem_assemble(c, kind, tol, maxrank, precision, seed, &hs[0]);
                c->k = k0;
                c->cc = cc0;
            } else {
                rc = bem_assemble_batch(c, kind, nf, kcc3, tol, maxrank, precision, seed, hs.data());
            }
            const double dt = now_ms() - t0;
            for (bem_hmat_t h : hs)
                if (h) bem_hmat_destroy(h);
            if (rc) return rc;
            best = std::min(best, dt);
        }
        *ms = best;
        return BEM_OK;
    };
    auto restore = [&](int rc) {
        c->cls_small[pi] = keep_small;
        c->cls_huge[pi] = keep_huge;
        c->aca_queue = keep_queue;
        drop_kept(c);
        return rc;
    };
    drop_kept(c);   // nothing kept before is adopted by a timed run, nothing a timed run keeps survives
    const int64_t inf = (int64_t)1 << 40;
    const std::vector<int64_t> smalls = {128, 256, 384, 512};
    const std::vector<int64_t> huges = pi ? std::vector<int64_t>{1024, 2048, 4096, inf} : std::vector<int64_t>{inf};
    int64_t bs = kSmallBlock[pi], bh = kHugeBlock[pi];
    bool bq = false;
    double bt = 0;
    int rc;
    if ((rc = timed(bs, bh, bq, &bt))) return restore(rc);   // the built-in schedule (also the warm-up)
    if ((rc = timed(bs, bh, bq, &bt))) return restore(rc);
    // coordinate descent: the one-wavefront bound, then the huge bound, then
    // the queue; a candidate replaces the current choice only when it is
    // kTuneGain faster (run-to-run noise does not move the schedule)
    constexpr double kTuneGain = 0.97;
    for (int64_t v : smalls) {
        double t;
        if (v == bs) continue;
        if ((rc = timed(v, bh, bq, &t))) return restore(rc);
        if (t < kTuneGain * bt) { bt = t; bs = v; }
    }
    for (int64_t v : huges) {
        double t;
        if (v == bh || v <= bs) continue;
        if ((rc = timed(bs, v, bq, &t))) return restore(rc);
        if (t < kTuneGain * bt) { bt = t; bh = v; }
    }
    if (nf > 1) {
        double t;
        if ((rc = timed(bs, bh, !bq, &t))) return restore(rc);
        if (t < kTuneGain * bt) { bt = t; bq = !bq; }
    }
    c->cls_small[pi] = bs;
    c->cls_huge[pi] = bh;
    c->aca_queue = bq;
    drop_kept(c);
    if (choice3) { choice3[0] = bs; choice3[1] = bh; choice3[2] = bq; }
    if (best_ms) *best_ms = bt;
    return BEM_OK;
}

int bem_set_fuse_ops(bem_ctx_t c, int on)
{
    if (!c) return BEM_ERR_ARG;
    c->fuse_ops = on != 0;
    return BEM_OK;
}

int bem_set_aca_schedule(bem_ctx_t c, int precision, int64_t small_bound, int64_t huge_bound, int queue)
{
    if (!c || (precision != BEM_PREC_FP64 && precision != BEM_PREC_FP32) || small_bound < 0 || huge_bound < small_bound)
        return set_err(c, BEM_ERR_ARG, "bem_set_aca_schedule: bad arguments");
    c->cls_small[precision] = small_bound;
    c->cls_huge[precision] = huge_bound;
    if (queue >= 0) c->aca_queue = queue != 0;
    return BEM_OK;
}

}  // extern "C"

namespace {

// near-field tiles and low-rank blocks of one H-MVM into the partial vectors
template <class S>
int hmvm_blocks(bem_ctx* c, bem_hmat* h, const bem_c64* xd)
{
    int rc;
    if (!c->tiles.empty()) {
        Timer t(c, 3);
        hipLaunchKernelGGL(k_mvm_dense<S>, dim3((unsigned)c->tiles.size()), dim3(kWG), 0, c->stream, c->d_tiles,
                           (const S*)h->dense, xd, c->d_part);
        if ((rc = launch_check(c, "k_mvm_dense"))) return rc;
    }
    if (h->nttask > 0) {
        Timer t(c, 4);
        hipLaunchKernelGGL(k_lr_t<S>, dim3(nblk(h->nttask, kWG / 64)), dim3(kWG), 0, c->stream, h->d_ttask,
                           h->nttask, c->d_col0, c->d_cols, h->d_rank, h->d_voff, h->d_aoff, h->d_toff,
                           (const S*)h->V, (const S*)h->A, xd, h->d_t);
        hipLaunchKernelGGL(k_lr_u<S>, dim3((unsigned)h->nutask), dim3(kWG), h->maxrank * sizeof(bem_c64), c->stream,
                           h->d_utask, c->d_rows, h->d_rank, h->d_uoff, h->d_toff, c->d_lr_part, (const S*)h->U,
                           h->d_t, c->d_part);
        if ((rc = launch_check(c, "k_lr"))) return rc;
    }
    return BEM_OK;
}

// The refinement operator of an FP32 H-matrix with a kept FP64 near field:
// the near-field tiles from the FP64 blocks, the low-rank blocks from the
// FP32 factors, all accumulated in FP64
int hmvm_blocks_near64(bem_ctx* c, bem_hmat* h, const bem_c64* xd)
{
    int rc;
    if (!c->tiles.empty()) {
        Timer t(c, 3);
        hipLaunchKernelGGL(k_mvm_dense<bem_c64>, dim3((unsigned)c->tiles.size()), dim3(kWG), 0, c->stream,
                           c->d_tiles, (const bem_c64*)h->dense64, xd, c->d_part);
        if ((rc = launch_check(c, "k_mvm_dense"))) return rc;
    }
    if (h->nttask > 0) {
        Timer t(c, 4);
        hipLaunchKernelGGL(k_lr_t<bem_c32>, dim3(nblk(h->nttask, kWG / 64)), dim3(kWG), 0, c->stream, h->d_ttask,
                           h->nttask, c->d_col0, c->d_cols, h->d_rank, h->d_voff, h->d_aoff, h->d_toff,
                           (const bem_c32*)h->V, (const bem_c32*)h->A, xd, h->d_t);
        hipLaunchKernelGGL(k_lr_u<bem_c32>, dim3((unsigned)h->nutask), dim3(kWG), h->maxrank * sizeof(bem_c64),
                           c->stream, h->d_utask, c->d_rows, h->d_rank, h->d_uoff, h->d_toff, c->d_lr_part,
                           (const bem_c32*)h->U, h->d_t, c->d_part);
        if ((rc = launch_check(c, "k_lr"))) return rc;
    }
    return BEM_OK;
}

// count complex entries at src (FP64 or FP32 on the device) into FP64 host memory
int download_c64(bem_hmat* h, bem_c64* dst, const void* src, int64_t count)
{
    bem_ctx* c = h->ctx;
    if (h->prec != BEM_PREC_FP32) {
        HIPC(c, hipMemcpyAsync(dst, src, count * sizeof(bem_c64), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        return BEM_OK;
    }
    std::vector<bem_c32> tmp(count);
    HIPC(c, hipMemcpyAsync(tmp.data(), src, count * sizeof(bem_c32), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    for (int64_t t = 0; t < count; t++) dst[t] = c64(tmp[t].re, tmp[t].im);
    return BEM_OK;
}

}  // namespace

extern "C" {

int bem_hmvm(bem_ctx_t c, bem_hmat_t h, const void* x, void* y, int accumulate)
{
    if (!c || !h || !x || !y || h->ctx != c) return set_err(c, BEM_ERR_ARG, "bem_hmvm: bad arguments");
    HIPC(c, hipSetDevice(c->device));
    const bem_c64* xd = (const bem_c64*)x;
    bem_c64* yd = (bem_c64*)y;
    int rc;
    if (h->prec == BEM_PREC_FP32) rc = hmvm_blocks<bem_c32>(c, h, xd);
    else rc = hmvm_blocks<bem_c64>(c, h, xd);
    if (rc) return rc;
    {
        Timer t(c, 5);
        hipLaunchKernelGGL(k_mvm_gather, dim3(nblk(16 * c->n, 256)), dim3(256), 0, c->stream, c->n, c->d_gptr,
                           c->d_gidx, c->d_part, yd, accumulate);
        if ((rc = launch_check(c, "k_mvm_gather"))) return rc;
    }
    return BEM_OK;
}

int bem_hmvm_refine(bem_ctx_t c, bem_hmat_t h, const void* x, void* y, int accumulate)
{
    if (!c || !h || !x || !y || h->ctx != c) return set_err(c, BEM_ERR_ARG, "bem_hmvm_refine: bad arguments");
    if (h->prec == BEM_PREC_FP64) return bem_hmvm(c, h, x, y, accumulate);
    if (!h->dense64)
        return set_err(c, BEM_ERR_STATE, "bem_hmvm_refine: FP32 operator without an FP64 near field "
                                         "(assemble after bem_set_near_fp64(ctx, 1))");
    HIPC(c, hipSetDevice(c->device));
    int rc = hmvm_blocks_near64(c, h, (const bem_c64*)x);
    if (rc) return rc;
    Timer t(c, 5);
    hipLaunchKernelGGL(k_mvm_gather, dim3(nblk(16 * c->n, 256)), dim3(256), 0, c->stream, c->n, c->d_gptr, c->d_gidx,
                       c->d_part, (bem_c64*)y, accumulate);
    return launch_check(c, "k_mvm_gather");
}

int bem_block_info(bem_hmat_t h, int64_t b, int64_t* rank)
{
    if (!h || !rank || b < 0 || b >= (int64_t)h->ctx->blocks.size()) return BEM_ERR_ARG;
    *rank = h->ctx->blocks[b].adm ? h->rank[b] : -1;
    return BEM_OK;
}

int bem_block_download(bem_hmat_t h, int64_t b, double* full_or_U, double* alpha, double* V)
{
    if (!h || b < 0 || b >= (int64_t)h->ctx->blocks.size() || !full_or_U) return BEM_ERR_ARG;
    bem_ctx* c = h->ctx;
    HIPC(c, hipSetDevice(c->device));
    const auto& B = c->blocks[b];
    const char* dense = (const char*)h->dense;
    const char *U = (const char*)h->U, *A = (const char*)h->A, *Vs = (const char*)h->V;
    int rc;
    if (!B.adm) return download_c64(h, (bem_c64*)full_or_U, dense + c->dense_off[b] * h->esz, B.rows * B.cols);
    int64_t r = h->rank[b];
    if ((rc = download_c64(h, (bem_c64*)full_or_U, U + h->uoff[b] * h->esz, B.rows * r))) return rc;
    if (alpha && (rc = download_c64(h, (bem_c64*)alpha, A + h->aoff[b] * h->esz, r))) return rc;
    if (V) {
        std::vector<bem_c64> rowmajor(r * B.cols);
        if ((rc = download_c64(h, rowmajor.data(), Vs + h->voff[b] * h->esz, r * B.cols))) return rc;
        bem_c64* Vo = (bem_c64*)V;  // reference layout: k x n column-major
        for (int64_t l = 0; l < r; l++)
            for (int64_t j = 0; j < B.cols; j++) Vo[j * r + l] = rowmajor[l * B.cols + j];
    }
    return BEM_OK;
}

int bem_hmat_stats(bem_hmat_t h, double* s)
{
    if (!h || !s) return BEM_ERR_ARG;
    std::memcpy(s, h->stats, sizeof h->stats);
    return BEM_OK;
}

int bem_precond_create(bem_ctx_t c, bem_hmat_t h, bem_prec_t* out)
{
    if (!c || !h || !out || h->ctx != c) return set_err(c, BEM_ERR_ARG, "bem_precond_create: bad arguments");
    *out = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    HIPC(c, hipSetDevice(c->device));
    // the diagonal leaf blocks must tile [0, n)
    std::vector<PrecBlock> blk;
    std::vector<int64_t> src;
    int64_t covered = 0, lu = 0;
    for (size_t b = 0; b < c->blocks.size(); b++) {
        const auto& B = c->blocks[b];
        if (B.adm || B.row0 != B.col0 || B.rows != B.cols) continue;
        if (B.rows > kPrecMax)
            return set_err(c, BEM_ERR_ARG, "bem_precond_create: diagonal block larger than 64 (leaf size above 64)");
        blk.push_back(PrecBlock{B.row0, lu, (int32_t)B.rows, 0});
        src.push_back(c->dense_off[b]);
        lu += B.rows * B.rows;
        covered += B.rows;
    }
    if (covered != c->n) return set_err(c, BEM_ERR_ARG, "bem_precond_create: diagonal leaf blocks do not cover the matrix");
    auto p = new bem_prec();
    p->ctx = c;
    p->nblk = (int64_t)blk.size();
    std::vector<void*> own;
    int32_t* d_flag;
    int64_t* d_src;
    int rc;
    auto fail = [&](int code) {
        free_list(c, own);
        pfree(c, p->d_blk); pfree(c, p->d_lu); pfree(c, p->d_piv);
        delete p;
        return code;
    };
    if ((rc = dupload(c, &p->d_blk, blk.data(), blk.size(), nullptr))) return fail(rc);
    if ((rc = dalloc(c, &p->d_lu, (size_t)lu, nullptr))) return fail(rc);
    if ((rc = dalloc(c, &p->d_piv, (size_t)c->n, nullptr))) return fail(rc);
    if ((rc = dupload(c, &d_src, src.data(), src.size(), &own))) return fail(rc);
    if ((rc = dalloc(c, &d_flag, 1, &own))) return fail(rc);
    HIPC(c, hipMemsetAsync(d_flag, 0, sizeof(int32_t), c->stream));
    if (h->prec == BEM_PREC_FP32)
        hipLaunchKernelGGL(k_blu_factor<bem_c32>, dim3((unsigned)p->nblk), dim3(kWG), 0, c->stream, p->d_blk, d_src,
                           (const bem_c32*)h->dense, p->d_lu, p->d_piv, d_flag);
    else
        hipLaunchKernelGGL(k_blu_factor<bem_c64>, dim3((unsigned)p->nblk), dim3(kWG), 0, c->stream, p->d_blk, d_src,
                           (const bem_c64*)h->dense, p->d_lu, p->d_piv, d_flag);
    if ((rc = launch_check(c, "k_blu_factor"))) return fail(rc);
    int32_t hflag = 0;
    HIPC(c, hipMemcpyAsync(&hflag, d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    free_list(c, own);
    if (hflag) {
        own.clear();
        return fail(set_err(c, BEM_ERR_NAN, "bem_precond_create: singular diagonal block"));
    }
    p->entries = (double)lu;
    p->setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = p;
    return BEM_OK;
}

// adds the device time of the H-LU applications recorded so far to apply_ms
// and releases their events: all of them (wait), or the finished ones
static void hlu_resolve_events(bem_prec* p, bool wait)
{
    size_t keep = 0;
    for (size_t i = 0; i < p->hlu_ev.size(); i++) {
        auto& e = p->hlu_ev[i];
        if (!wait && hipEventQuery(e.second) != hipSuccess) {
            p->hlu_ev[keep++] = e;
            continue;
        }
        float ms = 0;
        if (hipEventSynchronize(e.second) == hipSuccess && hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess)
            p->apply_ms += ms;
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    p->hlu_ev.resize(keep);
}

int bem_precond_create_hlu(bem_ctx_t c, bem_hmat_t h, int64_t nclusters, const int64_t* clusters2, int64_t max_rank,
                           double rel_err, bem_prec_t* out)
{
    if (!c || !h || !out || h->ctx != c || nclusters <= 0 || !clusters2 || rel_err < 0)
        return set_err(c, BEM_ERR_ARG, "bem_precond_create_hlu: bad arguments");
    *out = nullptr;
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipStreamSynchronize(c->stream));
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t nb = (int64_t)c->blocks.size();
    std::vector<int64_t> blocks5(5 * nb), rank(nb);
    // the four storage arrays in one copy each, then unpacked per block
    int64_t nd = 0, nu = 0, nv = 0, na = 0;
    for (int64_t b = 0; b < nb; b++) {
        const auto& B = c->blocks[b];
        blocks5[5 * b] = B.row0; blocks5[5 * b + 1] = B.rows;
        blocks5[5 * b + 2] = B.col0; blocks5[5 * b + 3] = B.cols;
        blocks5[5 * b + 4] = B.adm ? 1 : 0;
        rank[b] = B.adm ? (int64_t)h->rank[b] : -1;
        if (!B.adm) {
            nd = std::max(nd, c->dense_off[b] + B.rows * B.cols);
        } else {
            nu = std::max(nu, h->uoff[b] + B.rows * rank[b]);
            nv = std::max(nv, h->voff[b] + rank[b] * B.cols);
            na = std::max(na, h->aoff[b] + rank[b]);
        }
    }
    std::vector<char> hd((size_t)nd * h->esz), hu((size_t)nu * h->esz), hv((size_t)nv * h->esz), ha((size_t)na * h->esz);
    if (nd) HIPC(c, hipMemcpyAsync(hd.data(), h->dense, hd.size(), hipMemcpyDeviceToHost, c->stream));
    if (nu) HIPC(c, hipMemcpyAsync(hu.data(), h->U, hu.size(), hipMemcpyDeviceToHost, c->stream));
    if (nv) HIPC(c, hipMemcpyAsync(hv.data(), h->V, hv.size(), hipMemcpyDeviceToHost, c->stream));
    if (na) HIPC(c, hipMemcpyAsync(ha.data(), h->A, ha.size(), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    const bool f32 = h->prec == BEM_PREC_FP32;
    auto get = [&](const std::vector<char>& src, int64_t t) -> bem_c64 {
        if (f32) {
            const bem_c32 v = ((const bem_c32*)src.data())[t];
            return c64(v.re, v.im);
        }
        return ((const bem_c64*)src.data())[t];
    };
    std::vector<std::vector<bem_c64>> data(nb);
    std::vector<const double*> ptr(nb);
    for (int64_t b = 0; b < nb; b++) {
        const auto& B = c->blocks[b];
        const int64_t m = B.rows, n = B.cols, k = rank[b];
        if (k < 0) {
            data[b].resize((size_t)(m * n));
            for (int64_t t = 0; t < m * n; t++) data[b][t] = get(hd, c->dense_off[b] + t);
        } else {   // U (m x k col-major), alpha, V (k x n col-major; stored row-major)
            data[b].resize((size_t)(m * k + k + k * n));
            bem_c64* d = data[b].data();
            for (int64_t t = 0; t < m * k; t++) d[t] = get(hu, h->uoff[b] + t);
            for (int64_t l = 0; l < k; l++) d[m * k + l] = get(ha, h->aoff[b] + l);
            for (int64_t l = 0; l < k; l++)
                for (int64_t j = 0; j < n; j++) d[m * k + k + j * k + l] = get(hv, h->voff[b] + l * n + j);
        }
        ptr[b] = (const double*)data[b].data();
    }
    bem_hlu_t L = nullptr;
    int rc = bem_hlu_create(c->n, nclusters, clusters2, nb, blocks5.data(), rank.data(), ptr.data(), max_rank, rel_err,
                            &L);
    if (rc) return set_err(c, rc, std::string("bem_precond_create_hlu: ") + bem_hlu_error());
    for (auto& v : data) std::vector<bem_c64>().swap(v);
    double st[3];
    bem_hlu_stats(L, st);
    bemhlu::Flat F;
    try {
        bemhlu::flatten(L, F);
    } catch (const bemhlu::Error& e) {
        bem_hlu_destroy(L);
        return set_err(c, BEM_ERR_ARG, "bem_precond_create_hlu: " + e.msg);
    }
    bem_hlu_destroy(L);
    HluSplit S;
    hlu_split(F, S);
    std::string why;
    if (!hlu_split_ok(F, S, why)) return set_err(c, BEM_ERR_ARG, "bem_precond_create_hlu: " + why);
    auto p = new bem_prec();
    p->ctx = c;
    auto up = [&](void** dst, const void* src, size_t bytes) -> int {
        int r = dalloc_bytes(c, dst, bytes > 0 ? bytes : 16);
        if (r) return r;
        p->hlu_owned.push_back(*dst);
        if (bytes && src) HIPC(c, hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, c->stream));
        return BEM_OK;
    };
    p->hlu_nleaf = (int64_t)F.leaf.size();
    rc = up((void**)&p->hlu_leaf, F.leaf.data(), F.leaf.size() * sizeof(HluLeafD));
    for (int d = 0; d < 2 && !rc; d++)
        for (int o = 0; o < 2 && !rc; o++) {
            if (!rc) rc = up((void**)&p->hlu_tptr[d][o], S.tptr[d][o].data(), S.tptr[d][o].size() * sizeof(int64_t));
            if (!rc) rc = up((void**)&p->hlu_cptr[d][o], S.cptr[d][o].data(), S.cptr[d][o].size() * sizeof(int64_t));
            if (!rc)
                rc = up((void**)&p->hlu_titem[d][o], S.titem[d][o].data(), S.titem[d][o].size() * sizeof(HluItemD));
            if (!rc)
                rc = up((void**)&p->hlu_citem[d][o], S.citem[d][o].data(), S.citem[d][o].size() * sizeof(HluItemD));
        }
    p->hlu_panel = S.panel;
    for (int d = 0; d < 2; d++) {
        p->hlu_has_t[d] = S.has_t[d];
        p->hlu_has_c[d] = S.has_c[d];
    }
    p->hlu_gstride = S.gstride;
    if (!rc) rc = up((void**)&p->hlu_data, F.data.data(), F.data.size() * sizeof(bem_c64));
    if (!rc) rc = up((void**)&p->hlu_t, nullptr, (size_t)F.tsize * sizeof(bem_c64));   // written before read
    if (!rc)   // written before read
        rc = up((void**)&p->hlu_gpart, nullptr, (size_t)(S.maxleaves * S.gstride) * sizeof(bem_c64));
    if (!rc) {
        hipError_t e = hipStreamSynchronize(c->stream);   // the host buffers go out of scope
        if (e != hipSuccess) rc = set_err(c, BEM_ERR_HIP, std::string("bem_precond_create_hlu: ") + hipGetErrorString(e));
    }
    if (rc) {
        bem_precond_destroy(p);
        return rc;
    }
    p->entries = st[0];
    p->factor_ms = 1e3 * st[2];
    p->setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = p;
    return BEM_OK;
}

int bem_precond_stats(bem_prec_t p, double* out5)
{
    if (!p || !out5) return BEM_ERR_ARG;
    hlu_resolve_events(p, true);
    out5[0] = p->setup_ms;
    out5[1] = (double)p->applies;
    out5[2] = p->apply_ms;
    out5[3] = p->entries;
    out5[4] = p->factor_ms;
    return BEM_OK;
}

int bem_precond_destroy(bem_prec_t p)
{
    if (!p) return BEM_ERR_ARG;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    for (auto& e : p->hlu_ev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    free_list(p->ctx, p->hlu_owned);
    pfree(p->ctx, p->d_blk);
    pfree(p->ctx, p->d_lu);
    pfree(p->ctx, p->d_piv);
    delete p;
    return BEM_OK;
}

int bem_precond_apply(bem_ctx_t c, bem_prec_t p, const void* v, void* z)
{
    if (!c || !p || !v || !z || p->ctx != c || v == z) return set_err(c, BEM_ERR_ARG, "bem_precond_apply: bad arguments");
    HIPC(c, hipSetDevice(c->device));
    if (p->hlu_nleaf > 0) {   // H-LU: z = v, then the two sweeps in place
        if (p->hlu_ev.size() >= 256) hlu_resolve_events(p, false);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPC(c, hipEventCreate(&e0));
        if (hipEventCreate(&e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            return set_err(c, BEM_ERR_HIP, "bem_precond_apply: hipEventCreate");
        }
        p->hlu_ev.push_back({e0, e1});
        HIPC(c, hipEventRecord(e0, c->stream));
        HIPC(c, hipMemcpyAsync(z, v, sizeof(bem_c64) * (size_t)c->n, hipMemcpyDeviceToDevice, c->stream));
        const int64_t np = (int64_t)p->hlu_panel.size();
        for (int d = 0; d < 2; d++)
            for (int64_t i = 0; i < np; i++) {   // panels in the direction's order
                const int64_t q = d == 0 ? i : np - 1 - i;
                const int64_t k0 = p->hlu_panel[q].first, k1 = p->hlu_panel[q].second;
                const dim3 g((unsigned)(k1 - k0)), b(kHluThreads);
                if (p->hlu_has_t[d][q])
                    hipLaunchKernelGGL(k_hlu<1>, g, b, 0, c->stream, d, k0, k1, p->hlu_leaf, p->hlu_tptr[d][1],
                                       p->hlu_titem[d][1], p->hlu_cptr[d][1], p->hlu_citem[d][1], p->hlu_data,
                                       p->hlu_t, p->hlu_gpart, p->hlu_gstride, (bem_c64*)z);
                if (p->hlu_has_c[d][q])
                    hipLaunchKernelGGL(k_hlu<2>, g, b, 0, c->stream, d, k0, k1, p->hlu_leaf, p->hlu_tptr[d][1],
                                       p->hlu_titem[d][1], p->hlu_cptr[d][1], p->hlu_citem[d][1], p->hlu_data,
                                       p->hlu_t, p->hlu_gpart, p->hlu_gstride, (bem_c64*)z);
                hipLaunchKernelGGL(k_hlu<0>, dim3(1), b, 0, c->stream, d, k0, k1, p->hlu_leaf, p->hlu_tptr[d][0],
                                   p->hlu_titem[d][0], p->hlu_cptr[d][0], p->hlu_citem[d][0], p->hlu_data, p->hlu_t,
                                   p->hlu_gpart, p->hlu_gstride, (bem_c64*)z);
            }
        HIPC(c, hipEventRecord(e1, c->stream));
        p->applies++;
        return launch_check(c, "k_hlu");
    }
    p->applies++;   // asynchronous: its time is part of the GMRES iterations
    hipLaunchKernelGGL(k_blu_apply, dim3((unsigned)p->nblk), dim3(64), 0, c->stream, p->d_blk, p->d_lu, p->d_piv,
                       (const bem_c64*)v, (bem_c64*)z);
    return launch_check(c, "k_blu_apply");
}

}  // extern "C"

namespace {
// The operator of a lockstep solve: one H-matrix for all Q columns, or
// (several) operator g for the columns g R .. g R + R - 1; refine: products
// with the operators' refinement form (bem_hmvm_refine)
struct OpSet {
    std::vector<bem_hmat*> ops;
    int R = 0;
    bool refine = false;
    size_t active = 0;   // groups beyond the first `active` are padding (zero right-hand sides); 0: none
};
template <int Q>
int gmres_lockstep(bem_ctx* c, const OpSet& S, const bem_c64* B, bem_c64* X, const double* tol, int64_t m,
                   int64_t maxit, int64_t* its_out, double* rel_out);
}  // namespace

extern "C" {

int bem_gmres(bem_ctx_t c, bem_hmat_t h, const void* b, void* x, double tol, int64_t restart, int64_t maxit,
              int64_t* iters, double* rel_resid)
{
    return bem_gmres_prec(c, h, nullptr, b, x, tol, restart, maxit, iters, rel_resid);
}

int bem_set_precond_side(bem_ctx_t c, int left)
{
    if (!c) return BEM_ERR_ARG;
    c->prec_left = left != 0;
    return BEM_OK;
}

int bem_gmres_prec(bem_ctx_t c, bem_hmat_t h, bem_prec_t pc, const void* b, void* x, double tol, int64_t restart,
                   int64_t maxit, int64_t* iters, double* rel_resid)
{
    if (!c || !h || !b || !x || restart < 1 || maxit < 0 || h->ctx != c || (pc && pc->ctx != c))
        return set_err(c, BEM_ERR_ARG, "bem_gmres: bad arguments");
    HIPC(c, hipSetDevice(c->device));
    if (!pc && c->use_graphs) {   // the same iteration as below, one column per replayed graph
        int64_t its = 0;
        double rel = 0;
        OpSet S;
        S.ops = {h};
        int rc = gmres_lockstep<1>(c, S, (const bem_c64*)b, (bem_c64*)x, &tol, restart, maxit, &its, &rel);
        if (rc) return rc;
        if (iters) *iters = its;
        if (rel_resid) *rel_resid = rel;
        if (rel >= tol) return set_err(c, BEM_ERR_NOCONV, "bem_gmres: maxit reached above tolerance");
        return BEM_OK;
    }
    const int64_t n = c->n;
    const int64_t m = restart;
    const int nchunk = (int)((n + kDotChunk - 1) / kDotChunk);
    // workspace (complex slots): V (m+1) n, w n, r n, h, hacc, y, g, cs_s
    // (m+2 each), H (m+1) m, dot partials (m+1) nchunk, then doubles
    int64_t need = (m + 1) * n + 4 * n + 5 * (m + 2) + (m + 1) * m + (m + 1) * nchunk + (m + 2) + 640;
    if (c->gm_cap < need) {
        pfree(c, c->d_gm);
        c->d_gm = nullptr;
        c->gm_cap = 0;
        int rc;
        if ((rc = dalloc(c, &c->d_gm, need, nullptr))) return rc;
        c->gm_cap = need;
    }
    bem_c64* V = c->d_gm;
    bem_c64* w = V + (m + 1) * n;
    bem_c64* r = w + n;
    bem_c64* dh = r + n;
    bem_c64* dhacc = dh + (m + 2);
    bem_c64* dy = dhacc + (m + 2);
    bem_c64* dg = dy + (m + 2);
    bem_c64* dcs_s = dg + (m + 2);
    bem_c64* dH = dcs_s + (m + 2);
    bem_c64* dpart = dH + (m + 1) * m;
    bem_c64* zp = dpart + (m + 1) * nchunk;   // preconditioned vector
    bem_c64* pb = zp + n;                      // M^-1 b (left preconditioning)
    double* dcs_c = (double*)(pb + n);
    double* dn = dcs_c + 2 * (m + 2);      // norm partials (256) + result + residual estimate
    double* dnorm = dn + 512;
    double* dres = dnorm + 1;
    const bem_c64* bd = (const bem_c64*)b;
    bem_c64* xd = (bem_c64*)x;
    const int npart = (int)std::min<int64_t>(256, (n + kWG - 1) / kWG);
    auto norm_dev = [&](const bem_c64* v) {
        hipLaunchKernelGGL(k_norm2_part, dim3(npart), dim3(kWG), 0, c->stream, n, v, dn);
        hipLaunchKernelGGL(k_norm2_final, dim3(1), dim3(64), 0, c->stream, npart, dn, dnorm);
    };
    auto fetch = [&](double* dst, const double* src) -> int {
        HIPC(c, hipMemcpyAsync(dst, src, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        return BEM_OK;
    };
    int rc;
    // left preconditioning (gmresLUPreconditionedSolve, gmres.h:200-216): GMRES
    // on M^-1 A x = M^-1 b, its residual and tolerance those of that system
    const bool left = pc && c->prec_left;
    double bn = 0;
    if (left && (rc = bem_precond_apply(c, pc, bd, pb))) return rc;
    norm_dev(left ? pb : bd);
    if ((rc = launch_check(c, "norm")) || (rc = fetch(&bn, dnorm))) return rc;
    if (bn == 0) {
        HIPC(c, hipMemsetAsync(xd, 0, n * sizeof(bem_c64), c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        if (iters) *iters = 0;
        if (rel_resid) *rel_resid = 0;
        return BEM_OK;
    }
    std::vector<bem_c64> H((m + 1) * m), g(m + 2);
    int64_t it = 0;
    double rel = 1;
    while (true) {
        // r = b - A x (left: M^-1 (b - A x)), beta = |r|
        if ((rc = bem_hmvm(c, h, xd, r, 0))) return rc;
        hipLaunchKernelGGL(k_residual, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, bd, r);
        const bem_c64* rr = r;
        if (left) {
            if ((rc = bem_precond_apply(c, pc, r, zp))) return rc;
            rr = zp;
        }
        norm_dev(rr);
        double beta;
        if ((rc = launch_check(c, "residual")) || (rc = fetch(&beta, dnorm))) return rc;
        rel = beta / bn;
        if (rel < tol || it >= maxit) break;
        hipLaunchKernelGGL(k_scale_by_norm, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, rr, dnorm, V);
        std::fill(g.begin(), g.end(), c64(0, 0));
        g[0] = c64(beta, 0);
        HIPC(c, hipMemcpyAsync(dg, g.data(), (m + 2) * sizeof(bem_c64), hipMemcpyHostToDevice, c->stream));
        // Arnoldi column jj (basis vector jj + 1, Givens rotation jj); its
        // residual estimate is copied to the pinned ring slot jj & 1
        auto launch_col = [&](int64_t jj) -> int {
            int rc2;
            if (left) {   // w = M^-1 A v_jj
                if ((rc2 = bem_hmvm(c, h, V + jj * n, zp, 0))) return rc2;
                if ((rc2 = bem_precond_apply(c, pc, zp, w))) return rc2;
            } else if (pc) {   // w = A M^-1 v_jj
                if ((rc2 = bem_precond_apply(c, pc, V + jj * n, zp))) return rc2;
                if ((rc2 = bem_hmvm(c, h, zp, w, 0))) return rc2;
            } else if ((rc2 = bem_hmvm(c, h, V + jj * n, w, 0))) return rc2;
            const int cnt = (int)(jj + 1);
            for (int pass = 0; pass < 2; pass++) {   // classical Gram-Schmidt, twice
                const int ngroups = (cnt + kDotVecs - 1) / kDotVecs;
                hipLaunchKernelGGL(k_cgs_dots, dim3(nchunk, (ngroups + kWG / 64 - 1) / (kWG / 64)), dim3(kWG), 0,
                                   c->stream, n, cnt, V, w, dpart);
                hipLaunchKernelGGL(k_cgs_reduce, dim3(cnt), dim3(64), 0, c->stream, cnt, nchunk, dpart, dh,
                                   dhacc, pass == 0 ? 1 : 0);
                hipLaunchKernelGGL(k_orth_update, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, cnt, V, dh, w);
            }
            norm_dev(w);
            hipLaunchKernelGGL(k_scale_by_norm, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, w, dnorm,
                               V + (jj + 1) * n);
            hipLaunchKernelGGL(k_arnoldi, dim3(1), dim3(64), 0, c->stream, (int)jj, (int)m, dhacc, dnorm, dH, dcs_c,
                               dcs_s, dg, bn, dres);
            if ((rc2 = launch_check(c, "arnoldi"))) return rc2;
            HIPC(c, hipMemcpyAsync(&c->gm_res[jj & 1], dres, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPC(c, hipEventRecord(c->gm_ev[jj & 1], c->stream));
            return BEM_OK;
        };
        // One column is kept in flight ahead of the convergence check, so the
        // host round trip overlaps device work.  A column launched after the
        // converged one only writes basis vector jj + 1, column jj of H and
        // entries jj, jj + 1 of g, none of which the solution of the first
        // jj columns reads: the result is the same as without look-ahead.
        const int64_t cap = std::min<int64_t>(m, maxit - it);
        int64_t j = 0, nl = 0;
        if (cap > 0) {
            if ((rc = launch_col(0))) return rc;
            nl = 1;
        }
        while (j < nl) {
            if (nl < cap) {
                if ((rc = launch_col(nl))) return rc;
                nl++;
            }
            HIPC(c, hipEventSynchronize(c->gm_ev[j & 1]));
            rel = c->gm_res[j & 1];
            j++;
            if (rel < tol) break;
        }
        it += j;
        // y = R^-1 g on the host, x += V y
        HIPC(c, hipMemcpyAsync(H.data(), dH, (m + 1) * j * sizeof(bem_c64), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipMemcpyAsync(g.data(), dg, (j + 1) * sizeof(bem_c64), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        std::vector<bem_c64> y(j);
        for (int64_t l = j - 1; l >= 0; l--) {
            bem_c64 s = g[l];
            for (int64_t q = l + 1; q < j; q++) s = c64_sub(s, c64_mul(H[q * (m + 1) + l], y[q]));
            y[l] = c64_div(s, H[l * (m + 1) + l]);
        }
        if (j > 0) {
            HIPC(c, hipMemcpyAsync(dy, y.data(), j * sizeof(bem_c64), hipMemcpyHostToDevice, c->stream));
            if (pc && !left) {   // x += M^-1 (V y); w is free here
                hipLaunchKernelGGL(k_combine, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, (int)j, V, dy, w, 0);
                if ((rc = bem_precond_apply(c, pc, w, zp))) return rc;
                hipLaunchKernelGGL(k_axpy1, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, zp, xd);
            } else {
                hipLaunchKernelGGL(k_combine, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, (int)j, V, dy, xd, 1);
            }
            if ((rc = launch_check(c, "k_combine"))) return rc;
        }
    }
    HIPC(c, hipStreamSynchronize(c->stream));
    if (iters) *iters = it;
    if (rel_resid) *rel_resid = rel;
    if (rel >= tol) return set_err(c, BEM_ERR_NOCONV, "bem_gmres: maxit reached above tolerance");
    return BEM_OK;
}

}  // extern "C"

namespace {

int pow2_q(int64_t nrhs) { return nrhs <= 1 ? 1 : nrhs <= 2 ? 2 : nrhs <= 4 ? 4 : 8; }

// Q-interleaved partial vectors (context) of at least Q right-hand sides
int ensure_partq(bem_ctx* c, int Q)
{
    int rc;
    if (c->partq_cap < Q) {
        pfree(c, c->d_partq);
        c->d_partq = nullptr;
        c->partq_cap = 0;
        pfree(c, c->d_xq);
        c->d_xq = nullptr;
        if ((rc = dalloc(c, &c->d_partq, (size_t)std::max<int64_t>(c->part_total, 1) * Q, nullptr))) return rc;
        if ((rc = dalloc(c, &c->d_xq, (size_t)c->n * Q, nullptr))) return rc;
        c->partq_cap = Q;
    }
    return BEM_OK;
}

// ... and the operator's t vectors of at least Q right-hand sides
int ensure_multi(bem_ctx* c, bem_hmat* h, int Q)
{
    int rc;
    if ((rc = ensure_partq(c, Q))) return rc;
    if (h->tq_cap < Q) {
        pfree(c, h->d_tq);
        h->d_tq = nullptr;
        h->tq_cap = 0;
        if ((rc = dalloc(c, &h->d_tq, (size_t)std::max<int64_t>(h->t_total, 1) * Q, nullptr))) return rc;
        h->tq_cap = Q;
    }
    return BEM_OK;
}

// Y (+)= H X for Q interleaved right-hand sides; DS: storage of the near
// field read (dense), LS: storage of the low-rank factors
template <class DS, class LS, int Q>
int hmvm_q(bem_ctx* c, bem_hmat* h, const void* dense, const bem_c64* X, bem_c64* Y, int accumulate)
{
    int rc;
    bem_c64* part = Q == 1 ? c->d_part : c->d_partq;
    bem_c64* tv = Q == 1 ? h->d_t : h->d_tq;
    if (Q > 1 && (rc = ensure_multi(c, h, Q))) return rc;
    if (Q > 1) {
        part = c->d_partq;
        tv = h->d_tq;
    }
    const bool mf = c->use_mfma && Q >= 4;
    if (!c->tiles.empty()) {
        Timer t(c, 3);
        const dim3 g((unsigned)c->tiles.size());
        if (Q == 1)
            hipLaunchKernelGGL(k_mvm_dense<DS>, g, dim3(kWG), 0, c->stream, c->d_tiles, (const DS*)dense, X, part);
        else if (mf)
            hipLaunchKernelGGL((k_mvm_dense_mfma<DS, Q>), g, dim3(kWG), 0, c->stream, c->d_tiles, (const DS*)dense, X,
                               part);
        else
            hipLaunchKernelGGL((k_mvm_dense_q<DS, Q>), g, dim3(kWG), 0, c->stream, c->d_tiles, (const DS*)dense, X,
                               part);
        if ((rc = launch_check(c, "k_mvm_dense"))) return rc;
    }
    if (h->nttask > 0) {
        Timer t(c, 4);
        const dim3 gt(nblk(h->nttask, kWG / 64)), gu((unsigned)h->nutask);
        if (Q == 1) {
            hipLaunchKernelGGL(k_lr_t<LS>, gt, dim3(kWG), 0, c->stream, h->d_ttask, h->nttask, c->d_col0, c->d_cols,
                               h->d_rank, h->d_voff, h->d_aoff, h->d_toff, (const LS*)h->V, (const LS*)h->A, X, tv);
            hipLaunchKernelGGL(k_lr_u<LS>, gu, dim3(kWG), h->maxrank * sizeof(bem_c64), c->stream, h->d_utask,
                               c->d_rows, h->d_rank, h->d_uoff, h->d_toff, c->d_lr_part, (const LS*)h->U, tv, part);
        } else {
            hipLaunchKernelGGL((k_lr_t_q<LS, Q>), gt, dim3(kWG), 0, c->stream, h->d_ttask, h->nttask, c->d_col0,
                               c->d_cols, h->d_rank, h->d_voff, h->d_aoff, h->d_toff, (const LS*)h->V,
                               (const LS*)h->A, X, tv);
            if (mf)
                hipLaunchKernelGGL((k_lr_u_mfma<LS, Q>), gu, dim3(kWG), 0, c->stream, h->d_utask, c->d_rows, h->d_rank,
                                   h->d_uoff, h->d_toff, c->d_lr_part, (const LS*)h->U, tv, part);
            else
                hipLaunchKernelGGL((k_lr_u_q<LS, Q>), gu, dim3(kWG), 0, c->stream, h->d_utask, c->d_rows, h->d_rank,
                                   h->d_uoff, h->d_toff, c->d_lr_part, (const LS*)h->U, tv, part);
        }
        if ((rc = launch_check(c, "k_lr"))) return rc;
    }
    Timer t(c, 5);
    if (Q == 1)
        hipLaunchKernelGGL(k_mvm_gather, dim3(nblk(16 * c->n, 256)), dim3(256), 0, c->stream, c->n, c->d_gptr,
                           c->d_gidx, part, Y, accumulate);
    else
        hipLaunchKernelGGL(k_mvm_gather_q<Q>, dim3(nblk(16 * c->n, 256)), dim3(256), 0, c->stream, c->n, c->d_gptr,
                           c->d_gidx, part, Y, accumulate);
    return launch_check(c, "k_mvm_gather");
}

// the product with the operator's own storage (refine = false) or with its
// refinement operator (FP32 blocks with their FP64 near field)
template <int Q>
int hmvm_multi(bem_ctx* c, bem_hmat* h, bool refine, const bem_c64* X, bem_c64* Y, int accumulate)
{
    if (h->prec == BEM_PREC_FP64) return hmvm_q<bem_c64, bem_c64, Q>(c, h, h->dense, X, Y, accumulate);
    if (refine) {
        if (!h->dense64)
            return set_err(c, BEM_ERR_STATE, "refinement product: FP32 operator without an FP64 near field");
        return hmvm_q<bem_c64, bem_c32, Q>(c, h, h->dense64, X, Y, accumulate);
    }
    return hmvm_q<bem_c32, bem_c32, Q>(c, h, h->dense, X, Y, accumulate);
}

int hmvm_multi_any(bem_ctx* c, bem_hmat* h, int Q, bool refine, const bem_c64* X, bem_c64* Y, int acc);

// buffers of every operator of S for products of its column count (and
// the context's Q-wide partial vectors of the fused product)
int ensure_ops(bem_ctx* c, const OpSet& S, int Q)
{
    int rc;
    if (S.ops.size() > 1 && Q > 1 && (rc = ensure_partq(c, Q))) return rc;
    const int w = S.ops.size() == 1 ? Q : S.R;
    if (w == 1) return BEM_OK;
    for (bem_hmat* h : S.ops)
        if ((rc = ensure_multi(c, h, w))) return rc;
    return BEM_OK;
}

// Y = A X with every operator of S in one pass (k_*_ops): DS the near-field
// storage read, LS the low-rank storage; S.ops.size() R == Q
template <class DS, class LS, int R, int Q>
int hmvm_ops(bem_ctx* c, const OpSet& S, const bem_c64* X, bem_c64* Y)
{
    const int G = (int)S.ops.size();
    OpsArg O{};
    O.n = c->n;
    O.P = std::max<int64_t>(c->part_total, 1);
    int64_t mtt = 0, mut = 0;
    int maxrank = 0;
    for (int g = 0; g < G; g++) {
        const bem_hmat* h = S.ops[g];
        O.pad[g] = S.active && (size_t)g >= S.active;
        O.F[g] = sizeof(DS) == sizeof(bem_c64) && h->prec == BEM_PREC_FP32 ? h->dense64 : h->dense;
        O.tt[g] = h->d_ttask;
        O.ut[g] = h->d_utask;
        O.ntt[g] = O.pad[g] ? 0 : h->nttask;
        O.nut[g] = h->nutask;
        O.rank[g] = h->d_rank;
        O.voff[g] = h->d_voff;
        O.aoff[g] = h->d_aoff;
        O.toff[g] = h->d_toff;
        O.uoff[g] = h->d_uoff;
        O.U[g] = h->U;
        O.V[g] = h->V;
        O.A[g] = h->A;
        O.t[g] = R == 1 ? h->d_t : h->d_tq;
        mtt = std::max(mtt, h->nttask);
        mut = std::max(mut, h->nutask);
        maxrank = std::max(maxrank, h->maxrank);
    }
    int rc;
    hipLaunchKernelGGL((k_group_cols<R, Q>), dim3(nblk(c->n * Q, 256)), dim3(256), 0, c->stream, c->n, X, c->d_xq);
    if (!c->tiles.empty()) {
        Timer t(c, 3);
        hipLaunchKernelGGL((k_mvm_dense_ops<DS, R>), dim3((unsigned)c->tiles.size(), G), dim3(kWG), 0, c->stream,
                           c->d_tiles, O, c->d_xq, c->d_partq);
        if ((rc = launch_check(c, "k_mvm_dense_ops"))) return rc;
    }
    if (mut > 0) {
        Timer t(c, 4);
        if (mtt > 0)
            hipLaunchKernelGGL((k_lr_t_ops<LS, R>), dim3(nblk(mtt, kWG / 64), G), dim3(kWG), 0, c->stream, O,
                               c->d_col0, c->d_cols, c->d_xq);
        hipLaunchKernelGGL((k_lr_u_ops<LS, R>), dim3((unsigned)mut, G), dim3(kWG),
                           R == 1 ? maxrank * sizeof(bem_c64) : 0, c->stream, O, c->d_rows, c->d_lr_part, c->d_partq);
        if ((rc = launch_check(c, "k_lr_ops"))) return rc;
    }
    Timer t(c, 5);
    hipLaunchKernelGGL((k_mvm_gather_ops<R, Q>), dim3(nblk(16 * c->n, 256)), dim3(256), 0, c->stream, c->n, O.P,
                       c->d_gptr, c->d_gidx, c->d_partq, Y);
    return launch_check(c, "k_mvm_gather_ops");
}

// whether S's operators take the fused product: at most kMaxOps of one
// storage precision (and, for the refinement form of FP32 ones, each with
// its FP64 near field), no operator twice among the active groups (each
// group writes its operator's t vectors)
bool fusable(const OpSet& S)
{
    if (S.ops.size() < 2 || S.ops.size() > (size_t)kMaxOps) return false;
    const size_t na = S.active ? S.active : S.ops.size();
    for (size_t g = 0; g < S.ops.size(); g++) {
        const bem_hmat* h = S.ops[g];
        if (h->prec != S.ops[0]->prec) return false;
        if (S.refine && h->prec == BEM_PREC_FP32 && !h->dense64) return false;
        for (size_t e = 0; e < g && g < na; e++)
            if (S.ops[e] == h) return false;
    }
    return true;
}

template <int R, int Q>
int hmvm_ops_prec(bem_ctx* c, const OpSet& S, const bem_c64* X, bem_c64* Y)
{
    if (S.ops[0]->prec == BEM_PREC_FP64) return hmvm_ops<bem_c64, bem_c64, R, Q>(c, S, X, Y);
    if (S.refine) return hmvm_ops<bem_c64, bem_c32, R, Q>(c, S, X, Y);
    return hmvm_ops<bem_c32, bem_c32, R, Q>(c, S, X, Y);
}

template <int Q>
int hmvm_ops_any(bem_ctx* c, const OpSet& S, const bem_c64* X, bem_c64* Y)
{
    if constexpr (Q >= 2) {
        if (S.R == 1) return hmvm_ops_prec<1, Q>(c, S, X, Y);
    }
    if constexpr (Q >= 4) {
        if (S.R == 2) return hmvm_ops_prec<2, Q>(c, S, X, Y);
    }
    if constexpr (Q >= 8) {
        if (S.R == 4) return hmvm_ops_prec<4, Q>(c, S, X, Y);
    }
    return set_err(c, BEM_ERR_ARG, "fused product: unsupported column split");
}

// Y = A X for Q interleaved columns: one product of the single operator, or
// per operator g its R columns gathered into tx, multiplied into ty and
// scattered back (tx, ty: n R each, allocated by the caller)
template <int Q>
int apply_ops(bem_ctx* c, const OpSet& S, const bem_c64* X, bem_c64* Y, bem_c64* tx, bem_c64* ty)
{
    if (S.ops.size() == 1) return hmvm_multi<Q>(c, S.ops[0], S.refine, X, Y, 0);
    if (c->fuse_ops && fusable(S) && (int)S.ops.size() * S.R == Q) return hmvm_ops_any<Q>(c, S, X, Y);
    int rc;
    const int R = S.R;
    for (size_t g = 0; g < S.ops.size(); g++) {
        hipLaunchKernelGGL(k_cols, dim3(nblk(c->n * R, 256)), dim3(256), 0, c->stream, c->n, R, X, Q, (int)(g * R), tx,
                           R, 0);
        if ((rc = hmvm_multi_any(c, S.ops[g], R, S.refine, tx, ty, 0))) return rc;
        hipLaunchKernelGGL(k_cols, dim3(nblk(c->n * R, 256)), dim3(256), 0, c->stream, c->n, R, (const bem_c64*)ty, R,
                           0, Y, Q, (int)(g * R));
        if ((rc = launch_check(c, "k_cols"))) return rc;
    }
    return BEM_OK;
}

// |v_q| of Q interleaved vectors into nrm[0 .. Q) (device), using part
template <int Q>
int norms_q(bem_ctx* c, const bem_c64* v, double* part, double* nrm)
{
    const int npart = (int)std::min<int64_t>(256, (c->n + kWG - 1) / kWG);
    hipLaunchKernelGGL(k_norm2_part_q<Q>, dim3(npart), dim3(kWG), 0, c->stream, c->n, v, part);
    hipLaunchKernelGGL(k_norm2_final_q<Q>, dim3(1), dim3(64), 0, c->stream, npart, part, nrm);
    return launch_check(c, "norms");
}

// Restarted GMRES(m) in lockstep over Q interleaved right-hand sides B
// (start and solution X, in place): one product of the operator per Arnoldi
// column serves all Q; CGS2, Givens rotations and the residual estimates per
// right-hand side on the device as bem_gmres_prec.  Right-hand side q stops
// at the first column whose estimate is below tol[q] (its later columns are
// computed but not used); every restart recomputes b_q - A x_q.  With
// graphs on, one column's launches are captured once per call and replayed.
template <int Q>
int gmres_lockstep(bem_ctx* c, const OpSet& S, const bem_c64* B, bem_c64* X, const double* tol, int64_t m,
                   int64_t maxit, int64_t* its_out, double* rel_out)
{
    const int64_t n = c->n;
    const int nchunk = (int)((n + kDotChunk - 1) / kDotChunk);
    const int npart = (int)std::min<int64_t>(256, (n + kWG - 1) / kWG);
    std::vector<void*> own;
    auto fail = [&](int code) {
        free_list(c, own);
        return code;
    };
    bem_c64 *Vb, *w, *r, *xb, *dh, *dhacc, *dH, *dg, *dss, *dy, *dpart;
    double *dcc, *dnp, *dnrm, *dbn, *dres;
    int* jp;
    int rc;
    // buffers of the product exist before a graph is captured (no allocation inside)
    if ((rc = ensure_ops(c, S, Q))) return rc;
    bem_c64 *tx = nullptr, *ty = nullptr;
    if (S.ops.size() > 1 && ((rc = dalloc(c, &tx, (size_t)(n * S.R), &own)) || (rc = dalloc(c, &ty, (size_t)(n * S.R), &own))))
        return fail(rc);
    if ((rc = dalloc(c, &Vb, (size_t)((m + 1) * n * Q), &own)) || (rc = dalloc(c, &w, (size_t)(n * Q), &own)) ||
        (rc = dalloc(c, &r, (size_t)(n * Q), &own)) || (rc = dalloc(c, &xb, (size_t)(n * Q), &own)) ||
        (rc = dalloc(c, &dh, (size_t)((m + 2) * Q), &own)) || (rc = dalloc(c, &dhacc, (size_t)((m + 2) * Q), &own)) ||
        (rc = dalloc(c, &dH, (size_t)(Q * (m + 1) * m), &own)) || (rc = dalloc(c, &dg, (size_t)(Q * (m + 2)), &own)) ||
        (rc = dalloc(c, &dss, (size_t)(Q * (m + 2)), &own)) || (rc = dalloc(c, &dy, (size_t)(m * Q + 1), &own)) ||
        (rc = dalloc(c, &dpart, (size_t)((m + 1) * nchunk * Q), &own)) ||
        (rc = dalloc(c, &dcc, (size_t)(Q * (m + 2)), &own)) || (rc = dalloc(c, &dnp, (size_t)(npart * Q), &own)) ||
        (rc = dalloc(c, &dnrm, (size_t)Q, &own)) || (rc = dalloc(c, &dbn, (size_t)Q, &own)) ||
        (rc = dalloc(c, &dres, (size_t)Q, &own)) || (rc = dalloc(c, &jp, (size_t)1, &own)))
        return fail(rc);
    double bnh[Q], beta[Q];
    auto fetch = [&](double* dst, const double* src, int cnt) -> int {
        HIPC(c, hipMemcpyAsync(dst, src, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        return BEM_OK;
    };
    if ((rc = norms_q<Q>(c, B, dnp, dbn)) || (rc = fetch(bnh, dbn, Q))) return fail(rc);
    // one Arnoldi column for every right-hand side (column *jp)
    const int ng = (int)(((m + 1) + std::max(1, kDotVecs / Q) - 1) / std::max(1, kDotVecs / Q));
    auto iteration = [&]() -> int {
        int rc2;
        hipLaunchKernelGGL(k_basis_col<Q>, dim3(nblk(n * Q, 256)), dim3(256), 0, c->stream, n, jp, Vb, xb);
        if ((rc2 = apply_ops<Q>(c, S, xb, w, tx, ty))) return rc2;
        for (int pass = 0; pass < 2; pass++) {   // classical Gram-Schmidt, twice
            hipLaunchKernelGGL(k_cgs_dots_q<Q>, dim3(nchunk, (ng + kWG / 64 - 1) / (kWG / 64)), dim3(kWG), 0,
                               c->stream, n, jp, Vb, w, dpart);
            hipLaunchKernelGGL(k_cgs_reduce_q<Q>, dim3((unsigned)(m + 1)), dim3(64), 0, c->stream, jp, nchunk, dpart,
                               dh, dhacc, pass == 0 ? 1 : 0);
            hipLaunchKernelGGL(k_orth_update_q<Q>, dim3(nblk(n * Q, 256)), dim3(256), 0, c->stream, n, jp, Vb, dh, w);
        }
        if ((rc2 = norms_q<Q>(c, w, dnp, dnrm))) return rc2;
        hipLaunchKernelGGL(k_scale_next_q<Q>, dim3(nblk(n * Q, 256)), dim3(256), 0, c->stream, n, jp, w, dnrm, Vb);
        hipLaunchKernelGGL(k_arnoldi_q<Q>, dim3(1), dim3(64), 0, c->stream, jp, (int)m, dhacc, dnrm, dH, dcc, dss, dg,
                           dbn, dres);
        return launch_check(c, "arnoldi");
    };
    hipGraphExec_t gexec = nullptr;
    auto cleanup = [&](int code) {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        return fail(code);
    };
    if (c->use_graphs) {
        hipGraph_t graph = nullptr;
        c->capturing = true;
        hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        int rc2 = e == hipSuccess ? iteration() : BEM_OK;
        hipError_t e2 = e == hipSuccess ? hipStreamEndCapture(c->stream, &graph) : e;
        c->capturing = false;
        if (e != hipSuccess || e2 != hipSuccess || rc2) {
            if (graph) (void)hipGraphDestroy(graph);
            return cleanup(rc2 ? rc2 : set_err(c, BEM_ERR_HIP, std::string("GMRES graph capture: ") +
                                                                   hipGetErrorString(e != hipSuccess ? e : e2)));
        }
        e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            gexec = nullptr;
            return cleanup(set_err(c, BEM_ERR_HIP, std::string("GMRES graph instantiate: ") + hipGetErrorString(e)));
        }
    }
    auto launch_col = [&](int64_t jj) -> int {
        if (gexec) {
            HIPC(c, hipGraphLaunch(gexec, c->stream));
        } else {
            int rc2 = iteration();
            if (rc2) return rc2;
        }
        HIPC(c, hipMemcpyAsync(c->gm_resq + (jj & 1) * kArnMax, dres, Q * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
        HIPC(c, hipEventRecord(c->gm_ev[jj & 1], c->stream));
        return BEM_OK;
    };
    bool done[Q];
    int64_t its[Q];
    double rel[Q];
    for (int q = 0; q < Q; q++) {
        done[q] = !(bnh[q] > 0);
        its[q] = 0;
        rel[q] = 0;
    }
    std::vector<bem_c64> H((size_t)Q * (m + 1) * m), g((size_t)Q * (m + 2)), y((size_t)m * Q + 1);
    int64_t it = 0;
    while (true) {
        // r = b - A x per right-hand side
        if ((rc = apply_ops<Q>(c, S, X, r, tx, ty))) return cleanup(rc);
        hipLaunchKernelGGL(k_residual_n, dim3(nblk(n * Q, 256)), dim3(256), 0, c->stream, n * Q, B, r);
        if ((rc = norms_q<Q>(c, r, dnp, dnrm)) || (rc = fetch(beta, dnrm, Q))) return cleanup(rc);
        bool all = true;
        for (int q = 0; q < Q; q++) {
            rel[q] = bnh[q] > 0 ? beta[q] / bnh[q] : 0.0;
            if (rel[q] < tol[q]) done[q] = true;
            all = all && done[q];
        }
        if (all || it >= maxit) break;
        hipLaunchKernelGGL(k_scale_first_q<Q>, dim3(nblk(n * Q, 256)), dim3(256), 0, c->stream, n, r, dnrm, Vb);
        std::fill(g.begin(), g.end(), c64(0, 0));
        for (int q = 0; q < Q; q++) g[(size_t)q * (m + 2)] = c64(beta[q], 0);
        const int zero = 0;
        HIPC(c, hipMemcpyAsync(dg, g.data(), g.size() * sizeof(bem_c64), hipMemcpyHostToDevice, c->stream));
        HIPC(c, hipMemcpyAsync(jp, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
        // one column kept in flight ahead of the convergence check (the
        // columns after a right-hand side's last one are not used for it)
        const int64_t cap = std::min<int64_t>(m, maxit - it);
        int64_t conv[Q];
        for (int q = 0; q < Q; q++) conv[q] = -1;
        int64_t j = 0, nl = 0;
        if ((rc = launch_col(0))) return cleanup(rc);
        nl = 1;
        while (j < nl) {
            if (nl < cap) {
                if ((rc = launch_col(nl))) return cleanup(rc);
                nl++;
            }
            HIPC(c, hipEventSynchronize(c->gm_ev[j & 1]));
            const double* rs = c->gm_resq + (j & 1) * kArnMax;
            bool allc = true;
            for (int q = 0; q < Q; q++) {
                if (!done[q] && conv[q] < 0 && rs[q] < tol[q]) conv[q] = j;
                allc = allc && (done[q] || conv[q] >= 0);
            }
            j++;
            if (allc) break;
        }
        HIPC(c, hipStreamSynchronize(c->stream));
        it += j;
        HIPC(c, hipMemcpyAsync(H.data(), dH, H.size() * sizeof(bem_c64), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipMemcpyAsync(g.data(), dg, g.size() * sizeof(bem_c64), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        std::fill(y.begin(), y.end(), c64(0, 0));
        for (int q = 0; q < Q; q++) {
            const int64_t jq = done[q] ? 0 : (conv[q] >= 0 ? conv[q] + 1 : j);
            its[q] += jq;
            const bem_c64* Hq = H.data() + (size_t)q * m * (m + 1);
            const bem_c64* gq = g.data() + (size_t)q * (m + 2);
            for (int64_t l = jq - 1; l >= 0; l--) {   // y = R^-1 g
                bem_c64 s = gq[l];
                for (int64_t k2 = l + 1; k2 < jq; k2++) s = c64_sub(s, c64_mul(Hq[k2 * (m + 1) + l], y[k2 * Q + q]));
                y[l * Q + q] = c64_div(s, Hq[l * (m + 1) + l]);
            }
        }
        HIPC(c, hipMemcpyAsync(dy, y.data(), (size_t)j * Q * sizeof(bem_c64), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_combine_q<Q>, dim3(nblk(n * Q, 256)), dim3(256), 0, c->stream, n, (int)j, Vb, dy, X);
        if ((rc = launch_check(c, "k_combine_q"))) return cleanup(rc);
    }
    HIPC(c, hipStreamSynchronize(c->stream));
    for (int q = 0; q < Q; q++) {
        if (its_out) its_out[q] = its[q];
        if (rel_out) rel_out[q] = rel[q];
    }
    return cleanup(BEM_OK);
}

// n x nq (row stride ld) <-> n x Q interleaved copies around the Q-templated code
int repack(bem_ctx* c, int64_t nq, int64_t ld, int Q, const bem_c64* src, bem_c64* dst)
{
    hipLaunchKernelGGL(k_repack, dim3(nblk(c->n * Q, 256)), dim3(256), 0, c->stream, c->n, (int)nq, ld, Q, src, dst);
    return launch_check(c, "k_repack");
}
int unpack(bem_ctx* c, int64_t nq, int64_t ld, int Q, const bem_c64* src, bem_c64* dst)
{
    hipLaunchKernelGGL(k_repack, dim3(nblk(c->n * nq, 256)), dim3(256), 0, c->stream, c->n, (int)Q, (int64_t)Q,
                       (int)nq, src, dst);
    (void)ld;
    return launch_check(c, "k_repack");
}

int gmres_lockstep_any(bem_ctx* c, const OpSet& S, int Q, const bem_c64* B, bem_c64* X, const double* tol, int64_t m,
                       int64_t maxit, int64_t* its, double* rel)
{
    switch (Q) {
    case 1: return gmres_lockstep<1>(c, S, B, X, tol, m, maxit, its, rel);
    case 2: return gmres_lockstep<2>(c, S, B, X, tol, m, maxit, its, rel);
    case 4: return gmres_lockstep<4>(c, S, B, X, tol, m, maxit, its, rel);
    default: return gmres_lockstep<8>(c, S, B, X, tol, m, maxit, its, rel);
    }
}

int apply_ops_any(bem_ctx* c, const OpSet& S, int Q, const bem_c64* X, bem_c64* Y, bem_c64* tx, bem_c64* ty)
{
    switch (Q) {
    case 1: return apply_ops<1>(c, S, X, Y, tx, ty);
    case 2: return apply_ops<2>(c, S, X, Y, tx, ty);
    case 4: return apply_ops<4>(c, S, X, Y, tx, ty);
    default: return apply_ops<8>(c, S, X, Y, tx, ty);
    }
}

int hmvm_multi_any(bem_ctx* c, bem_hmat* h, int Q, bool refine, const bem_c64* X, bem_c64* Y, int acc)
{
    switch (Q) {
    case 1: return hmvm_multi<1>(c, h, refine, X, Y, acc);
    case 2: return hmvm_multi<2>(c, h, refine, X, Y, acc);
    case 4: return hmvm_multi<4>(c, h, refine, X, Y, acc);
    default: return hmvm_multi<8>(c, h, refine, X, Y, acc);
    }
}

int norms_any(bem_ctx* c, int Q, const bem_c64* v, double* part, double* nrm)
{
    switch (Q) {
    case 1: return norms_q<1>(c, v, part, nrm);
    case 2: return norms_q<2>(c, v, part, nrm);
    case 4: return norms_q<4>(c, v, part, nrm);
    default: return norms_q<8>(c, v, part, nrm);
    }
}

// Iterative refinement over Q interleaved right-hand sides (the loop of
// bem_gmres_ir): r = b - A_res x in FP64, inner lockstep GMRES on h with
// the per-right-hand-side relative tolerances min(0.5, 0.5 tol / rel_q)
int gmres_ir_q(bem_ctx* c, const OpSet& inner, const OpSet& res, int Q, const bem_c64* B, bem_c64* X, bool x_is_zero,
               double tol, int64_t restart, int64_t maxit, int64_t max_steps, int64_t* its_out, double* rel_out,
               int64_t* steps_out)
{
    const int64_t n = c->n;
    std::vector<void*> own;
    bem_c64 *r, *d, *tx = nullptr, *ty = nullptr;
    double *np, *nrm;
    int rc;
    auto done = [&](int code) {
        free_list(c, own);
        return code;
    };
    if ((rc = dalloc(c, &r, (size_t)(n * Q), &own)) || (rc = dalloc(c, &d, (size_t)(n * Q), &own)) ||
        (rc = dalloc(c, &np, (size_t)(256 * Q), &own)) || (rc = dalloc(c, &nrm, (size_t)Q, &own)))
        return done(rc);
    if (res.ops.size() > 1 &&
        ((rc = dalloc(c, &tx, (size_t)(n * res.R), &own)) || (rc = dalloc(c, &ty, (size_t)(n * res.R), &own))))
        return done(rc);
    std::vector<double> bn(Q), rn(Q), rel(Q, 0.0), tq(Q);
    auto fetch = [&](double* dst) -> int {
        HIPC(c, hipMemcpyAsync(dst, nrm, Q * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        return BEM_OK;
    };
    if ((rc = norms_any(c, Q, B, np, nrm)) || (rc = fetch(bn.data()))) return done(rc);
    std::vector<int64_t> its(Q, 0), inner_its(Q);
    int64_t st = 0, total = 0;
    bool zero = x_is_zero;
    if (zero) {
        hipLaunchKernelGGL(k_fill0, dim3(nblk(n * Q, 256)), dim3(256), 0, c->stream, n * Q, X);
        if ((rc = launch_check(c, "k_fill0"))) return done(rc);
    }
    while (true) {
        if (zero) {
            HIPC(c, hipMemcpyAsync(r, B, (size_t)(n * Q) * sizeof(bem_c64), hipMemcpyDeviceToDevice, c->stream));
        } else {
            if ((rc = apply_ops_any(c, res, Q, X, r, tx, ty))) return done(rc);
            hipLaunchKernelGGL(k_residual_n, dim3(nblk(n * Q, 256)), dim3(256), 0, c->stream, n * Q, B, r);
        }
        if ((rc = norms_any(c, Q, r, np, nrm)) || (rc = fetch(rn.data()))) return done(rc);
        bool all = true;
        for (int q = 0; q < Q; q++) {
            rel[q] = bn[q] > 0 ? rn[q] / bn[q] : 0.0;
            // converged right-hand sides get an inner tolerance of 1: no columns
            tq[q] = rel[q] < tol ? 1.0 : std::min(0.5, 0.5 * tol / rel[q]);
            all = all && rel[q] < tol;
        }
        if (all || st >= max_steps || total >= maxit) break;
        hipLaunchKernelGGL(k_fill0, dim3(nblk(n * Q, 256)), dim3(256), 0, c->stream, n * Q, d);
        rc = gmres_lockstep_any(c, inner, Q, r, d, tq.data(), restart, maxit - total, inner_its.data(), nullptr);
        if (rc != BEM_OK) return done(rc);
        int64_t mx = 0;
        for (int q = 0; q < Q; q++) {
            its[q] += inner_its[q];
            mx = std::max(mx, inner_its[q]);
        }
        total += std::max<int64_t>(mx, 1);
        hipLaunchKernelGGL(k_axpy1, dim3(nblk(n * Q, 256)), dim3(256), 0, c->stream, n * Q, (const bem_c64*)d, X);
        if ((rc = launch_check(c, "k_axpy1"))) return done(rc);
        zero = false;
        st++;
    }
    HIPC(c, hipStreamSynchronize(c->stream));
    for (int q = 0; q < Q; q++) {
        if (its_out) its_out[q] = its[q];
        if (rel_out) rel_out[q] = rel[q];
    }
    if (steps_out) *steps_out = st;
    return done(BEM_OK);
}

}  // namespace

extern "C" {

int bem_set_mfma(bem_ctx_t c, int on)
{
    if (!c) return BEM_ERR_ARG;
    c->use_mfma = on != 0;
    return BEM_OK;
}

int bem_set_graphs(bem_ctx_t c, int on)
{
    if (!c) return BEM_ERR_ARG;
    c->use_graphs = on != 0;
    return BEM_OK;
}

int bem_hmvm_multi(bem_ctx_t c, bem_hmat_t h, int64_t nrhs, const void* x, void* y, int accumulate, int refine)
{
    if (!c || !h || !x || !y || h->ctx != c || nrhs < 1 || nrhs > kArnMax)
        return set_err(c, BEM_ERR_ARG, "bem_hmvm_multi: bad arguments (1 <= nrhs <= 8)");
    HIPC(c, hipSetDevice(c->device));
    const int Q = pow2_q(nrhs);
    if (Q == nrhs) return hmvm_multi_any(c, h, Q, refine != 0, (const bem_c64*)x, (bem_c64*)y, accumulate);
    std::vector<void*> own;
    bem_c64 *xq, *yq;
    int rc;
    if ((rc = dalloc(c, &xq, (size_t)(c->n * Q), &own)) || (rc = dalloc(c, &yq, (size_t)(c->n * Q), &own)) ||
        (rc = repack(c, nrhs, nrhs, Q, (const bem_c64*)x, xq)) ||
        (accumulate && (rc = repack(c, nrhs, nrhs, Q, (const bem_c64*)y, yq))) ||
        (rc = hmvm_multi_any(c, h, Q, refine != 0, xq, yq, accumulate)) ||
        (rc = unpack(c, nrhs, nrhs, Q, yq, (bem_c64*)y))) {
        free_list(c, own);
        return rc;
    }
    HIPC(c, hipStreamSynchronize(c->stream));
    free_list(c, own);
    return BEM_OK;
}

int bem_gmres_multi(bem_ctx_t c, bem_hmat_t h, bem_hmat_t hres, int64_t nrhs, const void* b, void* x, int x_is_zero,
                    double tol, int64_t restart, int64_t maxit, int64_t max_steps, int64_t* iters, double* rel_resid,
                    int64_t* steps)
{
    if (!c || !h || !b || !x || h->ctx != c || (hres && hres->ctx != c) || nrhs < 1 || nrhs > kArnMax ||
        restart < 1 || maxit < 0 || max_steps < 0 || !(tol > 0))
        return set_err(c, BEM_ERR_ARG, "bem_gmres_multi: bad arguments (1 <= nrhs <= 8)");
    HIPC(c, hipSetDevice(c->device));
    const int Q = pow2_q(nrhs);
    const bool ir = hres != nullptr || max_steps > 0;
    if (ir && !hres) hres = h;
    if (ir && hres->prec == BEM_PREC_FP32 && !hres->dense64)
        return set_err(c, BEM_ERR_STATE, "bem_gmres_multi: the residual operator is FP32 without an FP64 near field");
    std::vector<void*> own;
    bem_c64 *bq = (bem_c64*)b, *xq = (bem_c64*)x;
    int rc = BEM_OK;
    if (Q != nrhs) {
        if ((rc = dalloc(c, &bq, (size_t)(c->n * Q), &own)) || (rc = dalloc(c, &xq, (size_t)(c->n * Q), &own)) ||
            (rc = repack(c, nrhs, nrhs, Q, (const bem_c64*)b, bq)) ||
            (rc = repack(c, nrhs, nrhs, Q, (const bem_c64*)x, xq))) {
            free_list(c, own);
            return rc;
        }
    }
    std::vector<int64_t> its(Q, 0);
    std::vector<double> rel(Q, 0.0);
    int64_t st = 0;
    if (ir) {
        OpSet in, rs;
        in.ops = {h};
        rs.ops = {hres};
        rs.refine = true;
        rc = gmres_ir_q(c, in, rs, Q, bq, xq, x_is_zero != 0, tol, restart, maxit, max_steps, its.data(), rel.data(),
                        &st);
    } else {
        if (x_is_zero) {
            hipLaunchKernelGGL(k_fill0, dim3(nblk(c->n * Q, 256)), dim3(256), 0, c->stream, c->n * Q, xq);
            rc = launch_check(c, "k_fill0");
        }
        std::vector<double> tq(Q, tol);
        OpSet in;
        in.ops = {h};
        if (!rc) rc = gmres_lockstep_any(c, in, Q, bq, xq, tq.data(), restart, maxit, its.data(), rel.data());
    }
    if (!rc && Q != nrhs) rc = unpack(c, nrhs, nrhs, Q, xq, (bem_c64*)x);
    if (!rc) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = set_err(c, BEM_ERR_HIP, std::string("bem_gmres_multi: ") + hipGetErrorString(e));
    }
    free_list(c, own);
    if (rc) return rc;
    bool conv = true;
    for (int64_t q = 0; q < nrhs; q++) {
        if (iters) iters[q] = its[q];
        if (rel_resid) rel_resid[q] = rel[q];
        conv = conv && rel[q] < tol;
    }
    if (steps) *steps = st;
    if (!conv) return set_err(c, BEM_ERR_NOCONV, "bem_gmres_multi: maxit or max_steps reached above tolerance");
    return BEM_OK;
}

int bem_gmres_batch(bem_ctx_t c, int64_t nops, const bem_hmat_t* hs, const bem_hmat_t* hres, int64_t nrhs,
                    const void* b, void* x, int x_is_zero, double tol, int64_t restart, int64_t maxit, int64_t max_steps,
                    int64_t* iters, double* rel_resid, int64_t* steps)
{
    if (!c || nops < 1 || !hs || !b || !x || (nrhs != 1 && nrhs != 2 && nrhs != 4) || nops * nrhs > kArnMax ||
        restart < 1 || maxit < 0 || max_steps < 0 || !(tol > 0))
        return set_err(c, BEM_ERR_ARG, "bem_gmres_batch: bad arguments (nrhs 1, 2 or 4; nops x nrhs <= 8)");
    for (int64_t o = 0; o < nops; o++) {
        if (!hs[o] || hs[o]->ctx != c || (hres && (!hres[o] || hres[o]->ctx != c)))
            return set_err(c, BEM_ERR_ARG, "bem_gmres_batch: operator of another context");
        const bem_hmat* hr = hres ? hres[o] : hs[o];
        if ((hres || max_steps > 0) && hr->prec == BEM_PREC_FP32 && !hr->dense64)
            return set_err(c, BEM_ERR_STATE, "bem_gmres_batch: FP32 residual operator without an FP64 near field");
    }
    HIPC(c, hipSetDevice(c->device));
    const int64_t nq = nops * nrhs;
    const int Q = pow2_q(nq);
    const bool ir = hres != nullptr || max_steps > 0;
    // operator groups of R = nrhs columns; the padding columns (zero
    // right-hand sides) reuse the first operator
    OpSet in, rs;
    in.R = rs.R = (int)nrhs;
    in.active = rs.active = (size_t)nops;
    rs.refine = true;
    for (int g = 0; g < Q / nrhs; g++) {
        const int64_t o = g < nops ? g : 0;
        in.ops.push_back(hs[o]);
        rs.ops.push_back(hres ? hres[o] : hs[o]);
    }
    if (Q == (int)nrhs) {   // one group: the multi-right-hand-side product
        in.ops.resize(1);
        rs.ops.resize(1);
    }
    std::vector<void*> own;
    bem_c64 *bq = (bem_c64*)b, *xq = (bem_c64*)x;
    int rc = BEM_OK;
    if (Q != nq) {
        if ((rc = dalloc(c, &bq, (size_t)(c->n * Q), &own)) || (rc = dalloc(c, &xq, (size_t)(c->n * Q), &own)) ||
            (rc = repack(c, nq, nq, Q, (const bem_c64*)b, bq)) || (rc = repack(c, nq, nq, Q, (const bem_c64*)x, xq))) {
            free_list(c, own);
            return rc;
        }
    }
    std::vector<int64_t> its(Q, 0);
    std::vector<double> rel(Q, 0.0);
    int64_t st = 0;
    if (ir) {
        rc = gmres_ir_q(c, in, rs, Q, bq, xq, x_is_zero != 0, tol, restart, maxit, max_steps, its.data(), rel.data(),
                        &st);
    } else {
        if (x_is_zero) {
            hipLaunchKernelGGL(k_fill0, dim3(nblk(c->n * Q, 256)), dim3(256), 0, c->stream, c->n * Q, xq);
            rc = launch_check(c, "k_fill0");
        }
        std::vector<double> tq(Q, tol);
        if (!rc) rc = gmres_lockstep_any(c, in, Q, bq, xq, tq.data(), restart, maxit, its.data(), rel.data());
    }
    if (!rc && Q != nq) rc = unpack(c, nq, nq, Q, xq, (bem_c64*)x);
    if (!rc) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = set_err(c, BEM_ERR_HIP, std::string("bem_gmres_batch: ") + hipGetErrorString(e));
    }
    free_list(c, own);
    if (rc) return rc;
    bool conv = true;
    for (int64_t q = 0; q < nq; q++) {
        if (iters) iters[q] = its[q];
        if (rel_resid) rel_resid[q] = rel[q];
        conv = conv && rel[q] < tol;
    }
    if (steps) *steps = st;
    if (!conv) return set_err(c, BEM_ERR_NOCONV, "bem_gmres_batch: maxit or max_steps reached above tolerance");
    return BEM_OK;
}

int bem_gmres_ir(bem_ctx_t c, bem_hmat_t h, bem_hmat_t hres, bem_prec_t pc, const void* b, void* x, int x_is_zero,
                 double tol, int64_t restart, int64_t maxit, int64_t max_steps, int64_t* iters, double* rel_resid,
                 int64_t* steps)
{
    if (!hres) hres = h;
    if (!c || !h || !b || !x || restart < 1 || maxit < 0 || max_steps < 0 || h->ctx != c || hres->ctx != c ||
        (pc && pc->ctx != c))
        return set_err(c, BEM_ERR_ARG, "bem_gmres_ir: bad arguments");
    if (hres->prec == BEM_PREC_FP32 && !hres->dense64)
        return set_err(c, BEM_ERR_STATE, "bem_gmres_ir: the residual operator is FP32 without an FP64 near field");
    HIPC(c, hipSetDevice(c->device));
    if (!pc) {   // the lockstep solver (graph-captured columns) with one right-hand side
        int64_t its = 0, st = 0;
        double rel = 0;
        OpSet in, rs;
        in.ops = {h};
        rs.ops = {hres};
        rs.refine = true;
        int rc = gmres_ir_q(c, in, rs, 1, (const bem_c64*)b, (bem_c64*)x, x_is_zero != 0, tol, restart, maxit,
                            max_steps, &its, &rel, &st);
        if (rc) return rc;
        if (iters) *iters = its;
        if (rel_resid) *rel_resid = rel;
        if (steps) *steps = st;
        if (rel >= tol) return set_err(c, BEM_ERR_NOCONV, "bem_gmres_ir: residual of the refinement operator above tol");
        return BEM_OK;
    }
    const int64_t n = c->n;
    std::vector<void*> own;
    bem_c64 *r = nullptr, *d = nullptr;
    double* nrm = nullptr;
    int rc;
    auto done = [&](int code) {
        free_list(c, own);
        return code;
    };
    if ((rc = dalloc(c, &r, (size_t)n, &own)) || (rc = dalloc(c, &d, (size_t)n, &own)) ||
        (rc = dalloc(c, &nrm, (size_t)257, &own)))
        return done(rc);
    const int npart = (int)std::min<int64_t>(256, (n + kWG - 1) / kWG);
    // |v| of a device vector, to the host
    auto norm = [&](const bem_c64* v, double* out) -> int {
        hipLaunchKernelGGL(k_norm2_part, dim3(npart), dim3(kWG), 0, c->stream, n, v, nrm);
        hipLaunchKernelGGL(k_norm2_final, dim3(1), dim3(64), 0, c->stream, npart, nrm, nrm + 256);
        int rc2 = launch_check(c, "norm");
        if (rc2) return rc2;
        HIPC(c, hipMemcpyAsync(out, nrm + 256, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        return BEM_OK;
    };
    double bn = 0, rn = 0;
    if ((rc = norm((const bem_c64*)b, &bn))) return done(rc);
    int64_t it = 0, st = 0;
    double rel = 0;
    if (bn > 0) {
        bool zero = x_is_zero != 0;
        while (true) {
            // r = b - A_res x, accumulated in FP64 (x = 0: r = b)
            if (zero) {
                HIPC(c, hipMemcpyAsync(r, b, (size_t)n * sizeof(bem_c64), hipMemcpyDeviceToDevice, c->stream));
            } else {
                if ((rc = bem_hmvm_refine(c, hres, x, r, 0))) return done(rc);
                hipLaunchKernelGGL(k_residual, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, (const bem_c64*)b, r);
            }
            if ((rc = norm(r, &rn))) return done(rc);
            rel = rn / bn;
            if (rel < tol || st >= max_steps || it >= maxit) break;
            // d = A^-1 r with GMRES on the inner operator to the relative
            // tolerance that brings |b - A_res x| under tol |b|, then x += d
            const double tau = std::min(0.5, 0.5 * tol / rel);
            hipLaunchKernelGGL(k_fill0, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, d);
            int64_t its = 0;
            double rr = 0;
            rc = bem_gmres_prec(c, h, pc, r, d, tau, restart, maxit - it, &its, &rr);
            if (rc != BEM_OK && rc != BEM_ERR_NOCONV) return done(rc);
            it += its;
            if (zero) {
                HIPC(c, hipMemcpyAsync(x, d, (size_t)n * sizeof(bem_c64), hipMemcpyDeviceToDevice, c->stream));
                zero = false;
            } else {
                hipLaunchKernelGGL(k_axpy1, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, (const bem_c64*)d,
                                   (bem_c64*)x);
            }
            if ((rc = launch_check(c, "k_axpy1"))) return done(rc);
            st++;
        }
        if (zero) {   // no step taken: x = 0
            hipLaunchKernelGGL(k_fill0, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, (bem_c64*)x);
            if ((rc = launch_check(c, "k_fill0"))) return done(rc);
        }
    } else {
        hipLaunchKernelGGL(k_fill0, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, (bem_c64*)x);
        if ((rc = launch_check(c, "k_fill0"))) return done(rc);
    }
    HIPC(c, hipStreamSynchronize(c->stream));
    if (iters) *iters = it;
    if (rel_resid) *rel_resid = rel;
    if (steps) *steps = st;
    done(BEM_OK);
    if (rel >= tol) return set_err(c, BEM_ERR_NOCONV, "bem_gmres_ir: residual of the refinement operator above tol");
    return BEM_OK;
}

int bem_reconstruct_g(bem_ctx_t c, const void* x, const void* g, double rho, double freq, void* phi, void* dphi,
                      void* pr)
{
    if (!c || !x || !g || !c->have_freq) return set_err(c, BEM_ERR_ARG, "bem_reconstruct: bad arguments");
    HIPC(c, hipSetDevice(c->device));
    double s = freq != 0 ? (rho * (2.0 * BEM_PI)) * freq : rho * (2.0 * BEM_PI);
    hipLaunchKernelGGL(k_reconstruct, dim3(nblk(c->n, 256)), dim3(256), 0, c->stream, c->n, (const bem_c64*)x,
                       c->d_subst, c->d_alpha, c->d_beta, (const bem_c64*)g, s, (bem_c64*)phi, (bem_c64*)dphi,
                       (bem_c64*)pr);
    int rc = launch_check(c, "k_reconstruct");
    if (rc) return rc;
    HIPC(c, hipStreamSynchronize(c->stream));
    return BEM_OK;
}

int bem_field(bem_ctx_t c, int64_t nobs, const double* obs3, const void* phi, const void* dphi, double* out2)
{
    if (!c || nobs < 0 || (nobs && (!obs3 || !out2)) || !phi || !dphi)
        return set_err(c, BEM_ERR_ARG, "bem_field: bad arguments");
    if (!c->have_freq || c->n <= 0) return set_err(c, BEM_ERR_STATE, "bem_field: mesh or frequency missing");
    if (nobs == 0) return BEM_OK;
    HIPC(c, hipSetDevice(c->device));
    const int64_t n = c->n;
    const int nchunk = (int)((n + kFieldChunk - 1) / kFieldChunk);
    std::vector<double> soa(3 * nobs);
    for (int64_t i = 0; i < nobs; i++)
        for (int d = 0; d < 3; d++) soa[d * nobs + i] = obs3[3 * i + d];
    std::vector<void*> own;
    double* d_obs;
    bem_c64 *d_part, *d_out;
    int rc;
    if ((rc = dupload(c, &d_obs, soa.data(), soa.size(), &own))) { free_list(c, own); return rc; }
    if ((rc = dalloc(c, &d_part, (size_t)nobs * nchunk, &own))) { free_list(c, own); return rc; }
    if ((rc = dalloc(c, &d_out, (size_t)nobs, &own))) { free_list(c, own); return rc; }
    {
        Timer t(c, 6);
        hipLaunchKernelGGL(k_field, dim3(nblk(nobs, kTile), nchunk), dim3(kWG), 0, c->stream, c->mesh, c->k, nobs,
                           d_obs, (const bem_c64*)phi, (const bem_c64*)dphi, nchunk, d_part);
    }
    hipLaunchKernelGGL(k_field_reduce, dim3(nblk(nobs, 256)), dim3(256), 0, c->stream, nobs, nchunk, d_part, d_out);
    if ((rc = launch_check(c, "k_field"))) { free_list(c, own); return rc; }
    hipError_t e = hipMemcpyAsync(out2, d_out, (size_t)nobs * sizeof(bem_c64), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    free_list(c, own);
    if (e != hipSuccess) return set_err(c, BEM_ERR_HIP, std::string("bem_field: ") + hipGetErrorString(e));
    return BEM_OK;
}

int bem_field_noalloc(bem_ctx_t c, int64_t nobs, const double* obs_dev, const void* phi, const void* dphi,
                      void* out_dev)
{
    if (!c || nobs < 0 || (nobs && (!obs_dev || !out_dev)) || !phi || !dphi)
        return set_err(c, BEM_ERR_ARG, "bem_field_noalloc: bad arguments");
    if (!c->have_freq || c->n <= 0) return set_err(c, BEM_ERR_STATE, "bem_field_noalloc: mesh or frequency missing");
    if (!c->scratch_dev) return set_err(c, BEM_ERR_STATE, "bem_field_noalloc: no scratch address for this device");
    if (nobs == 0) return BEM_OK;
    const int64_t nchunk = (c->n + kFieldChunk - 1) / kFieldChunk;
    const int64_t cap = (int64_t)(kFieldScratchBytes / sizeof(bem_c64));
    if (nobs > cap / nchunk)
        return set_err(c, BEM_ERR_ARG, "bem_field_noalloc: nobs x ceil(n / 1024) = " + std::to_string(nobs) + " x " +
                                           std::to_string(nchunk) + " partial sums exceed the fixed scratch of " +
                                           std::to_string(cap) + " (32 MiB)");
    HIPC(c, hipSetDevice(c->device));
    bem_c64* part = (bem_c64*)c->scratch_dev;
    hipLaunchKernelGGL(k_field, dim3(nblk(nobs, kTile), (unsigned)nchunk), dim3(kWG), 0, c->stream, c->mesh, c->k,
                       nobs, obs_dev, (const bem_c64*)phi, (const bem_c64*)dphi, (int)nchunk, part);
    hipLaunchKernelGGL(k_field_reduce, dim3(nblk(nobs, 256)), dim3(256), 0, c->stream, nobs, (int)nchunk, part,
                       (bem_c64*)out_dev);
    return launch_check(c, "k_field (noalloc)");
}

int bem_field_hmat(bem_ctx_t c, int64_t nobs, const double* obs3, int64_t nblocks, const int64_t* blocks5,
                   double tol, int64_t maxrank, uint64_t seed, const void* phi, const void* dphi, double* out2,
                   double* stats4)
{
    if (!c || nobs < 0 || (nobs && (!obs3 || !out2 || !blocks5)) || nblocks < 0 || !phi || !dphi || !(tol >= 0) ||
        maxrank < 0)
        return set_err(c, BEM_ERR_ARG, "bem_field_hmat: bad arguments");
    if (!c->have_freq || c->n <= 0) return set_err(c, BEM_ERR_STATE, "bem_field_hmat: mesh or frequency missing");
    if (c->mesh.rvtx) return set_err(c, BEM_ERR_ARG, "bem_field_hmat: not with an impedance plane (use bem_field)");
    if (nobs == 0) return BEM_OK;
    HIPC(c, hipSetDevice(c->device));
    const double t0 = now_ms();
    const int64_t n = c->n;
    // the partition, checked against both index ranges and for covering
    // every (observation, element) pair exactly once by area
    std::vector<bemsetup::Block> blocks((size_t)nblocks);
    double covered = 0;
    for (int64_t b = 0; b < nblocks; b++) {
        const int64_t* q = blocks5 + 5 * b;
        if (q[0] < 0 || q[1] < 1 || q[0] + q[1] > nobs || q[2] < 0 || q[3] < 1 || q[2] + q[3] > n)
            return set_err(c, BEM_ERR_ARG, "bem_field_hmat: block outside the observation or element range");
        blocks[b] = bemsetup::Block{q[0], q[1], q[2], q[3], q[4] != 0};
        covered += (double)q[1] * (double)q[3];
    }
    if (covered != (double)nobs * (double)n)
        return set_err(c, BEM_ERR_ARG, "bem_field_hmat: the blocks do not cover the observation x element range");
    std::vector<void*> own;
    auto fail = [&](int rc) {
        free_list(c, own);
        return rc;
    };
    int rc;
    // observation points SoA, block extents
    std::vector<double> soa(3 * nobs);
    for (int64_t i = 0; i < nobs; i++)
        for (int d = 0; d < 3; d++) soa[d * nobs + i] = obs3[3 * i + d];
    std::vector<int64_t> ext(4 * nblocks);
    for (int64_t b = 0; b < nblocks; b++) {
        ext[b] = blocks[b].row0;
        ext[nblocks + b] = blocks[b].rows;
        ext[2 * nblocks + b] = blocks[b].col0;
        ext[3 * nblocks + b] = blocks[b].cols;
    }
    double* d_obs;
    int64_t* d_ext;
    if ((rc = dupload(c, &d_obs, soa.data(), soa.size(), &own)) ||
        (rc = dupload(c, &d_ext, ext.data(), ext.size(), &own)))
        return fail(rc);
    const PartView pv{&blocks, d_ext, d_ext + nblocks, d_ext + 2 * nblocks, d_ext + 3 * nblocks, d_obs, nobs};
    // admissible blocks, largest first; the L kernel's pivot lists
    std::vector<int32_t> adm;
    for (int64_t b = 0; b < nblocks; b++)
        if (blocks[b].adm) adm.push_back((int32_t)b);
    std::stable_sort(adm.begin(), adm.end(), [&](int32_t x, int32_t y) {
        return blocks[x].rows + blocks[x].cols > blocks[y].rows + blocks[y].cols;
    });
    int64_t* d_pivoff = nullptr;
    int32_t* d_piv = nullptr;
    if (!adm.empty()) {
        bemsetup::Elements E;
        E.n = n;
        E.normal.resize(n);
        for (int64_t e = 0; e < n; e++)
            E.normal[e] = bemsetup::Vec3{c->h_normal[3 * e], c->h_normal[3 * e + 1], c->h_normal[3 * e + 2]};
        std::vector<int64_t> pivoff;
        std::vector<int32_t> piv;
        bemsetup::build_field_pivot_lists(E, blocks, seed, pivoff, piv);
        if (piv.empty()) piv.push_back(0);
        if ((rc = dupload(c, &d_pivoff, pivoff.data(), pivoff.size(), &own)) ||
            (rc = dupload(c, &d_piv, piv.data(), piv.size(), &own)))
            return fail(rc);
    }
    // ACA of both field operators (partialPivotACA with implicitPhiMatrixField,
    // partialPivotACAextra with implicitDPhiMatrixField, :2446-2453), first
    // with a modest rank capacity, the overflowing blocks again with their full one
    auto full_cap = [&](int32_t b) {
        int64_t m = std::min(blocks[b].rows, blocks[b].cols);
        if (maxrank > 0) m = std::min(m, maxrank);
        return (int32_t)m;
    };
    AcaScratch S[2][2];
    std::vector<int32_t> rank[2], over_of[2];
    std::vector<int64_t> evs[2];
    int64_t evaluated = 0;
    int maxr = 0;
    for (int kind = 0; kind < 2; kind++) {
        std::vector<int32_t> cap((size_t)nblocks, 0), flag, rank2, flag2;
        std::vector<int64_t> evs2;
        for (int32_t b : adm) cap[b] = std::min<int32_t>(full_cap(b), kCapFirst);
        if (adm.empty()) break;
        if ((rc = run_aca(c, kind, BEM_PREC_FP64, tol, maxrank, seed, adm, cap, kind ? d_pivoff : nullptr,
                          kind ? d_piv : nullptr, [] { return BEM_OK; }, S[kind][0], rank[kind], flag, evs[kind],
                          &pv)))
            return fail(rc);
        std::vector<int32_t> over;
        for (int32_t b : adm) {
            if (flag[b] == 2) return fail(set_err(c, BEM_ERR_NAN, "bem_field_hmat: non-finite ACA value"));
            if (flag[b] == 1) over.push_back(b);
        }
        if (!over.empty()) {
            std::vector<int32_t> cap2((size_t)nblocks, 0);
            for (int32_t b : over) cap2[b] = std::min<int32_t>(full_cap(b), kCapMax);
            if ((rc = run_aca(c, kind, BEM_PREC_FP64, tol, maxrank, seed, over, cap2, kind ? d_pivoff : nullptr,
                              kind ? d_piv : nullptr, [] { return BEM_OK; }, S[kind][1], rank2, flag2, evs2, &pv)))
                return fail(rc);
            for (int32_t b : over) {
                if (flag2[b])
                    return fail(set_err(c, flag2[b] == 2 ? BEM_ERR_NAN : BEM_ERR_ARG,
                                        "bem_field_hmat: ACA block " + std::to_string(b) +
                                            (flag2[b] == 2 ? " non-finite" : " needs rank > " + std::to_string(kCapMax))));
                rank[kind][b] = rank2[b];
                evs[kind][b] = evs2[b];
            }
        }
        over_of[kind] = std::vector<int32_t>((size_t)nblocks, 0);
        for (int32_t b : over) over_of[kind][b] = 1;
        for (int32_t b : adm) {
            evaluated += evs[kind][b];
            maxr = std::max(maxr, rank[kind][b]);
        }
    }
    // partial vectors: per admissible block its rows, per dense block one
    // per 64 x 64 tile; each row sums its partials in partition order
    std::vector<FieldLR> lr;
    std::vector<FieldTile> tiles;
    std::vector<std::vector<int64_t>> cover((size_t)nobs);
    int64_t ptot = 0, maxr2 = 1;
    for (int64_t b = 0; b < nblocks; b++) {
        const auto& B = blocks[b];
        if (B.adm) {
            FieldLR L{};
            const bem_c64* f[2][3];
            int32_t r[2];
            for (int kind = 0; kind < 2; kind++) {
                const AcaScratch& Sk = S[kind][over_of[kind][b]];
                f[kind][0] = (const bem_c64*)Sk.U + Sk.uoff[b];
                f[kind][1] = (const bem_c64*)Sk.V + Sk.voff[b];
                f[kind][2] = (const bem_c64*)Sk.A + Sk.aoff[b];
                r[kind] = rank[kind][b];
            }
            L.U0 = f[0][0]; L.V0 = f[0][1]; L.A0 = f[0][2];
            L.U1 = f[1][0]; L.V1 = f[1][1]; L.A1 = f[1][2];
            L.r0 = r[0];
            L.r1 = r[1];
            L.rows = B.rows;
            L.col0 = B.col0;
            L.cols = B.cols;
            L.part = ptot;
            maxr2 = std::max<int64_t>(maxr2, (int64_t)r[0] + r[1]);
            lr.push_back(L);
            for (int64_t i = 0; i < B.rows; i++) cover[B.row0 + i].push_back(ptot + i);
            ptot += B.rows;
            continue;
        }
        for (int64_t r0 = 0; r0 < B.rows; r0 += kTile) {
            const int32_t rr = (int32_t)std::min<int64_t>(kTile, B.rows - r0);
            for (int64_t c0 = 0; c0 < B.cols; c0 += kTile) {
                const int32_t cc = (int32_t)std::min<int64_t>(kTile, B.cols - c0);
                tiles.push_back(FieldTile{B.row0 + r0, B.col0 + c0, ptot, rr, cc});
                for (int64_t i = 0; i < rr; i++) cover[B.row0 + r0 + i].push_back(ptot + i);
                ptot += rr;
            }
        }
    }
    // (a row's partials: tiles of one dense block in column order, blocks in partition order)
    std::vector<int64_t> gptr((size_t)nobs + 1, 0), gidx;
    for (int64_t i = 0; i < nobs; i++) gptr[i + 1] = gptr[i] + (int64_t)cover[i].size();
    gidx.reserve(gptr[nobs]);
    for (auto& v : cover) gidx.insert(gidx.end(), v.begin(), v.end());
    FieldLR* d_lr = nullptr;
    FieldTile* d_tiles = nullptr;
    int64_t *d_gptr, *d_gidx;
    bem_c64 *d_part, *d_out;
    if ((!lr.empty() && (rc = dupload(c, &d_lr, lr.data(), lr.size(), &own))) ||
        (!tiles.empty() && (rc = dupload(c, &d_tiles, tiles.data(), tiles.size(), &own))) ||
        (rc = dupload(c, &d_gptr, gptr.data(), gptr.size(), &own)) ||
        (rc = dupload(c, &d_gidx, gidx.data(), std::max<size_t>(gidx.size(), 1), &own)) ||
        (rc = dalloc(c, &d_part, (size_t)std::max<int64_t>(ptot, 1), &own)) ||
        (rc = dalloc(c, &d_out, (size_t)nobs, &own)))
        return fail(rc);
    {
        Timer t(c, 6);
        if (!tiles.empty())
            hipLaunchKernelGGL(k_field_dense, dim3((unsigned)tiles.size()), dim3(kWG), 0, c->stream, c->mesh, c->k, nobs,
                               d_obs, d_tiles, (const bem_c64*)phi, (const bem_c64*)dphi, d_part);
        if (!lr.empty())
            hipLaunchKernelGGL(k_field_lr, dim3((unsigned)lr.size()), dim3(kWG), (size_t)maxr2 * sizeof(bem_c64),
                               c->stream, d_lr, (const bem_c64*)phi, (const bem_c64*)dphi, d_part);
    }
    hipLaunchKernelGGL(k_mvm_gather, dim3(nblk(16 * nobs, 256)), dim3(256), 0, c->stream, nobs, d_gptr, d_gidx, d_part,
                       d_out, 0);
    if ((rc = launch_check(c, "bem_field_hmat"))) return fail(rc);
    hipError_t e = hipMemcpyAsync(out2, d_out, (size_t)nobs * sizeof(bem_c64), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(set_err(c, BEM_ERR_HIP, std::string("bem_field_hmat: ") + hipGetErrorString(e)));
    if (stats4) {
        stats4[0] = now_ms() - t0;
        stats4[1] = (double)adm.size();
        stats4[2] = (double)maxr;
        stats4[3] = (double)evaluated;
    }
    return fail(BEM_OK);
}

int bem_reconstruct(bem_ctx_t c, const void* x, double rho, double freq, void* phi, void* dphi, void* pr)
{
    if (!c || !x || !c->have_freq) return set_err(c, BEM_ERR_ARG, "bem_reconstruct: bad arguments");
    return bem_reconstruct_g(c, x, c->d_f, rho, freq, phi, dphi, pr);
}

static int entries_impl(bem_ctx_t c, int mode, int64_t np, const int64_t* ij, double* out)
{
    if (!c || np < 0 || (np && (!ij || !out))) return set_err(c, BEM_ERR_ARG, "bem_entries: bad arguments");
    if (!c->have_freq) return set_err(c, BEM_ERR_STATE, "bem_entries: frequency not set");
    for (int64_t t = 0; t < 2 * np; t++)
        if (ij[t] < 0 || ij[t] >= c->n) return set_err(c, BEM_ERR_ARG, "bem_entries: index out of range");
    if (np == 0) return BEM_OK;
    HIPC(c, hipSetDevice(c->device));
    int64_t* d_ij;
    bem_c64* d_out;
    std::vector<void*> own;
    int rc;
    size_t per = mode == 2 ? 4 : 1;
    if ((rc = dupload(c, &d_ij, ij, 2 * np, &own))) return rc;
    if ((rc = dalloc(c, &d_out, per * np, &own))) { free_list(c, own); return rc; }
    hipLaunchKernelGGL(k_entries, dim3(nblk(np, 128)), dim3(128), 0, c->stream, c->mesh, dev_freq(c), mode, np, d_ij,
                       d_out);
    if ((rc = launch_check(c, "k_entries"))) { free_list(c, own); return rc; }
    hipError_t e = hipMemcpyAsync(out, d_out, per * np * sizeof(bem_c64), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    free_list(c, own);
    if (e != hipSuccess) return set_err(c, BEM_ERR_HIP, hipGetErrorString(e));
    return BEM_OK;
}

int bem_entries4(bem_ctx_t c, int64_t np, const int64_t* ij, double* out8) { return entries_impl(c, 2, np, ij, out8); }
int bem_entries(bem_ctx_t c, int kind, int64_t np, const int64_t* ij, double* out2)
{
    if (kind != 0 && kind != 1) return set_err(c, BEM_ERR_ARG, "bem_entries: bad kind");
    return entries_impl(c, kind, np, ij, out2);
}

int bem_malloc(bem_ctx_t c, size_t bytes, void** p)
{
    if (!c || !p) return BEM_ERR_ARG;
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipMalloc(p, bytes ? bytes : 1));
    return BEM_OK;
}
int bem_free(bem_ctx_t c, void* p)
{
    if (!c) return BEM_ERR_ARG;
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipStreamSynchronize(c->stream));
    if (p) HIPC(c, hipFree(p));
    return BEM_OK;
}
int bem_memcpy_htod(bem_ctx_t c, void* dst, const void* src, size_t bytes)
{
    if (!c) return BEM_ERR_ARG;
    HIPC(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return BEM_OK;
}
int bem_memcpy_dtoh(bem_ctx_t c, void* dst, const void* src, size_t bytes)
{
    if (!c) return BEM_ERR_ARG;
    HIPC(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return BEM_OK;
}
int bem_synchronize(bem_ctx_t c)
{
    if (!c) return BEM_ERR_ARG;
    HIPC(c, hipStreamSynchronize(c->stream));
    return BEM_OK;
}
int bem_profile(bem_ctx_t c, int enable, double* out)
{
    if (!c) return BEM_ERR_ARG;
    resolve_pending(c);
    if (out)
        for (int q = 0; q < 7; q++) {
            out[q] = c->prof_cnt[q] ? c->prof_ms[q] / (double)c->prof_cnt[q] : 0.0;
            out[7 + q] = (double)c->prof_cnt[q];
        }
    if (enable >= 0) {
        c->prof = enable != 0;
        for (int q = 0; q < 7; q++) { c->prof_ms[q] = 0; c->prof_cnt[q] = 0; }
    }
    return BEM_OK;
}

}  // extern "C"
