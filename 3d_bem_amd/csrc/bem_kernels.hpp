// Device-side entry generator of the collocation BEM operators (gfx950).
// Every function here follows the operation order of the host restatement
// so that FP64 entries are bit-identical to it (see include/bem_detmath.h).
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/bem_detmath.h"

// complex FP32 storage of the mixed-precision path (not part of the
// bit-exact FP64 contract)
typedef struct { float re, im; } bem_c32;

namespace bemdev {

// order-8 16-point triangle rule weights (trianglequadraturerules.cpp:91-108)
__constant__ static const double kTriW[16] = {
    0.14431560767779, 0.09509163426728, 0.09509163426728, 0.09509163426728, 0.10321737053472, 0.10321737053472,
    0.10321737053472, 0.03245849762320, 0.03245849762320, 0.03245849762320, 0.02723031417443, 0.02723031417443,
    0.02723031417443, 0.02723031417443, 0.02723031417443, 0.02723031417443};
// barycentric (xi1, xi2) of the same rule
__constant__ static const double kTriX[16][2] = {
    {0.33333333333333, 0.33333333333333}, {0.45929258829272, 0.45929258829272}, {0.45929258829272, 0.08141482341455},
    {0.08141482341455, 0.45929258829272}, {0.17056930775176, 0.17056930775176}, {0.17056930775176, 0.65886138449648},
    {0.65886138449648, 0.17056930775176}, {0.05054722831703, 0.05054722831703}, {0.05054722831703, 0.89890554336594},
    {0.89890554336594, 0.05054722831703}, {0.26311282963464, 0.72849239295540}, {0.72849239295540, 0.00839477740996},
    {0.00839477740996, 0.26311282963464}, {0.72849239295540, 0.26311282963464}, {0.26311282963464, 0.00839477740996},
    {0.00839477740996, 0.72849239295540}};
// 5-point Gauss-Legendre (linequadraturerules.cpp:63-86)
__constant__ static const double kGL5X[5] = {-0.906179845938664, -0.5384693101056831, 0.0, 0.5384693101056831,
                                             0.906179845938664};
__constant__ static const double kGL5W[5] = {0.23692688505618908, 0.47862867049936647, 0.5688888888888889,
                                             0.47862867049936647, 0.23692688505618908};
// positive half of 32-point Gauss-Legendre weights (boundaryelementsolver.cpp:906)
__constant__ static const double kHGW[16] = {
    0.0965400885147278,   0.09563872007927486, 0.09384439908080457, 0.09117387869576389, 0.08765209300440381,
    0.08331192422694675,  0.07819389578707031, 0.0723457941088485,  0.06582222277636185, 0.058684093478535544,
    0.050998059262376175, 0.04283589802222668, 0.03427386291302143, 0.02539206530926206, 0.01627439473090567,
    0.007018610009470096};

constexpr int kNearSub = 25;   // doubles per near-singular sub-triangle
constexpr int kSelf = 99;      // doubles per element of the self integral

// Mesh in HBM, structure of arrays, cluster order.  q* hold the 16
// quadrature points of every element, index q * n + j (coalesced over j).
struct DevMesh {
    int64_t n;
    const double *px, *py, *pz;   // collocation points
    const double *nx, *ny, *nz;   // unit normals
    const double* area;
    const double *qx, *qy, *qz;
    const double* vtx;            // vertices a, b, c: vtx[c * n + j], c = 0..8 (ax ay az bx .. cz)
    const double* edg;            // c, a - c, b - c: edg[c * n + j] (the FP32 path's point form)
    const int64_t* nr_ptr;        // near-singular CSR (n + 1)
    const int32_t* nr_row;        // row of each pair
    const int32_t* nr_col;
    const int32_t* nr_nsub;
    const int64_t* nr_off;
    const double* nr_geo;
    const double* self_geo;       // kSelf per element
    double near_lim;              // 4 * hbar
    // mirror images in the impedance plane (rvtx == nullptr: no plane), laid
    // out as vtx / edg / the normals
    const double* rvtx;
    const double* redg;
    const double *rnx, *rny, *rnz;
};

// frequency-dependent boundary-condition coefficients (bem_detmath.h)
struct DevFreq {
    double k;
    bem_c64 cc;
    const int32_t* subst;
    const bem_c64* cphi;
    const bem_c64* cd;
};

// regular branch of BemOperatorsConst (boundaryelementsolver.cpp:856-893);
// quadrature points read with stride qs from qx/qy/qz
template <int UNROLL = 2>
__device__ __forceinline__ bem_ops4 regular_ops(double k, double px, double py, double pz, double npx, double npy,
                                                double npz, double nqx, double nqy, double nqz, double area,
                                                const double* qx, const double* qy, const double* qz, int64_t qs)
{
    double upnq = fma(npz, nqz, fma(npy, nqy, npx * nqx));
    bem_ops4 o = bem_ops4_zero();
#pragma unroll UNROLL
    for (int q = 0; q < 16; q++)
        bem_point_accumulate(&o, k, px, py, pz, npx, npy, npz, nqx, nqy, nqz, upnq, qx[q * qs], qy[q * qs],
                             qz[q * qs], kTriW[q]);
    o.L = c64_scale(o.L, area);
    o.M = c64_scale(o.M, area);
    o.Mt = c64_scale(o.Mt, area);
    o.N = c64_scale(o.N, area);
    return o;
}

// The 16 quadrature points of element j formed on the fly from its vertices
// with the operation order of k_quadpoints (so the points, and every entry,
// are bit-identical to the stored-point path): 72 bytes read per element
// instead of 384.
struct ElemVtx { double ax, ay, az, bx, by, bz, cx, cy, cz; };

__device__ __forceinline__ ElemVtx load_vtx(const DevMesh& M, int64_t j)
{
    const double* v = M.vtx + j;
    const int64_t n = M.n;
    return ElemVtx{v[0], v[n], v[2 * n], v[3 * n], v[4 * n], v[5 * n], v[6 * n], v[7 * n], v[8 * n]};
}

__device__ __forceinline__ void quad_point(const ElemVtx& e, int q, double* qx, double* qy, double* qz)
{
    double x1 = kTriX[q][0], x2 = kTriX[q][1];
    double x3 = (1.0 - x2) - x1;
    *qx = (x1 * e.ax + x2 * e.bx) + x3 * e.cx;
    *qy = (x1 * e.ay + x2 * e.by) + x3 * e.cy;
    *qz = (x1 * e.az + x2 * e.bz) + x3 * e.cz;
}

// regular_ops with the points of element j formed from its vertices
template <int UNROLL = 1>
__device__ __forceinline__ bem_ops4 regular_ops_vtx(double k, double px, double py, double pz, double npx,
                                                    double npy, double npz, double nqx, double nqy, double nqz,
                                                    double area, const ElemVtx& e)
{
    double upnq = fma(npz, nqz, fma(npy, nqy, npx * nqx));
    bem_ops4 o = bem_ops4_zero();
#pragma unroll UNROLL
    for (int q = 0; q < 16; q++) {
        double qx, qy, qz;
        quad_point(e, q, &qx, &qy, &qz);
        bem_point_accumulate(&o, k, px, py, pz, npx, npy, npz, nqx, nqy, nqz, upnq, qx, qy, qz, kTriW[q]);
    }
    o.L = c64_scale(o.L, area);
    o.M = c64_scale(o.M, area);
    o.Mt = c64_scale(o.Mt, area);
    o.N = c64_scale(o.N, area);
    return o;
}

// BemOperatorsNearSing (boundaryelementsolver.cpp:1387-1514) on precomputed
// sub-triangle geometry
__device__ __noinline__ bem_ops4 near_ops(double k, double px, double py, double pz, double npx, double npy,
                                          double npz, double nqx, double nqy, double nqz, const double* geo,
                                          int nsub)
{
    double upnq = fma(npz, nqz, fma(npy, nqy, npx * nqx));
    bem_ops4 o = bem_ops4_zero();
    for (int s = 0; s < nsub; s++) {
        const double* g = geo + s * kNearSub;
        bem_ops4 t = bem_ops4_zero();
        for (int ia = 0; ia < 5; ia++) {
            double maxR = g[10 + 3 * ia], ca = g[11 + 3 * ia], sa = g[12 + 3 * ia];
            for (int ib = 0; ib < 5; ib++) {
                double rs = (maxR + kGL5X[ib] * maxR) / 2.0;
                double x = rs * ca, y = rs * sa;
                double qx = (g[0] + x * g[3]) + y * g[6];
                double qy = (g[1] + x * g[4]) + y * g[7];
                double qz = (g[2] + x * g[5]) + y * g[8];
                double fw = ((maxR / 2.0 * rs) * kGL5W[ia]) * kGL5W[ib];
                bem_point_accumulate(&t, k, px, py, pz, npx, npy, npz, nqx, nqy, nqz, upnq, qx, qy, qz, fw);
            }
        }
        double h = g[9];
        o.L = c64_axpy(h, t.L, o.L);
        o.M = c64_axpy(h, t.M, o.M);
        o.Mt = c64_axpy(h, t.Mt, o.Mt);
        o.N = c64_axpy(h, t.N, o.N);
    }
    return o;
}

// BemOperatorsSingularPolarInt (boundaryelementsolver.cpp:896-1029) on
// precomputed polar radii
__device__ __noinline__ bem_ops4 self_ops(double k, const double* g)
{
    bem_c64 Ls[3], Ns[3];
    for (int s = 0; s < 3; s++) {
        const double* gs = g + 33 * s;
        bem_c64 L = c64(0, 0), N = c64(0, 0);
        for (int q = 0; q < 16; q++) {
            double r1 = gs[1 + q], r2 = gs[17 + q];
            bem_c64 e1 = det_expi(k * r1), e2 = det_expi(k * r2);
            L = c64_axpy(kHGW[q], c64_add(e1, e2), L);
            N = c64_axpy(kHGW[q], c64_add(c64(e1.re / r1, e1.im / r1), c64(e2.re / r2, e2.im / r2)), N);
        }
        Ls[s] = c64_scale(L, gs[0]);
        Ns[s] = c64_scale(N, gs[0]);
    }
    bem_c64 LT = c64_add(c64_add(Ls[0], Ls[1]), Ls[2]);
    LT = c64(LT.re / (2.0 * BEM_PI), LT.im / (2.0 * BEM_PI));
    bem_c64 NT = c64_add(c64_add(Ns[0], Ns[1]), Ns[2]);
    NT = c64(NT.re / (4.0 * BEM_PI), NT.im / (4.0 * BEM_PI));
    bem_ops4 o = bem_ops4_zero();
    o.L = c64_mul(c64(0.0, 1.0 / (2.0 * k)), c64(1.0 - LT.re, -LT.im));
    o.N = c64(0.0 - NT.re, k / 2.0 - NT.im);
    return o;
}

__device__ __forceinline__ int64_t near_find(const DevMesh& M, int64_t i, int32_t j)
{
    int64_t lo = M.nr_ptr[i], hi = M.nr_ptr[i + 1];
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        int32_t c = M.nr_col[mid];
        if (c == j) return mid;
        if (c < j) lo = mid + 1; else hi = mid;
    }
    return -1;
}

__device__ __forceinline__ bool is_near(const DevMesh& M, int64_t i, int64_t j)
{
    double dx = M.px[i] - M.px[j], dy = M.py[i] - M.py[j], dz = M.pz[i] - M.pz[j];
    return sqrt(fma(dz, dz, fma(dy, dy, dx * dx))) < M.near_lim;
}

// BemOperatorsConst (boundaryelementsolver.cpp:844-894), any pair
__device__ bem_ops4 ops_general(const DevMesh& M, double k, int64_t i, int64_t j)
{
    if (i == j) return self_ops(k, M.self_geo + kSelf * i);
    if (is_near(M, i, j)) {
        int64_t t = near_find(M, i, (int32_t)j);
        if (t >= 0)
            return near_ops(k, M.px[i], M.py[i], M.pz[i], M.nx[i], M.ny[i], M.nz[i], M.nx[j], M.ny[j], M.nz[j],
                            M.nr_geo + M.nr_off[t], M.nr_nsub[t]);
    }
    return regular_ops(k, M.px[i], M.py[i], M.pz[i], M.nx[i], M.ny[i], M.nz[i], M.nx[j], M.ny[j], M.nz[j],
                       M.area[j], M.qx + j, M.qy + j, M.qz + j, M.n);
}

// implicitPhiMatrix2 / implicitDPhiMatrix2 (boundaryelementsolver.cpp:2004-2040)
__device__ __forceinline__ bem_c64 bc_entry(const DevFreq& F, int kind, const bem_ops4& o, int64_t i, int64_t j)
{
    bem_c64 X, Y;
    bem_xy(&o, F.cc, i == j, &X, &Y);
    return kind == 0 ? bem_phi_entry(X, Y, F.subst[j], F.cphi[j]) : bem_dphi_entry(X, Y, F.subst[j], F.cd[j]);
}

// BemOperatorsReflected (boundaryelementsolver.cpp:1516-1553): the regular
// rule over the mirror image of element j (its vertices and normal, the area
// of element j), for any pair including i == j
__device__ __forceinline__ bem_ops4 refl_ops(const DevMesh& M, double k, int64_t i, int64_t j)
{
    const double* v = M.rvtx + j;
    const int64_t n = M.n;
    const ElemVtx e{v[0], v[n], v[2 * n], v[3 * n], v[4 * n], v[5 * n], v[6 * n], v[7 * n], v[8 * n]};
    return regular_ops_vtx<1>(k, M.px[i], M.py[i], M.pz[i], M.nx[i], M.ny[i], M.nz[i], M.rnx[j], M.rny[j], M.rnz[j],
                              M.area[j], e);
}

// implicitPhiMatrixReflected / implicitDPhiMatrixReflected
// (boundaryelementsolver.cpp:1556-1643): the boundary-condition combination
// of the column, without the jump term of the diagonal
__device__ __forceinline__ void refl_xy(const DevMesh& M, const DevFreq& F, int64_t i, int64_t j, bem_c64* X,
                                        bem_c64* Y)
{
    bem_ops4 o = refl_ops(M, F.k, i, j);
    bem_xy(&o, F.cc, 0, X, Y);
}

__device__ __forceinline__ bem_c64 refl_entry(const DevMesh& M, const DevFreq& F, int kind, int64_t i, int64_t j)
{
    bem_c64 X, Y;
    refl_xy(M, F, i, j, &X, &Y);
    return kind == 0 ? bem_phi_entry(X, Y, F.subst[j], F.cphi[j]) : bem_dphi_entry(X, Y, F.subst[j], F.cd[j]);
}

// implicitPhiMatrixPlusReflected / implicitDPhiMatrixPlusReflected
// (boundaryelementsolver.cpp:2042-2050): free-space entry + mirror entry
__device__ __forceinline__ bem_c64 entry_general(const DevMesh& M, const DevFreq& F, int kind, int64_t i, int64_t j)
{
    bem_ops4 o = ops_general(M, F.k, i, j);
    bem_c64 e = bc_entry(F, kind, o, i, j);
    return M.rvtx ? c64_add(e, refl_entry(M, F, kind, i, j)) : e;
}

// ---------------------------------------------------------------- FP32 path
// Regular (far and near-field, non-singular) quadrature in complex FP32.
// Only the geometry that decides the phase stays in FP64: R = p - q, |R|^2,
// one Newton correction of the FP32 reciprocal square root, and the
// reduction of k r / 2 pi to a fraction of a revolution.  Everything after
// (hardware sin / cos of the revolution fraction, the kernel values and the
// 16-point sums) is FP32 with explicit fmaf.

struct bem_c32f4 { bem_c32 L, M, Mt, N; };

// FP32 copies of the frequency data; krev = k / (2 pi) in FP64
struct DevFreq32 {
    float k;
    double krev;
    bem_c32 cc;
    const bem_c32* cphi;
    const bem_c32* cd;
};

__device__ __forceinline__ bem_c32 c32(float re, float im)
{
    bem_c32 z;
    z.re = re;
    z.im = im;
    return z;
}

__device__ __forceinline__ bem_c32 c32_mul(bem_c32 a, bem_c32 b)
{
    return c32(fmaf(a.re, b.re, -(a.im * b.im)), fmaf(a.re, b.im, a.im * b.re));
}

__device__ __forceinline__ bem_c32 c32_fma(bem_c32 a, bem_c32 b, bem_c32 acc)
{
    return c32(fmaf(a.re, b.re, fmaf(-a.im, b.im, acc.re)), fmaf(a.re, b.im, fmaf(a.im, b.re, acc.im)));
}

// one quadrature point given R = p - q (FP64), same formulas as
// bem_point_accumulate
__device__ __forceinline__ void point_accumulate_f32_r(bem_c32f4* acc, const DevFreq32& F, double Rx, double Ry,
                                                       double Rz, float npx, float npy, float npz, float nqx,
                                                       float nqy, float nqz, float upnq, float w, bool needX = true,
                                                       bool needY = true);

// one quadrature point, same formulas as bem_point_accumulate
__device__ __forceinline__ void point_accumulate_f32(bem_c32f4* acc, const DevFreq32& F, double px, double py,
                                                     double pz, float npx, float npy, float npz, float nqx, float nqy,
                                                     float nqz, float upnq, double qx, double qy, double qz, float w)
{
    point_accumulate_f32_r(acc, F, px - qx, py - qy, pz - qz, npx, npy, npz, nqx, nqy, nqz, upnq, w);
}

// needX / needY: only the operators of X = L + cc Mt / Y = M + cc N are
// accumulated (the others keep their value); each one that is accumulated
// gets the same operations as with both
__device__ __forceinline__ void point_accumulate_f32_r(bem_c32f4* acc, const DevFreq32& F, double Rx, double Ry,
                                                       double Rz, float npx, float npy, float npz, float nqx,
                                                       float nqy, float nqz, float upnq, float w, bool needX,
                                                       bool needY)
{
    double rr = fma(Rz, Rz, fma(Ry, Ry, Rx * Rx));
    float inv_r = __builtin_amdgcn_rsqf((float)rr);
    double r0 = (double)((float)rr * inv_r);
    double r = fma(0.5 * fma(-r0, r0, rr), (double)inv_r, r0);
    double t = r * F.krev;
    float rev = (float)(t - rint(t));     // in [-1/2, 1/2]
    float s = __builtin_amdgcn_sinf(rev), c = __builtin_amdgcn_cosf(rev);
    float kr = F.k * (float)r;
    float fx = (float)Rx, fy = (float)Ry, fz = (float)Rz;
    float rnp = fmaf(fz, npz, fmaf(fy, npy, fx * npx)) * inv_r;
    float g = inv_r * (float)BEM_INV_4PI;
    float Gre = c * g, Gim = s * g;
    float Grre = Gre * inv_r, Grim = Gim * inv_r;
    float dGre = fmaf(-Grim, kr, -Grre);
    float dGim = fmaf(Grre, kr, -Grim);
    if (needX) {
        acc->L.re = fmaf(w, Gre, acc->L.re);
        acc->L.im = fmaf(w, Gim, acc->L.im);
        acc->Mt.re = fmaf(w, dGre * rnp, acc->Mt.re);
        acc->Mt.im = fmaf(w, dGim * rnp, acc->Mt.im);
    }
    if (needY) {
        float rnq = -fmaf(fz, nqz, fmaf(fy, nqy, fx * nqx)) * inv_r;
        float Grrre = Grre * inv_r, Grrim = Grim * inv_r;
        float a = fmaf(-kr, kr, 2.0f);
        float b = -2.0f * kr;
        float d2Gre = fmaf(Grrre, a, -(Grrim * b));
        float d2Gim = fmaf(Grrre, b, Grrim * a);
        float pp = rnp * rnq;
        float tt = -(upnq + pp) * inv_r;
        float Nre = fmaf(dGre, tt, d2Gre * pp);
        float Nim = fmaf(dGim, tt, d2Gim * pp);
        acc->M.re = fmaf(w, dGre * rnq, acc->M.re);
        acc->M.im = fmaf(w, dGim * rnq, acc->M.im);
        acc->N.re = fmaf(w, Nre, acc->N.re);
        acc->N.im = fmaf(w, Nim, acc->N.im);
    }
}

// order-8 16-point rule weights in FP32
__constant__ static const float kTriWf[16] = {
    0.14431560767779f, 0.09509163426728f, 0.09509163426728f, 0.09509163426728f, 0.10321737053472f, 0.10321737053472f,
    0.10321737053472f, 0.03245849762320f, 0.03245849762320f, 0.03245849762320f, 0.02723031417443f, 0.02723031417443f,
    0.02723031417443f, 0.02723031417443f, 0.02723031417443f, 0.02723031417443f};

template <int UNROLL = 2>
__device__ __forceinline__ bem_c32f4 regular_ops_f32(const DevFreq32& F, double px, double py, double pz, float npx,
                                                     float npy, float npz, float nqx, float nqy, float nqz, float area,
                                                     const double* qx, const double* qy, const double* qz, int64_t qs)
{
    float upnq = fmaf(npz, nqz, fmaf(npy, nqy, npx * nqx));
    bem_c32f4 o;
    o.L = o.M = o.Mt = o.N = c32(0.0f, 0.0f);
#pragma unroll UNROLL
    for (int q = 0; q < 16; q++)
        point_accumulate_f32(&o, F, px, py, pz, npx, npy, npz, nqx, nqy, nqz, upnq, qx[q * qs], qy[q * qs],
                             qz[q * qs], kTriWf[q]);
    o.L = c32(o.L.re * area, o.L.im * area);
    o.M = c32(o.M.re * area, o.M.im * area);
    o.Mt = c32(o.Mt.re * area, o.Mt.im * area);
    o.N = c32(o.N.re * area, o.N.im * area);
    return o;
}

template <int UNROLL = 1>
__device__ __forceinline__ bem_c32f4 regular_ops_f32_vtx(const DevFreq32& F, double px, double py, double pz,
                                                         float npx, float npy, float npz, float nqx, float nqy,
                                                         float nqz, float area, const ElemVtx& e)
{
    float upnq = fmaf(npz, nqz, fmaf(npy, nqy, npx * nqx));
    bem_c32f4 o;
    o.L = o.M = o.Mt = o.N = c32(0.0f, 0.0f);
#pragma unroll UNROLL
    for (int q = 0; q < 16; q++) {
        double qx, qy, qz;
        quad_point(e, q, &qx, &qy, &qz);
        point_accumulate_f32(&o, F, px, py, pz, npx, npy, npz, nqx, nqy, nqz, upnq, qx, qy, qz, kTriWf[q]);
    }
    o.L = c32(o.L.re * area, o.L.im * area);
    o.M = c32(o.M.re * area, o.M.im * area);
    o.Mt = c32(o.Mt.re * area, o.Mt.im * area);
    o.N = c32(o.N.re * area, o.N.im * area);
    return o;
}

// regular_ops_f32 with R = d - xi1 (a - c) - xi2 (b - c), d = p - c: the
// point q = c + xi1 (a - c) + xi2 (b - c) (the rule's point, rounded
// differently from the FP64 path's form) at two FP64 fma per coordinate
template <int UNROLL = 1>
__device__ __forceinline__ bem_c32f4 regular_ops_f32_edge_of(const DevFreq32& F, const DevMesh& M, int64_t i,
                                                             int64_t j, const double* edg, const double* qnx,
                                                             const double* qny, const double* qnz,
                                                             bool needX = true, bool needY = true)
{
    const int64_t n = M.n;
    const double* g = edg + j;
    const double dx = M.px[i] - g[0], dy = M.py[i] - g[n], dz = M.pz[i] - g[2 * n];
    const double ax = g[3 * n], ay = g[4 * n], az = g[5 * n], bx = g[6 * n], by = g[7 * n], bz = g[8 * n];
    const float npx = (float)M.nx[i], npy = (float)M.ny[i], npz = (float)M.nz[i];
    const float nqx = (float)qnx[j], nqy = (float)qny[j], nqz = (float)qnz[j];
    const float upnq = fmaf(npz, nqz, fmaf(npy, nqy, npx * nqx));
    bem_c32f4 o;
    o.L = o.M = o.Mt = o.N = c32(0.0f, 0.0f);
#pragma unroll UNROLL
    for (int q = 0; q < 16; q++) {
        const double x1 = kTriX[q][0], x2 = kTriX[q][1];
        point_accumulate_f32_r(&o, F, fma(-x1, ax, fma(-x2, bx, dx)), fma(-x1, ay, fma(-x2, by, dy)),
                               fma(-x1, az, fma(-x2, bz, dz)), npx, npy, npz, nqx, nqy, nqz, upnq, kTriWf[q], needX,
                               needY);
    }
    const float area = (float)M.area[j];
    o.L = c32(o.L.re * area, o.L.im * area);
    o.M = c32(o.M.re * area, o.M.im * area);
    o.Mt = c32(o.Mt.re * area, o.Mt.im * area);
    o.N = c32(o.N.re * area, o.N.im * area);
    return o;
}

template <int UNROLL = 1>
__device__ __forceinline__ bem_c32f4 regular_ops_f32_edge(const DevFreq32& F, const DevMesh& M, int64_t i, int64_t j,
                                                          bool needX = true, bool needY = true)
{
    return regular_ops_f32_edge_of<UNROLL>(F, M, i, j, M.edg, M.nx, M.ny, M.nz, needX, needY);
}

// which of X and Y an off-diagonal entry of column j uses (bc_entry_f32):
// Phi takes X (subst) or Y; DPhi takes both, except cd = 0, where the
// other one is multiplied by zero.  Returns the X bit 1 and the Y bit 2.
__device__ __forceinline__ int bc_needs_f32(const DevFreq32& F, const int32_t* subst, int kind, int64_t j)
{
    if (kind == 0) return subst[j] ? 1 : 2;
    const bem_c32 cd = F.cd[j];
    if (cd.re == 0.0f && cd.im == 0.0f) return subst[j] ? 2 : 1;
    return 3;
}

// the FP32 form of refl_ops (mirror image of element j, edge form)
__device__ __forceinline__ bem_c32f4 refl_ops_f32(const DevFreq32& F, const DevMesh& M, int64_t i, int64_t j)
{
    return regular_ops_f32_edge_of<1>(F, M, i, j, M.redg, M.rnx, M.rny, M.rnz);
}

// off-diagonal Phi / DPhi entry from the FP32 operators (bem_phi_entry,
// bem_dphi_entry)
__device__ __forceinline__ bem_c32 bc_entry_f32(const DevFreq32& F, const int32_t* subst, int kind,
                                                const bem_c32f4& o, int64_t j)
{
    bem_c32 X = c32_fma(F.cc, o.Mt, o.L);
    bem_c32 Y = c32_fma(F.cc, o.N, o.M);
    if (kind == 0) return c32_mul(F.cphi[j], subst[j] ? X : Y);
    if (subst[j]) {
        bem_c32 t = c32_mul(F.cd[j], X);
        return c32(t.re - Y.re, t.im - Y.im);
    }
    return c32_fma(F.cd[j], Y, X);
}

}  // namespace bemdev
