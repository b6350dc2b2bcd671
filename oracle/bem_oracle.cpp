/*
 * CPU restatement of the reference's H-matrix BEM hot path -- TEST
 * INFRASTRUCTURE ONLY (the checker).  The product never links or calls it.
 * See bem_oracle.h for the list of deliberate deviations D1-D5.
 *
 * Built by oracle/Makefile with the reference's release flags
 * (3d_BEM.pro:17-18: -O3 -fopenmp -fno-fast-math -frounding-math) plus
 * -ffp-contract=off so that the operation order written here is the one
 * executed.
 */
#include "bem_oracle.h"
#include "../include/bem_detmath.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <vector>
#include <omp.h>

namespace {

typedef bem_c64 cplx;
const double TINY = 10.0 * std::numeric_limits<double>::epsilon(); /* global.h:53 */
const double EPS_BARY = 0.05;                                       /* global.h:43 */

/* order-8 16-point symmetric triangle rule, (xi1, xi2, w) with sum w = 1:
   trianglequadraturerules.cpp:91-108 (values as printed there) */
const double TRI16[16][3] = {
    {0.33333333333333, 0.33333333333333, 0.14431560767779},
    {0.45929258829272, 0.45929258829272, 0.09509163426728},
    {0.45929258829272, 0.08141482341455, 0.09509163426728},
    {0.08141482341455, 0.45929258829272, 0.09509163426728},
    {0.17056930775176, 0.17056930775176, 0.10321737053472},
    {0.17056930775176, 0.65886138449648, 0.10321737053472},
    {0.65886138449648, 0.17056930775176, 0.10321737053472},
    {0.05054722831703, 0.05054722831703, 0.03245849762320},
    {0.05054722831703, 0.89890554336594, 0.03245849762320},
    {0.89890554336594, 0.05054722831703, 0.03245849762320},
    {0.26311282963464, 0.72849239295540, 0.02723031417443},
    {0.72849239295540, 0.00839477740996, 0.02723031417443},
    {0.00839477740996, 0.26311282963464, 0.02723031417443},
    {0.72849239295540, 0.26311282963464, 0.02723031417443},
    {0.26311282963464, 0.00839477740996, 0.02723031417443},
    {0.00839477740996, 0.72849239295540, 0.02723031417443}};

/* 5-point Gauss-Legendre on [-1,1] (linequadraturerules.cpp:63-86) */
const double GL5X[5] = {-0.906179845938664, -0.5384693101056831, 0.0, 0.5384693101056831, 0.906179845938664};
const double GL5W[5] = {0.23692688505618908, 0.47862867049936647, 0.5688888888888889, 0.47862867049936647, 0.23692688505618908};

/* positive half of 32-point Gauss-Legendre (boundaryelementsolver.cpp:905-906) */
const double HGX[16] = {0.04830766568773832, 0.1444719615827965, 0.23928736225213706, 0.33186860228212767,
                        0.42135127613063533, 0.5068999089322294, 0.5877157572407623, 0.6630442669302152,
                        0.7321821187402897, 0.7944837959679424, 0.84936761373257, 0.8963211557660521,
                        0.9349060759377397, 0.9647622555875064, 0.9856115115452684, 0.9972638618494816};
const double HGW[16] = {0.0965400885147278, 0.09563872007927486, 0.09384439908080457, 0.09117387869576389,
                        0.08765209300440381, 0.08331192422694675, 0.07819389578707031, 0.0723457941088485,
                        0.06582222277636185, 0.058684093478535544, 0.050998059262376175, 0.04283589802222668,
                        0.03427386291302143, 0.02539206530926206, 0.01627439473090567, 0.007018610009470096};

struct V3 { double x, y, z; };
inline V3 v3(double x, double y, double z) { V3 v; v.x = x; v.y = y; v.z = z; return v; }
inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 scl(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
inline double dot(V3 a, V3 b) { return fma(a.z, b.z, fma(a.y, b.y, a.x * b.x)); }
inline V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalized(V3 a) { double n = norm(a); return v3(a.x / n, a.y / n, a.z / n); }

struct Tri { V3 a, b, c; };

struct Cluster {
    int64_t start = 0, size = 0;
    int depth = 0;
    bool leaf = false;
    V3 bmin, bmax;
    int son1 = -1, son2 = -1;
};

struct Block { int64_t row0, rows, col0, cols; bool adm; int rc, cc; };

} // namespace

struct orc_problem {
    int64_t n = 0;
    std::vector<int64_t> perm;       /* perm[new] = old */
    std::vector<Tri> tri;            /* cluster order */
    std::vector<V3> normal, colloc;
    std::vector<double> area;
    double hbar = 0;
    std::vector<cplx> a, b, g;       /* Robin coefficients */
    std::vector<Cluster> clusters;   /* depth-first, root first */
    std::vector<Block> blocks;       /* minPartition order */
    /* frequency dependent */
    double freq = 0, k = 0;
    cplx cc = {0, 0};
    std::vector<cplx> s;             /* source term vector */
    std::vector<int> subst;
    std::vector<cplx> cphi, cd;
    /* impedance plane: mirror images of the elements (cluster order) */
    bool plane = false;
    std::vector<Tri> rtri;
    std::vector<V3> rnormal;
};

struct orc_hmat {
    const orc_problem* p = nullptr;
    int kind = 0;
    struct Blk {
        int64_t rank = -1;                /* -1 dense */
        std::vector<cplx> F;              /* dense m x n col-major */
        std::vector<cplx> U, alpha, V;    /* U m x k col-major, V k x n row-major */
    };
    std::vector<Blk> blk;
    int64_t evaluated = 0;
};

namespace {

/* ------------------------------------------------------------ geometry */

/* global::normalVectorOfTriangle (global.cpp:903-908) */
V3 tri_normal(const Tri& t) { return normalized(cross(sub(t.b, t.a), sub(t.c, t.a))); }
/* global::midpointOfTriangle (global.cpp:910-913) */
V3 tri_mid(const Tri& t) { V3 s = add(add(t.a, t.b), t.c); return v3(s.x / 3.0, s.y / 3.0, s.z / 3.0); }
/* global::areaOfTriangle (global.cpp:1285-1290) */
double tri_area(const Tri& t) { return norm(cross(sub(t.a, t.b), sub(t.a, t.c))) / 2.0; }

enum Region { IN_INTERIOR, ON_NODE1, ON_NODE2, ON_NODE3, ON_SIDE21, ON_SIDE32, ON_SIDE13 };

/* Nearest point of triangle t to point p (Eberly, "Distance Between Point and
   Triangle in 3D") followed by the barycentric region snapping with
   epsilon = 0.05: global::projectPointOnTriangleFaster (global.cpp:955-1283) */
Region project_point(const Tri& tr, V3 p, V3* out)
{
    V3 diff = sub(tr.a, p);
    V3 e0 = sub(tr.b, tr.a), e1 = sub(tr.c, tr.a);
    double a00 = dot(e0, e0), a01 = dot(e0, e1), a11 = dot(e1, e1);
    double b0 = dot(diff, e0), b1 = dot(diff, e1);
    double det = std::max(a00 * a11 - a01 * a01, 0.0);
    double s = a01 * b1 - a11 * b0;
    double t = a01 * b0 - a00 * b1;
    if (s + t <= det) {
        if (s < 0.0) {
            if (t < 0.0) {
                if (b0 < 0.0) {
                    t = 0.0;
                    s = (-b0 >= a00) ? 1.0 : -b0 / a00;
                } else {
                    s = 0.0;
                    if (b1 >= 0.0) t = 0.0;
                    else if (-b1 >= a11) t = 1.0;
                    else t = -b1 / a11;
                }
            } else {
                s = 0.0;
                if (b1 >= 0.0) t = 0.0;
                else if (-b1 >= a11) t = 1.0;
                else t = -b1 / a11;
            }
        } else if (t < 0.0) {
            t = 0.0;
            if (b0 >= 0.0) s = 0.0;
            else if (-b0 >= a00) s = 1.0;
            else s = -b0 / a00;
        } else {
            s /= det;
            t /= det;
        }
    } else {
        if (s < 0.0) {
            double tmp0 = a01 + b0, tmp1 = a11 + b1;
            if (tmp1 > tmp0) {
                double numer = tmp1 - tmp0, denom = a00 - 2.0 * a01 + a11;
                if (numer >= denom) { s = 1.0; t = 0.0; }
                else { s = numer / denom; t = 1.0 - s; }
            } else {
                s = 0.0;
                if (tmp1 <= 0.0) t = 1.0;
                else if (b1 >= 0.0) t = 0.0;
                else t = -b1 / a11;
            }
        } else if (t < 0.0) {
            double tmp0 = a01 + b1, tmp1 = a00 + b0;
            if (tmp1 > tmp0) {
                double numer = tmp1 - tmp0, denom = a00 - 2.0 * a01 + a11;
                if (numer >= denom) { t = 1.0; s = 0.0; }
                else { t = numer / denom; s = 1.0 - t; }
            } else {
                t = 0.0;
                if (tmp1 <= 0.0) s = 1.0;
                else if (b0 >= 0.0) s = 0.0;
                else s = -b0 / a00;
            }
        } else {
            double numer = a11 + b1 - a01 - b0;
            if (numer <= 0.0) { s = 0.0; t = 1.0; }
            else {
                double denom = a00 - 2.0 * a01 + a11;
                if (numer >= denom) { s = 1.0; t = 0.0; }
                else { s = numer / denom; t = 1.0 - s; }
            }
        }
    }
    double w1 = 1.0 - s - t, w2 = s, w3 = t;
    int sig = (w1 > EPS_BARY) + (w2 > EPS_BARY) + (w3 > EPS_BARY);
    Region reg;
    if (sig == 3) {
        reg = IN_INTERIOR;
    } else if (sig == 2) {
        if (w1 > EPS_BARY && w2 > EPS_BARY) {
            reg = ON_SIDE21; w1 += w3 / 2.0; w2 += w3 / 2.0; w3 = 0.0;
        } else if (w1 > EPS_BARY && w3 > EPS_BARY) {
            reg = ON_SIDE13; w1 += w2 / 2.0; w3 += w2 / 2.0; w2 = 0.0;
        } else {
            reg = ON_SIDE32; w2 += w1 / 2.0; w3 += w1 / 2.0; w1 = 0.0;
        }
    } else {
        if (w1 > EPS_BARY) { reg = ON_NODE1; w1 = 1.0; w2 = 0.0; w3 = 0.0; }
        else if (w2 > EPS_BARY) { reg = ON_NODE2; w1 = 0.0; w2 = 1.0; w3 = 0.0; }
        else { reg = ON_NODE3; w1 = 0.0; w2 = 0.0; w3 = 1.0; }
    }
    *out = add(add(scl(w1, tr.a), scl(w2, tr.b)), scl(w3, tr.c));
    return reg;
}

/* ----------------------------------------------------- operator kernels */

/* regular branch of BemOperatorsConst (boundaryelementsolver.cpp:856-893) */
bem_ops4 ops_regular(const orc_problem& P, int64_t i, int64_t j)
{
    bem_ops4 o = bem_ops4_zero();
    V3 p = P.colloc[i], np = P.normal[i], nq = P.normal[j];
    double upnq = dot(np, nq);
    const Tri& t = P.tri[j];
    for (int q = 0; q < 16; q++) {
        double x1 = TRI16[q][0], x2 = TRI16[q][1], w = TRI16[q][2];
        double x3 = (1.0 - x2) - x1;
        double qx = (x1 * t.a.x + x2 * t.b.x) + x3 * t.c.x;
        double qy = (x1 * t.a.y + x2 * t.b.y) + x3 * t.c.y;
        double qz = (x1 * t.a.z + x2 * t.b.z) + x3 * t.c.z;
        bem_point_accumulate(&o, P.k, p.x, p.y, p.z, np.x, np.y, np.z, nq.x, nq.y, nq.z, upnq, qx, qy, qz, w);
    }
    double A = P.area[j];
    o.L = c64_scale(o.L, A); o.M = c64_scale(o.M, A); o.Mt = c64_scale(o.Mt, A); o.N = c64_scale(o.N, A);
    return o;
}

/* BemOperatorsNearSing (boundaryelementsolver.cpp:1387-1514): split the
   source triangle at the projection of the collocation point, 5x5 polar
   Gauss-Legendre on each sub-triangle (nq x nq with line rule x/w when given:
   the test suite's convergence check of the same polar decomposition). */
bem_ops4 ops_near_rule(const orc_problem& P, int64_t i, int64_t j, int nrule, const double* LX, const double* LW);
bem_ops4 ops_near(const orc_problem& P, int64_t i, int64_t j) { return ops_near_rule(P, i, j, 5, GL5X, GL5W); }
bem_ops4 ops_near_rule(const orc_problem& P, int64_t i, int64_t j, int nrule, const double* LX, const double* LW)
{
    bem_ops4 o = bem_ops4_zero();
    V3 p = P.colloc[i], np = P.normal[i], nq = P.normal[j];
    double upnq = dot(np, nq);
    const Tri& T = P.tri[j];
    V3 pp;
    Region reg = project_point(T, p, &pp);
    Tri st[3];
    int ns = 0;
    switch (reg) {
    case IN_INTERIOR: st[0] = {pp, T.a, T.b}; st[1] = {pp, T.b, T.c}; st[2] = {pp, T.c, T.a}; ns = 3; break;
    case ON_NODE1: st[0] = T; ns = 1; break;
    case ON_NODE2: st[0] = {T.b, T.c, T.a}; ns = 1; break;
    case ON_NODE3: st[0] = {T.c, T.a, T.b}; ns = 1; break;
    case ON_SIDE21: st[0] = {pp, T.c, T.a}; st[1] = {pp, T.b, T.c}; ns = 2; break;
    case ON_SIDE32: st[0] = {pp, T.c, T.a}; st[1] = {pp, T.a, T.b}; ns = 2; break;
    case ON_SIDE13: st[0] = {pp, T.a, T.b}; st[1] = {pp, T.b, T.c}; ns = 2; break;
    }
    for (int sidx = 0; sidx < ns; sidx++) {
        const Tri& c = st[sidx];
        V3 e13 = sub(c.c, c.a), e23 = sub(c.c, c.b), e12 = sub(c.b, c.a);
        double s13sq = dot(e13, e13), s23sq = dot(e23, e23), s12sq = dot(e12, e12);
        double s13 = std::sqrt(s13sq), s23 = std::sqrt(s23sq), s12 = std::sqrt(s12sq);
        double angP = std::acos((s13sq + s12sq - s23sq) / (2.0 * s13 * s12));
        double ang2 = std::acos((s23sq + s12sq - s13sq) / (2.0 * s23 * s12));
        V3 c0 = normalized(e12);
        V3 c1 = normalized(sub(e13, scl(dot(c0, e13), c0)));
        bem_ops4 t = bem_ops4_zero();
        for (int ia = 0; ia < nrule; ia++) {
            double angle1 = (angP + LX[ia] * angP) / 2.0;
            double maxR = s12 * std::sin(ang2) / (std::sin(angle1 + ang2));
            double ca = std::cos(angle1), sa = std::sin(angle1);
            for (int ib = 0; ib < nrule; ib++) {
                double rs = (maxR + LX[ib] * maxR) / 2.0;
                double x = rs * ca, y = rs * sa;
                double qx = (c.a.x + x * c0.x) + y * c1.x;
                double qy = (c.a.y + x * c0.y) + y * c1.y;
                double qz = (c.a.z + x * c0.z) + y * c1.z;
                double fw = ((maxR / 2.0 * rs) * LW[ia]) * LW[ib];
                bem_point_accumulate(&t, P.k, p.x, p.y, p.z, np.x, np.y, np.z, nq.x, nq.y, nq.z, upnq, qx, qy, qz, fw);
            }
        }
        double h = angP / 2.0;
        o.L = c64_axpy(h, t.L, o.L);
        o.M = c64_axpy(h, t.M, o.M);
        o.Mt = c64_axpy(h, t.Mt, o.Mt);
        o.N = c64_axpy(h, t.N, o.N);
    }
    return o;
}

/* polar radii of the self integral for one sub-triangle (node1, node2,
   centroid) of BemOperatorsSingularPolarInt (boundaryelementsolver.cpp:910-1015) */
void self_subtri(V3 n1, V3 n2, V3 n3, double* A_out, double* c_out, double* B_out)
{
    V3 C = sub(n3, n1), Bs = sub(n3, n2), As = sub(n2, n1);
    double bsq = dot(Bs, Bs), csq = dot(C, C), asq = dot(As, As);
    double b = std::sqrt(bsq), c = std::sqrt(csq);
    if (c < b) std::swap(b, c);
    double A = std::acos((bsq + csq - asq) / (2.0 * b * c));
    double B = std::atan((b * std::sin(A)) / (c - b * std::cos(A)));
    *A_out = A; *c_out = c; *B_out = B;
}

/* BemOperatorsSingularPolarInt (boundaryelementsolver.cpp:896-1029):
   L = (i/2k)(1 - S_L/2pi), N = ik/2 - S_N/4pi, M = Mt = 0 */
bem_ops4 ops_self(const orc_problem& P, int64_t i)
{
    const Tri& T = P.tri[i];
    V3 cp = P.colloc[i];
    V3 sn[3][3] = {{T.a, T.b, cp}, {T.b, T.c, cp}, {T.c, T.a, cp}};
    cplx Ls[3], Ns[3];
    double Aprev = 0, Acum = 0;
    for (int s = 0; s < 3; s++) {
        double A, c, B;
        self_subtri(sn[s][0], sn[s][1], sn[s][2], &A, &c, &B);
        double lower, upper, off;
        if (s == 0) { lower = 0.5 * (A - 0); upper = 0.5 * (A + 0); off = 0; }
        else if (s == 1) { off = Aprev; lower = 0.5 * A; upper = 0.5 * (A + off + off); }
        else { off = Acum; lower = 0.5 * A; upper = 0.5 * (A + off + off); }
        cplx L = c64(0, 0), N = c64(0, 0);
        double csB = c * std::sin(B);
        for (int q = 0; q < 16; q++) {
            double xv1 = HGX[q] * lower + upper;
            double xv2 = -HGX[q] * lower + upper;
            double r1, r2;
            if (s == 0) { r1 = csB / (std::sin(xv1 + B)); r2 = csB / (std::sin(xv2 + B)); }
            else { r1 = csB / (std::sin(xv1 - off + B)); r2 = csB / (std::sin(xv2 - off + B)); }
            cplx e1 = det_expi(P.k * r1), e2 = det_expi(P.k * r2);
            L = c64_axpy(HGW[q], c64_add(e1, e2), L);
            N = c64_axpy(HGW[q], c64_add(c64(e1.re / r1, e1.im / r1), c64(e2.re / r2, e2.im / r2)), N);
        }
        Ls[s] = c64_scale(L, lower);
        Ns[s] = c64_scale(N, lower);
        if (s == 0) { Aprev = A; }
        else if (s == 1) { Acum = Aprev + A; }
    }
    cplx LT = c64_add(c64_add(Ls[0], Ls[1]), Ls[2]);
    LT = c64(LT.re / (2.0 * BEM_PI), LT.im / (2.0 * BEM_PI));
    cplx NT = c64_add(c64_add(Ns[0], Ns[1]), Ns[2]);
    NT = c64(NT.re / (4.0 * BEM_PI), NT.im / (4.0 * BEM_PI));
    bem_ops4 o = bem_ops4_zero();
    o.L = c64_mul(c64(0.0, 1.0 / (2.0 * P.k)), c64(1.0 - LT.re, -LT.im));
    o.N = c64(0.0 - NT.re, P.k / 2.0 - NT.im);
    return o;
}

/* BemOperatorsReflected (boundaryelementsolver.cpp:1516-1553): the regular
   rule over the mirror image of element j, with the image's normal and the
   area of element j */
bem_ops4 ops_reflected(const orc_problem& P, int64_t i, int64_t j)
{
    bem_ops4 o = bem_ops4_zero();
    V3 p = P.colloc[i], np = P.normal[i], nq = P.rnormal[j];
    double upnq = dot(np, nq);
    const Tri& t = P.rtri[j];
    for (int q = 0; q < 16; q++) {
        double x1 = TRI16[q][0], x2 = TRI16[q][1], w = TRI16[q][2];
        double x3 = (1.0 - x2) - x1;
        double qx = (x1 * t.a.x + x2 * t.b.x) + x3 * t.c.x;
        double qy = (x1 * t.a.y + x2 * t.b.y) + x3 * t.c.y;
        double qz = (x1 * t.a.z + x2 * t.b.z) + x3 * t.c.z;
        bem_point_accumulate(&o, P.k, p.x, p.y, p.z, np.x, np.y, np.z, nq.x, nq.y, nq.z, upnq, qx, qy, qz, w);
    }
    double A = P.area[j];
    o.L = c64_scale(o.L, A); o.M = c64_scale(o.M, A); o.Mt = c64_scale(o.Mt, A); o.N = c64_scale(o.N, A);
    return o;
}

/* BemOperatorsConst (boundaryelementsolver.cpp:844-894) */
bem_ops4 ops4(const orc_problem& P, int64_t i, int64_t j)
{
    if (i == j) return ops_self(P, i);
    if (norm(sub(P.colloc[i], P.colloc[j])) < 4.0 * P.hbar) return ops_near(P, i, j);
    return ops_regular(P, i, j);
}

/* implicitPhiMatrix2 / implicitDPhiMatrix2 (boundaryelementsolver.cpp:2004-2040) */
cplx entry(const orc_problem& P, int kind, int64_t i, int64_t j)
{
    bem_ops4 o = ops4(P, i, j);
    cplx X, Y;
    bem_xy(&o, P.cc, i == j, &X, &Y);
    cplx e = kind == 0 ? bem_phi_entry(X, Y, P.subst[j], P.cphi[j]) : bem_dphi_entry(X, Y, P.subst[j], P.cd[j]);
    if (!P.plane) return e;
    /* implicitPhiMatrixPlusReflected / implicitDPhiMatrixPlusReflected
       (:2042-2050): the free-space entry plus the mirror image's entry
       (:1556-1643, same boundary-condition coefficients, no jump term) */
    o = ops_reflected(P, i, j);
    bem_xy(&o, P.cc, 0, &X, &Y);
    return c64_add(e, kind == 0 ? bem_phi_entry(X, Y, P.subst[j], P.cphi[j])
                                : bem_dphi_entry(X, Y, P.subst[j], P.cd[j]));
}

/* ------------------------------------------------------ cluster tree */

struct Box { V3 mn, mx; };
Box empty_box()
{
    double inf = (double)std::numeric_limits<float>::infinity();
    Box b; b.mn = v3(inf, inf, inf); b.mx = v3(-inf, -inf, -inf); return b;
}
void box_add(Box& b, V3 p)
{
    b.mn.x = std::min(b.mn.x, p.x); b.mn.y = std::min(b.mn.y, p.y); b.mn.z = std::min(b.mn.z, p.z);
    b.mx.x = std::max(b.mx.x, p.x); b.mx.y = std::max(b.mx.y, p.y); b.mx.z = std::max(b.mx.z, p.z);
}
double comp(V3 v, int d) { return d == 0 ? v.x : (d == 1 ? v.y : v.z); }
void setc(V3& v, int d, double x) { if (d == 0) v.x = x; else if (d == 1) v.y = x; else v.z = x; }

struct TreeBuild {
    const std::vector<V3>& mid;
    int64_t leaf;
    std::vector<Cluster>& out;
    std::vector<std::vector<int64_t>> idx; /* per cluster, element indices (input order) */
    /* ClusterTree::splitCluster (clustertree.h:264-310): leaf if |I| < leaf or
       |I| <= 1; else split the midpoint bounding box at the middle of its
       longest side, son1 = midpoints with coordinate >= cut */
    int make(std::vector<int64_t>&& I, int depth)
    {
        Box bx = empty_box();
        for (int64_t e : I) box_add(bx, mid[e]);
        int id = (int)out.size();
        out.push_back(Cluster());
        idx.push_back(std::move(I));
        out[id].bmin = bx.mn; out[id].bmax = bx.mx; out[id].depth = depth;
        out[id].size = (int64_t)idx[id].size();
        if (out[id].size < leaf || out[id].size <= 1) { out[id].leaf = true; return id; }
        V3 side = sub(bx.mx, bx.mn);
        int d = 0;
        double best = std::fabs(side.x);
        if (std::fabs(side.y) > best) { d = 1; best = std::fabs(side.y); }
        if (std::fabs(side.z) > best) { d = 2; }
        double cut = comp(bx.mn, d) + 0.5 * comp(side, d);
        std::vector<int64_t> s1, s2;
        for (int64_t e : idx[id]) {
            if (comp(mid[e], d) >= cut) s1.push_back(e); else s2.push_back(e);
        }
        int a = make(std::move(s1), depth + 1);
        out[id].son1 = a;
        int b = make(std::move(s2), depth + 1);
        out[id].son2 = b;
        return id;
    }
};

/* a cluster tree on the points mid, then depth-first reordering
   (clustertree.cpp:265-283): the depth-first leaf order gives perm
   (perm[new] = old); every cluster becomes a contiguous range */
void build_tree(const std::vector<V3>& mid, int64_t leaf, std::vector<Cluster>& clusters, std::vector<int64_t>& perm)
{
    const int64_t n = (int64_t)mid.size();
    std::vector<int64_t> all(n);
    for (int64_t e = 0; e < n; e++) all[e] = e;
    TreeBuild tb{mid, leaf, clusters, {}};
    tb.make(std::move(all), 0);
    int64_t pos = 0;
    struct Rec {
        std::vector<Cluster>& cl; std::vector<int64_t>& perm; TreeBuild* tb; int64_t* pos;
        void go(int c)
        {
            Cluster& C = cl[c];
            if (C.leaf) {
                C.start = *pos;
                for (int64_t e : tb->idx[c]) perm.push_back(e), (*pos)++;
            } else {
                go(C.son1); go(C.son2);
                C.start = cl[C.son1].start;
            }
        }
    } rec{clusters, perm, &tb, &pos};
    rec.go(0);
}

/* HMatrix::clusterDistance (hmatrix.cpp:1070-1129) */
double cluster_distance(const Cluster& c1, const Cluster& c2)
{
    V3 a0 = c1.bmin, a1 = c1.bmax, b0 = c2.bmin, b1 = c2.bmax;
    bool ov[3];
    double dd[3];
    for (int d = 0; d < 3; d++) {
        double mn1 = comp(a0, d), mx1 = comp(a1, d), mn2 = comp(b0, d), mx2 = comp(b1, d);
        ov[d] = (mn1 < mx2 && mx1 > mn2) || (mn2 < mx1 && mx2 > mn1);
        dd[d] = std::min(std::fabs(mn1 - mx2), std::fabs(mn2 - mx1));
    }
    int nov = ov[0] + ov[1] + ov[2];
    if (nov == 3) return 0;
    if (nov == 2) {
        if (!ov[0]) return dd[0];
        if (!ov[1]) return dd[1];
        return dd[2];
    }
    if (nov == 1) {
        if (ov[0]) return std::sqrt(dd[1] * dd[1] + dd[2] * dd[2]);
        if (ov[1]) return std::sqrt(dd[0] * dd[0] + dd[2] * dd[2]);
        return std::sqrt(dd[0] * dd[0] + dd[1] * dd[1]);
    }
    return std::sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
}

/* HMatrix::clusterSize (hmatrix.cpp:1131-1135) */
double cluster_size(const Cluster& c) { return norm(sub(c.bmax, c.bmin)); }

/* HMatrix::createBlockClusterChildren + isBlockAdmissible (hmatrix.cpp:958-1068),
   rows from the tree R, columns from the tree C */
void make_blocks2(const std::vector<Cluster>& R, const std::vector<Cluster>& C, int rc, int cc, double eta,
                  std::vector<Block>& out)
{
    const Cluster& r = R[rc];
    const Cluster& c = C[cc];
    double dist = cluster_distance(r, c);
    bool adm = false;
    if (dist != 0) adm = std::min(cluster_size(r), cluster_size(c)) / dist < eta;
    if (adm || r.leaf || c.leaf) {
        Block b;
        b.row0 = r.start; b.rows = r.size; b.col0 = c.start; b.cols = c.size; b.adm = adm; b.rc = rc; b.cc = cc;
        out.push_back(b);
        return;
    }
    make_blocks2(R, C, r.son1, c.son1, eta, out);
    make_blocks2(R, C, r.son1, c.son2, eta, out);
    make_blocks2(R, C, r.son2, c.son1, eta, out);
    make_blocks2(R, C, r.son2, c.son2, eta, out);
}

void make_blocks(orc_problem& P, int rc, int cc, double eta)
{
    const Cluster& r = P.clusters[rc];
    const Cluster& c = P.clusters[cc];
    double dist = cluster_distance(r, c);
    bool adm = false;
    if (dist != 0) adm = std::min(cluster_size(r), cluster_size(c)) / dist < eta;
    if (adm || r.leaf || c.leaf) {
        Block b;
        b.row0 = r.start; b.rows = r.size; b.col0 = c.start; b.cols = c.size; b.adm = adm; b.rc = rc; b.cc = cc;
        P.blocks.push_back(b);
        return;
    }
    make_blocks(P, r.son1, c.son1, eta);
    make_blocks(P, r.son1, c.son2, eta);
    make_blocks(P, r.son2, c.son1, eta);
    make_blocks(P, r.son2, c.son2, eta);
}

/* ------------------------------------------------ deterministic reductions */

const int NLANE = 64;
/* lane-strided partial sums over one 64-wide wavefront + pairwise tree (the
   device reduction order: lane l takes elements l, l+64, ..., then
   part[l] += part[l+s] for s = 32, 16, ..., 1) */
double det_sum_abs2(const cplx* v, int64_t n)
{
    double part[NLANE];
    for (int l = 0; l < NLANE; l++) part[l] = 0.0;
    for (int64_t i = 0; i < n; i++) part[i % NLANE] = part[i % NLANE] + c64_abs2(v[i]);
    for (int s = NLANE / 2; s > 0; s >>= 1)
        for (int l = 0; l < s; l++) part[l] = part[l] + part[l + s];
    return part[0];
}
/* sum_i conj(a_i) b_i with a stride on b */
cplx det_dot_conj(const cplx* a, const cplx* b, int64_t n)
{
    cplx part[NLANE];
    for (int l = 0; l < NLANE; l++) part[l] = c64(0, 0);
    for (int64_t i = 0; i < n; i++) part[i % NLANE] = c64_fma(c64_conj(a[i]), b[i], part[i % NLANE]);
    for (int s = NLANE / 2; s > 0; s >>= 1)
        for (int l = 0; l < s; l++) part[l] = c64_add(part[l], part[l + s]);
    return part[0];
}
/* first index of the largest |v_i|^2 (Eigen maxCoeff keeps the first) */
int64_t argmax_abs2(const cplx* v, int64_t n, double* maxval)
{
    int64_t best = 0;
    double bv = c64_abs2(v[0]);
    for (int64_t i = 1; i < n; i++) {
        double x = c64_abs2(v[i]);
        if (x > bv) { bv = x; best = i; }
    }
    *maxval = bv;
    return best;
}

/* ------------------------------------------------------------- ACA */

struct ACAState {
    int64_t m, n, cap, rank = 0;
    std::vector<cplx> U, alpha, V;
    cplx S = {0, 0};
};

/* the implicit matrix an ACA reads: entry (row, column), global indices */
using EntryFn = std::function<cplx(int64_t, int64_t)>;

void eval_row(const EntryFn& E, const Block& B, int64_t i, cplx* row)
{
    for (int64_t j = 0; j < B.cols; j++) row[j] = E(B.row0 + i, B.col0 + j);
}
void eval_col(const EntryFn& E, const Block& B, int64_t j, cplx* col)
{
    for (int64_t i = 0; i < B.rows; i++) col[i] = E(B.row0 + i, B.col0 + j);
}
/* row_j -= sum_l (U[i,l] alpha_l) V[l,j] */
void sub_row(ACAState& st, int64_t i, cplx* row)
{
    if (st.rank == 0) return;
    std::vector<cplx> w(st.rank);
    for (int64_t l = 0; l < st.rank; l++) w[l] = c64_mul(st.U[l * st.m + i], st.alpha[l]);
    for (int64_t j = 0; j < st.n; j++) {
        cplx s = c64(0, 0);
        for (int64_t l = 0; l < st.rank; l++) s = c64_fma(w[l], st.V[l * st.n + j], s);
        row[j] = c64_sub(row[j], s);
    }
}
/* col_i -= sum_l U[i,l] (alpha_l V[l,j]) */
void sub_col(ACAState& st, int64_t j, cplx* col)
{
    if (st.rank == 0) return;
    std::vector<cplx> z(st.rank);
    for (int64_t l = 0; l < st.rank; l++) z[l] = c64_mul(st.alpha[l], st.V[l * st.n + j]);
    for (int64_t i = 0; i < st.m; i++) {
        cplx s = c64(0, 0);
        for (int64_t l = 0; l < st.rank; l++) s = c64_fma(st.U[l * st.m + i], z[l], s);
        col[i] = c64_sub(col[i], s);
    }
}
/* append the cross (col, alpha, row) and update the Frobenius-norm estimate
   (hmatrix.cpp:560-575); returns the gain |alpha|^2 |row|^2 |col|^2 */
double append_cross(ACAState& st, const cplx* col, const cplx* row, cplx alpha)
{
    int64_t r = st.rank;
    std::memcpy(&st.U[r * st.m], col, sizeof(cplx) * st.m);
    std::memcpy(&st.V[r * st.n], row, sizeof(cplx) * st.n);
    st.alpha[r] = alpha;
    st.rank = r + 1;
    double nu = det_sum_abs2(row, st.n);
    double nc = det_sum_abs2(col, st.m);
    double gain = (c64_abs2(alpha) * nu) * nc;
    cplx S = st.S;
    S.re = S.re + gain;
    cplx ca2 = c64_scale(c64_conj(alpha), 2.0);
    for (int64_t l = 0; l < r; l++) {
        cplx d1 = det_dot_conj(col, &st.U[l * st.m], st.m);           /* col^H U_l */
        cplx d2 = det_dot_conj(row, &st.V[l * st.n], st.n);          /* V_l row^H */
        cplx t = c64_mul(c64_mul(c64_mul(ca2, st.alpha[l]), d1), d2);
        S = c64_add(S, t);
    }
    st.S = S;
    return gain;
}

/* BlockCluster::partialPivotACA (hmatrix.cpp:471-593) */
int64_t aca_plain(const EntryFn& E, const Block& B, int64_t bidx, double tol, int64_t maxrank, uint64_t seed,
                  ACAState& st)
{
    int64_t m = B.rows, n = B.cols, ev = 0;
    std::vector<cplx> row(n), col(m);
    int64_t j = bem_rand_bounded(seed, 0, (uint64_t)bidx, 0, n);
    eval_col(E, B, j, col.data()); ev += m;
    int64_t i = 0;
    double tiny2 = TINY * TINY;
    while (st.rank < maxrank) {
        if (st.rank >= 1) col[i] = c64(0, 0);
        double mv;
        i = argmax_abs2(col.data(), m, &mv);
        if (mv <= tiny2 && st.rank >= 1) break;
        eval_row(E, B, i, row.data()); ev += n;
        sub_row(st, i, row.data());
        double mr;
        j = argmax_abs2(row.data(), n, &mr);
        if (mr <= tiny2 && st.rank >= 1) break;
        eval_col(E, B, j, col.data()); ev += m;
        cplx alpha = c64_div(c64(1.0, 0.0), row[j]);
        sub_col(st, j, col.data());
        double gain = append_cross(st, col.data(), row.data(), alpha);
        double target = (tol * tol) * st.S.re;
        if (tol > 0 && gain < target) break;
    }
    return ev;
}

/* BoundaryElementSolver::getNormalFilteredPivotIndices
   (boundaryelementsolver.cpp:2052-2172): greedy normal-orientation groups
   (cone cos(0.44), at most 20 groups) of rows and columns; one random member
   per (row group, column group). */
std::vector<std::vector<int64_t>> normal_groups(const orc_problem& P, int64_t start, int64_t count)
{
    const int maxc = 20;
    double fv = std::cos(0.44);
    std::vector<std::vector<int64_t>> G(maxc);
    for (int64_t e = 0; e < count; e++) G[0].push_back(start + e);
    int64_t nun = count;
    int ci = 0;
    while (nun > 1 && ci < maxc - 1) {
        G[ci + 1].clear();
        std::vector<int64_t> keep;
        V3 ref = P.normal[G[ci][0]];
        keep.push_back(G[ci][0]);
        for (size_t t = 1; t < G[ci].size(); t++) {
            int64_t e = G[ci][t];
            if (std::fabs(dot(ref, P.normal[e])) < fv) G[ci + 1].push_back(e);
            else keep.push_back(e);
        }
        G[ci] = keep;
        nun = (int64_t)G[ci + 1].size();
        if (nun >= 1) ci++;
    }
    G.resize(ci + 1);
    return G;
}

std::vector<std::pair<int64_t, int64_t>> pivot_list(const orc_problem& P, const Block& B, int64_t bidx, uint64_t seed)
{
    auto R = normal_groups(P, B.row0, B.rows);
    auto C = normal_groups(P, B.col0, B.cols);
    std::vector<std::pair<int64_t, int64_t>> out;
    uint64_t draw = 0;
    for (size_t a = 0; a < R.size(); a++)
        for (size_t b = 0; b < C.size(); b++) {
            int64_t ri = bem_rand_bounded(seed, 1, (uint64_t)bidx, draw++, (int64_t)R[a].size());
            int64_t cj = bem_rand_bounded(seed, 1, (uint64_t)bidx, draw++, (int64_t)C[b].size());
            out.push_back({R[a][ri] - B.row0, C[b][cj] - B.col0});
        }
    return out;
}

/* getNormalFilteredPivotIndicesForField (boundaryelementsolver.cpp:2721-2807):
   orientation groups of the block's boundary columns only; as many random
   rows as groups, drawn first; then for every (row, group) a random member
   of the group.  Draws from stream 2 of the counter-based generator. */
std::vector<std::pair<int64_t, int64_t>> field_pivot_list(const orc_problem& P, const Block& B, int64_t bidx,
                                                          uint64_t seed)
{
    auto C = normal_groups(P, B.col0, B.cols);
    uint64_t draw = 0;
    std::vector<int64_t> rows(C.size());
    for (auto& r : rows) r = bem_rand_bounded(seed, 2, (uint64_t)bidx, draw++, B.rows);
    std::vector<std::pair<int64_t, int64_t>> out;
    for (int64_t r : rows)
        for (auto& g : C) {
            int64_t cj = bem_rand_bounded(seed, 2, (uint64_t)bidx, draw++, (int64_t)g.size());
            out.push_back({r, g[cj] - B.col0});
        }
    return out;
}

/* implicitPhiMatrixField (kind 0: M, :2647-2683) / implicitDPhiMatrixField
   (kind 1: L, :2685-2719) of observation point p and element j: the 7-point
   rule of BemOperatorField (:2319-2353) */
cplx field_entry(const orc_problem& P, int kind, V3 p, int64_t j)
{
    static const double rule[7][3] = BEM_TRI7;
    const Tri& t = P.tri[j];
    const V3 nq = P.normal[j];
    cplx L = c64(0, 0), M = c64(0, 0);
    for (int q = 0; q < 7; q++) {
        double x1 = rule[q][0], x2 = rule[q][1];
        double x3 = (1.0 - x2) - x1;
        double qx = (x1 * t.a.x + x2 * t.b.x) + x3 * t.c.x;
        double qy = (x1 * t.a.y + x2 * t.b.y) + x3 * t.c.y;
        double qz = (x1 * t.a.z + x2 * t.b.z) + x3 * t.c.z;
        bem_field_point_accumulate(&L, &M, P.k, p.x, p.y, p.z, nq.x, nq.y, nq.z, qx, qy, qz, rule[q][2]);
    }
    return c64_scale(kind == 0 ? M : L, P.area[j]);
}

/* BlockCluster::partialPivotACAextra (hmatrix.cpp:595-745) over the pivot
   list piv (pairs relative to the block origin) */
int64_t aca_extra(const EntryFn& E, const std::vector<std::pair<int64_t, int64_t>>& piv, const Block& B, double tol,
                  int64_t maxrank, ACAState& st)
{
    int64_t m = B.rows, n = B.cols, ev = 0;
    std::vector<cplx> row(n), col(m), colabs(m);
    double tiny2 = TINY * TINY;
    for (size_t pi = 0; pi < piv.size() && st.rank < maxrank; pi++) {
        int64_t i = piv[pi].first, j = piv[pi].second;
        cplx a = E(B.row0 + i, B.col0 + j); ev++;
        if (c64_abs2(a) < tiny2) continue;
        cplx approx = c64(0, 0);
        for (int64_t l = 0; l < st.rank; l++)
            approx = c64_add(approx, c64_mul(c64_mul(st.U[l * m + i], st.alpha[l]), st.V[l * n + j]));
        if (c64_abs2(c64_sub(a, approx)) < (tol * tol) * c64_abs2(a)) continue;
        eval_col(E, B, j, col.data()); ev += m;
        sub_col(st, j, col.data());
        bool first = true;
        while (st.rank < maxrank) {
            std::memcpy(colabs.data(), col.data(), sizeof(cplx) * m);
            if (!first) colabs[i] = c64(0, 0);
            double mv;
            i = argmax_abs2(colabs.data(), m, &mv);
            first = false;
            if (c64_abs2(col[i]) <= tiny2 && st.rank >= 1) break;
            eval_row(E, B, i, row.data()); ev += n;
            sub_row(st, i, row.data());
            double mr;
            j = argmax_abs2(row.data(), n, &mr);
            if (mr <= tiny2 && st.rank >= 1) break;
            eval_col(E, B, j, col.data()); ev += m;
            cplx alpha = c64_div(c64(1.0, 0.0), row[j]);
            sub_col(st, j, col.data());
            cplx oldS = st.S;
            double gain = append_cross(st, col.data(), row.data(), alpha);
            double target = (tol * tol) * st.S.re;
            if (tol > 0 && gain < target) {
                if (gain < 0.1 * target) { st.rank--; st.S = oldS; }
                break;
            }
        }
    }
    return ev;
}

/* ---------------------------------------------------------- MVM */

/* HArithm::MVM (harithm.cpp:3-28): y += H x; block products computed in
   parallel, added to y in partition order */
void hmvm(const orc_hmat& H, const cplx* x, cplx* y)
{
    const orc_problem& P = *H.p;
    int64_t nb = (int64_t)P.blocks.size();
    std::vector<std::vector<cplx>> tmp(nb);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t b = 0; b < nb; b++) {
        const Block& B = P.blocks[b];
        const auto& K = H.blk[b];
        int64_t m = B.rows, n = B.cols;
        std::vector<cplx>& t = tmp[b];
        t.assign(m, c64(0, 0));
        const cplx* xs = x + B.col0;
        if (K.rank < 0) {
            for (int64_t j = 0; j < n; j++)
                for (int64_t i = 0; i < m; i++) t[i] = c64_fma(K.F[j * m + i], xs[j], t[i]);
        } else {
            int64_t k = K.rank;
            std::vector<cplx> tv(k);
            for (int64_t l = 0; l < k; l++) {
                cplx s = c64(0, 0);
                for (int64_t j = 0; j < n; j++) s = c64_fma(K.V[l * n + j], xs[j], s);
                tv[l] = c64_mul(K.alpha[l], s);
            }
            for (int64_t l = 0; l < k; l++)
                for (int64_t i = 0; i < m; i++) t[i] = c64_fma(K.U[l * m + i], tv[l], t[i]);
        }
    }
    for (int64_t b = 0; b < nb; b++) {
        const Block& B = P.blocks[b];
        for (int64_t i = 0; i < B.rows; i++) y[B.row0 + i] = c64_add(y[B.row0 + i], tmp[b][i]);
    }
}

/* ---------------------------------------------------------- GMRES */

cplx vdot(const std::vector<cplx>& a, const cplx* b, int64_t n) /* a^H b */
{
    cplx s = c64(0, 0);
    for (int64_t i = 0; i < n; i++) s = c64_fma(c64_conj(a[i]), b[i], s);
    return s;
}
double vnorm(const cplx* a, int64_t n)
{
    double s = 0;
    for (int64_t i = 0; i < n; i++) s += c64_abs2(a[i]);
    return std::sqrt(s);
}

/* restarted GMRES(m), CGS2 Arnoldi, complex Givens (deviation D3 from
   Eigen::GMRES used at gmres.h:170-180) */
int64_t gmres(const orc_hmat& H, const cplx* b, cplx* x, double tol, int64_t restart, int64_t maxit, double* est)
{
    int64_t n = H.p->n;
    double bn = vnorm(b, n);
    if (bn == 0) { for (int64_t i = 0; i < n; i++) x[i] = c64(0, 0); *est = 0; return 0; }
    std::vector<std::vector<cplx>> Vb(restart + 1, std::vector<cplx>(n));
    std::vector<cplx> Hm((restart + 1) * restart), cs_s(restart), g(restart + 1), w(n);
    std::vector<double> cs_c(restart);
    int64_t it = 0;
    double rel = 1.0;
    while (true) {
        /* r = b - A x */
        std::vector<cplx> r(n, c64(0, 0));
        hmvm(H, x, r.data());
        for (int64_t i = 0; i < n; i++) r[i] = c64_sub(b[i], r[i]);
        double beta = vnorm(r.data(), n);
        rel = beta / bn;
        if (rel < tol || it >= maxit) break;
        for (int64_t i = 0; i < n; i++) Vb[0][i] = c64_scale(r[i], 1.0 / beta);
        for (auto& z : g) z = c64(0, 0);
        g[0] = c64(beta, 0);
        int64_t j = 0;
        for (; j < restart && it < maxit; j++, it++) {
            for (int64_t i = 0; i < n; i++) w[i] = c64(0, 0);
            hmvm(H, Vb[j].data(), w.data());
            std::vector<cplx> h(j + 2, c64(0, 0));
            for (int pass = 0; pass < 2; pass++) {
                std::vector<cplx> hh(j + 1);
                for (int64_t l = 0; l <= j; l++) hh[l] = vdot(Vb[l], w.data(), n);
                for (int64_t l = 0; l <= j; l++) {
                    for (int64_t i = 0; i < n; i++) w[i] = c64_fms(hh[l], Vb[l][i], w[i]);
                    h[l] = c64_add(h[l], hh[l]);
                }
            }
            double hn = vnorm(w.data(), n);
            h[j + 1] = c64(hn, 0);
            for (int64_t i = 0; i < n; i++) Vb[j + 1][i] = hn > 0 ? c64_scale(w[i], 1.0 / hn) : c64(0, 0);
            for (int64_t l = 0; l < j; l++) { /* apply previous rotations */
                cplx a = h[l], bb = h[l + 1];
                h[l] = c64_add(c64_scale(a, cs_c[l]), c64_mul(cs_s[l], bb));
                h[l + 1] = c64_sub(c64_scale(bb, cs_c[l]), c64_mul(c64_conj(cs_s[l]), a));
            }
            double na = std::sqrt(c64_abs2(h[j])), nb = std::sqrt(c64_abs2(h[j + 1]));
            double t = std::sqrt(na * na + nb * nb);
            double c;
            cplx s;
            if (na == 0) { c = 0; s = c64(1, 0); }
            else {
                c = na / t;
                cplx ph = c64_scale(h[j], 1.0 / na);
                s = c64_scale(c64_mul(ph, c64_conj(h[j + 1])), 1.0 / t);
            }
            cs_c[j] = c; cs_s[j] = s;
            h[j] = c64_add(c64_scale(h[j], c), c64_mul(s, h[j + 1]));
            h[j + 1] = c64(0, 0);
            cplx gj = g[j];
            g[j] = c64_scale(gj, c);
            g[j + 1] = c64_neg(c64_mul(c64_conj(s), gj));
            for (int64_t l = 0; l <= j; l++) Hm[j * (restart + 1) + l] = h[l];
            rel = std::sqrt(c64_abs2(g[j + 1])) / bn;
            if (rel < tol) { j++; it++; break; }
        }
        /* y = R^-1 g, x += V y */
        std::vector<cplx> y(j);
        for (int64_t l = j - 1; l >= 0; l--) {
            cplx s = g[l];
            for (int64_t q = l + 1; q < j; q++) s = c64_sub(s, c64_mul(Hm[q * (restart + 1) + l], y[q]));
            y[l] = c64_div(s, Hm[l * (restart + 1) + l]);
        }
        for (int64_t l = 0; l < j; l++)
            for (int64_t i = 0; i < n; i++) x[i] = c64_fma(Vb[l][i], y[l], x[i]);
        if (rel < tol) {
            /* confirm with the true residual on the next pass */
        }
    }
    *est = rel;
    return it;
}

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

} // namespace

/* ================================================================ C ABI */

extern "C" {

orc_problem* orc_new(int64_t n, const double* nodes, const double* bc, int64_t leaf, double eta)
{
    if (n <= 0) return nullptr;
    auto P = new orc_problem();
    P->n = n;
    std::vector<Tri> tin(n);
    std::vector<V3> mid(n);
    for (int64_t e = 0; e < n; e++) {
        const double* v = nodes + 9 * e;
        tin[e].a = v3(v[0], v[1], v[2]); tin[e].b = v3(v[3], v[4], v[5]); tin[e].c = v3(v[6], v[7], v[8]);
        mid[e] = tri_mid(tin[e]);
    }
    /* ClusterTree(BoundaryElements*) (clustertree.h:77-86) */
    build_tree(mid, leaf, P->clusters, P->perm);
    P->tri.resize(n); P->normal.resize(n); P->colloc.resize(n); P->area.resize(n);
    P->a.resize(n); P->b.resize(n); P->g.resize(n);
    for (int64_t e = 0; e < n; e++) {
        int64_t o = P->perm[e];
        P->tri[e] = tin[o];
        P->normal[e] = tri_normal(tin[o]);
        P->colloc[e] = tri_mid(tin[o]);
        P->area[e] = tri_area(tin[o]);
        const double* q = bc + 6 * o;
        P->a[e] = c64(q[0], q[1]); P->b[e] = c64(q[2], q[3]); P->g[e] = c64(q[4], q[5]);
    }
    /* calculateTrianglesAreaAndAverageDimension (boundaryelements.cpp:35-46) */
    double sum = 0;
    for (int64_t e = 0; e < n; e++) sum += std::sqrt(P->area[e]);
    P->hbar = (sum / (double)n) / 2.0;
    make_blocks(*P, 0, 0, eta);
    P->s.assign(n, c64(0, 0));
    return P;
}

void orc_free(orc_problem* p) { delete p; }
int64_t orc_n(const orc_problem* p) { return p->n; }
int64_t orc_nblocks(const orc_problem* p) { return (int64_t)p->blocks.size(); }
void orc_get_perm(const orc_problem* p, int64_t* perm) { std::copy(p->perm.begin(), p->perm.end(), perm); }
void orc_get_blocks(const orc_problem* p, int64_t* out)
{
    for (size_t b = 0; b < p->blocks.size(); b++) {
        const Block& B = p->blocks[b];
        out[5 * b + 0] = B.row0; out[5 * b + 1] = B.rows; out[5 * b + 2] = B.col0; out[5 * b + 3] = B.cols;
        out[5 * b + 4] = B.adm ? 1 : 0;
    }
}
void orc_get_geometry(const orc_problem* p, double* nodes9, double* normal3, double* colloc3, double* area, double* hbar)
{
    for (int64_t e = 0; e < p->n; e++) {
        const Tri& t = p->tri[e];
        double v[9] = {t.a.x, t.a.y, t.a.z, t.b.x, t.b.y, t.b.z, t.c.x, t.c.y, t.c.z};
        if (nodes9) std::memcpy(nodes9 + 9 * e, v, sizeof v);
        if (normal3) { normal3[3 * e] = p->normal[e].x; normal3[3 * e + 1] = p->normal[e].y; normal3[3 * e + 2] = p->normal[e].z; }
        if (colloc3) { colloc3[3 * e] = p->colloc[e].x; colloc3[3 * e + 1] = p->colloc[e].y; colloc3[3 * e + 2] = p->colloc[e].z; }
        if (area) area[e] = p->area[e];
    }
    if (hbar) *hbar = p->hbar;
}
int64_t orc_nclusters(const orc_problem* p) { return (int64_t)p->clusters.size(); }
void orc_get_clusters(const orc_problem* p, double* o)
{
    for (size_t c = 0; c < p->clusters.size(); c++) {
        const Cluster& C = p->clusters[c];
        double v[10] = {(double)C.start, (double)C.size, (double)C.depth, C.leaf ? 1.0 : 0.0,
                        C.bmin.x, C.bmin.y, C.bmin.z, C.bmax.x, C.bmax.y, C.bmax.z};
        std::memcpy(o + 10 * c, v, sizeof v);
    }
}

/* calculateBoundarySolution (boundaryelementsolver.cpp:30-59) and
   calculateBoundaryConditionAndSourceTermVector (:3386-3429) */
int orc_set_frequency(orc_problem* p, double freq, double c, int coupling)
{
    p->freq = freq;
    p->k = (2.0 * freq * BEM_PI) / c;
    if (coupling == 0) p->cc = c64(0, 0);
    else if (coupling == 1) p->cc = std::fabs(p->k) < TINY ? c64(0, 1) : c64(0.0, 1.0 / p->k);
    else p->cc = c64_div(c64(0, 1), c64(p->k + 1.0, 0));
    int64_t n = p->n;
    p->subst.resize(n); p->cphi.resize(n); p->cd.resize(n);
    for (int64_t j = 0; j < n; j++) {
        double aa = std::hypot(p->a[j].re, p->a[j].im), bb = std::hypot(p->b[j].re, p->b[j].im);
        p->subst[j] = aa < bb;
        if (p->subst[j]) {
            p->cphi[j] = c64_div(c64(-1.0, 0.0), p->b[j]);
            p->cd[j] = c64_div(c64_neg(p->a[j]), p->b[j]);
        } else {
            p->cphi[j] = c64_div(c64(1.0, 0.0), p->a[j]);
            p->cd[j] = c64_div(p->b[j], p->a[j]);
        }
    }
    return 0;
}

/* sourceTerm (boundaryelementsolver.cpp:3431-3441, including its division
   by (ikr - 1)) for point sources; plane wave phi_inc = A e^{ik d.x}:
   s = phi_inc + cc ik (d.n) phi_inc */
void orc_set_sources(orc_problem* p, int64_t npoint, const double* pts, int has_plane, const double* pl)
{
    int64_t n = p->n;
    for (int64_t i = 0; i < n; i++) {
        cplx acc = c64(0, 0);
        V3 x = p->colloc[i], nn = p->normal[i];
        for (int64_t s = 0; s < npoint; s++) {
            const double* q = pts + 6 * s;
            V3 R = sub(x, v3(q[0], q[1], q[2]));
            double r = norm(R);
            double rup = dot(R, nn) / r;
            cplx G = c64_mul(c64(q[3], q[4]), c64_scale(det_expi(p->k * r), 1.0 / (4.0 * BEM_PI * r)));
            cplx dG = c64_div(c64_scale(G, 1.0 / r), c64(-1.0, p->k * r));
            acc = c64_add(acc, c64_add(G, c64_mul(p->cc, c64_scale(dG, rup))));
        }
        if (has_plane) {
            V3 d = v3(pl[0], pl[1], pl[2]);
            cplx phi = c64_mul(c64(pl[3], pl[4]), det_expi(p->k * dot(d, x)));
            cplx dphi = c64_mul(c64(0.0, p->k * dot(d, nn)), phi);
            acc = c64_add(acc, c64_add(phi, c64_mul(p->cc, dphi)));
        }
        p->s[i] = acc;
    }
}

void orc_get_rhs_parts(const orc_problem* p, double* s2, double* a2, double* b2, double* f2, int32_t* subst)
{
    for (int64_t i = 0; i < p->n; i++) {
        if (s2) { s2[2 * i] = p->s[i].re; s2[2 * i + 1] = p->s[i].im; }
        if (a2) { a2[2 * i] = p->a[i].re; a2[2 * i + 1] = p->a[i].im; }
        if (b2) { b2[2 * i] = p->b[i].re; b2[2 * i + 1] = p->b[i].im; }
        if (f2) { f2[2 * i] = p->g[i].re; f2[2 * i + 1] = p->g[i].im; }
        if (subst) subst[i] = p->subst.empty() ? -1 : p->subst[i];
    }
}
void orc_get_params(const orc_problem* p, double* o) { o[0] = p->k; o[1] = p->cc.re; o[2] = p->cc.im; o[3] = p->hbar; }

void orc_field(const orc_problem* p, int64_t nobs, const double* obs3, const double* phi2, const double* dphi2,
               double* out2)
{
    static const double rule[7][3] = BEM_TRI7;
    const int64_t n = p->n, chunk = 1024;
    const cplx* phi = (const cplx*)phi2;
    const cplx* dphi = (const cplx*)dphi2;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nobs; i++) {
        const double px = obs3[3 * i], py = obs3[3 * i + 1], pz = obs3[3 * i + 2];
        cplx total = c64(0, 0);
        for (int64_t j0 = 0; j0 < n; j0 += chunk) {
            const int64_t j1 = std::min(n, j0 + chunk);
            cplx part[4] = {c64(0, 0), c64(0, 0), c64(0, 0), c64(0, 0)};
            for (int64_t j = j0; j < j1; j++) {
                auto term = [&](const Tri& t, const V3& nq) {
                    cplx L = c64(0, 0), M = c64(0, 0);
                    for (int q = 0; q < 7; q++) {
                        double x1 = rule[q][0], x2 = rule[q][1];
                        double x3 = (1.0 - x2) - x1;
                        double qx = (x1 * t.a.x + x2 * t.b.x) + x3 * t.c.x;
                        double qy = (x1 * t.a.y + x2 * t.b.y) + x3 * t.c.y;
                        double qz = (x1 * t.a.z + x2 * t.b.z) + x3 * t.c.z;
                        bem_field_point_accumulate(&L, &M, p->k, px, py, pz, nq.x, nq.y, nq.z, qx, qy, qz,
                                                   rule[q][2]);
                    }
                    L = c64_scale(L, p->area[j]);
                    M = c64_scale(M, p->area[j]);
                    return bem_field_term(L, M, phi[j], dphi[j]);
                };
                cplx v = term(p->tri[j], p->normal[j]);
                /* the mirror image's term (:2305-2311) */
                if (p->plane) v = c64_add(v, term(p->rtri[j], p->rnormal[j]));
                part[j & 3] = c64_add(part[j & 3], v);
            }
            cplx cs = c64_add(c64_add(part[0], part[1]), c64_add(part[2], part[3]));
            total = j0 == 0 ? cs : c64_add(total, cs);
        }
        out2[2 * i] = total.re;
        out2[2 * i + 1] = total.im;
    }
}

/* calculateFieldSolutionFast (boundaryelementsolver.cpp:2397-2476): the
   boundary term M phi - L dphi at the points obs3 (input order) through the
   observation x boundary H-matrix.  The points' own cluster tree (leaf),
   partition against the boundary tree (eta); per admissible block ACA of
   the M kernel and, with the field pivots, of the L kernel, each block's
   U alpha V x applied right away; dense blocks row by row; every block's
   term added to its rows in partition order.  out2 in input order. */
void orc_field_hmat(const orc_problem* p, int64_t nobs, const double* obs3, const double* phi2, const double* dphi2,
                    double tol, int64_t maxrank, uint64_t seed, int64_t leaf, double eta, double* out2, double* stats3)
{
    const cplx* phi = (const cplx*)phi2;
    const cplx* dphi = (const cplx*)dphi2;
    std::vector<V3> pts(nobs);
    for (int64_t i = 0; i < nobs; i++) pts[i] = v3(obs3[3 * i], obs3[3 * i + 1], obs3[3 * i + 2]);
    std::vector<Cluster> fcl;
    std::vector<int64_t> fperm;
    build_tree(pts, leaf, fcl, fperm);
    std::vector<V3> obs(nobs);
    for (int64_t i = 0; i < nobs; i++) obs[i] = pts[fperm[i]];
    std::vector<Block> blocks;
    make_blocks2(fcl, p->clusters, 0, 0, eta, blocks);
    const int64_t nb = (int64_t)blocks.size();
    std::vector<std::vector<cplx>> term(nb);
    int64_t nadm = 0, maxr = 0, evaluated = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : nadm, evaluated) reduction(max : maxr)
    for (int64_t b = 0; b < nb; b++) {
        const Block& B = blocks[b];
        auto& T = term[b];
        T.assign(B.rows, c64(0, 0));
        if (!B.adm) {
            for (int64_t i = 0; i < B.rows; i++) {
                cplx sm = c64(0, 0), sl = c64(0, 0);
                for (int64_t j = 0; j < B.cols; j++) {
                    sm = c64_fma(field_entry(*p, 0, obs[B.row0 + i], B.col0 + j), phi[B.col0 + j], sm);
                    sl = c64_fma(field_entry(*p, 1, obs[B.row0 + i], B.col0 + j), dphi[B.col0 + j], sl);
                }
                T[i] = c64_sub(sm, sl);
            }
            continue;
        }
        nadm++;
        int64_t mr = std::min(B.rows, B.cols);
        if (maxrank > 0) mr = std::min(mr, maxrank);
        for (int kind = 0; kind < 2; kind++) {
            const EntryFn E = [&](int64_t i, int64_t j) { return field_entry(*p, kind, obs[i], j); };
            ACAState st;
            st.m = B.rows; st.n = B.cols; st.cap = mr;
            st.U.assign(mr * B.rows, c64(0, 0)); st.alpha.assign(mr, c64(0, 0)); st.V.assign(mr * B.cols, c64(0, 0));
            evaluated += kind == 0 ? aca_plain(E, B, b, tol, mr, seed, st)
                                   : aca_extra(E, field_pivot_list(*p, B, b, seed), B, tol, mr, st);
            maxr = std::max(maxr, st.rank);
            const cplx* x = (kind == 0 ? phi : dphi) + B.col0;
            std::vector<cplx> t(st.rank);
            for (int64_t l = 0; l < st.rank; l++) {
                cplx d = c64(0, 0);
                for (int64_t j = 0; j < B.cols; j++) d = c64_fma(st.V[l * B.cols + j], x[j], d);
                t[l] = c64_mul(st.alpha[l], d);
            }
            for (int64_t i = 0; i < B.rows; i++) {
                cplx a = c64(0, 0);
                for (int64_t l = 0; l < st.rank; l++) a = c64_fma(st.U[l * B.rows + i], t[l], a);
                T[i] = kind == 0 ? a : c64_sub(T[i], a);
            }
        }
    }
    std::vector<cplx> f(nobs, c64(0, 0));
    for (int64_t b = 0; b < nb; b++)
        for (int64_t i = 0; i < blocks[b].rows; i++) f[blocks[b].row0 + i] = c64_add(f[blocks[b].row0 + i], term[b][i]);
    for (int64_t i = 0; i < nobs; i++) {
        out2[2 * fperm[i]] = f[i].re;
        out2[2 * fperm[i] + 1] = f[i].im;
    }
    if (stats3) {
        stats3[0] = (double)nadm;
        stats3[1] = (double)maxr;
        stats3[2] = (double)evaluated;
    }
}

/* the points' cluster tree and the observation x boundary partition of
   orc_field_hmat: perm (perm[new] = old), blocks5 (row0, rows, col0, cols,
   admissible) up to cap blocks; returns the number of blocks */
int64_t orc_field_partition(const orc_problem* p, int64_t nobs, const double* obs3, int64_t leaf, double eta,
                            int64_t* perm, int64_t* blocks5, int64_t cap)
{
    std::vector<V3> pts(nobs);
    for (int64_t i = 0; i < nobs; i++) pts[i] = v3(obs3[3 * i], obs3[3 * i + 1], obs3[3 * i + 2]);
    std::vector<Cluster> fcl;
    std::vector<int64_t> fperm;
    build_tree(pts, leaf, fcl, fperm);
    std::copy(fperm.begin(), fperm.end(), perm);
    std::vector<Block> blocks;
    make_blocks2(fcl, p->clusters, 0, 0, eta, blocks);
    for (int64_t b = 0; b < (int64_t)blocks.size() && b < cap; b++) {
        const Block& B = blocks[b];
        int64_t* q = blocks5 + 5 * b;
        q[0] = B.row0; q[1] = B.rows; q[2] = B.col0; q[3] = B.cols; q[4] = B.adm;
    }
    return (int64_t)blocks.size();
}

/* MeshFunctions::reflectTriangles (meshfunctions.cpp:349-378) through
   BoundaryElements::reflectGeometry (boundaryelements.cpp:151-155): the
   plane normal normalised, x' = R [x; 1] with R = [I - 2 n n^T | -2 n d],
   d = -n.point, each row summed left to right; nodes reversed (a' = R c,
   b' = R b, c' = R a) and the normal recomputed from them; areas unchanged
   (copied before the reflection) */
void orc_set_plane(orc_problem* p, const double* point3, const double* normal3)
{
    p->plane = normal3 && (normal3[0] != 0 || normal3[1] != 0 || normal3[2] != 0);
    p->rtri.clear();
    p->rnormal.clear();
    if (!p->plane) return;
    V3 u = normalized(v3(normal3[0], normal3[1], normal3[2]));
    double d = -dot(u, v3(point3[0], point3[1], point3[2]));
    double R[3][4] = {{1.0 - 2.0 * (u.x * u.x), -2.0 * u.x * u.y, -2.0 * u.x * u.z, -2.0 * u.x * d},
                      {-2.0 * u.x * u.y, 1.0 - 2.0 * (u.y * u.y), -2.0 * u.y * u.z, -2.0 * u.y * d},
                      {-2.0 * u.x * u.z, -2.0 * u.y * u.z, 1.0 - 2.0 * (u.z * u.z), -2.0 * u.z * d}};
    auto mirror = [&](V3 x) {
        double o[3];
        for (int r = 0; r < 3; r++) o[r] = ((R[r][0] * x.x + R[r][1] * x.y) + R[r][2] * x.z) + R[r][3];
        return v3(o[0], o[1], o[2]);
    };
    p->rtri.resize(p->n);
    p->rnormal.resize(p->n);
    for (int64_t e = 0; e < p->n; e++) {
        const Tri& t = p->tri[e];
        Tri r;
        r.a = mirror(t.c);
        r.b = mirror(t.b);
        r.c = mirror(t.a);
        p->rtri[e] = r;
        p->rnormal[e] = tri_normal(r);
    }
}

void orc_entries4(const orc_problem* p, int64_t np, const int64_t* ij, double* out)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t t = 0; t < np; t++) {
        bem_ops4 o = ops4(*p, ij[2 * t], ij[2 * t + 1]);
        double v[8] = {o.L.re, o.L.im, o.M.re, o.M.im, o.Mt.re, o.Mt.im, o.N.re, o.N.im};
        std::memcpy(out + 8 * t, v, sizeof v);
    }
}
void orc_entries4_polar(const orc_problem* p, int64_t np, const int64_t* ij, int nq, const double* x,
                        const double* w, double* out)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < np; t++) {
        bem_ops4 o = ops_near_rule(*p, ij[2 * t], ij[2 * t + 1], nq, x, w);
        double v[8] = {o.L.re, o.L.im, o.M.re, o.M.im, o.Mt.re, o.Mt.im, o.N.re, o.N.im};
        std::memcpy(out + 8 * t, v, sizeof v);
    }
}
void orc_entries(const orc_problem* p, int kind, int64_t np, const int64_t* ij, double* out)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t t = 0; t < np; t++) {
        cplx z = entry(*p, kind, ij[2 * t], ij[2 * t + 1]);
        out[2 * t] = z.re; out[2 * t + 1] = z.im;
    }
}
void orc_dense_matrix(const orc_problem* p, int kind, double* out)
{
    int64_t n = p->n;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = 0; i < n; i++) {
            cplx z = entry(*p, kind, i, j);
            out[2 * (j * n + i)] = z.re; out[2 * (j * n + i) + 1] = z.im;
        }
}

/* HMatrix::assembleBlocks / assembleBlocksExtra (hmatrix.cpp:793-846) */
orc_hmat* orc_assemble(const orc_problem* p, int kind, double tol, int64_t maxrank, uint64_t seed, int nthreads)
{
    auto H = new orc_hmat();
    H->p = p;
    H->kind = kind;
    int64_t nb = (int64_t)p->blocks.size();
    H->blk.resize(nb);
    int64_t ev = 0;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : ev) num_threads(nthreads)
    for (int64_t b = 0; b < nb; b++) {
        const Block& B = p->blocks[b];
        auto& K = H->blk[b];
        if (!B.adm) {
            /* BlockCluster::fullRankAssembly (hmatrix.cpp:350-374) */
            K.rank = -1;
            K.F.resize(B.rows * B.cols);
            for (int64_t j = 0; j < B.cols; j++)
                for (int64_t i = 0; i < B.rows; i++) K.F[j * B.rows + i] = entry(*p, kind, B.row0 + i, B.col0 + j);
            ev += B.rows * B.cols;
        } else {
            int64_t mr = std::min(B.rows, B.cols);
            if (maxrank > 0) mr = std::min(mr, maxrank);
            ACAState st;
            st.m = B.rows; st.n = B.cols; st.cap = mr;
            st.U.assign(mr * B.rows, c64(0, 0)); st.alpha.assign(mr, c64(0, 0)); st.V.assign(mr * B.cols, c64(0, 0));
            const EntryFn E = [&](int64_t i, int64_t j) { return entry(*p, kind, i, j); };
            ev += kind == 0 ? aca_plain(E, B, b, tol, mr, seed, st) : aca_extra(E, pivot_list(*p, B, b, seed), B, tol, mr, st);
            if (st.rank == 0) { /* zero rank-1 block (hmatrix.cpp:587-592) */
                K.rank = 1;
                K.U.assign(B.rows, c64(0, 0)); K.alpha.assign(1, c64(0, 0)); K.V.assign(B.cols, c64(0, 0));
            } else {
                K.rank = st.rank;
                K.U.assign(st.U.begin(), st.U.begin() + st.rank * B.rows);
                K.alpha.assign(st.alpha.begin(), st.alpha.begin() + st.rank);
                K.V.assign(st.V.begin(), st.V.begin() + st.rank * B.cols);
            }
        }
    }
    H->evaluated = ev;
    return H;
}
void orc_hmat_free(orc_hmat* h) { delete h; }
int64_t orc_hmat_rank(const orc_hmat* h, int64_t b) { return h->blk[b].rank; }
int64_t orc_hmat_evaluated(const orc_hmat* h) { return h->evaluated; }
void orc_hmat_block(const orc_hmat* h, int64_t b, double* U, double* alpha, double* V)
{
    const auto& K = h->blk[b];
    if (K.rank < 0) { std::memcpy(U, K.F.data(), sizeof(cplx) * K.F.size()); return; }
    if (U) std::memcpy(U, K.U.data(), sizeof(cplx) * K.U.size());
    if (alpha) std::memcpy(alpha, K.alpha.data(), sizeof(cplx) * K.alpha.size());
    if (V) std::memcpy(V, K.V.data(), sizeof(cplx) * K.V.size());
}
void orc_hmvm(const orc_hmat* h, const double* x, double* y) { hmvm(*h, (const cplx*)x, (cplx*)y); }

int64_t orc_normal_groups(const orc_problem* p, int64_t start, int64_t count, int32_t* gid)
{
    auto G = normal_groups(*p, start, count);
    for (size_t g = 0; g < G.size(); g++)
        for (int64_t e : G[g]) gid[e - start] = (int32_t)g;
    return (int64_t)G.size();
}

int64_t orc_pivot_list(const orc_problem* p, int64_t b, uint64_t seed, int64_t* out, int64_t cap)
{
    auto piv = pivot_list(*p, p->blocks[b], b, seed);
    int64_t c = 0;
    for (auto& q : piv) {
        if (c < cap) { out[2 * c] = q.first; out[2 * c + 1] = q.second; }
        c++;
    }
    return c;
}

int64_t orc_gmres(const orc_hmat* h, const double* b, double* x, double tol, int64_t restart, int64_t maxit, double* est)
{
    return gmres(*h, (const cplx*)b, (cplx*)x, tol, restart, maxit, est);
}

/* reconstruct phi/dphi from the unknown (boundaryelementsolver.cpp:729-740)
   and phiSolutionToSoundPressure (:2251-2261) */
static void reconstruct(const orc_problem* p, const cplx* x, double rho, double freq, double* phi, double* dphi, double* pr)
{
    double s = freq != 0 ? (rho * (2.0 * BEM_PI)) * freq : rho * (2.0 * BEM_PI);
    for (int64_t i = 0; i < p->n; i++) {
        cplx ph = x[i], dp = x[i];
        if (p->subst[i]) dp = c64_div(c64_sub(p->g[i], c64_mul(p->a[i], ph)), p->b[i]);
        else ph = c64_div(c64_sub(p->g[i], c64_mul(p->b[i], dp)), p->a[i]);
        cplx pp = c64_mul(c64(0.0, s), ph);
        phi[2 * i] = ph.re; phi[2 * i + 1] = ph.im;
        dphi[2 * i] = dp.re; dphi[2 * i + 1] = dp.im;
        pr[2 * i] = pp.re; pr[2 * i + 1] = pp.im;
    }
}

int orc_solve(orc_problem* p, double acatol, int64_t maxrank, double gtol, int64_t restart, uint64_t seed, double rho,
              double freq, int nthreads, double* phi, double* dphi, double* pressure, double* stats)
{
    int64_t n = p->n;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    double t0 = now();
    orc_hmat* Hphi = orc_assemble(p, 0, acatol, maxrank, seed, nthreads);
    double t1 = now();
    std::vector<cplx> rhs(p->s);
    hmvm(*Hphi, p->g.data(), rhs.data());
    double t2 = now();
    int64_t mrp = 0;
    for (auto& K : Hphi->blk) mrp = std::max(mrp, K.rank);
    int64_t evp = Hphi->evaluated;
    delete Hphi;
    double t3 = now();
    orc_hmat* Hd = orc_assemble(p, 1, acatol, maxrank, seed, nthreads);
    double t4 = now();
    std::vector<cplx> x(n, c64(0, 0));
    double est = 0;
    int64_t its = gmres(*Hd, rhs.data(), x.data(), gtol, restart, 100000, &est);
    double t5 = now();
    std::vector<cplx> r(n, c64(0, 0));
    hmvm(*Hd, x.data(), r.data());
    for (int64_t i = 0; i < n; i++) r[i] = c64_sub(r[i], rhs[i]);
    double res = vnorm(r.data(), n) / vnorm(rhs.data(), n);
    reconstruct(p, x.data(), rho, freq, phi, dphi, pressure);
    double t6 = now();
    int64_t mrd = 0, nnear = 0, nfar = 0;
    for (auto& K : Hd->blk) mrd = std::max(mrd, K.rank);
    for (auto& B : p->blocks) (B.adm ? nfar : nnear)++;
    if (stats) {
        stats[0] = t1 - t0; stats[1] = t2 - t1; stats[2] = t4 - t3; stats[3] = t5 - t4; stats[4] = t6 - t0;
        stats[5] = (double)its; stats[6] = res; stats[7] = (double)mrp; stats[8] = (double)mrd;
        stats[9] = (double)evp; stats[10] = (double)Hd->evaluated; stats[11] = (double)nnear; stats[12] = (double)nfar;
        stats[13] = est;
    }
    delete Hd;
    return 0;
}

/* regularSolveWithLU (boundaryelementsolver.cpp:140-331): dense Phi, DPhi,
   b = s + Phi f, partial-pivoting LU solve */
int orc_dense_solve(orc_problem* p, double rho, double freq, double* phi, double* dphi, double* pressure)
{
    int64_t n = p->n;
    std::vector<cplx> A(n * n), Ph(n * n);
    orc_dense_matrix(p, 1, (double*)A.data());
    orc_dense_matrix(p, 0, (double*)Ph.data());
    std::vector<cplx> bvec(p->s);
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = 0; i < n; i++) bvec[i] = c64_fma(Ph[j * n + i], p->g[j], bvec[i]);
    Ph.clear(); Ph.shrink_to_fit();
    /* right-looking LU with partial pivoting, column-major */
    std::vector<int64_t> piv(n);
    for (int64_t k = 0; k < n; k++) {
        int64_t pr = k;
        double best = c64_abs2(A[k * n + k]);
        for (int64_t i = k + 1; i < n; i++) {
            double v = c64_abs2(A[k * n + i]);
            if (v > best) { best = v; pr = i; }
        }
        piv[k] = pr;
        if (pr != k)
            for (int64_t j = 0; j < n; j++) std::swap(A[j * n + k], A[j * n + pr]);
        cplx inv = c64_div(c64(1, 0), A[k * n + k]);
        for (int64_t i = k + 1; i < n; i++) A[k * n + i] = c64_mul(A[k * n + i], inv);
#pragma omp parallel for schedule(static)
        for (int64_t j = k + 1; j < n; j++) {
            cplx akj = A[j * n + k];
            for (int64_t i = k + 1; i < n; i++) A[j * n + i] = c64_fms(A[k * n + i], akj, A[j * n + i]);
        }
    }
    for (int64_t k = 0; k < n; k++) if (piv[k] != k) std::swap(bvec[k], bvec[piv[k]]);
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = j + 1; i < n; i++) bvec[i] = c64_fms(A[j * n + i], bvec[j], bvec[i]);
    for (int64_t j = n - 1; j >= 0; j--) {
        bvec[j] = c64_div(bvec[j], A[j * n + j]);
        for (int64_t i = 0; i < j; i++) bvec[i] = c64_fms(A[j * n + i], bvec[j], bvec[i]);
    }
    reconstruct(p, bvec.data(), rho, freq, phi, dphi, pressure);
    return 0;
}

void orc_tables(double* tri, double* gl5, double* half)
{
    for (int q = 0; q < 16; q++) for (int c = 0; c < 3; c++) tri[3 * q + c] = TRI16[q][c];
    for (int q = 0; q < 5; q++) { gl5[2 * q] = GL5X[q]; gl5[2 * q + 1] = GL5W[q]; }
    for (int q = 0; q < 16; q++) { half[2 * q] = HGX[q]; half[2 * q + 1] = HGW[q]; }
}
void orc_sincos(int64_t n, const double* x, double* s, double* c)
{
    for (int64_t i = 0; i < n; i++) det_sincos(x[i], &s[i], &c[i]);
}
int orc_project(const double* t9, const double* p3, double* out3)
{
    Tri t;
    t.a = v3(t9[0], t9[1], t9[2]); t.b = v3(t9[3], t9[4], t9[5]); t.c = v3(t9[6], t9[7], t9[8]);
    V3 o;
    int r = (int)project_point(t, v3(p3[0], p3[1], p3[2]), &o);
    out3[0] = o.x; out3[1] = o.y; out3[2] = o.z;
    return r;
}

} /* extern "C" */
