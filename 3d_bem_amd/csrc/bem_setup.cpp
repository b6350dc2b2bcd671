// Host-side per-mesh set-up (see bem_setup.hpp).  Compiled with
// -ffp-contract=off: the expressions below are evaluated exactly as written,
// which the device kernels rely on to reproduce the CPU restatement bit for
// bit.
#include "bem_setup.hpp"
#include "../../include/bem_detmath.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <unordered_map>

namespace bemsetup {
namespace {

inline Vec3 v3(double x, double y, double z) { return Vec3{x, y, z}; }
inline Vec3 operator-(Vec3 a, Vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec3 operator+(Vec3 a, Vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline Vec3 operator*(double s, Vec3 a) { return v3(s * a.x, s * a.y, s * a.z); }
inline double dot(Vec3 a, Vec3 b) { return fma(a.z, b.z, fma(a.y, b.y, a.x * b.x)); }
inline Vec3 cross(Vec3 a, Vec3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline double norm(Vec3 a) { return std::sqrt(dot(a, a)); }
inline Vec3 unit(Vec3 a) { double n = norm(a); return v3(a.x / n, a.y / n, a.z / n); }
inline double comp(Vec3 v, int d) { return d == 0 ? v.x : (d == 1 ? v.y : v.z); }

struct Tri { Vec3 a, b, c; };
inline Tri tri_of(const double* v) { return Tri{v3(v[0], v[1], v[2]), v3(v[3], v[4], v[5]), v3(v[6], v[7], v[8])}; }
inline Vec3 centroid(const Tri& t) { Vec3 s = (t.a + t.b) + t.c; return v3(s.x / 3.0, s.y / 3.0, s.z / 3.0); }

const double GL5X[5] = {-0.906179845938664, -0.5384693101056831, 0.0, 0.5384693101056831, 0.906179845938664};
const double HGX[16] = {0.04830766568773832, 0.1444719615827965, 0.23928736225213706, 0.33186860228212767,
                        0.42135127613063533, 0.5068999089322294, 0.5877157572407623, 0.6630442669302152,
                        0.7321821187402897, 0.7944837959679424, 0.84936761373257, 0.8963211557660521,
                        0.9349060759377397, 0.9647622555875064, 0.9856115115452684, 0.9972638618494816};

// Region of the nearest point of a triangle (Eberly's point-triangle distance
// algorithm) snapped to an edge or vertex when a barycentric coordinate is
// below 0.05: global::projectPointOnTriangleFaster (global.cpp:955-1283).
enum Region { kInterior, kNode1, kNode2, kNode3, kSide21, kSide32, kSide13 };

Region nearest_point(const Tri& T, Vec3 p, Vec3* proj)
{
    Vec3 d = T.a - p, e0 = T.b - T.a, e1 = T.c - T.a;
    double a00 = dot(e0, e0), a01 = dot(e0, e1), a11 = dot(e1, e1), b0 = dot(d, e0), b1 = dot(d, e1);
    double det = std::max(a00 * a11 - a01 * a01, 0.0);
    double s = a01 * b1 - a11 * b0, t = a01 * b0 - a00 * b1;
    auto edge_t = [&](void) {  // s = 0 edge, parameter t
        if (b1 >= 0.0) t = 0.0;
        else if (-b1 >= a11) t = 1.0;
        else t = -b1 / a11;
    };
    auto edge_s = [&](void) {  // t = 0 edge, parameter s
        if (b0 >= 0.0) s = 0.0;
        else if (-b0 >= a00) s = 1.0;
        else s = -b0 / a00;
    };
    if (s + t <= det) {
        if (s < 0.0) {
            if (t < 0.0) {
                if (b0 < 0.0) { t = 0.0; s = (-b0 >= a00) ? 1.0 : -b0 / a00; }
                else { s = 0.0; edge_t(); }
            } else { s = 0.0; edge_t(); }
        } else if (t < 0.0) { t = 0.0; edge_s(); }
        else { s /= det; t /= det; }
    } else {
        double denom = a00 - 2.0 * a01 + a11;
        if (s < 0.0) {
            double q0 = a01 + b0, q1 = a11 + b1;
            if (q1 > q0) {
                double num = q1 - q0;
                if (num >= denom) { s = 1.0; t = 0.0; } else { s = num / denom; t = 1.0 - s; }
            } else {
                s = 0.0;
                if (q1 <= 0.0) t = 1.0;
                else if (b1 >= 0.0) t = 0.0;
                else t = -b1 / a11;
            }
        } else if (t < 0.0) {
            double q0 = a01 + b1, q1 = a00 + b0;
            if (q1 > q0) {
                double num = q1 - q0;
                if (num >= denom) { t = 1.0; s = 0.0; } else { t = num / denom; s = 1.0 - t; }
            } else {
                t = 0.0;
                if (q1 <= 0.0) s = 1.0;
                else if (b0 >= 0.0) s = 0.0;
                else s = -b0 / a00;
            }
        } else {
            double num = a11 + b1 - a01 - b0;
            if (num <= 0.0) { s = 0.0; t = 1.0; }
            else if (num >= denom) { s = 1.0; t = 0.0; }
            else { s = num / denom; t = 1.0 - s; }
        }
    }
    const double eps = 0.05;  // global.h:43
    double w1 = 1.0 - s - t, w2 = s, w3 = t;
    int big = (w1 > eps) + (w2 > eps) + (w3 > eps);
    Region r;
    if (big == 3) r = kInterior;
    else if (big == 2) {
        if (w1 > eps && w2 > eps) { r = kSide21; w1 += w3 / 2.0; w2 += w3 / 2.0; w3 = 0.0; }
        else if (w1 > eps && w3 > eps) { r = kSide13; w1 += w2 / 2.0; w3 += w2 / 2.0; w2 = 0.0; }
        else { r = kSide32; w2 += w1 / 2.0; w3 += w1 / 2.0; w1 = 0.0; }
    } else {
        if (w1 > eps) { r = kNode1; w1 = 1.0; w2 = 0.0; w3 = 0.0; }
        else if (w2 > eps) { r = kNode2; w1 = 0.0; w2 = 1.0; w3 = 0.0; }
        else { r = kNode3; w1 = 0.0; w2 = 0.0; w3 = 1.0; }
    }
    *proj = (w1 * T.a + w2 * T.b) + w3 * T.c;
    return r;
}

// geometry of one polar sub-triangle (origin n1) into 25 doubles
void near_subtri(const Tri& c, double* out)
{
    Vec3 e13 = c.c - c.a, e23 = c.c - c.b, e12 = c.b - c.a;
    double q13 = dot(e13, e13), q23 = dot(e23, e23), q12 = dot(e12, e12);
    double l13 = std::sqrt(q13), l23 = std::sqrt(q23), l12 = std::sqrt(q12);
    double angP = std::acos((q13 + q12 - q23) / (2.0 * l13 * l12));
    double ang2 = std::acos((q23 + q12 - q13) / (2.0 * l23 * l12));
    Vec3 ax0 = unit(e12);
    Vec3 ax1 = unit(e13 - dot(ax0, e13) * ax0);
    out[0] = c.a.x; out[1] = c.a.y; out[2] = c.a.z;
    out[3] = ax0.x; out[4] = ax0.y; out[5] = ax0.z;
    out[6] = ax1.x; out[7] = ax1.y; out[8] = ax1.z;
    out[9] = angP / 2.0;
    for (int ia = 0; ia < 5; ia++) {
        double ang = (angP + GL5X[ia] * angP) / 2.0;
        out[10 + 3 * ia] = l12 * std::sin(ang2) / (std::sin(ang + ang2));
        out[11 + 3 * ia] = std::cos(ang);
        out[12 + 3 * ia] = std::sin(ang);
    }
}

}  // namespace

void prepare_elements(Elements& E)
{
    E.normal.resize(E.n);
    E.colloc.resize(E.n);
    E.area.resize(E.n);
    double sum = 0;
    for (int64_t e = 0; e < E.n; e++) {
        Tri t = tri_of(&E.nodes[9 * e]);
        E.normal[e] = unit(cross(t.b - t.a, t.c - t.a));
        E.colloc[e] = centroid(t);
        E.area[e] = norm(cross(t.a - t.b, t.a - t.c)) / 2.0;
    }
    for (int64_t e = 0; e < E.n; e++) sum += std::sqrt(E.area[e]);
    E.hbar = E.n > 0 ? (sum / (double)E.n) / 2.0 : 0.0;
}

void reflect_elements(int64_t n, const double* nodes9, const double point[3], const double normal[3],
                      std::vector<double>& rnodes9, std::vector<Vec3>& rnormal)
{
    const Vec3 u = unit(v3(normal[0], normal[1], normal[2]));
    const double d = -dot(u, v3(point[0], point[1], point[2]));
    const double R[3][4] = {{1.0 - 2.0 * (u.x * u.x), -2.0 * u.x * u.y, -2.0 * u.x * u.z, -2.0 * u.x * d},
                            {-2.0 * u.x * u.y, 1.0 - 2.0 * (u.y * u.y), -2.0 * u.y * u.z, -2.0 * u.y * d},
                            {-2.0 * u.x * u.z, -2.0 * u.y * u.z, 1.0 - 2.0 * (u.z * u.z), -2.0 * u.z * d}};
    auto mirror = [&](const double* x, double* out) {
        for (int r = 0; r < 3; r++) out[r] = ((R[r][0] * x[0] + R[r][1] * x[1]) + R[r][2] * x[2]) + R[r][3];
    };
    rnodes9.resize(9 * n);
    rnormal.resize(n);
    for (int64_t e = 0; e < n; e++) {
        const double* v = nodes9 + 9 * e;
        double* o = &rnodes9[9 * e];
        mirror(v + 6, o);
        mirror(v + 3, o + 3);
        mirror(v, o + 6);
        Tri t = tri_of(o);
        rnormal[e] = unit(cross(t.b - t.a, t.c - t.a));
    }
}

ClusterTree build_cluster_tree(int64_t n, const double* nodes9, int64_t leaf)
{
    std::vector<double> mid(3 * n);
    for (int64_t e = 0; e < n; e++) {
        const Vec3 c = centroid(tri_of(nodes9 + 9 * e));
        mid[3 * e] = c.x;
        mid[3 * e + 1] = c.y;
        mid[3 * e + 2] = c.z;
    }
    return build_cluster_tree_points(n, mid.data(), leaf);
}

ClusterTree build_cluster_tree_points(int64_t n, const double* pts3, int64_t leaf)
{
    ClusterTree T;
    std::vector<Vec3> mid(n);
    for (int64_t e = 0; e < n; e++) mid[e] = v3(pts3[3 * e], pts3[3 * e + 1], pts3[3 * e + 2]);
    std::vector<std::vector<int64_t>> members;
    const double inf = (double)std::numeric_limits<float>::infinity();  // Cuboid() (cuboid.h:17-20)
    // explicit stack instead of recursion; children are created son1 first so
    // that the depth-first numbering equals the recursive one
    struct Item { int parent; int which; std::vector<int64_t> idx; int depth; };
    std::vector<Item> stack;
    std::vector<int64_t> all(n);
    for (int64_t e = 0; e < n; e++) all[e] = e;
    stack.push_back(Item{-1, 0, std::move(all), 0});
    while (!stack.empty()) {
        Item it = std::move(stack.back());
        stack.pop_back();
        Cluster C;
        C.bmin = v3(inf, inf, inf);
        C.bmax = v3(-inf, -inf, -inf);
        for (int64_t e : it.idx) {
            C.bmin = v3(std::min(C.bmin.x, mid[e].x), std::min(C.bmin.y, mid[e].y), std::min(C.bmin.z, mid[e].z));
            C.bmax = v3(std::max(C.bmax.x, mid[e].x), std::max(C.bmax.y, mid[e].y), std::max(C.bmax.z, mid[e].z));
        }
        C.size = (int64_t)it.idx.size();
        C.depth = it.depth;
        int id = (int)T.clusters.size();
        if (it.parent >= 0) (it.which == 1 ? T.clusters[it.parent].son1 : T.clusters[it.parent].son2) = id;
        C.leaf = C.size < leaf || C.size <= 1;
        T.clusters.push_back(C);
        members.push_back(it.idx);
        if (C.leaf) continue;
        Vec3 side = C.bmax - C.bmin;
        int d = 0;
        double best = std::fabs(side.x);
        if (std::fabs(side.y) > best) { d = 1; best = std::fabs(side.y); }
        if (std::fabs(side.z) > best) d = 2;
        double cut = comp(C.bmin, d) + 0.5 * comp(side, d);
        std::vector<int64_t> s1, s2;
        for (int64_t e : it.idx) (comp(mid[e], d) >= cut ? s1 : s2).push_back(e);
        stack.push_back(Item{id, 2, std::move(s2), it.depth + 1});
        stack.push_back(Item{id, 1, std::move(s1), it.depth + 1});
    }
    // depth-first leaf order = new element order; clusters become ranges
    T.perm.reserve(n);
    std::vector<int> order{0};
    std::vector<std::pair<int, bool>> st{{0, false}};
    while (!st.empty()) {
        auto [c, visited] = st.back();
        st.pop_back();
        Cluster& C = T.clusters[c];
        if (C.leaf) {
            C.start = (int64_t)T.perm.size();
            for (int64_t e : members[c]) T.perm.push_back(e);
        } else if (!visited) {
            st.push_back({c, true});
            st.push_back({C.son2, false});
            st.push_back({C.son1, false});
        } else {
            C.start = T.clusters[C.son1].start;
        }
    }
    return T;
}

namespace {
// HMatrix::clusterDistance (hmatrix.cpp:1070-1129)
double box_distance(const Cluster& a, const Cluster& b)
{
    bool ov[3];
    double dd[3];
    for (int d = 0; d < 3; d++) {
        double lo1 = comp(a.bmin, d), hi1 = comp(a.bmax, d), lo2 = comp(b.bmin, d), hi2 = comp(b.bmax, d);
        ov[d] = (lo1 < hi2 && hi1 > lo2) || (lo2 < hi1 && hi2 > lo1);
        dd[d] = std::min(std::fabs(lo1 - hi2), std::fabs(lo2 - hi1));
    }
    int k = ov[0] + ov[1] + ov[2];
    if (k == 3) return 0;
    if (k == 2) return !ov[0] ? dd[0] : (!ov[1] ? dd[1] : dd[2]);
    if (k == 1) {
        if (ov[0]) return std::sqrt(dd[1] * dd[1] + dd[2] * dd[2]);
        if (ov[1]) return std::sqrt(dd[0] * dd[0] + dd[2] * dd[2]);
        return std::sqrt(dd[0] * dd[0] + dd[1] * dd[1]);
    }
    return std::sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
}
}  // namespace

std::vector<Block> build_partition(const ClusterTree& T, double eta)
{
    return build_partition(T, T, eta);
}

std::vector<Block> build_partition(const ClusterTree& TR, const ClusterTree& TC, double eta)
{
    std::vector<Block> out;
    if (TR.clusters.empty() || TC.clusters.empty()) return out;
    std::vector<std::pair<int, int>> st{{0, 0}};
    while (!st.empty()) {
        auto [rc, cc] = st.back();
        st.pop_back();
        const Cluster& R = TR.clusters[rc];
        const Cluster& C = TC.clusters[cc];
        double dist = box_distance(R, C);
        bool adm = dist != 0 && std::min(norm(R.bmax - R.bmin), norm(C.bmax - C.bmin)) / dist < eta;
        if (adm || R.leaf || C.leaf) {
            out.push_back(Block{R.start, R.size, C.start, C.size, adm});
            continue;
        }
        // push in reverse so that 11, 12, 21, 22 come out in that order
        st.push_back({R.son2, C.son2});
        st.push_back({R.son2, C.son1});
        st.push_back({R.son1, C.son2});
        st.push_back({R.son1, C.son1});
    }
    return out;
}

NearList build_near(const Elements& E)
{
    NearList L;
    int64_t n = E.n;
    L.rowptr.assign(n + 1, 0);
    double lim = 4.0 * E.hbar;
    // uniform grid with cell = lim; a pair closer than lim lies in adjacent cells
    double cell = lim > 0 ? lim : 1.0;
    Vec3 lo = E.colloc.empty() ? v3(0, 0, 0) : E.colloc[0];
    for (auto& p : E.colloc) lo = v3(std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z));
    auto key = [&](int64_t ix, int64_t iy, int64_t iz) { return (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791); };
    std::vector<std::array<int64_t, 3>> cidx(n);
    std::unordered_map<int64_t, std::vector<int64_t>> grid;
    for (int64_t e = 0; e < n; e++) {
        Vec3 p = E.colloc[e];
        std::array<int64_t, 3> c = {(int64_t)std::floor((p.x - lo.x) / cell), (int64_t)std::floor((p.y - lo.y) / cell),
                                    (int64_t)std::floor((p.z - lo.z) / cell)};
        cidx[e] = c;
        grid[key(c[0], c[1], c[2])].push_back(e);
    }
    std::vector<std::vector<int32_t>> rows(n);
    for (int64_t i = 0; i < n; i++) {
        auto c = cidx[i];
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    auto it = grid.find(key(c[0] + dx, c[1] + dy, c[2] + dz));
                    if (it == grid.end()) continue;
                    for (int64_t j : it->second) {
                        if (j == i) continue;
                        auto cj = cidx[j];
                        if (cj[0] != c[0] + dx || cj[1] != c[1] + dy || cj[2] != c[2] + dz) continue;
                        if (norm(E.colloc[i] - E.colloc[j]) < lim) rows[i].push_back((int32_t)j);
                    }
                }
        std::sort(rows[i].begin(), rows[i].end());
        L.rowptr[i + 1] = L.rowptr[i] + (int64_t)rows[i].size();
    }
    int64_t nnz = L.rowptr[n];
    L.col.resize(nnz);
    L.nsub.resize(nnz);
    L.geo_off.resize(nnz);
    for (int64_t i = 0; i < n; i++) std::copy(rows[i].begin(), rows[i].end(), L.col.begin() + L.rowptr[i]);
    // geometry: region first (sizes), then fill in parallel
    std::vector<int> reg(nnz);
    std::vector<Vec3> proj(nnz);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++)
        for (int64_t t = L.rowptr[i]; t < L.rowptr[i + 1]; t++) {
            Tri T = tri_of(&E.nodes[9 * (int64_t)L.col[t]]);
            reg[t] = nearest_point(T, E.colloc[i], &proj[t]);
            L.nsub[t] = reg[t] == kInterior ? 3 : (reg[t] <= kNode3 ? 1 : 2);
        }
    int64_t off = 0;
    for (int64_t t = 0; t < nnz; t++) { L.geo_off[t] = off; off += (int64_t)L.nsub[t] * NEAR_SUB_DOUBLES; }
    L.geo.resize(off);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t t = 0; t < nnz; t++) {
        Tri T = tri_of(&E.nodes[9 * (int64_t)L.col[t]]);
        Vec3 pp = proj[t];
        Tri st[3];
        switch (reg[t]) {
        case kInterior: st[0] = Tri{pp, T.a, T.b}; st[1] = Tri{pp, T.b, T.c}; st[2] = Tri{pp, T.c, T.a}; break;
        case kNode1: st[0] = T; break;
        case kNode2: st[0] = Tri{T.b, T.c, T.a}; break;
        case kNode3: st[0] = Tri{T.c, T.a, T.b}; break;
        case kSide21: st[0] = Tri{pp, T.c, T.a}; st[1] = Tri{pp, T.b, T.c}; break;
        case kSide32: st[0] = Tri{pp, T.c, T.a}; st[1] = Tri{pp, T.a, T.b}; break;
        default: st[0] = Tri{pp, T.a, T.b}; st[1] = Tri{pp, T.b, T.c}; break;  // side 13
        }
        for (int s = 0; s < L.nsub[t]; s++) near_subtri(st[s], &L.geo[L.geo_off[t] + s * NEAR_SUB_DOUBLES]);
    }
    return L;
}

std::vector<double> build_self(const Elements& E)
{
    std::vector<double> out((size_t)E.n * SELF_DOUBLES);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E.n; e++) {
        Tri T = tri_of(&E.nodes[9 * e]);
        Vec3 cp = E.colloc[e];
        Vec3 sub[3][3] = {{T.a, T.b, cp}, {T.b, T.c, cp}, {T.c, T.a, cp}};
        double* o = &out[(size_t)e * SELF_DOUBLES];
        double Afirst = 0, Asum = 0;
        for (int s = 0; s < 3; s++) {
            Vec3 Cs = sub[s][2] - sub[s][0], Bs = sub[s][2] - sub[s][1], As = sub[s][1] - sub[s][0];
            double bq = dot(Bs, Bs), cq = dot(Cs, Cs), aq = dot(As, As);
            double b = std::sqrt(bq), c = std::sqrt(cq);
            if (c < b) std::swap(b, c);
            double A = std::acos((bq + cq - aq) / (2.0 * b * c));
            double B = std::atan((b * std::sin(A)) / (c - b * std::cos(A)));
            double lower, upper, shift;
            if (s == 0) { lower = 0.5 * (A - 0); upper = 0.5 * (A + 0); shift = 0; }
            else { shift = s == 1 ? Afirst : Asum; lower = 0.5 * A; upper = 0.5 * (A + shift + shift); }
            double csB = c * std::sin(B);
            o[33 * s] = lower;
            for (int q = 0; q < 16; q++) {
                double x1 = HGX[q] * lower + upper, x2 = -HGX[q] * lower + upper;
                if (s == 0) {
                    o[33 * s + 1 + q] = csB / (std::sin(x1 + B));
                    o[33 * s + 17 + q] = csB / (std::sin(x2 + B));
                } else {
                    o[33 * s + 1 + q] = csB / (std::sin(x1 - shift + B));
                    o[33 * s + 17 + q] = csB / (std::sin(x2 - shift + B));
                }
            }
            if (s == 0) Afirst = A;
            else if (s == 1) Asum = Afirst + A;
        }
    }
    return out;
}

namespace {
std::vector<std::vector<int64_t>> orientation_groups(const Elements& E, int64_t start, int64_t count)
{
    const int maxg = 20;
    const double cone = std::cos(0.44);
    std::vector<std::vector<int64_t>> G(1);
    for (int64_t e = 0; e < count; e++) G[0].push_back(start + e);
    int64_t rest = count;
    int g = 0;
    while (rest > 1 && g < maxg - 1) {
        std::vector<int64_t> keep{G[g][0]}, moved;
        Vec3 ref = E.normal[G[g][0]];
        for (size_t t = 1; t < G[g].size(); t++) {
            int64_t e = G[g][t];
            (std::fabs(dot(ref, E.normal[e])) < cone ? moved : keep).push_back(e);
        }
        G[g] = std::move(keep);
        rest = (int64_t)moved.size();
        if (rest >= 1) { G.push_back(std::move(moved)); g++; }
    }
    return G;
}
}  // namespace

void build_pivot_lists(const Elements& E, const std::vector<Block>& blocks, uint64_t seed,
                       std::vector<int64_t>& offsets, std::vector<int32_t>& pairs)
{
    size_t nb = blocks.size();
    std::vector<std::vector<int32_t>> per(nb);
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t b = 0; b < (int64_t)nb; b++) {
        const Block& B = blocks[b];
        if (!B.adm) continue;
        auto R = orientation_groups(E, B.row0, B.rows);
        auto C = orientation_groups(E, B.col0, B.cols);
        uint64_t draw = 0;
        for (auto& rg : R)
            for (auto& cg : C) {
                int64_t ri = bem_rand_bounded(seed, 1, (uint64_t)b, draw++, (int64_t)rg.size());
                int64_t cj = bem_rand_bounded(seed, 1, (uint64_t)b, draw++, (int64_t)cg.size());
                per[b].push_back((int32_t)(rg[ri] - B.row0));
                per[b].push_back((int32_t)(cg[cj] - B.col0));
            }
    }
    offsets.assign(nb + 1, 0);
    for (size_t b = 0; b < nb; b++) offsets[b + 1] = offsets[b] + (int64_t)per[b].size() / 2;
    pairs.resize(2 * offsets[nb]);
    for (size_t b = 0; b < nb; b++) std::copy(per[b].begin(), per[b].end(), pairs.begin() + 2 * offsets[b]);
}

void build_field_pivot_lists(const Elements& E, const std::vector<Block>& blocks, uint64_t seed,
                             std::vector<int64_t>& offsets, std::vector<int32_t>& pairs)
{
    size_t nb = blocks.size();
    std::vector<std::vector<int32_t>> per(nb);
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t b = 0; b < (int64_t)nb; b++) {
        const Block& B = blocks[b];
        if (!B.adm) continue;
        auto C = orientation_groups(E, B.col0, B.cols);
        // one random row per column group, drawn first (:2773-2777); then
        // for every (row, group) a random member of the group (:2781-2789)
        uint64_t draw = 0;
        std::vector<int64_t> rows(C.size());
        for (auto& r : rows) r = bem_rand_bounded(seed, kFieldPivotStream, (uint64_t)b, draw++, B.rows);
        for (int64_t r : rows)
            for (auto& cg : C) {
                int64_t cj = bem_rand_bounded(seed, kFieldPivotStream, (uint64_t)b, draw++, (int64_t)cg.size());
                per[b].push_back((int32_t)r);
                per[b].push_back((int32_t)(cg[cj] - B.col0));
            }
    }
    offsets.assign(nb + 1, 0);
    for (size_t b = 0; b < nb; b++) offsets[b + 1] = offsets[b] + (int64_t)per[b].size() / 2;
    pairs.resize(2 * offsets[nb]);
    for (size_t b = 0; b < nb; b++) std::copy(per[b].begin(), per[b].end(), pairs.begin() + 2 * offsets[b]);
}

}  // namespace bemsetup
