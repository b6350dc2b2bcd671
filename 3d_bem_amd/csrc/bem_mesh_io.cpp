// GMSH .msh reader for the boundary mesh (mshReader::readMsh,
// mshreader.cpp:10-428): triangles (element type 2) and quadrilaterals
// (type 3, split into two triangles along the shorter diagonal as
// global::quadrilateralsToTriangles, global.cpp) filtered by the reference's
// include / exclude lists; and the edge-length refinement of the triangles
// (ElementSection::refineTriangles / recursiveRefineTriangle,
// SolvingScript/elementSection.cpp:688-800, called by
// SolvingScriptReader::refineElements, solvingscriptreader.cpp:1104-1128).  The reference parses the ASCII format 2; format
// 4.1 (entity blocks) is read as well, with an element's geometrical group =
// its entity tag and its physical group = the entity's first physical tag.
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/bem_hip.h"

namespace {

struct Range { uint64_t lo, hi; };

// ElementSection::lineLength (elementSection.cpp:1001-1005)
double line_length(const double* p, const double* q)
{
    const double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
    return std::sqrt((dx * dx + dy * dy) + dz * dz);
}

std::string trim(const std::string& s)
{
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

std::string lower(std::string s)
{
    for (char& ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

std::vector<std::string> split_list(const char* csv)
{
    std::vector<std::string> out;
    if (!csv) return out;
    std::stringstream ss(csv);
    std::string item;
    while (std::getline(ss, item, ',')) {
        item = trim(item);
        if (!item.empty()) out.push_back(item);
    }
    return out;
}

// "a...b" / "a..b" (either side may be empty) or a single number: a range of
// geometrical groups; anything else: a physical group name
// (mshReader::sortOutIncludesAndExcludes)
bool parse_range(const std::string& s, Range* r)
{
    size_t dots = s.find("..");
    auto is_num = [](const std::string& t) {
        return !t.empty() && std::all_of(t.begin(), t.end(), [](char ch) { return std::isdigit((unsigned char)ch); });
    };
    if (dots != std::string::npos) {
        std::string a = trim(s.substr(0, dots));
        size_t e = dots + 2;
        if (e < s.size() && s[e] == '.') e++;
        std::string b = trim(s.substr(e));
        if ((!a.empty() && !is_num(a)) || (!b.empty() && !is_num(b))) return false;
        if (a.empty() && b.empty()) return false;
        uint64_t lo = a.empty() ? 1 : std::stoull(a);
        uint64_t hi = b.empty() ? std::numeric_limits<uint64_t>::max() : std::stoull(b);
        *r = Range{std::min(lo, hi), std::max(lo, hi)};
        return true;
    }
    if (is_num(s)) {
        uint64_t v = std::stoull(s);
        *r = Range{v, v};
        return true;
    }
    return false;
}

struct Filter {
    bool includeAll = true;
    std::vector<uint64_t> phys;   // include or exclude list of physical groups
    std::vector<Range> geo;       // include or exclude list of geometrical groups
    bool included(uint64_t p, uint64_t g) const
    {
        bool inPhys = std::find(phys.begin(), phys.end(), p) != phys.end();
        bool inGeo = false;
        for (const Range& r : geo) inGeo = inGeo || (g >= r.lo && g <= r.hi);
        return includeAll ? !(inPhys || inGeo) : (inPhys || inGeo);
    }
};

Filter make_filter(const std::vector<std::string>& inc, const std::vector<std::string>& exc,
                   const std::map<std::string, uint64_t>& names)
{
    Filter f;
    std::vector<std::string> include;
    for (const auto& s : inc)
        if (lower(s) != "all") include.push_back(s);
    f.includeAll = include.empty();
    const auto& use = f.includeAll ? exc : include;
    for (const auto& s : use) {
        Range r;
        if (parse_range(s, &r)) {
            f.geo.push_back(r);
            continue;
        }
        auto it = names.find(lower(s));
        if (it != names.end()) f.phys.push_back(it->second);
    }
    return f;
}

struct Elem { int type; uint64_t phys, geo; uint64_t node[4]; };

int read_file(const char* path, const char* inc, const char* exc, std::vector<double>& tri, std::string& err)
{
    std::ifstream in(path);
    if (!in) { err = std::string("cannot open ") + path; return BEM_ERR_ARG; }
    std::string line;
    double version = 0;
    std::unordered_map<uint64_t, std::array<double, 3>> nodes;
    std::map<std::string, uint64_t> names;
    std::map<uint64_t, uint64_t> surfPhys;   // 4.1: surface entity tag -> first physical tag
    std::vector<Elem> elems;
    while (std::getline(in, line)) {
        line = trim(line);
        if (line == "$MeshFormat") {
            std::getline(in, line);
            version = std::stod(line);
        } else if (line == "$PhysicalNames") {
            size_t cnt;
            in >> cnt;
            std::getline(in, line);
            for (size_t i = 0; i < cnt; i++) {
                std::getline(in, line);
                std::stringstream ss(line);
                uint64_t dim, num;
                ss >> dim >> num;
                std::string rest;
                std::getline(ss, rest);
                rest = trim(rest);
                if (rest.size() >= 2 && rest.front() == '"' && rest.back() == '"') rest = rest.substr(1, rest.size() - 2);
                names[lower(rest)] = num;
            }
        } else if (line == "$Entities" && version >= 4) {
            size_t np, nc, ns, nv;
            in >> np >> nc >> ns >> nv;
            std::getline(in, line);
            for (size_t i = 0; i < np + nc + ns + nv; i++) {
                std::getline(in, line);
                if (i < np + nc || i >= np + nc + ns) continue;   // surfaces only
                std::stringstream ss(line);
                uint64_t tag;
                double bb[6];
                size_t nphys;
                ss >> tag;
                for (double& b : bb) ss >> b;
                ss >> nphys;
                long long p0 = 0;
                if (nphys > 0) ss >> p0;
                surfPhys[tag] = (uint64_t)(p0 < 0 ? -p0 : p0);
            }
        } else if (line == "$Nodes") {
            if (version >= 4) {
                size_t nblocks, nnodes, mn, mx;
                in >> nblocks >> nnodes >> mn >> mx;
                for (size_t b = 0; b < nblocks; b++) {
                    int dim, tag, param;
                    size_t cnt;
                    in >> dim >> tag >> param >> cnt;
                    std::vector<uint64_t> ids(cnt);
                    for (auto& id : ids) in >> id;
                    for (size_t q = 0; q < cnt; q++) {
                        std::array<double, 3> x;
                        in >> x[0] >> x[1] >> x[2];
                        if (param) {   // parametric coordinates follow (dim of the entity)
                            double skip;
                            for (int d = 0; d < dim; d++) in >> skip;
                        }
                        nodes[ids[q]] = x;
                    }
                }
            } else {
                size_t cnt;
                in >> cnt;
                for (size_t q = 0; q < cnt; q++) {
                    uint64_t id;
                    std::array<double, 3> x;
                    in >> id >> x[0] >> x[1] >> x[2];
                    nodes[id] = x;
                }
            }
            if (!in) { err = "malformed $Nodes section"; return BEM_ERR_ARG; }
        } else if (line == "$Elements") {
            if (version >= 4) {
                size_t nblocks, nel, mn, mx;
                in >> nblocks >> nel >> mn >> mx;
                for (size_t b = 0; b < nblocks; b++) {
                    int dim, type;
                    uint64_t tag;
                    size_t cnt;
                    in >> dim >> tag >> type >> cnt;
                    int nn = type == 2 ? 3 : type == 3 ? 4 : type == 15 ? 1 : type == 1 ? 2 : type == 4 ? 4 : -1;
                    std::getline(in, line);
                    for (size_t q = 0; q < cnt; q++) {
                        std::getline(in, line);
                        if (nn < 0 || (type != 2 && type != 3)) continue;
                        std::stringstream ss(line);
                        uint64_t id;
                        Elem e{type, surfPhys.count(tag) ? surfPhys[tag] : 0, tag, {0, 0, 0, 0}};
                        ss >> id;
                        for (int k = 0; k < nn; k++) ss >> e.node[k];
                        elems.push_back(e);
                    }
                }
            } else {
                size_t cnt;
                in >> cnt;
                std::getline(in, line);
                for (size_t q = 0; q < cnt; q++) {
                    std::getline(in, line);
                    std::stringstream ss(line);
                    uint64_t id, type, ntags;
                    ss >> id >> type >> ntags;
                    std::vector<uint64_t> tags(ntags);
                    for (auto& t : tags) ss >> t;
                    if (type != 2 && type != 3) continue;
                    Elem e{(int)type, ntags >= 1 ? tags[0] : 0, ntags >= 2 ? tags[1] : 0, {0, 0, 0, 0}};
                    for (int k = 0; k < (type == 2 ? 3 : 4); k++) ss >> e.node[k];
                    elems.push_back(e);
                }
            }
            if (!in) { err = "malformed $Elements section"; return BEM_ERR_ARG; }
        }
    }
    if (version == 0) { err = "no $MeshFormat section"; return BEM_ERR_ARG; }
    Filter f = make_filter(split_list(inc), split_list(exc), names);
    tri.clear();
    auto put = [&](uint64_t a, uint64_t b, uint64_t c) -> bool {
        for (uint64_t id : {a, b, c}) {
            auto it = nodes.find(id);
            if (it == nodes.end()) return false;
            tri.insert(tri.end(), it->second.begin(), it->second.end());
        }
        return true;
    };
    auto dist = [&](uint64_t a, uint64_t b) {
        const auto &p = nodes[a], &q = nodes[b];
        return line_length(p.data(), q.data());
    };
    for (const Elem& e : elems) {
        if (!f.included(e.phys, e.geo)) continue;
        bool ok;
        if (e.type == 2) {
            ok = put(e.node[0], e.node[1], e.node[2]);
        } else {   // the shorter diagonal: (1, 2, 4) + (2, 3, 4) or (4, 1, 3) + (1, 2, 3)
            ok = true;
            for (int k = 0; k < 4; k++) ok = ok && nodes.count(e.node[k]);
            if (ok && dist(e.node[0], e.node[2]) > dist(e.node[1], e.node[3]))
                ok = put(e.node[0], e.node[1], e.node[3]) && put(e.node[1], e.node[2], e.node[3]);
            else if (ok)
                ok = put(e.node[3], e.node[0], e.node[2]) && put(e.node[0], e.node[1], e.node[2]);
        }
        if (!ok) { err = "element with an undefined node"; return BEM_ERR_ARG; }
    }
    return BEM_OK;
}

// indexOfMaxOfThree (global.cpp:593-617): 1, 2 or 3, the later index on ties
int index_of_max3(double a, double b, double c)
{
    if (a > b) return a > c ? 1 : 3;
    return b > c ? 2 : 3;
}

// recursiveRefineTriangle: bisect the longest edge (edge 1 = node1-node2,
// 2 = node2-node3, 3 = node3-node1) at its midpoint until every longest edge
// is at most L; the children keep the parent's orientation and come out
// depth first, child 1 before child 2
void refine_triangle(const double* t, int longest, double L, std::vector<double>& out)
{
    const double* n1 = t;
    const double* n2 = t + 3;
    const double* n3 = t + 6;
    const double* e0 = longest == 1 ? n1 : (longest == 2 ? n2 : n3);
    const double* e1 = longest == 1 ? n2 : (longest == 2 ? n3 : n1);
    if (line_length(e0, e1) <= L) {
        out.insert(out.end(), t, t + 9);
        return;
    }
    double mid[3];
    for (int d = 0; d < 3; d++) mid[d] = (e0[d] + e1[d]) / 2;
    double c1[9], c2[9];
    auto set = [](double* dst, const double* a, const double* b, const double* c) {
        for (int d = 0; d < 3; d++) {
            dst[d] = a[d];
            dst[3 + d] = b[d];
            dst[6 + d] = c[d];
        }
    };
    switch (longest) {
    case 1: set(c1, n1, mid, n3); set(c2, mid, n2, n3); break;
    case 2: set(c1, n1, n2, mid); set(c2, mid, n3, n1); break;
    default: set(c1, n1, n2, mid); set(c2, mid, n2, n3); break;
    }
    for (const double* c : {(const double*)c1, (const double*)c2})
        refine_triangle(c, index_of_max3(line_length(c, c + 3), line_length(c + 3, c + 6), line_length(c + 6, c)), L,
                        out);
}

thread_local std::string g_msh_err;

}  // namespace

extern "C" {

int bem_read_msh(const char* path, const char* include_csv, const char* exclude_csv, int64_t cap, double* nodes9,
                 int64_t* ntri)
{
    if (!path || !ntri || cap < 0) return BEM_ERR_ARG;
    std::vector<double> tri;
    int rc;
    try {
        rc = read_file(path, include_csv, exclude_csv, tri, g_msh_err);
    } catch (const std::exception& e) {
        g_msh_err = e.what();
        rc = BEM_ERR_ARG;
    }
    if (rc) return rc;
    *ntri = (int64_t)(tri.size() / 9);
    if (!nodes9) return BEM_OK;   // count only
    if (*ntri > cap) return BEM_ERR_ARG;
    if (!tri.empty()) std::memcpy(nodes9, tri.data(), tri.size() * sizeof(double));
    return BEM_OK;
}

const char* bem_read_msh_error(void) { return g_msh_err.c_str(); }

int bem_refine_triangles(int64_t n, const double* tri9, double edge_length, int64_t cap, double* out9, int64_t* nout)
{
    if (n < 0 || (n && !tri9) || !nout || !(edge_length > 0) || cap < 0) return BEM_ERR_ARG;
    std::vector<double> out;
    for (int64_t e = 0; e < n; e++) {
        const double* t = tri9 + 9 * e;
        const int mx = index_of_max3(line_length(t, t + 3), line_length(t + 3, t + 6), line_length(t + 6, t));
        refine_triangle(t, mx, edge_length, out);
    }
    *nout = (int64_t)(out.size() / 9);
    if (!out9) return BEM_OK;   // count only
    if (*nout > cap) return BEM_ERR_ARG;
    if (!out.empty()) std::memcpy(out9, out.data(), out.size() * sizeof(double));
    return BEM_OK;
}

}  // extern "C"
