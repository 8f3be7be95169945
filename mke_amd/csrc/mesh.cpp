// Sphere mesh generator and loader (host C++).
//
// sphere_mesh_text() follows `mke_mesh_sphere --type full --coord local
// --iter k`: icosahedron (utils/mesh_builder.cpp:52-113), orthogonal
// projection (mesh_builder.cpp:115-125), k midpoint subdivisions with edge
// memo (mesh_builder.cpp:143-208), south-pole boundary
// (utils/icosahedron.cpp:407-421), Cuthill-McKee order without the final
// reversal (mesh_builder.cpp:316-402, 536-585) and the writer
// (icosahedron.cpp:268-405).  The floating-point steps keep the reference's
// order; this file must be built with -ffp-contract=off.
//
// load_mesh_text() follows Mesh::load (src/mesh.cpp:323-519) for P1 meshes.
#include <cstdint>
#include "mesh.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_map>

namespace phg {
namespace {

struct P3 {
    double x, y, z;
};
struct Tri {
    int v1, v2, v3;
};

void project(P3& p)
{
    double l = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    p.x /= l;
    p.y /= l;
    p.z /= l;
}

void build_icosahedron(std::vector<Tri>& tris, std::vector<P3>& pts)
{
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    const double p[12][3] = {{t, 1.0, 0.0},  {-t, 1.0, 0.0},  {t, -1.0, 0.0},
                             {-t, -1.0, 0.0}, {1.0, 0.0, t},   {1.0, 0.0, -t},
                             {-1.0, 0.0, t},  {-1.0, 0.0, -t}, {0.0, t, 1.0},
                             {0.0, -t, 1.0},  {0.0, t, -1.0},  {0.0, -t, -1.0}};
    const int f[20][3] = {{0, 8, 4},  {1, 10, 7}, {2, 9, 11}, {7, 3, 1},
                          {0, 5, 10}, {3, 9, 6},  {3, 11, 9}, {8, 6, 4},
                          {2, 4, 9},  {3, 7, 11}, {4, 2, 0},  {9, 4, 6},
                          {2, 11, 5}, {0, 10, 8}, {5, 0, 2},  {10, 5, 7},
                          {1, 6, 8},  {1, 8, 10}, {6, 1, 3},  {11, 7, 5}};
    for (auto& q : p) pts.push_back({q[0], q[1], q[2]});
    for (auto& q : f) tris.push_back({q[0], q[1], q[2]});
}

void divide(std::vector<Tri>& tris, std::vector<P3>& pts)
{
    std::unordered_map<unsigned long long, int> memo;
    memo.reserve(tris.size() * 2);
    std::vector<P3> np(pts);
    np.reserve(pts.size() + tris.size() * 3 / 2 + 8);
    std::vector<Tri> nt;
    nt.reserve(tris.size() * 4);
    auto key = [](int a, int b) {
        unsigned lo = (unsigned)std::min(a, b), hi = (unsigned)std::max(a, b);
        return ((unsigned long long)lo << 32) | hi;
    };
    auto mid = [&](int a, int b) {
        auto it = memo.find(key(a, b));
        if (it != memo.end()) return it->second;
        const P3& A = pts[a];
        const P3& B = pts[b];
        P3 m{(A.x + B.x) / 2, (A.y + B.y) / 2, (A.z + B.z) / 2};
        project(m);
        int idx = (int)np.size();
        np.push_back(m);
        memo.emplace(key(a, b), idx);
        return idx;
    };
    for (const Tri& t : tris) {
        int p1 = mid(t.v1, t.v2);
        int p2 = mid(t.v1, t.v3);
        int p3 = mid(t.v2, t.v3);
        nt.push_back({t.v1, p1, p2});
        nt.push_back({t.v3, p2, p3});
        nt.push_back({t.v2, p3, p1});
        nt.push_back({p1, p2, p3});
    }
    tris.swap(nt);
    pts.swap(np);
}

// Graph::rcm + Mesh::full_order: returns old point id -> new point id.
std::vector<int> cm_order(const std::vector<Tri>& tris, int n_points,
                          const std::vector<char>& bnd)
{
    std::vector<int> inner, p2io(n_points);
    int no = 0;
    for (int i = 0; i < n_points; ++i) {
        if (!bnd[i]) {
            p2io[i] = (int)inner.size();
            inner.push_back(i);
        } else {
            p2io[i] = no++;
        }
    }
    std::vector<int> deg_ptr(n_points + 1, 0);
    for (const Tri& t : tris) {
        deg_ptr[t.v1 + 1]++;
        deg_ptr[t.v2 + 1]++;
        deg_ptr[t.v3 + 1]++;
    }
    for (int i = 0; i < n_points; ++i) deg_ptr[i + 1] += deg_ptr[i];
    std::vector<int> adj(deg_ptr[n_points]), fill(deg_ptr.begin(), deg_ptr.end() - 1);
    for (int tid = 0; tid < (int)tris.size(); ++tid) {
        adj[fill[tris[tid].v1]++] = tid;
        adj[fill[tris[tid].v2]++] = tid;
        adj[fill[tris[tid].v3]++] = tid;
    }
    int rs = (int)inner.size();
    std::vector<std::vector<int>> data(rs);
    for (int i = 0; i < rs; ++i) {
        int p = inner[i];
        auto& d = data[i];
        for (int e = deg_ptr[p]; e < deg_ptr[p + 1]; ++e) {
            const Tri& t = tris[adj[e]];
            for (int v : {t.v1, t.v2, t.v3})
                if (!bnd[v]) d.push_back(p2io[v]);
        }
        std::sort(d.begin(), d.end());
        d.erase(std::unique(d.begin(), d.end()), d.end());
    }
    std::vector<int> deg(rs);
    for (int i = 0; i < rs; ++i) deg[i] = (int)data[i].size();
    for (int i = 0; i < rs; ++i)
        std::stable_sort(data[i].begin(), data[i].end(),
                         [&](int a, int b) { return deg[a] < deg[b]; });
    std::vector<int> order(rs, -1), queue;
    queue.reserve(rs * 4);
    int k = 0, start = 0;
    for (;;) {
        queue.clear();
        queue.push_back(start);
        for (size_t head = 0; head < queue.size(); ++head) {
            int v = queue[head];
            if (order[v] != -1) continue;
            order[v] = k++;
            for (int nb : data[v])
                if (order[nb] < 0) queue.push_back(nb);
        }
        start = -1;
        for (int i = 0; i < rs; ++i)
            if (order[i] == -1) {
                start = i;
                break;
            }
        if (start < 0) break;
    }
    std::vector<int> full(n_points);
    int outer_next = rs;
    for (int i = 0; i < n_points; ++i) full[i] = bnd[i] ? outer_next++ : order[p2io[i]];
    return full;
}

double chart_u(double, double, double z) { return std::asin(z); }

double chart_v(double x, double y, double)
{
    double r = std::sqrt(x * x + y * y);
    if (r < 1e-15) return 0.0;
    double v1 = std::acos(x / r);
    if (y < 0) v1 = 2.0 * M_PI - v1;
    return v1;
}

const char* kHeader =
    "# zone1 ; zone2 ; zone 3; ... \n"
    "# u = theta  [-pi/2, pi/2]\n"
    "# v = lambda [0, 2pi]\n"
    "#predef: zone 1 stdsphere1\n"
    "#predef: zone 2 stdsphere2\n"
    "# zone 1\n# x = cos(u)*cos(v) \n# y = cos(u)*sin(v) \n# z = sin(u) \n"
    "# zone 2\n# x = -cos(u)*cos(v) \n# y = -cos(u)*sin(v) \n# z = sin(u) \n"
    "#predef: zone 3 stdsphere3\n"
    "#predef: zone 4 stdsphere4\n"
    "# zone 3\n# x = cos(u)*cos(v) \n# y = cos(u)*sin(v) \n# z = -sin(u) \n"
    "# zone 4\n# x = cos(u)*cos(v) \n# y = -cos(u)*sin(v) \n# z = sin(u) \n";

}  // namespace

std::string sphere_mesh_text(int level)
{
    std::vector<Tri> tris;
    std::vector<P3> pts;
    build_icosahedron(tris, pts);
    for (auto& p : pts) project(p);
    for (int i = 0; i < level; ++i) divide(tris, pts);
    std::vector<char> bnd(pts.size(), 0);
    for (size_t i = 0; i < pts.size(); ++i) bnd[i] = std::fabs(pts[i].z + 1.0) < 1e-14;
    std::vector<int> order = cm_order(tris, (int)pts.size(), bnd);
    std::vector<P3> np(pts.size());
    for (size_t i = 0; i < pts.size(); ++i) np[order[i]] = pts[i];
    pts.swap(np);
    for (Tri& t : tris) t = {order[t.v1], order[t.v2], order[t.v3]};

    std::string out;
    out.reserve(pts.size() * 200 + tris.size() * 32 + 4096);
    char buf[512];
    snprintf(buf, sizeof buf,
             "#cmd:mke_mesh_sphere --type full --coord local --iter %d \n#build:$Id$\n"
             "#points:%zu\n#triangles:%zu\n",
             level, pts.size(), tris.size());
    out += buf;
    out += kHeader;
    for (const P3& p : pts) {
        int n = snprintf(buf, sizeof buf, "%.16f %.16f %.16f", p.x, p.y, p.z);
        out.append(buf, n);
        const double c[4][3] = {{p.x, p.y, p.z}, {-p.x, -p.y, p.z},
                                {p.x, -p.z, p.y}, {p.x, p.z, -p.y}};
        for (auto& q : c) {
            n = snprintf(buf, sizeof buf, " ; %.16f %.16f ", chart_u(q[0], q[1], q[2]),
                         chart_v(q[0], q[1], q[2]));
            out.append(buf, n);
        }
        out += '\n';
    }
    snprintf(buf, sizeof buf, "# triangles %zu \n", tris.size());
    out += buf;
    const double half_pi = M_PI / 2;
    for (const Tri& t : tris) {
        int n = snprintf(buf, sizeof buf, "%d %d %d ", t.v1 + 1, t.v2 + 1, t.v3 + 1);
        out.append(buf, n);
        const P3 &a = pts[t.v1], &b = pts[t.v2], &c = pts[t.v3];
        double v1 = chart_v(a.x, a.y, a.z), v2 = chart_v(b.x, b.y, b.z);
        if (a.z > 0.9 || b.z > 0.9 || c.z > 0.9) {
            out += " ; 4 ";
        } else if (a.z < -0.9 || b.z < -0.9 || c.z < -0.9) {
            out += " ; 3 ";
        } else if (v1 < half_pi || v2 < half_pi || v2 < half_pi || v1 > 1.5 * M_PI ||
                   v2 > 1.5 * M_PI || v2 > 1.5 * M_PI) {
            // the reference tests v2 twice and never v3 (icosahedron.cpp:380-393)
            out += " ; 2 ";
        }
        out += '\n';
    }
    out += "# boundary \n";
    for (size_t i = 0; i < pts.size(); ++i)
        if (std::fabs(pts[i].z + 1.0) < 1e-14) {
            snprintf(buf, sizeof buf, "%zu \n", i + 1);
            out += buf;
        }
    return out;
}

namespace {

struct LineReader {
    const char* p;
    const char* end;
    int lineno = 0;
    std::string cur;
    bool next()
    {
        if (p >= end) return false;
        const char* nl = (const char*)memchr(p, '\n', end - p);
        const char* e = nl ? nl : end;
        cur.assign(p, e);
        p = nl ? nl + 1 : end;
        ++lineno;
        return true;
    }
};

// sscanf("%lf...") over a ';'-separated token: returns how many numbers parsed
int parse_doubles(const char* s, const char* e, double* out, int maxn)
{
    int n = 0;
    while (n < maxn) {
        while (s < e && (*s == ' ' || *s == '\t' || *s == '\r')) ++s;
        if (s >= e) break;
        char* stop;
        double v = strtod(s, &stop);
        if (stop == s || stop > e) break;
        out[n++] = v;
        s = stop;
    }
    return n;
}

}  // namespace

bool load_mesh_text(const char* text, size_t len, SphereMesh& m, std::string& err)
{
    m = SphereMesh();
    LineReader r{text, text + len};
    auto bad = [&](int line) {
        char b[64];
        snprintf(b, sizeof b, "bad file format, line = %d", line);
        err = b;
        return false;
    };
    if (!r.next()) return bad(1);
    while (!r.cur.empty() && r.cur[0] == '#') {  // comment block (mesh.cpp:338-363)
        if (!r.next()) return bad(r.lineno + 1);
    }
    // points (mesh.cpp:365-402)
    std::vector<std::vector<double>> charts;
    for (;;) {
        if (!r.cur.empty() && r.cur[0] == '#') break;
        const char* s = r.cur.data();
        const char* e = s + r.cur.size();
        double pr[3] = {0, 0, 0};
        bool pr_set = false;
        std::vector<double> ch;
        while (s < e) {
            const char* semi = (const char*)memchr(s, ';', e - s);
            const char* te = semi ? semi : e;
            double v[3];
            int k = parse_doubles(s, te, v, 3);
            if (k == 3) {
                pr[0] = v[0];
                pr[1] = v[1];
                pr[2] = v[2];
                pr_set = true;
            } else if (k == 2) {
                ch.push_back(v[0]);
                ch.push_back(v[1]);
            } else {
                bool blank = true;
                for (const char* q = s; q < te; ++q) blank &= (*q == ' ' || *q == '\r' || *q == '\t');
                if (!blank) return bad(r.lineno);
            }
            s = semi ? semi + 1 : e;
        }
        if (ch.empty() && !pr_set) return bad(r.lineno);
        if (!pr_set) {
            pr[0] = ch[0];
            pr[1] = ch[1];
            pr[2] = 0;
        }
        m.pr.insert(m.pr.end(), pr, pr + 3);
        charts.push_back(std::move(ch));
        if (!r.next()) return bad(r.lineno + 1);
    }
    m.n_points = (int)charts.size();
    // triangles (mesh.cpp:404-450): the '#' line that ended the points is skipped
    if (!r.next()) return bad(r.lineno + 1);
    for (;;) {
        if (!r.cur.empty() && r.cur[0] == '#') break;
        int n1, n2, n3, z = 1;
        if (r.cur.empty() || sscanf(r.cur.c_str(), "%d%d%d ; %d", &n1, &n2, &n3, &z) < 3)
            return bad(r.lineno);
        --n1;
        --n2;
        --n3;
        --z;
        if (n1 >= m.n_points || n2 >= m.n_points || n3 >= m.n_points) return bad(r.lineno);
        if (n1 < 0 || n2 < 0 || n3 < 0 || z < 0) return bad(r.lineno);
        m.tri.push_back(n1);
        m.tri.push_back(n2);
        m.tri.push_back(n3);
        m.tri_zone.push_back(z);
        if (!r.next()) break;
    }
    m.n_triangles = (int)m.tri_zone.size();
    m.is_boundary.assign(m.n_points, 0);
    // boundary (mesh.cpp:452-479)
    if (r.next()) {
        for (;;) {
            if (!r.cur.empty() && r.cur[0] == '#') break;
            int n;
            if (sscanf(r.cur.c_str(), "%d", &n) != 1) {
                bool blank = true;
                for (char c : r.cur) blank &= (c == ' ' || c == '\r' || c == '\t');
                if (blank && r.p >= r.end) break;
                return bad(r.lineno);
            }
            --n;
            if (n >= m.n_points || n < 0) return bad(r.lineno);
            m.is_boundary[n] = 1;
            if (!r.next()) break;
        }
    }
    int nz = 0;
    for (auto& c : charts) nz = std::max(nz, (int)c.size() / 2);
    for (int z : m.tri_zone)
        if (z >= nz) {
            err = "triangle zone without chart coordinates";
            return false;
        }
    m.n_zones = nz;
    m.coords.assign((size_t)m.n_points * nz * 2, 0.0);
    for (int i = 0; i < m.n_points; ++i)
        std::copy(charts[i].begin(), charts[i].end(), m.coords.begin() + (size_t)i * nz * 2);
    // adjacency in triangle-id order (mesh.cpp:312-320)
    m.adj_ptr.assign(m.n_points + 1, 0);
    for (int v : m.tri) m.adj_ptr[v + 1]++;
    for (int i = 0; i < m.n_points; ++i) m.adj_ptr[i + 1] += m.adj_ptr[i];
    m.adj.resize(m.adj_ptr[m.n_points]);
    std::vector<int> fill(m.adj_ptr.begin(), m.adj_ptr.end() - 1);
    for (int t = 0; t < m.n_triangles; ++t)
        for (int a = 0; a < 3; ++a) m.adj[fill[m.tri[3 * t + a]]++] = t;
    // inner / outer / p2io (mesh.cpp:487-507)
    m.p2io.assign(m.n_points, 0);
    for (int i = 0; i < m.n_points; ++i) {
        if (!m.is_boundary[i]) {
            m.p2io[i] = (int)m.inner.size();
            m.inner.push_back(i);
        } else {
            m.p2io[i] = (int)m.outer.size();
            m.outer.push_back(i);
        }
    }
    for (int i = 0; i < m.n_points; ++i)
        if (m.adj_ptr[i + 1] == m.adj_ptr[i]) {
            err = "point without triangles";
            return false;
        }
    return true;
}

// ---------------------------------------------------------------- binary --
namespace {

const char MESH_MAGIC[8] = {'P', 'H', 'G', 'M', 'E', 'S', 'H', '1'};

uint64_t fnv1a(const unsigned char* p, size_t n, uint64_t h)
{
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

template <typename T>
void put(std::string& out, const std::vector<T>& v)
{
    out.append(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T));
}

template <typename T>
bool take(const std::string& in, size_t& at, std::vector<T>& v, size_t count)
{
    const size_t bytes = count * sizeof(T);
    if (in.size() - at < bytes) return false;
    v.resize(count);
    if (bytes) std::memcpy(v.data(), in.data() + at, bytes);
    at += bytes;
    return true;
}

}  // namespace

bool save_mesh_binary(const SphereMesh& m, const char* path, std::string& err)
{
    std::string body;
    const int32_t head[5] = {m.n_points, m.n_triangles, m.n_zones, (int32_t)m.inner.size(),
                             (int32_t)m.outer.size()};
    body.append(reinterpret_cast<const char*>(head), sizeof head);
    put(body, m.pr);
    put(body, m.coords);
    put(body, m.tri);
    put(body, m.tri_zone);
    put(body, m.adj_ptr);
    put(body, m.adj);
    put(body, m.is_boundary);
    put(body, m.inner);
    put(body, m.outer);
    put(body, m.p2io);
    const uint64_t sum = fnv1a(reinterpret_cast<const unsigned char*>(body.data()), body.size(),
                               1469598103934665603ull);
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) {
        err = "cannot write " + tmp;
        return false;
    }
    bool ok = std::fwrite(MESH_MAGIC, 1, 8, f) == 8 &&
              std::fwrite(&sum, sizeof sum, 1, f) == 1 &&
              std::fwrite(body.data(), 1, body.size(), f) == body.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        err = std::string("cannot write ") + path;
        return false;
    }
    return true;
}

bool load_mesh_binary(const char* path, SphereMesh& m, std::string& err)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        err = std::string("cannot open ") + path;
        return false;
    }
    std::string in;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) in.append(buf, got);
    std::fclose(f);
    err = std::string("bad mesh image ") + path;
    if (in.size() < 8 + 8 + 20 || std::memcmp(in.data(), MESH_MAGIC, 8) != 0) return false;
    uint64_t sum = 0;
    std::memcpy(&sum, in.data() + 8, sizeof sum);
    if (fnv1a(reinterpret_cast<const unsigned char*>(in.data()) + 16, in.size() - 16,
              1469598103934665603ull) != sum)
        return false;
    int32_t head[5];
    std::memcpy(head, in.data() + 16, sizeof head);
    const long V = head[0], T = head[1], nz = head[2], ni = head[3], no = head[4];
    if (V < 0 || T < 0 || nz < 1 || ni < 0 || no < 0 || ni + no != V) return false;
    SphereMesh r;
    r.n_points = (int)V;
    r.n_triangles = (int)T;
    r.n_zones = (int)nz;
    size_t at = 16 + sizeof head;
    if (!take(in, at, r.pr, (size_t)V * 3) || !take(in, at, r.coords, (size_t)V * nz * 2) ||
        !take(in, at, r.tri, (size_t)T * 3) || !take(in, at, r.tri_zone, (size_t)T) ||
        !take(in, at, r.adj_ptr, (size_t)V + 1) ||
        !take(in, at, r.adj, (size_t)(V + 1 > 0 && !r.adj_ptr.empty() ? r.adj_ptr[V] : 0)) ||
        !take(in, at, r.is_boundary, (size_t)V) || !take(in, at, r.inner, (size_t)ni) ||
        !take(in, at, r.outer, (size_t)no) || !take(in, at, r.p2io, (size_t)V) || at != in.size())
        return false;
    // structural checks: indices in range, adjacency consistent
    if (r.adj_ptr[0] != 0 || r.adj_ptr[V] != 3 * T) return false;
    for (int v : r.tri)
        if (v < 0 || v >= V) return false;
    for (long i = 0; i < V; ++i)
        if (r.adj_ptr[i + 1] < r.adj_ptr[i]) return false;
    for (int t : r.adj)
        if (t < 0 || t >= T) return false;
    for (int t : r.tri_zone)
        if (t < 0 || t >= nz) return false;
    for (int p : r.inner)
        if (p < 0 || p >= V || r.is_boundary[p]) return false;
    for (int p : r.outer)
        if (p < 0 || p >= V || !r.is_boundary[p]) return false;
    for (long i = 0; i < V; ++i) {
        const int q = r.p2io[i];
        if (q < 0 || q >= (r.is_boundary[i] ? no : ni)) return false;
    }
    m = std::move(r);
    err.clear();
    return true;
}

std::string print_function_text(const SphereMesh& m, const double* u, bool sphere_xyz)
{
    std::string out;
    char line[256];
    std::snprintf(line, sizeof line, "# points %d\n", m.n_points);
    out += line;
    const int nz = m.n_zones;
    for (int i = 0; i < m.n_points; ++i) {
        const double a = m.coords[(size_t)i * nz * 2], b = m.coords[(size_t)i * nz * 2 + 1];
        const double x = sphere_xyz ? std::cos(a) * std::cos(b) : a;
        const double y = sphere_xyz ? std::cos(a) * std::sin(b) : b;
        const double z = sphere_xyz ? std::sin(a) : 0.0;
        std::snprintf(line, sizeof line, "%.16f %.16f %.16f %.16f %.16f %.16f \n", x, y, z, u[i],
                      a, b);
        out += line;
    }
    std::snprintf(line, sizeof line, "# triangles %d\n", m.n_triangles);
    out += line;
    for (int t = 0; t < m.n_triangles; ++t) {
        std::snprintf(line, sizeof line, "%d %d %d\n", m.tri[3 * t] + 1, m.tri[3 * t + 1] + 1,
                      m.tri[3 * t + 2] + 1);
        out += line;
    }
    out += "# end \n";
    return out;
}

}  // namespace phg
