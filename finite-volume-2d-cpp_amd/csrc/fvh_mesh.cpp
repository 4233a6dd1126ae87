// fvh_mesh.cpp -- mesh input, reference numbering and geometry, RCM order.
//
// Reproduces what fvhyper::mesh::read_file (fvhyper/src/mesh.cpp:803-855)
// leaves in the mesh object on one rank: nodes, cells and boundary elements
// in file order; edges numbered by first appearance walking the cells in
// order (mesh.cpp:130-154); the second owner of an edge is the last other
// cell that lists it, 0 if none (mesh.cpp:157-187); geometry with the same
// operation order (mesh.cpp:190-260); one ghost cell per boundary element
// appended after the real cells (mesh.cpp:623-669).  Compiled with FP
// contraction off so the geometry is bit-identical to the reference's.
#include "fvh_internal.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <queue>
#include <unordered_map>

namespace fvh {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

namespace {

// Open-addressing map from an undirected node pair to an edge id.
class EdgeTable {
  public:
    explicit EdgeTable(size_t n) {
        size_t cap = 64;
        while (cap < 2 * n + 64) cap <<= 1;
        keys_.assign(cap, kEmpty);
        vals_.assign(cap, -1);
        mask_ = cap - 1;
    }
    static uint64_t key(int32_t a, int32_t b) {
        uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
        return ((uint64_t)lo << 32) | hi;
    }
    int32_t find(uint64_t k) const {
        for (size_t i = slot(k);; i = (i + 1) & mask_) {
            if (keys_[i] == k) return vals_[i];
            if (keys_[i] == kEmpty) return -1;
        }
    }
    // Inserts k -> v unless present; returns the stored id.
    int32_t insert(uint64_t k, int32_t v) {
        for (size_t i = slot(k);; i = (i + 1) & mask_) {
            if (keys_[i] == k) return vals_[i];
            if (keys_[i] == kEmpty) { keys_[i] = k; vals_[i] = v; return v; }
        }
    }

  private:
    static constexpr uint64_t kEmpty = ~0ull;
    size_t slot(uint64_t k) const {
        k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull; k ^= k >> 33;
        return (size_t)k & mask_;
    }
    std::vector<uint64_t> keys_;
    std::vector<int32_t> vals_;
    size_t mask_;
};

}  // namespace

// SU2 ASCII, 2-D: NDIME, NELEM (element lines "type nodes... [index]", 5 =
// triangle, 9 = quadrilateral), NPOIN ("x y [index]"), NMARK with
// MARKER_TAG / MARKER_ELEMS and line elements "3 a b".  Cells, nodes and
// boundary elements keep file order; marker tags become the physical names
// (the reference converts such files to MSH 4.1 with gmsh first).
int32_t read_su2(const char* path, fvh_mesh& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) { set_error(std::string("cannot open ") + path); return FVH_E_IO; }
    std::vector<std::string> lines;
    {
        std::string cur;
        int ch;
        while ((ch = std::fgetc(f)) != EOF) {
            if (ch == '\n') { lines.push_back(cur); cur.clear(); }
            else if (ch != '\r') cur.push_back((char)ch);
        }
        if (!cur.empty()) lines.push_back(cur);
        std::fclose(f);
    }
    // strip comments and blank lines
    std::vector<std::string> t;
    for (auto& l : lines) {
        const size_t p = l.find('%');
        std::string s2 = p == std::string::npos ? l : l.substr(0, p);
        const size_t a = s2.find_first_not_of(" \t"), b = s2.find_last_not_of(" \t");
        if (a != std::string::npos) t.push_back(s2.substr(a, b - a + 1));
    }
    auto fail = [&](const std::string& msg) { set_error("SU2 " + std::string(path) + ": " + msg); return FVH_E_MESH; };
    auto is_key = [](const std::string& l, const char* k) {
        std::string u;
        for (char c : l) {
            if (c == '=') break;
            if (c != ' ' && c != '\t') u.push_back((char)std::toupper((unsigned char)c));
        }
        return u == k && l.find('=') != std::string::npos;
    };
    auto value = [](const std::string& l) {
        std::string v = l.substr(l.find('=') + 1);
        const size_t a = v.find_first_not_of(" \t");
        return a == std::string::npos ? std::string() : v.substr(a);
    };
    auto fields = [](const std::string& l) {
        std::vector<std::string> r;
        size_t i = 0;
        while (i < l.size()) {
            while (i < l.size() && (l[i] == ' ' || l[i] == '\t')) ++i;
            const size_t j = i;
            while (i < l.size() && l[i] != ' ' && l[i] != '\t') ++i;
            if (i > j) r.push_back(l.substr(j, i - j));
        }
        return r;
    };
    std::vector<double> xs, ys;
    std::vector<int32_t> cells, bnd, tags;
    std::vector<std::string> names;
    size_t i = 0;
    while (i < t.size()) {
        const std::string& l = t[i];
        if (is_key(l, "NDIME")) {
            if (std::stoi(value(l)) != 2) return fail("only 2-D meshes");
            ++i;
        } else if (is_key(l, "NELEM")) {
            const long n = std::stol(value(l));
            if (n < 0 || i + 1 + (size_t)n > t.size()) return fail("truncated NELEM block");
            for (long k = 0; k < n; ++k) {
                const auto fs = fields(t[i + 1 + k]);
                const int kind = fs.empty() ? -1 : std::stoi(fs[0]);
                if (kind == 5 && fs.size() >= 4) {
                    cells.insert(cells.end(), {std::stoi(fs[1]), std::stoi(fs[2]), std::stoi(fs[3]), -1});
                } else if (kind == 9 && fs.size() >= 5) {
                    cells.insert(cells.end(), {std::stoi(fs[1]), std::stoi(fs[2]), std::stoi(fs[3]), std::stoi(fs[4])});
                } else {
                    return fail("element type " + std::to_string(kind) + " is not a triangle or quadrilateral");
                }
            }
            i += 1 + n;
        } else if (is_key(l, "NPOIN")) {
            const long n = std::stol(fields(value(l))[0]);
            if (n < 0 || i + 1 + (size_t)n > t.size()) return fail("truncated NPOIN block");
            for (long k = 0; k < n; ++k) {
                const auto fs = fields(t[i + 1 + k]);
                if (fs.size() < 2) return fail("bad point line");
                xs.push_back(std::strtod(fs[0].c_str(), nullptr));
                ys.push_back(std::strtod(fs[1].c_str(), nullptr));
            }
            i += 1 + n;
        } else if (is_key(l, "MARKER_TAG")) {
            names.push_back(value(l));
            if (i + 1 >= t.size() || !is_key(t[i + 1], "MARKER_ELEMS")) return fail("MARKER_TAG without MARKER_ELEMS");
            const long n = std::stol(value(t[i + 1]));
            if (n < 0 || i + 2 + (size_t)n > t.size()) return fail("truncated marker block");
            for (long k = 0; k < n; ++k) {
                const auto fs = fields(t[i + 2 + k]);
                if (fs.size() < 3 || std::stoi(fs[0]) != 3) return fail("marker element is not a line");
                bnd.push_back(std::stoi(fs[1]));
                bnd.push_back(std::stoi(fs[2]));
                tags.push_back((int32_t)names.size() - 1);
            }
            i += 2 + n;
        } else {
            ++i;
        }
    }
    for (int32_t v : cells)
        if (v >= (int32_t)xs.size()) return fail("element node out of range");
    return build_mesh(out, (int32_t)xs.size(), xs.data(), ys.data(), (int32_t)(cells.size() / 4), cells.data(),
                      (int32_t)(bnd.size() / 2), bnd.data(), tags.data(), names);
}

void GeoBar::resize(int32_t n_edges, int32_t n_cells) {
    for (auto* v : {&e_nx, &e_ny, &e_len, &e_w, &e_cx, &e_cy}) v->assign(n_edges, 0.0);
    for (auto* v : {&c_area, &c_cx, &c_cy}) v->assign(n_cells, 0.0);
}

void mesh_geometry_vjp(const fvh_mesh& m, GeoBar& b, double* gxy) {
    const int32_t E = m.n_edges();
    const double* x = m.x.data();
    const double* y = m.y.data();
    // Green-Gauss weight w = dif / dij of each edge (cells i, j: first and second owner)
    for (int32_t e = 0; e < E; ++e) {
        const double lw = b.e_w[e];
        if (lw == 0.0) continue;
        const int32_t i = m.edge_cells[2 * (int64_t)e], j = m.edge_cells[2 * (int64_t)e + 1];
        const double dxif = m.e_cx[e] - m.c_cx[i], dyif = m.e_cy[e] - m.c_cy[i];
        const double dif = std::sqrt(dxif * dxif + dyif * dyif);
        const double dxij = m.c_cx[i] - m.c_cx[j], dyij = m.c_cy[i] - m.c_cy[j];
        const double dij = std::sqrt(dxij * dxij + dyij * dyij);
        if (!(dij > 0.0)) continue;
        const double ldif = lw / dij, ldij = -lw * dif / (dij * dij);
        if (dif > 0.0) {
            b.e_cx[e] += ldif * dxif / dif;
            b.e_cy[e] += ldif * dyif / dif;
            b.c_cx[i] -= ldif * dxif / dif;
            b.c_cy[i] -= ldif * dyif / dif;
        }
        b.c_cx[i] += ldij * dxij / dij;
        b.c_cy[i] += ldij * dyij / dij;
        b.c_cx[j] -= ldij * dxij / dij;
        b.c_cy[j] -= ldij * dyij / dij;
    }
    // ghost cells: area of the owner, centre x_f + |x_f - x_owner| n
    for (size_t k = 0; k < m.bnd_edge.size(); ++k) {
        const int32_t g = m.n_real + (int32_t)k, e = m.bnd_edge[k];
        const int32_t o = m.edge_cells[2 * (int64_t)e];
        b.c_area[o] += b.c_area[g];
        const double lgx = b.c_cx[g], lgy = b.c_cy[g];
        if (lgx == 0.0 && lgy == 0.0) continue;
        const double dx = m.e_cx[e] - m.c_cx[o], dy = m.e_cy[e] - m.c_cy[o];
        const double dist = std::sqrt(dx * dx + dy * dy);
        b.e_cx[e] += lgx;
        b.e_cy[e] += lgy;
        b.e_nx[e] += dist * lgx;
        b.e_ny[e] += dist * lgy;
        const double ldist = lgx * m.e_nx[e] + lgy * m.e_ny[e];
        if (dist > 0.0) {
            b.e_cx[e] += ldist * dx / dist;
            b.e_cy[e] += ldist * dy / dist;
            b.c_cx[o] -= ldist * dx / dist;
            b.c_cy[o] -= ldist * dy / dist;
        }
    }
    // edges: centre (x_a + x_b) / 2; t = x_b - x_a; len = |t|; n = s (-t_y, t_x) / len
    for (int32_t e = 0; e < E; ++e) {
        const int32_t a = m.edge_nodes[2 * (int64_t)e], bb = m.edge_nodes[2 * (int64_t)e + 1];
        const double tx = x[bb] - x[a], ty = y[bb] - y[a];
        const double l = m.e_len[e];
        const double nx = m.e_nx[e], ny = m.e_ny[e];
        // n = s u / l with u = (-t_y, t_x): lam_u = (s / l)(lam_n - (lam_n . n) n)
        const double s = (nx * (-ty) + ny * tx) < 0.0 ? -1.0 : 1.0;
        const double ln = b.e_nx[e] * nx + b.e_ny[e] * ny;
        const double lux = s / l * (b.e_nx[e] - ln * nx);
        const double luy = s / l * (b.e_ny[e] - ln * ny);
        const double ltx = luy + b.e_len[e] * tx / l;
        const double lty = -lux + b.e_len[e] * ty / l;
        gxy[2 * (int64_t)bb] += ltx + 0.5 * b.e_cx[e];
        gxy[2 * (int64_t)bb + 1] += lty + 0.5 * b.e_cy[e];
        gxy[2 * (int64_t)a] += -ltx + 0.5 * b.e_cx[e];
        gxy[2 * (int64_t)a + 1] += -lty + 0.5 * b.e_cy[e];
    }
    // real cells: centre = mean of the vertices; area 0.5 |D| per triangle
    auto tri = [&](int32_t p, int32_t q, int32_t r, double lA) {
        const double D = x[p] * (y[q] - y[r]) + x[q] * (y[r] - y[p]) + x[r] * (y[p] - y[q]);
        const double lD = 0.5 * (D < 0.0 ? -1.0 : 1.0) * lA;
        gxy[2 * (int64_t)p] += lD * (y[q] - y[r]);
        gxy[2 * (int64_t)q] += lD * (y[r] - y[p]);
        gxy[2 * (int64_t)r] += lD * (y[p] - y[q]);
        gxy[2 * (int64_t)p + 1] += lD * (x[r] - x[q]);
        gxy[2 * (int64_t)q + 1] += lD * (x[p] - x[r]);
        gxy[2 * (int64_t)r + 1] += lD * (x[q] - x[p]);
    };
    for (int32_t c = 0; c < m.n_real; ++c) {
        const int32_t* cn = &m.cell_nodes[4 * (int64_t)c];
        const int n = m.cell_tri[c] ? 3 : 4;
        for (int j = 0; j < n; ++j) {
            gxy[2 * (int64_t)cn[j]] += b.c_cx[c] / (double)n;
            gxy[2 * (int64_t)cn[j] + 1] += b.c_cy[c] / (double)n;
        }
        if (b.c_area[c] != 0.0) {
            tri(cn[0], cn[1], cn[2], b.c_area[c]);
            if (n == 4) tri(cn[0], cn[2], cn[3], b.c_area[c]);
        }
    }
}

int32_t build_mesh(fvh_mesh& m, int32_t n_nodes, const double* x, const double* y,
                   int32_t n_cells, const int32_t* cells, int32_t n_bnd,
                   const int32_t* bnd, const int32_t* bnd_tag,
                   const std::vector<std::string>& names) {
    if (n_nodes < 0 || n_cells < 0 || n_bnd < 0) { set_error("negative count"); return FVH_E_ARG; }
    m = fvh_mesh();
    m.x.assign(x, x + n_nodes);
    m.y.assign(y, y + n_nodes);
    m.names = names;
    const int64_t NC = (int64_t)n_cells + n_bnd;
    m.cell_nodes.assign(4 * NC, 0);
    m.cell_tri.assign(NC, 1);
    for (int32_t c = 0; c < n_cells; ++c) {
        const int32_t* cn = cells + 4 * (int64_t)c;
        const bool tri = cn[3] < 0;
        for (int k = 0; k < (tri ? 3 : 4); ++k) {
            if (cn[k] < 0 || cn[k] >= n_nodes) {
                set_error("cell " + std::to_string(c) + " references node " + std::to_string(cn[k]) +
                          " outside [0," + std::to_string(n_nodes) + ")");
                return FVH_E_MESH;
            }
            m.cell_nodes[4 * (int64_t)c + k] = cn[k];
        }
        m.cell_tri[c] = tri ? 1 : 0;
    }
    // edges by first appearance (cell order, local order 0-1, 1-2, .., last-0)
    EdgeTable tab(4 * (size_t)n_cells);
    m.edge_nodes.reserve(4 * (size_t)n_cells);
    m.edge_cells.reserve(4 * (size_t)n_cells);
    int32_t E = 0;
    for (int32_t c = 0; c < n_cells; ++c) {
        const int sz = m.cell_tri[c] ? 3 : 4;
        const int32_t* cn = &m.cell_nodes[4 * (int64_t)c];
        for (int i = 0; i < sz; ++i) {
            const int32_t a = cn[i], b = cn[(i + 1) % sz];
            if (tab.insert(EdgeTable::key(a, b), E) == E) {
                m.edge_nodes.push_back(a);
                m.edge_nodes.push_back(b);
                m.edge_cells.push_back(c);
                m.edge_cells.push_back(0);
                ++E;
            }
        }
    }
    // second owner: last cell other than the first owner that lists the edge
    for (int32_t c = 0; c < n_cells; ++c) {
        const int sz = m.cell_tri[c] ? 3 : 4;
        const int32_t* cn = &m.cell_nodes[4 * (int64_t)c];
        for (int i = 0; i < sz; ++i) {
            const int32_t e = tab.find(EdgeTable::key(cn[i], cn[(i + 1) % sz]));
            if (m.edge_cells[2 * (int64_t)e] != c) m.edge_cells[2 * (int64_t)e + 1] = c;
        }
    }
    m.n_real = n_cells;
    // geometry
    m.c_area.assign(NC, 0.0);
    m.c_cx.assign(NC, 0.0);
    m.c_cy.assign(NC, 0.0);
    for (int32_t c = 0; c < n_cells; ++c) {
        const int n = m.cell_tri[c] ? 3 : 4;
        const double dn = (double)n;
        double sx = 0.0, sy = 0.0;
        for (int j = 0; j < n; ++j) {
            const int32_t v = m.cell_nodes[4 * (int64_t)c + j];
            sx += x[v] / dn;
            sy += y[v] / dn;
        }
        m.c_cx[c] = sx;
        m.c_cy[c] = sy;
    }
    m.e_len.resize(E); m.e_nx.resize(E); m.e_ny.resize(E); m.e_cx.resize(E); m.e_cy.resize(E);
    for (int32_t e = 0; e < E; ++e) {
        const int32_t a = m.edge_nodes[2 * (int64_t)e], b = m.edge_nodes[2 * (int64_t)e + 1];
        const double cx = (x[b] + x[a]) * 0.5;
        const double cy = (y[b] + y[a]) * 0.5;
        const double tx = x[b] - x[a];
        const double ty = y[b] - y[a];
        const double l = std::sqrt(tx * tx + ty * ty);
        double nx = -ty / l;
        double ny = tx / l;
        const int32_t c0 = m.edge_cells[2 * (int64_t)e];
        // orient away from the first owner
        if (nx * (cx - m.c_cx[c0]) + ny * (cy - m.c_cy[c0]) < 0) {
            nx = -nx;
            ny = -ny;
        }
        m.e_len[e] = l; m.e_nx[e] = nx; m.e_ny[e] = ny; m.e_cx[e] = cx; m.e_cy[e] = cy;
    }
    auto tri_area2 = [&](int32_t p, int32_t q, int32_t r) {
        return x[p] * (y[q] - y[r]) + x[q] * (y[r] - y[p]) + x[r] * (y[p] - y[q]);
    };
    for (int32_t c = 0; c < n_cells; ++c) {
        const int32_t* cn = &m.cell_nodes[4 * (int64_t)c];
        double A = 0.5 * std::fabs(tri_area2(cn[0], cn[1], cn[2]));
        if (!m.cell_tri[c]) A = A + 0.5 * std::fabs(tri_area2(cn[0], cn[2], cn[3]));
        m.c_area[c] = A;
    }
    // boundary ghost cells
    m.bnd_edge.resize(n_bnd);
    m.bnd_tag.resize(n_bnd);
    int32_t g = n_cells;
    for (int32_t b = 0; b < n_bnd; ++b) {
        const int32_t a = bnd[2 * (int64_t)b], bb = bnd[2 * (int64_t)b + 1];
        const int32_t e = (a >= 0 && bb >= 0) ? tab.find(EdgeTable::key(a, bb)) : -1;
        if (e < 0) {
            set_error("invalid edge ref (" + std::to_string(std::min(a, bb)) + ", " +
                      std::to_string(std::max(a, bb)) + ")");
            return FVH_E_MESH;
        }
        if (bnd_tag[b] < 0 || bnd_tag[b] >= (int32_t)names.size()) {
            set_error("boundary element " + std::to_string(b) + " has no physical name");
            return FVH_E_MESH;
        }
        const int32_t owner = m.edge_cells[2 * (int64_t)e];
        const double dx = m.e_cx[e] - m.c_cx[owner];
        const double dy = m.e_cy[e] - m.c_cy[owner];
        const double dist = std::sqrt(dx * dx + dy * dy);
        m.c_area[g] = m.c_area[owner];
        m.c_cx[g] = m.e_cx[e] + dist * m.e_nx[e];
        m.c_cy[g] = m.e_cy[e] + dist * m.e_ny[e];
        m.cell_tri[g] = 1;
        m.cell_nodes[4 * (int64_t)g] = std::min(a, bb);
        m.cell_nodes[4 * (int64_t)g + 1] = std::max(a, bb);
        m.edge_cells[2 * (int64_t)e + 1] = g;
        m.bnd_edge[b] = e;
        m.bnd_tag[b] = bnd_tag[b];
        ++g;
    }
    m.n_cells = g;
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// gmsh MSH 4.1 ASCII, the subset fvhyper reads (mesh.cpp:264-617), in one pass.
// ---------------------------------------------------------------------------
namespace {

struct LineReader {
    const char* p;
    const char* end;
    size_t lineno = 0;
    bool next(std::vector<const char*>& tok, std::vector<size_t>& len, const char*& line, size_t& line_len) {
        if (p >= end) return false;
        const char* s = p;
        const char* e = (const char*)memchr(s, '\n', end - s);
        if (!e) e = end;
        p = e + 1;
        ++lineno;
        line = s;
        line_len = e - s;
        if (line_len && s[line_len - 1] == '\r') --line_len;
        tok.clear();
        len.clear();
        size_t i = 0;
        while (i < line_len) {
            while (i < line_len && (s[i] == ' ' || s[i] == '\t')) ++i;
            if (i >= line_len) break;
            size_t j = i;
            while (j < line_len && s[j] != ' ' && s[j] != '\t') ++j;
            tok.push_back(s + i);
            len.push_back(j - i);
            i = j;
        }
        return true;
    }
};

bool to_i64(const char* s, size_t n, int64_t& v) {
    char buf[32];
    if (n == 0 || n >= sizeof buf) return false;
    memcpy(buf, s, n);
    buf[n] = 0;
    char* endp = nullptr;
    errno = 0;
    v = std::strtoll(buf, &endp, 10);
    // like std::stoi, trailing non-digits (e.g. the ".5" of "0.5") end the number
    return endp != buf && errno == 0;
}

bool to_f64(const char* s, size_t n, double& v) {
    char buf[64];
    if (n == 0 || n >= sizeof buf) return false;
    memcpy(buf, s, n);
    buf[n] = 0;
    char* endp = nullptr;
    v = std::strtod(buf, &endp);
    return endp != buf;
}

}  // namespace

int32_t read_msh(const char* path, fvh_mesh& out) {
    std::FILE* f = std::fopen(path, "rb");
    if (!f) { set_error(std::string("cannot open ") + path); return FVH_E_IO; }
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<char> buf(sz > 0 ? sz : 1);
    if (sz > 0 && std::fread(buf.data(), 1, sz, f) != (size_t)sz) {
        std::fclose(f);
        set_error(std::string("cannot read ") + path);
        return FVH_E_IO;
    }
    std::fclose(f);

    LineReader rd{buf.data(), buf.data() + (sz > 0 ? sz : 0)};
    std::vector<const char*> tok;
    std::vector<size_t> len;
    const char* line;
    size_t line_len;
    auto fail = [&](const std::string& what) {
        set_error(std::string(path) + ":" + std::to_string(rd.lineno) + ": " + what);
        return FVH_E_FORMAT;
    };
    auto I = [&](size_t k, int64_t& v) { return k < tok.size() && to_i64(tok[k], len[k], v); };

    std::map<int64_t, std::string> phys_names;          // physical tag -> name
    std::unordered_map<int64_t, int64_t> entity_phys;   // curve entity -> physical tag
    std::vector<double> xs, ys;
    std::vector<int32_t> node_of_tag;                    // node tag -> file-order index
    std::unordered_map<int64_t, int32_t> node_of_tag_sparse;
    bool dense_tags = true;
    std::vector<int32_t> cells;
    std::vector<int32_t> bnd;
    std::vector<int64_t> bnd_phys;

    auto node_index = [&](int64_t tag, int32_t& idx) -> bool {
        if (dense_tags) {
            if (tag < 0 || tag >= (int64_t)node_of_tag.size() || node_of_tag[tag] < 0) return false;
            idx = node_of_tag[tag];
            return true;
        }
        auto it = node_of_tag_sparse.find(tag);
        if (it == node_of_tag_sparse.end()) return false;
        idx = it->second;
        return true;
    };

    std::string section;
    while (rd.next(tok, len, line, line_len)) {
        if (line_len && line[0] == '$') {
            section.assign(line + 1, line_len - 1);
            if (section == "MeshFormat") {
                if (!rd.next(tok, len, line, line_len) || tok.empty()) return fail("missing format line");
                if (std::string(tok[0], len[0]).rfind("4", 0) != 0) return fail("only MSH 4.x is read");
                int64_t ft = 0;
                if (I(1, ft) && ft != 0) return fail("binary MSH files are not supported");
            } else if (section == "PhysicalNames") {
                int64_t n = 0;
                if (!rd.next(tok, len, line, line_len) || !I(0, n)) return fail("bad PhysicalNames count");
                for (int64_t k = 0; k < n; ++k) {
                    int64_t tag = 0;
                    if (!rd.next(tok, len, line, line_len) || tok.size() < 3 || !I(1, tag))
                        return fail("bad PhysicalNames entry");
                    std::string nm(tok[2], len[2]);
                    nm.erase(std::remove(nm.begin(), nm.end(), '"'), nm.end());
                    phys_names[tag] = nm;
                }
            } else if (section == "Entities") {
                int64_t np = 0, nc = 0, ns = 0, nv = 0;
                if (!rd.next(tok, len, line, line_len) || !I(0, np) || !I(1, nc) || !I(2, ns) || !I(3, nv))
                    return fail("bad Entities counts");
                for (int64_t k = 0; k < np + nc + ns + nv; ++k) {
                    if (!rd.next(tok, len, line, line_len)) return fail("truncated Entities");
                    if (k >= np && k < np + nc) {
                        // field 8 is the first physical tag (mesh.cpp:311)
                        int64_t tag = 0, ph = 0;
                        if (I(0, tag) && I(8, ph)) entity_phys[tag] = ph;
                    }
                }
            } else if (section == "PartitionedEntities") {
                int64_t nparts = 0, nghost = 0;
                if (!rd.next(tok, len, line, line_len) || !I(0, nparts)) return fail("bad PartitionedEntities");
                if (!rd.next(tok, len, line, line_len) || !I(0, nghost)) return fail("bad PartitionedEntities");
                for (int64_t k = 0; k < nghost; ++k)
                    if (!rd.next(tok, len, line, line_len)) return fail("truncated PartitionedEntities");
                int64_t np = 0, nc = 0, ns = 0, nv = 0;
                if (!rd.next(tok, len, line, line_len) || !I(0, np) || !I(1, nc) || !I(2, ns) || !I(3, nv))
                    return fail("bad PartitionedEntities counts");
                for (int64_t k = 0; k < np + nc + ns + nv; ++k) {
                    if (!rd.next(tok, len, line, line_len)) return fail("truncated PartitionedEntities");
                    if (k >= np && k < np + nc) {
                        int64_t tag = 0, pdim = 0, ptag = 0;
                        if (I(0, tag) && I(1, pdim) && I(2, ptag) && pdim == 1) {
                            auto it = entity_phys.find(ptag);
                            if (it == entity_phys.end()) return fail("partition curve with unknown parent");
                            entity_phys[tag] = it->second;
                        }
                    }
                }
            } else if (section == "Nodes") {
                int64_t nblocks = 0, nnodes = 0, mintag = 0, maxtag = 0;
                if (!rd.next(tok, len, line, line_len) || !I(0, nblocks) || !I(1, nnodes) || !I(2, mintag) ||
                    !I(3, maxtag))
                    return fail("bad Nodes header");
                dense_tags = maxtag >= 0 && maxtag <= 4 * nnodes + 1024;
                if (dense_tags) node_of_tag.assign(maxtag + 1, -1);
                xs.reserve(nnodes);
                ys.reserve(nnodes);
                for (int64_t bk = 0; bk < nblocks; ++bk) {
                    int64_t n = 0;
                    if (!rd.next(tok, len, line, line_len) || !I(3, n)) return fail("bad node block header");
                    std::vector<int64_t> tags(n);
                    for (int64_t k = 0; k < n; ++k)
                        if (!rd.next(tok, len, line, line_len) || !I(0, tags[k])) return fail("bad node tag");
                    for (int64_t k = 0; k < n; ++k) {
                        double vx, vy;
                        if (!rd.next(tok, len, line, line_len) || tok.size() < 2 || !to_f64(tok[0], len[0], vx) ||
                            !to_f64(tok[1], len[1], vy))
                            return fail("bad node coordinates");
                        const int32_t idx = (int32_t)xs.size();
                        const int64_t t = tags[k] - 1;   // tags are 1-based
                        if (dense_tags) {
                            if (t < 0 || t >= (int64_t)node_of_tag.size()) return fail("node tag out of range");
                            node_of_tag[t] = idx;
                        } else {
                            node_of_tag_sparse[t] = idx;
                        }
                        xs.push_back(vx);
                        ys.push_back(vy);
                    }
                }
            } else if (section == "Elements") {
                int64_t nblocks = 0;
                if (!rd.next(tok, len, line, line_len) || !I(0, nblocks)) return fail("bad Elements header");
                for (int64_t bk = 0; bk < nblocks; ++bk) {
                    int64_t dim = 0, etag = 0, n = 0;
                    if (!rd.next(tok, len, line, line_len) || !I(0, dim) || !I(1, etag) || !I(3, n))
                        return fail("bad element block header");
                    int64_t ph = -1;
                    if (dim == 1) {
                        auto it = entity_phys.find(etag);
                        if (it == entity_phys.end()) return fail("line block on a curve without physical tag");
                        ph = it->second;
                    }
                    for (int64_t k = 0; k < n; ++k) {
                        if (!rd.next(tok, len, line, line_len)) return fail("truncated element block");
                        if (dim != 1 && dim != 2) continue;
                        const size_t nn = tok.size() - 1;
                        int32_t v[4];
                        if (dim == 1) {
                            if (nn != 2) continue;
                        } else {
                            if (nn == 2) continue;
                            if (nn != 3 && nn != 4) return fail("only 3- and 4-node cells are supported");
                        }
                        for (size_t j = 0; j < nn; ++j) {
                            int64_t t = 0;
                            if (!I(1 + j, t) || !node_index(t - 1, v[j])) return fail("element references an unknown node");
                        }
                        if (dim == 1) {
                            bnd.push_back(v[0]);
                            bnd.push_back(v[1]);
                            bnd_phys.push_back(ph);
                        } else {
                            cells.push_back(v[0]);
                            cells.push_back(v[1]);
                            cells.push_back(v[2]);
                            cells.push_back(nn == 4 ? v[3] : -1);
                        }
                    }
                }
            } else if (section == "GhostElements") {
                int64_t n = 0;
                if (rd.next(tok, len, line, line_len) && I(0, n) && n > 0)
                    return fail("partitioned (ghost-element) meshes need several ranks; one device reads whole meshes");
            }
        }
    }
    // a file without nodes or cells (empty, or not a mesh) is refused rather
    // than read as an empty mesh
    if (xs.empty() || cells.empty()) return fail("no nodes or no 2-D cells: not a mesh");
    // physical names of the boundary groups, in order of first use
    std::vector<std::string> names;
    std::map<int64_t, int32_t> slot_of_phys;
    std::vector<int32_t> tags(bnd_phys.size());
    for (size_t b = 0; b < bnd_phys.size(); ++b) {
        auto it = slot_of_phys.find(bnd_phys[b]);
        if (it == slot_of_phys.end()) {
            auto pn = phys_names.find(bnd_phys[b]);
            names.push_back(pn == phys_names.end() ? std::string() : pn->second);
            it = slot_of_phys.emplace(bnd_phys[b], (int32_t)names.size() - 1).first;
        }
        tags[b] = it->second;
    }
    return build_mesh(out, (int32_t)xs.size(), xs.data(), ys.data(), (int32_t)(cells.size() / 4),
                      cells.data(), (int32_t)(bnd.size() / 2), bnd.data(), tags.data(), names);
}

// ---------------------------------------------------------------------------
// Reverse Cuthill-McKee over the face graph of the real cells.
// ---------------------------------------------------------------------------
std::vector<int32_t> rcm_order(const fvh_mesh& m) {
    const int32_t N = m.n_real;
    std::vector<int32_t> off(N + 1, 0), adj;
    const int32_t E = m.n_edges();
    for (int32_t e = 0; e < E; ++e) {
        const int32_t a = m.edge_cells[2 * (int64_t)e], b = m.edge_cells[2 * (int64_t)e + 1];
        if (a != b && a < N && b < N) { ++off[a + 1]; ++off[b + 1]; }
    }
    for (int32_t i = 0; i < N; ++i) off[i + 1] += off[i];
    adj.resize(off[N]);
    {
        std::vector<int32_t> pos(off.begin(), off.end() - 1);
        for (int32_t e = 0; e < E; ++e) {
            const int32_t a = m.edge_cells[2 * (int64_t)e], b = m.edge_cells[2 * (int64_t)e + 1];
            if (a != b && a < N && b < N) { adj[pos[a]++] = b; adj[pos[b]++] = a; }
        }
    }
    auto deg = [&](int32_t v) { return off[v + 1] - off[v]; };
    auto by_deg = [&](int32_t a, int32_t b) { return deg(a) != deg(b) ? deg(a) < deg(b) : a < b; };
    for (int32_t v = 0; v < N; ++v) std::sort(adj.begin() + off[v], adj.begin() + off[v + 1], by_deg);

    std::vector<int32_t> order;
    order.reserve(N);
    std::vector<uint8_t> done(N, 0);
    std::vector<int32_t> level(N, -1);
    std::vector<int32_t> bfs;
    // BFS restricted to the component of r among not-done cells; returns the
    // eccentricity and fills bfs with the visit order.
    auto bfs_levels = [&](int32_t r) {
        for (int32_t v : bfs) level[v] = -1;
        bfs.clear();
        bfs.push_back(r);
        level[r] = 0;
        for (size_t h = 0; h < bfs.size(); ++h) {
            const int32_t v = bfs[h];
            for (int32_t k = off[v]; k < off[v + 1]; ++k) {
                const int32_t w = adj[k];
                if (!done[w] && level[w] < 0) { level[w] = level[v] + 1; bfs.push_back(w); }
            }
        }
        return level[bfs.back()];
    };
    int32_t scan = 0;
    while ((int32_t)order.size() < N) {
        // lowest-degree not-done cell (ties: lowest index) seeds the component
        int32_t r = -1;
        for (int32_t v = scan; v < N; ++v)
            if (!done[v] && (r < 0 || deg(v) < deg(r))) r = v;
        while (scan < N && done[scan]) ++scan;
        // pseudo-peripheral cell (George and Liu)
        int32_t ecc = bfs_levels(r);
        for (;;) {
            int32_t best = -1;
            for (int32_t v : bfs)
                if (level[v] == ecc && (best < 0 || by_deg(v, best))) best = v;
            const int32_t e2 = bfs_levels(best);
            if (e2 <= ecc) break;
            r = best;
            ecc = e2;
        }
        for (int32_t v : bfs) level[v] = -1;
        bfs.clear();
        // Cuthill-McKee sweep
        const size_t start = order.size();
        order.push_back(r);
        done[r] = 1;
        for (size_t h = start; h < order.size(); ++h) {
            const int32_t v = order[h];
            for (int32_t k = off[v]; k < off[v + 1]; ++k) {
                const int32_t w = adj[k];
                if (!done[w]) { done[w] = 1; order.push_back(w); }
            }
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

std::vector<int32_t> rcb_order(const fvh_mesh& m, int32_t tile, const std::vector<uint8_t>* split,
                               std::vector<int32_t>* leaf_sizes) {
    const int32_t N = m.n_real;
    std::vector<int32_t> idx(N);
    for (int32_t i = 0; i < N; ++i) idx[i] = i;
    if (tile < 1) tile = 1;
    if (leaf_sizes) leaf_sizes->clear();
    // the longer side of the centroids' bounding box of idx[lo, hi)
    auto by_x_of = [&](int32_t lo, int32_t hi) {
        double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
        for (int32_t k = lo; k < hi; ++k) {
            const int32_t c = idx[k];
            x0 = std::min(x0, m.c_cx[c]); x1 = std::max(x1, m.c_cx[c]);
            y0 = std::min(y0, m.c_cy[c]); y1 = std::max(y1, m.c_cy[c]);
        }
        return (x1 - x0) >= (y1 - y0);
    };
    // idx[lo, hi) split at mid: the cells below mid along the longer side first
    auto partition = [&](int32_t lo, int32_t mid, int32_t hi) {
        const std::vector<double>& key = by_x_of(lo, hi) ? m.c_cx : m.c_cy;
        std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi, [&](int32_t a, int32_t b) {
            return key[a] != key[b] ? key[a] < key[b] : a < b;
        });
    };
    // explicit stack: (lo, hi) ranges, left part processed first
    std::vector<std::pair<int32_t, int32_t>> stack;
    std::vector<int32_t> out;
    out.reserve(N);
    stack.push_back({0, N});
    int32_t leaf = 0;
    auto emit = [&](int32_t lo, int32_t hi) {
        std::sort(idx.begin() + lo, idx.begin() + hi);
        out.insert(out.end(), idx.begin() + lo, idx.begin() + hi);
        if (leaf_sizes) leaf_sizes->push_back(hi - lo);
    };
    while (!stack.empty()) {
        auto [lo, hi] = stack.back();
        stack.pop_back();
        const int32_t n = hi - lo;
        if (n <= tile) {
            // a leaf marked in split is bisected once more, into two leaves
            const bool bisect = split && leaf < (int32_t)split->size() && (*split)[leaf] && n >= 2;
            ++leaf;
            if (bisect) {
                const int32_t mid = lo + n / 2;
                partition(lo, mid, hi);
                emit(lo, mid);
                emit(mid, hi);
            } else {
                emit(lo, hi);
            }
            continue;
        }
        const int32_t tiles = (n + tile - 1) / tile;
        const int32_t mid = lo + (tiles / 2) * tile;
        partition(lo, mid, hi);
        stack.push_back({mid, hi});   // right after left
        stack.push_back({lo, mid});
    }
    return out;
}

}  // namespace fvh

// ---------------------------------------------------------------------------
// C ABI: mesh
// ---------------------------------------------------------------------------
extern "C" {

const char* fvh_last_error(void) { return fvh::last_error(); }
int32_t fvh_abi_version(void) { return FVH_ABI_VERSION; }

int32_t fvh_mesh_read_su2(const char* path, fvh_mesh** out) {
    if (!path || !out) { fvh::set_error("null argument"); return FVH_E_ARG; }
    try {
        fvh_mesh* m = new fvh_mesh();
        int32_t rc = fvh::read_su2(path, *m);
        if (rc != FVH_OK) { delete m; return rc; }
        *out = m;
        return FVH_OK;
    } catch (const std::exception& e) {
        fvh::set_error(e.what());
        return FVH_E_ARG;
    }
}

int32_t fvh_mesh_read_msh(const char* path, fvh_mesh** out) {
    if (!path || !out) { fvh::set_error("null argument"); return FVH_E_ARG; }
    try {
        fvh_mesh* m = new fvh_mesh();
        int32_t rc = fvh::read_msh(path, *m);
        if (rc != FVH_OK) { delete m; return rc; }
        *out = m;
        return FVH_OK;
    } catch (const std::exception& e) {
        fvh::set_error(e.what());
        return FVH_E_ARG;
    }
}

int32_t fvh_mesh_from_arrays(int32_t n_nodes, const double* x, const double* y,
                             int32_t n_cells, const int32_t* cells,
                             int32_t n_bnd, const int32_t* bnd, const int32_t* bnd_tag,
                             int32_t n_names, const char* const* names, fvh_mesh** out) {
    if (!out || (n_nodes && (!x || !y)) || (n_cells && !cells) || (n_bnd && (!bnd || !bnd_tag)) ||
        (n_names && !names)) {
        fvh::set_error("null argument");
        return FVH_E_ARG;
    }
    try {
        std::vector<std::string> nm;
        for (int32_t k = 0; k < n_names; ++k) nm.emplace_back(names[k] ? names[k] : "");
        fvh_mesh* m = new fvh_mesh();
        int32_t rc = fvh::build_mesh(*m, n_nodes, x, y, n_cells, cells, n_bnd, bnd, bnd_tag, nm);
        if (rc != FVH_OK) { delete m; return rc; }
        *out = m;
        return FVH_OK;
    } catch (const std::exception& e) {
        fvh::set_error(e.what());
        return FVH_E_ARG;
    }
}

void fvh_mesh_free(fvh_mesh* m) { delete m; }

int32_t fvh_mesh_counts(const fvh_mesh* m, int64_t* c) {
    if (!m || !c) { fvh::set_error("null argument"); return FVH_E_ARG; }
    c[0] = (int64_t)m->x.size();
    c[1] = m->n_real;
    c[2] = m->n_cells;
    c[3] = m->n_edges();
    c[4] = (int64_t)m->bnd_edge.size();
    c[5] = (int64_t)m->names.size();
    return FVH_OK;
}

int64_t fvh_mesh_export(const fvh_mesh* m, const char* field, void* dst, int64_t n) {
    if (!m || !field) { fvh::set_error("null argument"); return FVH_E_ARG; }
    const std::string f(field);
    const std::vector<double>* dv = nullptr;
    std::vector<int32_t> tmp;
    const std::vector<int32_t>* iv = nullptr;
    if (f == "nodesX") dv = &m->x;
    else if (f == "nodesY") dv = &m->y;
    else if (f == "edgesLengths") dv = &m->e_len;
    else if (f == "edgesNormalsX") dv = &m->e_nx;
    else if (f == "edgesNormalsY") dv = &m->e_ny;
    else if (f == "edgesCentersX") dv = &m->e_cx;
    else if (f == "edgesCentersY") dv = &m->e_cy;
    else if (f == "cellsAreas") dv = &m->c_area;
    else if (f == "cellsCentersX") dv = &m->c_cx;
    else if (f == "cellsCentersY") dv = &m->c_cy;
    else if (f == "edgesNodes") iv = &m->edge_nodes;
    else if (f == "edgesCells") iv = &m->edge_cells;
    else if (f == "cellsNodes") iv = &m->cell_nodes;
    else if (f == "boundaryEdges") iv = &m->bnd_edge;
    else if (f == "boundaryTags") iv = &m->bnd_tag;
    else if (f == "cellsIsTriangle") { tmp.assign(m->cell_tri.begin(), m->cell_tri.end()); iv = &tmp; }
    else if (f == "rcmPermutation" || f == "tilePermutation") {
        std::vector<int32_t> order = f == "rcmPermutation" ? fvh::rcm_order(*m) : fvh::rcb_order(*m, 256);
        tmp.assign(m->n_cells, 0);
        for (int32_t k = 0; k < (int32_t)order.size(); ++k) tmp[order[k]] = k;
        for (int32_t g = m->n_real; g < m->n_cells; ++g) tmp[g] = g;
        iv = &tmp;
    } else {
        fvh::set_error("unknown mesh field " + f);
        return FVH_E_ARG;
    }
    const int64_t cnt = dv ? (int64_t)dv->size() : (int64_t)iv->size();
    if (dst) {
        if (n < cnt) { fvh::set_error("destination too small for " + f); return FVH_E_ARG; }
        if (dv) std::memcpy(dst, dv->data(), sizeof(double) * cnt);
        else std::memcpy(dst, iv->data(), sizeof(int32_t) * cnt);
    }
    return cnt;
}

int32_t fvh_mesh_name(const fvh_mesh* m, int32_t idx, char* buf, int32_t len) {
    if (!m || !buf || len <= 0 || idx < 0 || idx >= (int32_t)m->names.size()) {
        fvh::set_error("bad name query");
        return FVH_E_ARG;
    }
    std::snprintf(buf, (size_t)len, "%s", m->names[idx].c_str());
    return FVH_OK;
}

}  // extern "C"
