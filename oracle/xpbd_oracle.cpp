// TEST INFRASTRUCTURE ONLY — CPU oracle for the XPBD cloth hot path.
//
// A literal restatement of the reference's CPU solver (src/xpbd_cpu.rs) and
// topology preparation (src/mesh_prep.rs), templated on float (parity target)
// and double (error reference).  Every function cites the reference lines it
// follows.  Arithmetic follows glam's scalar Vec3 operation order
// (dot = (x*x + y*y) + z*z, length = sqrt(dot), component-wise '/' and '*'),
// with no FMA contraction (built with -ffp-contract=off), because Rust never
// contracts a multiply-add unless mul_add is written.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library.  The product (bevy_xpbd_amd) never links it.
#include "xpbd_oracle.h"

#include <algorithm>
#include <array>
#include <cctype>
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Reference constants (cloth_compute.rs:82-90; mesh_prep.rs:89-101).
constexpr float JACOBI_CORRECTION_CAP = 0.28f;
constexpr float PREDICT_MAX_SPEED = 12.0f;
constexpr float GRAB_MAX_PULL_PER_SUBSTEP = 0.065f;
constexpr float COLLISION_APPLY_CLAMP = 0.35f;
constexpr float DEFAULT_BEND_COMPLIANCE = 3.8f;
constexpr float DEFAULT_STRETCH_COMPLIANCE = 2e-8f;
constexpr float DEFAULT_SKIP2_COMPLIANCE = 2.5e-8f;
constexpr float DEFAULT_CROSS_DIAG_COMPLIANCE = 6e-9f;
constexpr uint32_t COLLISION_GRID_MAX_CELLS = 32768;
constexpr uint32_t COLLISION_GRID_MAX_AXIS = 96;
constexpr uint32_t NONE_U32 = 0xFFFFFFFFu;

template <class T>
struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline T length_squared(V3<T> a) { return dot(a, a); }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
// f32::min / f32::max ignore a NaN operand (IEEE minNum/maxNum).
inline float fmin_rs(float a, float b) { return std::fmin(a, b); }
inline float fmax_rs(float a, float b) { return std::fmax(a, b); }

template <class T> inline V3<T> load3(const T* p, size_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
template <class T> inline void store3(T* p, size_t i, V3<T> v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}
template <class T> inline V3<T> load3f(const float* p, size_t i) {
    return {T(p[3 * i]), T(p[3 * i + 1]), T(p[3 * i + 2])};
}

float distf(V3<float> a, V3<float> b) { return length(a - b); }

// ---------------------------------------------------------------------------
// Topology (mesh_prep.rs)
// ---------------------------------------------------------------------------
struct Constraint {
    uint32_t i, j;
    float rest, comp;
};

struct OMesh {
    std::vector<float> positions;       // 3N (initial state)
    std::vector<float> rest_positions;  // 3N
    std::vector<float> normals;         // 3N
    std::vector<float> uvs;             // 2N
    std::vector<uint32_t> indices;
    std::vector<uint32_t> tri_edge_neighbors;
    std::vector<float> inv_mass;
    std::vector<uint32_t> neighbor_other;
    std::vector<float> neighbor_rest_len;
    std::vector<float> neighbor_compliance;
    std::vector<uint32_t> neighbor_constraint_id;
    std::vector<uint32_t> neighbor_offsets;
    std::vector<uint32_t> constraint_i, constraint_j;
    std::vector<float> constraint_rest, constraint_comp;
    std::vector<uint32_t> batch_offsets;
    std::vector<uint32_t> ground;
    std::vector<uint32_t> cloth_offsets;
};

V3<float> pos_at(const std::vector<float>& p, size_t i) { return load3(p.data(), i); }

// partition_constraints_for_gs_batches, mesh_prep.rs:153-221.
void partition_constraints_for_gs_batches(size_t n_particles, std::vector<Constraint>& constraints,
                                          std::vector<uint32_t>& offs) {
    const size_t e_count = constraints.size();
    if (e_count == 0) {
        offs.assign({0, 0});
        return;
    }
    std::vector<uint32_t> deg(n_particles, 0);
    for (const auto& c : constraints) {
        deg[c.i] += 1;
        deg[c.j] += 1;
    }
    std::vector<size_t> process_order(e_count);
    for (size_t e = 0; e < e_count; ++e) process_order[e] = e;
    // sort_by_key((Reverse(d), min(i,j), max(i,j))) — stable.
    std::stable_sort(process_order.begin(), process_order.end(), [&](size_t a, size_t b) {
        const auto& ca = constraints[a];
        const auto& cb = constraints[b];
        uint32_t da = std::max(deg[ca.i], deg[ca.j]);
        uint32_t db = std::max(deg[cb.i], deg[cb.j]);
        if (da != db) return da > db;
        uint32_t mina = std::min(ca.i, ca.j), minb = std::min(cb.i, cb.j);
        if (mina != minb) return mina < minb;
        return std::max(ca.i, ca.j) < std::max(cb.i, cb.j);
    });
    std::vector<uint32_t> color_of_edge(e_count, 0);
    std::vector<std::vector<uint32_t>> occupied(n_particles);
    std::set<uint32_t> forb;
    for (size_t e : process_order) {
        const uint32_t i = constraints[e].i, j = constraints[e].j;
        forb.clear();
        for (uint32_t c : occupied[i]) forb.insert(c);
        for (uint32_t c : occupied[j]) forb.insert(c);
        uint32_t c = 0;
        while (forb.count(c)) c += 1;
        color_of_edge[e] = c;
        occupied[i].push_back(c);
        occupied[j].push_back(c);
    }
    std::vector<size_t> order(e_count);
    for (size_t e = 0; e < e_count; ++e) order[e] = e;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        if (color_of_edge[a] != color_of_edge[b]) return color_of_edge[a] < color_of_edge[b];
        return a < b;
    });
    std::vector<Constraint> sorted(e_count);
    std::vector<uint32_t> sorted_colors(e_count);
    for (size_t k = 0; k < e_count; ++k) {
        sorted[k] = constraints[order[k]];
        sorted_colors[k] = color_of_edge[order[k]];
    }
    offs.assign({0});
    for (size_t k = 1; k < e_count; ++k)
        if (sorted_colors[k] != sorted_colors[k - 1]) offs.push_back(uint32_t(k));
    offs.push_back(uint32_t(e_count));
    constraints.swap(sorted);
}

// Colouring + CSR cursor scatter: finalize_cloth_mesh, mesh_prep.rs:584-635.
void assemble_constraints(OMesh& m, size_t n, std::vector<Constraint> constraints) {
    partition_constraints_for_gs_batches(n, constraints, m.batch_offsets);
    m.neighbor_offsets.assign(n + 1, 0);
    std::vector<uint32_t> counts(n, 0);
    for (const auto& c : constraints) {
        counts[c.i] += 1;
        counts[c.j] += 1;
    }
    for (size_t i = 0; i < n; ++i) m.neighbor_offsets[i + 1] = m.neighbor_offsets[i] + counts[i];
    const size_t nn = m.neighbor_offsets[n];
    m.neighbor_other.assign(nn, 0);
    m.neighbor_rest_len.assign(nn, 0.f);
    m.neighbor_compliance.assign(nn, 0.f);
    m.neighbor_constraint_id.assign(nn, 0);
    std::vector<uint32_t> cursors(m.neighbor_offsets.begin(), m.neighbor_offsets.begin() + n);
    m.constraint_i.clear();
    m.constraint_j.clear();
    m.constraint_rest.clear();
    m.constraint_comp.clear();
    for (size_t eid = 0; eid < constraints.size(); ++eid) {
        const auto& c = constraints[eid];
        m.constraint_i.push_back(c.i);
        m.constraint_j.push_back(c.j);
        m.constraint_rest.push_back(c.rest);
        m.constraint_comp.push_back(c.comp);
        uint32_t pi = cursors[c.i]++;
        m.neighbor_other[pi] = c.j;
        m.neighbor_rest_len[pi] = c.rest;
        m.neighbor_compliance[pi] = c.comp;
        m.neighbor_constraint_id[pi] = uint32_t(eid);
        uint32_t pj = cursors[c.j]++;
        m.neighbor_other[pj] = c.i;
        m.neighbor_rest_len[pj] = c.rest;
        m.neighbor_compliance[pj] = c.comp;
        m.neighbor_constraint_id[pj] = uint32_t(eid);
    }
}

// find_tri_neighbors, mesh_prep.rs:661-695.
std::vector<uint32_t> find_tri_neighbors(const std::vector<uint32_t>& indices) {
    const size_t num_tris = indices.size() / 3;
    struct Edge {
        uint32_t id0, id1, edge_nr;
    };
    std::vector<Edge> edges;
    edges.reserve(num_tris * 3);
    for (size_t i = 0; i < num_tris; ++i)
        for (uint32_t j = 0; j < 3; ++j) {
            uint32_t id0 = indices[3 * i + j];
            uint32_t id1 = indices[3 * i + (j + 1) % 3];
            edges.push_back({std::min(id0, id1), std::max(id0, id1), uint32_t(3 * i + j)});
        }
    std::stable_sort(edges.begin(), edges.end(), [](const Edge& a, const Edge& b) {
        if (a.id0 != b.id0) return a.id0 < b.id0;
        return a.id1 < b.id1;
    });
    std::vector<uint32_t> neighbors(3 * num_tris, NONE_U32);
    size_t i = 0;
    while (i + 1 < edges.size()) {
        const Edge& e0 = edges[i];
        const Edge& e1 = edges[i + 1];
        if (e0.id0 == e1.id0 && e0.id1 == e1.id1) {
            neighbors[e0.edge_nr] = e1.edge_nr;
            neighbors[e1.edge_nr] = e0.edge_nr;
        }
        i += 1;
    }
    return neighbors;
}

// unique_edges_and_rests, mesh_prep.rs:697-722.
std::vector<Constraint> unique_edges_and_rests(const std::vector<uint32_t>& indices,
                                               const std::vector<float>& positions,
                                               const std::vector<uint32_t>& neighbors) {
    const size_t num_tris = indices.size() / 3;
    std::set<std::pair<uint32_t, uint32_t>> seen;
    std::vector<Constraint> out;
    for (size_t i = 0; i < num_tris; ++i)
        for (uint32_t j = 0; j < 3; ++j) {
            uint32_t id0 = indices[3 * i + j];
            uint32_t id1 = indices[3 * i + (j + 1) % 3];
            uint32_t nb = neighbors[3 * i + j];
            if (nb != NONE_U32 && id0 >= id1) continue;
            uint32_t a = std::min(id0, id1), b = std::max(id0, id1);
            if (seen.insert({a, b}).second)
                out.push_back({a, b, distf(pos_at(positions, a), pos_at(positions, b)), 0.f});
        }
    return out;
}

// bending_pairs, mesh_prep.rs:724-747.
std::vector<Constraint> bending_pairs(const std::vector<uint32_t>& indices,
                                      const std::vector<uint32_t>& neighbors,
                                      const std::vector<float>& positions) {
    const size_t num_tris = indices.size() / 3;
    std::set<std::pair<uint32_t, uint32_t>> seen;
    std::vector<Constraint> out;
    for (size_t i = 0; i < num_tris; ++i)
        for (uint32_t j = 0; j < 3; ++j) {
            uint32_t nb = neighbors[3 * i + j];
            if (nb == NONE_U32) continue;
            uint32_t id2 = indices[3 * i + (j + 2) % 3];
            uint32_t ni = nb / 3, nj = nb % 3;
            uint32_t id3 = indices[3 * ni + (nj + 2) % 3];
            uint32_t a = std::min(id2, id3), b = std::max(id2, id3);
            if (seen.insert({a, b}).second)
                out.push_back({a, b, distf(pos_at(positions, a), pos_at(positions, b)), 0.f});
        }
    return out;
}

// finalize_cloth_mesh, mesh_prep.rs:532-659.
OMesh* finalize_cloth_mesh(std::vector<float> positions, std::vector<float> normals,
                           std::vector<float> uvs, std::vector<uint32_t> indices,
                           const std::vector<Constraint>& extras) {
    OMesh* m = new OMesh();
    const size_t n = positions.size() / 3;
    const size_t num_tris = indices.size() / 3;
    m->inv_mass.assign(n, 0.f);
    for (size_t t = 0; t < num_tris; ++t) {
        uint32_t id0 = indices[3 * t], id1 = indices[3 * t + 1], id2 = indices[3 * t + 2];
        V3<float> e0 = pos_at(positions, id1) - pos_at(positions, id0);
        V3<float> e1 = pos_at(positions, id2) - pos_at(positions, id0);
        V3<float> c = cross(e0, e1);
        float a = 0.5f * length(c);
        float p_inv = a > 0.f ? 1.0f / a / 3.0f : 0.f;
        m->inv_mass[id0] += p_inv;
        m->inv_mass[id1] += p_inv;
        m->inv_mass[id2] += p_inv;
    }
    m->tri_edge_neighbors = find_tri_neighbors(indices);
    float max_y = -INFINITY;
    for (size_t i = 0; i < n; ++i) max_y = fmax_rs(max_y, positions[3 * i + 1]);
    const float eps = 1e-5f;
    for (size_t i = 0; i < n; ++i)
        if (positions[3 * i + 1] >= max_y - eps) m->inv_mass[i] = 0.f;
    m->rest_positions = positions;
    auto edge_pairs = unique_edges_and_rests(indices, positions, m->tri_edge_neighbors);
    auto bend = bending_pairs(indices, m->tri_edge_neighbors, positions);
    std::vector<Constraint> constraints;
    for (auto c : edge_pairs) {
        c.comp = DEFAULT_STRETCH_COMPLIANCE;
        constraints.push_back(c);
    }
    for (auto c : bend) {
        c.comp = DEFAULT_BEND_COMPLIANCE;
        constraints.push_back(c);
    }
    constraints.insert(constraints.end(), extras.begin(), extras.end());
    assemble_constraints(*m, n, std::move(constraints));
    m->positions = std::move(positions);
    m->normals = std::move(normals);
    m->uvs = std::move(uvs);
    m->indices = std::move(indices);
    m->ground.assign(n, 0);
    m->cloth_offsets = {0, uint32_t(n)};
    return m;
}

// dedup_sorted_coords, mesh_prep.rs:429-439.
void dedup_sorted_coords(std::vector<float>& values, float eps) {
    std::stable_sort(values.begin(), values.end(), [](float a, float b) { return a < b; });
    size_t write = 0;
    for (size_t read = 0; read < values.size(); ++read) {
        if (write == 0 || std::fabs(values[read] - values[write - 1]) > eps) {
            values[write] = values[read];
            write += 1;
        }
    }
    values.resize(write);
}

// axis_aligned_grid_skip_two_distance_constraints, mesh_prep.rs:444-501.
bool axis_aligned_grid_skip_two(const std::vector<float>& positions, std::vector<Constraint>& out) {
    const float EPS = 2e-4f;
    const size_t n = positions.size() / 3;
    if (n == 0) return false;
    std::vector<float> xs(n), ys(n);
    for (size_t i = 0; i < n; ++i) {
        xs[i] = positions[3 * i];
        ys[i] = positions[3 * i + 1];
    }
    dedup_sorted_coords(xs, EPS);
    dedup_sorted_coords(ys, EPS);
    const size_t cols = xs.size(), rows = ys.size();
    if (cols < 3 || rows < 3 || cols * rows != n) return false;
    std::vector<int64_t> grid_pid(cols * rows, -1);
    for (size_t pid = 0; pid < n; ++pid) {
        float px = positions[3 * pid], py = positions[3 * pid + 1];
        size_t ix = SIZE_MAX, iy = SIZE_MAX;
        for (size_t k = 0; k < cols; ++k)
            if (std::fabs(xs[k] - px) <= EPS) {
                ix = k;
                break;
            }
        if (ix == SIZE_MAX) return false;
        for (size_t k = 0; k < rows; ++k)
            if (std::fabs(ys[k] - py) <= EPS) {
                iy = k;
                break;
            }
        if (iy == SIZE_MAX) return false;
        size_t slot = iy * cols + ix;
        if (grid_pid[slot] >= 0) return false;
        grid_pid[slot] = int64_t(pid);
    }
    for (auto s : grid_pid)
        if (s < 0) return false;
    auto pid_at = [&](size_t ix, size_t iy) { return uint32_t(grid_pid[iy * cols + ix]); };
    for (size_t iy = 0; iy < rows; ++iy)
        for (size_t ix = 0; ix + 2 < cols; ++ix) {
            uint32_t i = pid_at(ix, iy), j = pid_at(ix + 2, iy);
            float rest = distf(pos_at(positions, i), pos_at(positions, j));
            out.push_back({std::min(i, j), std::max(i, j), rest, DEFAULT_SKIP2_COMPLIANCE});
        }
    for (size_t iy = 0; iy + 2 < rows; ++iy)
        for (size_t ix = 0; ix < cols; ++ix) {
            uint32_t i = pid_at(ix, iy), j = pid_at(ix, iy + 2);
            float rest = distf(pos_at(positions, i), pos_at(positions, j));
            out.push_back({std::min(i, j), std::max(i, j), rest, DEFAULT_SKIP2_COMPLIANCE});
        }
    return true;
}

// Strict Rust-style parsers (str::parse::<f32> / ::<usize>): whole token or fail.
bool parse_f32(const std::string& s, float& out) {
    if (s.empty()) return false;
    char* end = nullptr;
    errno = 0;
    float v = std::strtof(s.c_str(), &end);
    if (end != s.c_str() + s.size()) return false;
    // Rust rejects hex floats and "infinity"-style spellings it does not know;
    // OBJ files here hold plain decimals.
    out = v;
    return true;
}
bool parse_usize(const std::string& s, size_t& out) {
    if (s.empty()) return false;
    size_t k = 0;
    if (s[0] == '+') k = 1;
    if (k == s.size()) return false;
    size_t v = 0;
    for (; k < s.size(); ++k) {
        if (s[k] < '0' || s[k] > '9') return false;
        v = v * 10 + size_t(s[k] - '0');
    }
    out = v;
    return true;
}
std::vector<std::string> split_ws(const std::string& line) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && std::isspace((unsigned char)line[i])) ++i;
        size_t s = i;
        while (i < line.size() && !std::isspace((unsigned char)line[i])) ++i;
        if (i > s) out.push_back(line.substr(s, i - s));
    }
    return out;
}

// parse_welded_obj, mesh_prep.rs:223-357.
OMesh* parse_welded_obj(const std::string& text) {
    std::vector<std::array<float, 3>> verts_pos, verts_nrm;
    std::vector<std::array<float, 2>> verts_uv;
    std::vector<std::vector<std::string>> faces;
    size_t start = 0;
    while (start <= text.size()) {
        size_t nl = text.find('\n', start);
        if (nl == std::string::npos) nl = text.size();
        std::string line = text.substr(start, nl - start);
        start = nl + 1;
        auto parts = split_ws(line);
        if (parts.empty() || parts[0][0] == '#') continue;
        const std::string& head = parts[0];
        auto getf = [&](size_t k) {
            float v = 0.f;
            if (k < parts.size() && parse_f32(parts[k], v)) return v;
            return 0.f;
        };
        if (head == "v") verts_pos.push_back({getf(1), getf(2), getf(3)});
        else if (head == "vt") verts_uv.push_back({getf(1), getf(2)});
        else if (head == "vn") verts_nrm.push_back({getf(1), getf(2), getf(3)});
        else if (head == "f") faces.emplace_back(parts.begin() + 1, parts.end());
    }
    std::unordered_map<std::string, uint32_t> cache;
    std::vector<float> positions, normals, uvs;
    std::vector<uint32_t> indices;
    std::vector<std::pair<uint32_t, uint32_t>> obj_quad_cross;
    auto corner_index = [&](const std::string& fs) -> uint32_t {
        auto it = cache.find(fs);
        if (it != cache.end()) return it->second;
        uint32_t idx = uint32_t(positions.size() / 3);
        cache.emplace(fs, idx);
        std::vector<std::string> parts;
        size_t s = 0;
        while (true) {
            size_t p = fs.find('/', s);
            if (p == std::string::npos) {
                parts.push_back(fs.substr(s));
                break;
            }
            parts.push_back(fs.substr(s, p - s));
            s = p + 1;
        }
        size_t vi = 1;
        if (!(parts.size() > 0 && parse_usize(parts[0], vi))) vi = 1;
        vi = vi > 0 ? vi - 1 : 0;
        bool has_uv = false, has_n = false;
        size_t uvi = 0, ni = 0;
        if (parts.size() > 1 && !parts[1].empty() && parse_usize(parts[1], uvi)) {
            has_uv = true;
            uvi = uvi > 0 ? uvi - 1 : 0;
        }
        if (parts.size() > 2 && parse_usize(parts[2], ni)) {
            has_n = true;
            ni = ni > 0 ? ni - 1 : 0;
        }
        std::array<float, 3> p = vi < verts_pos.size() ? verts_pos[vi] : std::array<float, 3>{0.f, 0.f, 0.f};
        positions.insert(positions.end(), p.begin(), p.end());
        if (has_uv) {
            std::array<float, 2> uv = uvi < verts_uv.size() ? verts_uv[uvi] : std::array<float, 2>{0.f, 0.f};
            uvs.insert(uvs.end(), uv.begin(), uv.end());
        } else {
            uvs.insert(uvs.end(), {0.f, 0.f});
        }
        if (has_n) {
            std::array<float, 3> nv = ni < verts_nrm.size() ? verts_nrm[ni] : std::array<float, 3>{0.f, 1.f, 0.f};
            normals.insert(normals.end(), nv.begin(), nv.end());
        } else {
            normals.insert(normals.end(), {0.f, 1.f, 0.f});
        }
        return idx;
    };
    for (const auto& face : faces) {
        if (face.size() < 3) continue;
        if (face.size() == 4) {
            uint32_t v0 = corner_index(face[0]);
            uint32_t v1 = corner_index(face[1]);
            uint32_t v2 = corner_index(face[2]);
            uint32_t v3 = corner_index(face[3]);
            indices.insert(indices.end(), {v0, v1, v2, v0, v2, v3});
            obj_quad_cross.push_back({std::min(v1, v3), std::max(v1, v3)});
            continue;
        }
        uint32_t c0 = corner_index(face[0]);
        for (size_t i = 1; i + 1 < face.size(); ++i) {
            indices.push_back(c0);
            uint32_t a = corner_index(face[i]);
            indices.push_back(a);
            uint32_t b = corner_index(face[i + 1]);
            indices.push_back(b);
        }
    }
    std::vector<Constraint> extras;
    axis_aligned_grid_skip_two(positions, extras);  // unwrap_or_default
    for (auto [ia, ib] : obj_quad_cross)
        extras.push_back({ia, ib, distf(pos_at(positions, ia), pos_at(positions, ib)),
                          DEFAULT_CROSS_DIAG_COMPLIANCE});
    return finalize_cloth_mesh(std::move(positions), std::move(normals), std::move(uvs),
                               std::move(indices), extras);
}

// grid_cloth_hanging, mesh_prep.rs:363-406 (+ skip-2 :504-530, cross diagonals :409-426).
OMesh* grid_cloth_hanging(uint32_t quad_cols, uint32_t quad_rows, float cell_size) {
    const uint32_t vtx_cols = quad_cols + 1, vtx_rows = quad_rows + 1;
    std::vector<float> positions, normals, uvs;
    const float width = float(quad_cols) * cell_size;
    const float x0 = -0.5f * width;
    for (uint32_t iy = 0; iy < vtx_rows; ++iy)
        for (uint32_t ix = 0; ix < vtx_cols; ++ix) {
            float x = x0 + float(ix) * cell_size;
            float y = float(iy) * cell_size;
            positions.insert(positions.end(), {x, y, 0.f});
            normals.insert(normals.end(), {0.f, 0.f, 1.f});
            uvs.insert(uvs.end(), {float(ix) / float(quad_cols), float(iy) / float(quad_rows)});
        }
    std::vector<uint32_t> indices;
    const size_t cols = vtx_cols;
    for (size_t iy = 0; iy < quad_rows; ++iy)
        for (size_t ix = 0; ix < quad_cols; ++ix) {
            uint32_t i0 = uint32_t(iy * cols + ix), i1 = i0 + 1, i2 = uint32_t(i0 + cols), i3 = i2 + 1;
            indices.insert(indices.end(), {i0, i1, i2, i1, i3, i2});
        }
    std::vector<Constraint> extras;
    const size_t rows = vtx_rows;
    for (size_t iy = 0; iy < rows; ++iy)
        for (size_t ix = 0; ix + 2 < cols; ++ix) {
            uint32_t i = uint32_t(iy * cols + ix), j = i + 2;
            extras.push_back({i, j, distf(pos_at(positions, i), pos_at(positions, j)), DEFAULT_SKIP2_COMPLIANCE});
        }
    for (size_t iy = 0; iy + 2 < rows; ++iy)
        for (size_t ix = 0; ix < cols; ++ix) {
            uint32_t i = uint32_t(iy * cols + ix), j = uint32_t((iy + 2) * cols + ix);
            extras.push_back({i, j, distf(pos_at(positions, i), pos_at(positions, j)), DEFAULT_SKIP2_COMPLIANCE});
        }
    for (size_t iy = 0; iy < quad_rows; ++iy)
        for (size_t ix = 0; ix < quad_cols; ++ix) {
            size_t i0 = iy * cols + ix, i3 = (iy + 1) * cols + ix + 1;
            extras.push_back({uint32_t(std::min(i0, i3)), uint32_t(std::max(i0, i3)),
                              distf(pos_at(positions, i0), pos_at(positions, i3)),
                              DEFAULT_CROSS_DIAG_COMPLIANCE});
        }
    return finalize_cloth_mesh(std::move(positions), std::move(normals), std::move(uvs),
                               std::move(indices), extras);
}

// SplitMix64 (Steele, Lea, Flood 2014): the seeded RNG of the benchmark builder.
struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float unit() { return float(uint32_t(next() >> 32) >> 8) * (1.0f / 16777216.0f); }
};

// Benchmark sheet: axial stretch, both diagonals as shear, distance-2 axial
// bend; uniform inverse mass; explicit pins; seeded out-of-plane jitter.
// No reference counterpart (the configurations are new); the constraint
// assembly afterwards follows finalize_cloth_mesh's colouring + CSR order.
void build_sheet(const om_sheet_desc& d, std::vector<float>& rest, std::vector<float>& init,
                 std::vector<float>& normals, std::vector<float>& uvs, std::vector<uint32_t>& indices,
                 std::vector<float>& inv_mass, std::vector<Constraint>& cons) {
    const uint32_t C = d.cols, R = d.rows;
    const float h = d.spacing;
    const float x0 = -0.5f * (float(C - 1) * h);
    const float z0 = -0.5f * (float(R - 1) * h);
    const float ca = std::cos(d.angle), sa = std::sin(d.angle);
    const size_t n = size_t(C) * R;
    rest.resize(3 * n);
    normals.resize(3 * n);
    uvs.resize(2 * n);
    for (uint32_t iy = 0; iy < R; ++iy)
        for (uint32_t ix = 0; ix < C; ++ix) {
            size_t i = size_t(iy) * C + ix;
            float a = x0 + float(ix) * h;
            if (d.layout == OM_LAYOUT_HANGING_XY) {
                rest[3 * i] = a;
                rest[3 * i + 1] = d.height + float(iy) * h;
                rest[3 * i + 2] = 0.f;
                normals[3 * i] = 0.f;
                normals[3 * i + 1] = 0.f;
                normals[3 * i + 2] = 1.f;
            } else {
                float b = z0 + float(iy) * h;
                rest[3 * i] = a * ca - b * sa;
                rest[3 * i + 1] = d.height;
                rest[3 * i + 2] = a * sa + b * ca;
                normals[3 * i] = 0.f;
                normals[3 * i + 1] = 1.f;
                normals[3 * i + 2] = 0.f;
            }
            uvs[2 * i] = float(ix) / float(C - 1);
            uvs[2 * i + 1] = float(iy) / float(R - 1);
        }
    init = rest;
    if (d.jitter > 0.f) {
        SplitMix64 rng{d.seed};
        const int axis = d.layout == OM_LAYOUT_HANGING_XY ? 2 : 1;
        for (size_t i = 0; i < n; ++i) init[3 * i + axis] = init[3 * i + axis] + (rng.unit() * 2.0f - 1.0f) * d.jitter;
    }
    indices.clear();
    for (uint32_t iy = 0; iy + 1 < R; ++iy)
        for (uint32_t ix = 0; ix + 1 < C; ++ix) {
            uint32_t i0 = iy * C + ix, i1 = i0 + 1, i2 = i0 + C, i3 = i2 + 1;
            indices.insert(indices.end(), {i0, i1, i2, i1, i3, i2});
        }
    inv_mass.assign(n, d.inv_mass);
    auto pin = [&](uint32_t ix, uint32_t iy) { inv_mass[size_t(iy) * C + ix] = 0.f; };
    if (d.pins == OM_PIN_TWO_CORNERS) {
        pin(0, R - 1);
        pin(C - 1, R - 1);
    } else if (d.pins == OM_PIN_TOP_ROW) {
        for (uint32_t ix = 0; ix < C; ++ix) pin(ix, R - 1);
    } else if (d.pins == OM_PIN_CENTER) {
        pin(C / 2, R / 2);
    }
    cons.clear();
    auto add = [&](uint32_t i, uint32_t j, float comp) {
        uint32_t a = std::min(i, j), b = std::max(i, j);
        cons.push_back({a, b, distf(pos_at(rest, a), pos_at(rest, b)), comp});
    };
    for (uint32_t iy = 0; iy < R; ++iy)
        for (uint32_t ix = 0; ix + 1 < C; ++ix) add(iy * C + ix, iy * C + ix + 1, d.stretch_compliance);
    for (uint32_t iy = 0; iy + 1 < R; ++iy)
        for (uint32_t ix = 0; ix < C; ++ix) add(iy * C + ix, (iy + 1) * C + ix, d.stretch_compliance);
    for (uint32_t iy = 0; iy + 1 < R; ++iy)
        for (uint32_t ix = 0; ix + 1 < C; ++ix) {
            uint32_t i0 = iy * C + ix;
            add(i0, i0 + C + 1, d.shear_compliance);
            add(i0 + 1, i0 + C, d.shear_compliance);
        }
    for (uint32_t iy = 0; iy < R; ++iy)
        for (uint32_t ix = 0; ix + 2 < C; ++ix) add(iy * C + ix, iy * C + ix + 2, d.bend_compliance);
    for (uint32_t iy = 0; iy + 2 < R; ++iy)
        for (uint32_t ix = 0; ix < C; ++ix) add(iy * C + ix, (iy + 2) * C + ix, d.bend_compliance);
}

OMesh* make_batch(const om_sheet_desc* sheets, uint32_t count) {
    OMesh* m = new OMesh();
    std::vector<Constraint> all;
    m->cloth_offsets.push_back(0);
    for (uint32_t c = 0; c < count; ++c) {
        std::vector<float> rest, init, nrm, uv, im;
        std::vector<uint32_t> idx;
        std::vector<Constraint> cons;
        build_sheet(sheets[c], rest, init, nrm, uv, idx, im, cons);
        const uint32_t base = uint32_t(m->inv_mass.size());
        for (auto& k : cons) {
            k.i += base;
            k.j += base;
        }
        all.insert(all.end(), cons.begin(), cons.end());
        for (auto& t : idx) t += base;
        m->indices.insert(m->indices.end(), idx.begin(), idx.end());
        m->positions.insert(m->positions.end(), init.begin(), init.end());
        m->rest_positions.insert(m->rest_positions.end(), rest.begin(), rest.end());
        m->normals.insert(m->normals.end(), nrm.begin(), nrm.end());
        m->uvs.insert(m->uvs.end(), uv.begin(), uv.end());
        m->inv_mass.insert(m->inv_mass.end(), im.begin(), im.end());
        m->ground.insert(m->ground.end(), im.size(), sheets[c].ground ? 1u : 0u);
        m->cloth_offsets.push_back(uint32_t(m->inv_mass.size()));
    }
    assemble_constraints(*m, m->inv_mass.size(), std::move(all));
    return m;
}

// ---------------------------------------------------------------------------
// Solver (xpbd_cpu.rs)
// ---------------------------------------------------------------------------
template <class T>
struct P {
    T dt, inv_dt, omega;
    uint32_t iters;
    V3<T> gravity;
    int32_t grab_idx;
    bool grab_active;
    V3<T> grab_target;
    T grab_stiffness, drag, thickness, coll_scale;
    bool collision;
    uint32_t solver, coll_mode;
};
template <class T>
P<T> promote(const om_params* p) {
    P<T> q;
    q.dt = T(p->dt);
    q.inv_dt = T(p->inv_dt);
    q.omega = T(p->jacobi_omega);
    q.iters = p->inner_iterations;
    q.gravity = v3<T>(T(p->gravity[0]), T(p->gravity[1]), T(p->gravity[2]));
    q.grab_idx = p->grab_idx;
    q.grab_active = p->grab_active != 0;
    q.grab_target = v3<T>(T(p->grab_target[0]), T(p->grab_target[1]), T(p->grab_target[2]));
    q.grab_stiffness = T(p->grab_stiffness);
    q.drag = T(p->linear_drag_per_sec);
    q.thickness = T(p->thickness);
    q.coll_scale = T(p->coll_scale);
    q.collision = p->collision != 0;
    q.solver = p->solver;
    q.coll_mode = p->coll_mode;
    return q;
}

// clamp_correction, xpbd_cpu.rs:28-36.
template <class T>
V3<T> clamp_correction(V3<T> dx) {
    T ml = length(dx);
    T cap = T(JACOBI_CORRECTION_CAP);
    if (ml > cap && ml > T(0)) return dx * (cap / ml);
    return dx;
}

// predict_particle, xpbd_cpu.rs:38-65.
template <class T>
void predict_particle(size_t i, V3<T> sim, V3<T> vel, T inv_mass, const P<T>& p, V3<T>& pos_out,
                      V3<T>& v_out) {
    if (inv_mass <= T(0)) {
        pos_out = sim;
        v_out = vel;
        return;
    }
    V3<T> v = vel;
    v = v + p.gravity * p.dt;
    T speed = length(v);
    T max_v = T(PREDICT_MAX_SPEED);
    if (speed > max_v && speed > T(0)) v = v * (max_v / speed);
    V3<T> pos = sim + v * p.dt;
    if (p.grab_active && int32_t(i) == p.grab_idx) {
        V3<T> pull = (p.grab_target - pos) * p.grab_stiffness;
        T pl = length(pull);
        T gm = T(GRAB_MAX_PULL_PER_SUBSTEP);
        if (pl > gm && pl > T(0)) pull = pull * (gm / pl);
        pos = pos + pull;
    }
    pos_out = pos;
    v_out = v;
}

// self_collision_resolve, xpbd_cpu.rs:67-137: the pair kernel for (lo, hi), lo < hi.
template <class T>
inline bool collision_pair(const T* pos, const om_solver_in* in, size_t i, size_t j, T thickness,
                           T coll_scale, V3<T>& di, V3<T>& dj) {
    const T w_i = T(in->inv_mass[i]), w_j = T(in->inv_mass[j]);
    if (w_i <= T(0) && w_j <= T(0)) return false;
    const T thickness_sq = thickness * thickness;
    V3<T> p_i = load3(pos, i), p_j = load3(pos, j);
    V3<T> d = p_j - p_i;
    T dist2 = length_squared(d);
    if (dist2 > thickness_sq || dist2 < T(1e-18)) return false;
    V3<T> r0 = load3f<T>(in->rest_pos, i), r1 = load3f<T>(in->rest_pos, j);
    V3<T> rest_d = r1 - r0;
    T rest2 = length_squared(rest_d);
    if (dist2 > rest2) return false;
    T min_d = thickness;
    if (rest2 < thickness_sq) min_d = std::sqrt(rest2);
    T dist = std::sqrt(dist2);
    T corr = (min_d - dist) * T(0.5) * coll_scale;
    if (corr <= T(0)) return false;
    d = (d / dist) * corr;
    T inv_w = T(1.0) / std::max(w_i + w_j, T(1e-8));
    di = -d * w_i * inv_w;
    dj = d * w_j * inv_w;
    return true;
}

template <class T>
void apply_collision_accum(T* pos, const om_solver_in* in, const std::vector<V3<T>>& accum) {
    const T max_d = T(COLLISION_APPLY_CLAMP);
    for (size_t i = 0; i < in->n; ++i) {
        if (T(in->inv_mass[i]) <= T(0)) continue;
        V3<T> c = accum[i];
        auto cl = [&](T x) { return x < -max_d ? -max_d : (x > max_d ? max_d : x); };
        store3(pos, i, load3(pos, i) + v3<T>(cl(c.x), cl(c.y), cl(c.z)));
    }
}

// Cloth c's particle range; self-collision pairs never cross cloths (each
// cloth is its own self_collision_resolve call in the reference's terms).
inline void cloth_range(const om_solver_in* in, uint32_t c, size_t& b, size_t& e) {
    if (in->num_cloths <= 1 || !in->cloth_offsets) {
        b = 0;
        e = in->n;
    } else {
        b = in->cloth_offsets[c];
        e = in->cloth_offsets[c + 1];
    }
}
inline uint32_t cloth_count(const om_solver_in* in) {
    return (in->num_cloths <= 1 || !in->cloth_offsets) ? 1u : in->num_cloths;
}

template <class T>
void self_collision_brute(const om_solver_in* in, T* pos, T thickness, T coll_scale) {
    const size_t n = in->n;
    std::vector<V3<T>> accum(n, v3<T>(0, 0, 0));
    for (uint32_t c = 0; c < cloth_count(in); ++c) {
        size_t b, e;
        cloth_range(in, c, b, e);
        for (size_t i = b; i < e; ++i)
            for (size_t j = i + 1; j < e; ++j) {
                V3<T> di, dj;
                if (!collision_pair(pos, in, i, j, thickness, coll_scale, di, dj)) continue;
                if (T(in->inv_mass[i]) > T(0)) accum[i] = accum[i] + di;
                if (T(in->inv_mass[j]) > T(0)) accum[j] = accum[j] + dj;
            }
    }
    apply_collision_accum(pos, in, accum);
}

// Hash-grid broad phase with the brute-force accumulation order: each
// particle sums its contributions over partners in ascending id, which is the
// order the O(N^2) loop above adds them.  Cells are 2x the contact distance so
// every pair within `thickness` lies in the 27-cell neighbourhood.
template <class T>
void self_collision_grid_range(const om_solver_in* in, const T* pos, size_t b, size_t e, T thickness,
                               T coll_scale, std::vector<V3<T>>& accum) {
    const size_t m = e - b;
    if (m == 0) return;
    double lo[3] = {INFINITY, INFINITY, INFINITY};
    for (size_t i = b; i < e; ++i)
        for (int a = 0; a < 3; ++a) lo[a] = std::fmin(lo[a], double(pos[3 * i + a]));
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(lo[a])) lo[a] = 0.0;
    const double cell = 2.0 * double(thickness) > 0.0 ? 2.0 * double(thickness) : 1.0;
    auto cell_of = [&](size_t i, int64_t c[3]) {
        for (int a = 0; a < 3; ++a) {
            double q = std::floor((double(pos[3 * i + a]) - lo[a]) / cell);
            if (!std::isfinite(q)) q = 0.0;
            c[a] = int64_t(std::min(std::max(q, 0.0), 2097151.0));
        }
    };
    auto key_of = [](const int64_t c[3]) { return (uint64_t(c[2]) << 42) | (uint64_t(c[1]) << 21) | uint64_t(c[0]); };
    std::vector<std::pair<uint64_t, uint32_t>> keyed(m);
    for (size_t i = b; i < e; ++i) {
        int64_t c[3];
        cell_of(i, c);
        keyed[i - b] = {key_of(c), uint32_t(i)};
    }
    std::sort(keyed.begin(), keyed.end());
    std::unordered_map<uint64_t, std::pair<uint32_t, uint32_t>> ranges;
    ranges.reserve(m * 2);
    for (size_t k = 0; k < m;) {
        size_t q = k;
        while (q < m && keyed[q].first == keyed[k].first) ++q;
        ranges[keyed[k].first] = {uint32_t(k), uint32_t(q)};
        k = q;
    }
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t ii = int64_t(b); ii < int64_t(e); ++ii) {
        const size_t i = size_t(ii);
        if (T(in->inv_mass[i]) <= T(0)) continue;
        int64_t c[3];
        cell_of(i, c);
        std::vector<uint32_t> cand;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int64_t nc[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
                    if (nc[0] < 0 || nc[1] < 0 || nc[2] < 0) continue;
                    auto it = ranges.find(key_of(nc));
                    if (it == ranges.end()) continue;
                    for (uint32_t k = it->second.first; k < it->second.second; ++k)
                        if (keyed[k].second != i) cand.push_back(keyed[k].second);
                }
        std::sort(cand.begin(), cand.end());
        V3<T> acc = v3<T>(0, 0, 0);
        for (uint32_t j : cand) {
            V3<T> dlo, dhi;
            size_t a = std::min<size_t>(i, j), bb = std::max<size_t>(i, j);
            if (!collision_pair(pos, in, a, bb, thickness, coll_scale, dlo, dhi)) continue;
            acc = acc + (i == a ? dlo : dhi);
        }
        accum[i] = acc;
    }
}

template <class T>
void self_collision_grid(const om_solver_in* in, T* pos, T thickness, T coll_scale) {
    std::vector<V3<T>> accum(in->n, v3<T>(0, 0, 0));
    for (uint32_t c = 0; c < cloth_count(in); ++c) {
        size_t b, e;
        cloth_range(in, c, b, e);
        self_collision_grid_range(in, pos, b, e, thickness, coll_scale, accum);
    }
    apply_collision_accum(pos, in, accum);
}

template <class T>
void self_collision(const om_solver_in* in, T* pos, T thickness, T coll_scale, uint32_t mode) {
    if (mode == OM_COLL_BRUTE) self_collision_brute(in, pos, thickness, coll_scale);
    else self_collision_grid(in, pos, thickness, coll_scale);
}

// jacobi_edges_pass, xpbd_cpu.rs:140-184.
template <class T>
void jacobi_edges_pass(const om_solver_in* in, const T* jac_in, T* lambda, T* delta_lambda, T inv_dt_sq) {
#pragma omp parallel for schedule(static)
    for (int64_t ee = 0; ee < int64_t(in->e); ++ee) {
        const size_t e = size_t(ee);
        const size_t i = in->ci[e], j = in->cj[e];
        const T w_i = T(in->inv_mass[i]), w_j = T(in->inv_mass[j]);
        if (w_j <= T(0) && w_i <= T(0)) {
            delta_lambda[e] = T(0);
            continue;
        }
        const T rest = T(in->crest[e]);
        const T compliance = T(in->ccomp[e]);
        V3<T> gv = load3(jac_in, i) - load3(jac_in, j);
        T len = length(gv);
        if (len < T(1e-8)) {
            delta_lambda[e] = T(0);
            continue;
        }
        gv = gv / len;
        T c = len - rest;
        T alpha_t = compliance * inv_dt_sq;
        T wsum = w_i + w_j + alpha_t;
        if (wsum < T(1e-8)) {
            delta_lambda[e] = T(0);
            continue;
        }
        T lam = lambda[e];
        T dlam = (-c - alpha_t * lam) / wsum;
        delta_lambda[e] = dlam;
        lambda[e] = lam + dlam;
    }
}

// jacobi_gather_particle, xpbd_cpu.rs:187-226.
template <class T>
V3<T> jacobi_gather_particle(const om_solver_in* in, size_t i, const T* jac_in, const T* delta_lambda, T omega) {
    const T w_i = T(in->inv_mass[i]);
    if (w_i <= T(0)) return load3(jac_in, i);
    V3<T> p_i = load3(jac_in, i);
    V3<T> acc = v3<T>(0, 0, 0);
    for (uint32_t k = in->noffs[i]; k < in->noffs[i + 1]; ++k) {
        const size_t j = in->nother[k];
        const T w_j = T(in->inv_mass[j]);
        if (w_j <= T(0) && w_i <= T(0)) continue;
        V3<T> gv = p_i - load3(jac_in, j);
        T len = length(gv);
        if (len < T(1e-8)) continue;
        gv = gv / len;
        T dlam = delta_lambda[in->neid[k]];
        acc = acc + gv * w_i * dlam;
    }
    V3<T> delta = omega * acc;
    delta = clamp_correction(delta);
    return p_i + delta;
}

template <class T>
void predict_all(const om_solver_in* in, T* sim, T* prev, T* vel, const P<T>& p) {
    for (size_t i = 0; i < in->n; ++i) {
        const T w = T(in->inv_mass[i]);
        if (w <= T(0)) {
            store3(prev, i, load3(sim, i));
            continue;
        }
        V3<T> np, nv;
        predict_particle(i, load3(sim, i), load3(vel, i), w, p, np, nv);
        store3(prev, i, load3(sim, i));
        store3(sim, i, np);
        store3(vel, i, nv);
    }
}

// xpbd_substep_integrate (Jacobi), xpbd_cpu.rs:229-310.
template <class T>
void integrate_jacobi(const om_solver_in* in, T* sim, T* prev, T* vel, const P<T>& p) {
    const size_t n = in->n;
    std::vector<T> lambda(in->e), delta_lambda(in->e);
    std::vector<T> jac_a(3 * n), jac_b(3 * n);
    const T inv_dt_sq = p.inv_dt * p.inv_dt;
    predict_all(in, sim, prev, vel, p);
    std::copy(sim, sim + 3 * n, jac_b.begin());
    for (uint32_t k = 0; k < p.iters; ++k) {
        std::fill(lambda.begin(), lambda.end(), T(0));
        std::fill(delta_lambda.begin(), delta_lambda.end(), T(0));
        const T* read = (k % 2 == 0) ? jac_b.data() : jac_a.data();
        T* write = (k % 2 == 0) ? jac_a.data() : jac_b.data();
        jacobi_edges_pass(in, read, lambda.data(), delta_lambda.data(), inv_dt_sq);
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < int64_t(n); ++ii)
            store3(write, size_t(ii), jacobi_gather_particle(in, size_t(ii), read, delta_lambda.data(), p.omega));
    }
    const T* final_jac = (p.iters % 2 == 0) ? jac_b.data() : jac_a.data();
    std::copy(final_jac, final_jac + 3 * n, sim);
}

// xpbd_substep_integrate (Gauss-Seidel), xpbd_cpu.rs:312-391.
template <class T>
void integrate_gs(const om_solver_in* in, T* sim, T* prev, T* vel, const P<T>& p) {
    const size_t n = in->n;
    std::vector<T> lambda(in->e);
    std::vector<T> jac_work(3 * n);
    predict_all(in, sim, prev, vel, p);
    std::copy(sim, sim + 3 * n, jac_work.begin());
    for (uint32_t kk = 0; kk < p.iters; ++kk) {
        std::fill(lambda.begin(), lambda.end(), T(0));
        const T omega = p.omega;
        const T inv_dt_sq = p.inv_dt * p.inv_dt;
        for (size_t e = 0; e < in->e; ++e) {
            const size_t i = in->ci[e], j = in->cj[e];
            const T w_i = T(in->inv_mass[i]), w_j = T(in->inv_mass[j]);
            if (w_j <= T(0) && w_i <= T(0)) continue;
            const T rest = T(in->crest[e]), compliance = T(in->ccomp[e]);
            V3<T> p_i = load3(jac_work.data(), i), p_j = load3(jac_work.data(), j);
            V3<T> gv = p_i - p_j;
            T len = length(gv);
            if (len < T(1e-8)) continue;
            gv = gv / len;
            T c = len - rest;
            T alpha_t = compliance * inv_dt_sq;
            T wsum = w_i + w_j + alpha_t;
            if (wsum < T(1e-8)) continue;
            T lam = lambda[e];
            T dlam = (-c - alpha_t * lam) / wsum;
            lambda[e] = lam + dlam;
            V3<T> dx_i = clamp_correction(omega * gv * w_i * dlam);
            V3<T> dx_j = clamp_correction(-(omega * gv * w_j * dlam));
            if (w_i > T(0)) store3(jac_work.data(), i, p_i + dx_i);
            if (w_j > T(0)) store3(jac_work.data(), j, p_j + dx_j);
        }
    }
    std::copy(jac_work.begin(), jac_work.end(), sim);
}

// Ground plane y = 0 (no reference counterpart): after collision, before the
// velocity update, a free vertex flagged for the ground is lifted to y = 0.
template <class T>
void ground_plane(const om_solver_in* in, T* sim) {
    if (!in->ground) return;
    for (size_t i = 0; i < in->n; ++i)
        if (in->ground[i] && T(in->inv_mass[i]) > T(0) && sim[3 * i + 1] < T(0)) sim[3 * i + 1] = T(0);
}

// xpbd_substep_post_velocity, xpbd_cpu.rs:393-409.
template <class T>
void post_velocity(const om_solver_in* in, const T* sim, const T* prev, T* vel, T inv_dt, T dt_sub, T drag) {
    const T damp = std::exp(-drag * dt_sub);
    for (size_t i = 0; i < in->n; ++i) {
        if (T(in->inv_mass[i]) <= T(0)) continue;
        store3(vel, i, (load3(sim, i) - load3(prev, i)) * inv_dt * damp);
    }
}

// xpbd_substep_{no_collision,with_self_collision}, xpbd_cpu.rs:411-578.
template <class T>
void substep(const om_solver_in* in, T* pos, T* vel, const P<T>& p) {
    std::vector<T> prev(3 * size_t(in->n));
    if (p.solver == OM_SOLVER_GS) integrate_gs(in, pos, prev.data(), vel, p);
    else integrate_jacobi(in, pos, prev.data(), vel, p);
    if (p.collision) self_collision(in, pos, p.thickness, p.coll_scale, p.coll_mode);
    ground_plane(in, pos);
    post_velocity(in, pos, prev.data(), vel, p.inv_dt, p.dt, p.drag);
}

// xpbd_frame_*, xpbd_cpu.rs:580-722 with SUBSTEPS as a parameter and the
// GPU path's collision stride rule (cloth_jacobi.rs:699-704).
template <class T>
void frames(const om_solver_in* in, T* pos, T* vel, const om_params* op, uint32_t substeps,
            uint32_t stride, uint32_t nframes) {
    P<T> p = promote<T>(op);
    const bool coll_enabled = op->collision != 0 && op->coll_scale > 1e-12f;
    const uint32_t s = std::max(stride, 1u);
    for (uint32_t f = 0; f < nframes; ++f)
        for (uint32_t si = 0; si < std::max(substeps, 1u); ++si) {
            P<T> q = p;
            q.collision = coll_enabled && (si % s == s - 1);
            substep(in, pos, vel, q);
        }
}

template <class T>
double residual(const om_solver_in* in, const T* pos) {
    double c2 = 0.0, r2 = 0.0;
    for (size_t e = 0; e < in->e; ++e) {
        V3<double> a = {double(pos[3 * in->ci[e]]), double(pos[3 * in->ci[e] + 1]), double(pos[3 * in->ci[e] + 2])};
        V3<double> b = {double(pos[3 * in->cj[e]]), double(pos[3 * in->cj[e] + 1]), double(pos[3 * in->cj[e] + 2])};
        double c = length(a - b) - double(in->crest[e]);
        c2 += c * c;
        r2 += double(in->crest[e]) * double(in->crest[e]);
    }
    return r2 > 0.0 ? std::sqrt(c2) / std::sqrt(r2) : 0.0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

void* om_grid_cloth_hanging(uint32_t quad_cols, uint32_t quad_rows, float cell_size) {
    if (quad_cols < 1 || quad_rows < 1 || !(cell_size > 0.f) || !std::isfinite(cell_size)) return nullptr;
    return grid_cloth_hanging(quad_cols, quad_rows, cell_size);
}

void* om_parse_welded_obj(const char* text) { return parse_welded_obj(std::string(text ? text : "")); }

// finalize_cloth_mesh's constraint half (mesh_prep.rs:580-635) on a caller's
// constraint list: the greedy colouring and colour sort
// (partition_constraints_for_gs_batches, mesh_prep.rs:153-221), then the CSR
// by cursor scatter.  Positions are the initial and rest pose; no triangles.
void* om_assemble(uint32_t n, const float* pos3, const float* inv_mass, uint32_t e, const uint32_t* ci,
                  const uint32_t* cj, const float* rest, const float* comp) {
    if (n == 0 || !pos3 || !inv_mass || (e && (!ci || !cj || !rest || !comp))) return nullptr;
    for (uint32_t k = 0; k < e; ++k)
        if (ci[k] >= n || cj[k] >= n || ci[k] == cj[k]) return nullptr;
    auto* m = new OMesh();
    m->positions.assign(pos3, pos3 + 3 * size_t(n));
    m->rest_positions = m->positions;
    m->inv_mass.assign(inv_mass, inv_mass + n);
    m->ground.assign(n, 0);
    std::vector<Constraint> cons(e);
    for (uint32_t k = 0; k < e; ++k) cons[k] = {ci[k], cj[k], rest[k], comp[k]};
    assemble_constraints(*m, n, std::move(cons));
    return m;
}

void* om_sheet(const om_sheet_desc* d) {
    if (!d || d->cols < 2 || d->rows < 2) return nullptr;
    return make_batch(d, 1);
}

void* om_batch(const om_sheet_desc* sheets, uint32_t count) {
    if (!sheets || count == 0) return nullptr;
    for (uint32_t c = 0; c < count; ++c)
        if (sheets[c].cols < 2 || sheets[c].rows < 2) return nullptr;
    return make_batch(sheets, count);
}

static const void* field_ptr(const OMesh* m, int f, uint64_t& count, size_t& elem) {
    elem = 4;
    switch (f) {
        case OM_POSITIONS: count = m->positions.size(); return m->positions.data();
        case OM_REST_POSITIONS: count = m->rest_positions.size(); return m->rest_positions.data();
        case OM_NORMALS: count = m->normals.size(); return m->normals.data();
        case OM_UVS: count = m->uvs.size(); return m->uvs.data();
        case OM_INDICES: count = m->indices.size(); return m->indices.data();
        case OM_TRI_EDGE_NEIGHBORS: count = m->tri_edge_neighbors.size(); return m->tri_edge_neighbors.data();
        case OM_INV_MASS: count = m->inv_mass.size(); return m->inv_mass.data();
        case OM_NEIGHBOR_OTHER: count = m->neighbor_other.size(); return m->neighbor_other.data();
        case OM_NEIGHBOR_REST_LEN: count = m->neighbor_rest_len.size(); return m->neighbor_rest_len.data();
        case OM_NEIGHBOR_COMPLIANCE: count = m->neighbor_compliance.size(); return m->neighbor_compliance.data();
        case OM_NEIGHBOR_CONSTRAINT_ID: count = m->neighbor_constraint_id.size(); return m->neighbor_constraint_id.data();
        case OM_NEIGHBOR_OFFSETS: count = m->neighbor_offsets.size(); return m->neighbor_offsets.data();
        case OM_CONSTRAINT_I: count = m->constraint_i.size(); return m->constraint_i.data();
        case OM_CONSTRAINT_J: count = m->constraint_j.size(); return m->constraint_j.data();
        case OM_CONSTRAINT_REST: count = m->constraint_rest.size(); return m->constraint_rest.data();
        case OM_CONSTRAINT_COMP: count = m->constraint_comp.size(); return m->constraint_comp.data();
        case OM_BATCH_OFFSETS: count = m->batch_offsets.size(); return m->batch_offsets.data();
        case OM_GROUND: count = m->ground.size(); return m->ground.data();
        case OM_CLOTH_OFFSETS: count = m->cloth_offsets.size(); return m->cloth_offsets.data();
        default: count = 0; return nullptr;
    }
}

uint64_t om_count(const void* mesh, int field) {
    uint64_t c = 0;
    size_t e = 0;
    if (!mesh) return 0;
    field_ptr(static_cast<const OMesh*>(mesh), field, c, e);
    return c;
}

void om_copy(const void* mesh, int field, void* dst) {
    uint64_t c = 0;
    size_t e = 0;
    if (!mesh || !dst) return;
    const void* p = field_ptr(static_cast<const OMesh*>(mesh), field, c, e);
    if (p && c) std::memcpy(dst, p, c * e);
}

void om_free(void* mesh) { delete static_cast<OMesh*>(mesh); }

// derive_collision_grid, mesh_prep.rs:16-85.
void om_derive_collision_grid(const float* rest, uint32_t n, float thickness, float* origin3,
                              float* inv_cell_out, uint32_t* dims3, uint32_t* num_cells, uint32_t* radix_digits) {
    if (n == 0) {
        origin3[0] = origin3[1] = origin3[2] = 0.f;
        *inv_cell_out = 1.f;
        dims3[0] = dims3[1] = dims3[2] = 1;
        *num_cells = 1;
        *radix_digits = 1;
        return;
    }
    V3<float> mn = v3(FLT_MAX, FLT_MAX, FLT_MAX), mx = v3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (uint32_t i = 0; i < n; ++i) {
        V3<float> v = v3(rest[4 * i], rest[4 * i + 1], rest[4 * i + 2]);
        mn = v3(fmin_rs(mn.x, v.x), fmin_rs(mn.y, v.y), fmin_rs(mn.z, v.z));
        mx = v3(fmax_rs(mx.x, v.x), fmax_rs(mx.y, v.y), fmax_rs(mx.z, v.z));
    }
    const float margin = std::fma(thickness, 6.0f, thickness * 0.5f);
    mn = mn - v3(margin, margin, margin);
    mx = mx + v3(margin, margin, margin);
    V3<float> ext = mx - mn;
    ext = v3(fmax_rs(ext.x, 1e-4f), fmax_rs(ext.y, 1e-4f), fmax_rs(ext.z, 1e-4f));
    const float nf = float(n);
    float target = std::sqrt(nf) * 6.25f;
    target = target < 256.f ? 256.f : (target > float(COLLISION_GRID_MAX_CELLS) ? float(COLLISION_GRID_MAX_CELLS) : target);
    const float vol = fmax_rs(ext.x * ext.y * ext.z, 1e-9f);
    const float crude = std::pow(vol / target, 1.0f / 3.0f);
    float cell_side = fmax_rs(fmax_rs(crude, thickness * 2.5f), 1e-5f);
    // Rust `f as u32` saturates: NaN and negatives -> 0, too large -> u32::MAX.
    auto sat_u32 = [](float f) -> uint32_t {
        if (!(f > 0.f)) return 0u;
        if (f >= 4294967296.0f) return 0xFFFFFFFFu;
        return uint32_t(f);
    };
    auto clampu = [](uint32_t v, uint32_t lo, uint32_t hi) { return v < lo ? lo : (v > hi ? hi : v); };
    auto satmul = [](uint32_t a, uint32_t b) {
        uint64_t r = uint64_t(a) * b;
        return r > 0xFFFFFFFFull ? 0xFFFFFFFFu : uint32_t(r);
    };
    uint32_t gx, gy, gz;
    for (;;) {
        gx = clampu(sat_u32(std::ceil(ext.x / cell_side)), 1, COLLISION_GRID_MAX_AXIS);
        gy = clampu(sat_u32(std::ceil(ext.y / cell_side)), 1, COLLISION_GRID_MAX_AXIS);
        gz = clampu(sat_u32(std::ceil(ext.z / cell_side)), 1, COLLISION_GRID_MAX_AXIS);
        uint32_t total = satmul(satmul(gx, gy), gz);
        if (total <= COLLISION_GRID_MAX_CELLS && total >= 1) break;
        cell_side *= 1.25f;
        float mxe = fmax_rs(fmax_rs(ext.x, ext.y), ext.z);
        if (cell_side > mxe * 2.5f) {
            gx = gy = gz = 1;
            cell_side = fmax_rs(mxe, thickness * 2.5f);
            break;
        }
    }
    *inv_cell_out = 1.0f / cell_side;
    uint32_t total = satmul(satmul(gx, gy), gz);
    uint32_t m = total == 0 ? 0 : total - 1;
    uint32_t digits;
    if (m == 0) digits = 1;
    else {
        uint32_t bits = 32 - uint32_t(__builtin_clz(m));
        digits = std::min(std::max((bits + 7) / 8, 1u), 4u);
    }
    origin3[0] = mn.x;
    origin3[1] = mn.y;
    origin3[2] = mn.z;
    dims3[0] = gx;
    dims3[1] = gy;
    dims3[2] = gz;
    *num_cells = total;
    *radix_digits = digits;
}

void oracle_substep_f32(const om_solver_in* in, float* pos, float* vel, const om_params* p) {
    substep(in, pos, vel, promote<float>(p));
}
void oracle_substep_f64(const om_solver_in* in, double* pos, double* vel, const om_params* p) {
    substep(in, pos, vel, promote<double>(p));
}
void oracle_frames_f32(const om_solver_in* in, float* pos, float* vel, const om_params* p, uint32_t substeps,
                       uint32_t stride, uint32_t nframes) {
    frames(in, pos, vel, p, substeps, stride, nframes);
}
void oracle_frames_f64(const om_solver_in* in, double* pos, double* vel, const om_params* p, uint32_t substeps,
                       uint32_t stride, uint32_t nframes) {
    frames(in, pos, vel, p, substeps, stride, nframes);
}
void oracle_self_collision_f32(const om_solver_in* in, float* pos, float thickness, float coll_scale, uint32_t mode) {
    self_collision(in, pos, thickness, coll_scale, mode);
}
void oracle_self_collision_f64(const om_solver_in* in, double* pos, double thickness, double coll_scale,
                               uint32_t mode) {
    self_collision(in, pos, thickness, coll_scale, mode);
}

// The build's collision cell key (no reference counterpart; the reference's
// collision_flat_packed, cloth_sim_jacobi.wgsl:243-253, clamps into a capped
// rest-pose grid instead).  Per substep: axis-aligned bounds of the current
// positions; cell = cell_size * 65/64; the cell doubles until every axis has
// fewer than 16383 cells and the cell count fits in 32 bits.
static void cell_grid_capped(const float* pos4, size_t b, size_t e, float cell_size, uint64_t max_cells, om_grid* g) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = b; i < e; ++i)
        for (int a = 0; a < 3; ++a) {
            float v = pos4[4 * i + a] + 0.0f;  // -0 folds to +0
            if (v < lo[a]) lo[a] = v;          // NaN never compares: ignored
            if (v > hi[a]) hi[a] = v;
        }
    for (int a = 0; a < 3; ++a)
        if (!(lo[a] <= hi[a])) lo[a] = hi[a] = 0.f;
    float cell = cell_size * 1.015625f;
    uint32_t dims[3] = {1, 1, 1};
    float inv = 1.0f / cell;
    bool ok = false;
    for (int it = 0; it < 64; ++it) {
        inv = 1.0f / cell;
        float ex[3];
        bool fits = true;
        for (int a = 0; a < 3; ++a) {
            ex[a] = (hi[a] - lo[a]) * inv;
            if (!(ex[a] < 16383.0f)) fits = false;
        }
        if (fits) {
            for (int a = 0; a < 3; ++a) dims[a] = uint32_t(std::floor(ex[a])) + 1u;
            const uint64_t total = uint64_t(dims[0]) * dims[1] * dims[2];
            if (total <= max_cells) {
                ok = true;
                break;
            }
        }
        cell = cell * 2.0f;
    }
    if (!ok) dims[0] = dims[1] = dims[2] = 1;
    for (int a = 0; a < 3; ++a) {
        g->lo[a] = lo[a];
        g->hi[a] = hi[a];
        g->dims[a] = dims[a];
    }
    g->cell = cell;
    g->inv_cell = inv;
}

// The build's collision cell key (no reference counterpart; the reference's
// collision_flat_packed, cloth_sim_jacobi.wgsl:243-253, clamps into a capped
// rest-pose grid instead).  Per substep: axis-aligned bounds of the current
// positions; cell = cell_size * 65/64; the cell doubles until every axis has
// fewer than 16383 cells and the cell count fits the key (2^27 cells for
// one cloth, so its keys sort in 3 passes of 9-bit digits).
void oracle_cell_grid(const float* pos4, uint32_t n, float cell_size, om_grid* g) {
    cell_grid_capped(pos4, 0, n, cell_size, 1ull << 27, g);
}

void oracle_cell_keys(const float* pos4, uint32_t n, const om_grid* g, uint32_t* keys) {
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t c[3];
        for (int a = 0; a < 3; ++a) {
            float q = (pos4[4 * size_t(i) + a] - g->lo[a]) * g->inv_cell;
            q = std::floor(q);
            float top = float(g->dims[a] - 1);
            q = std::fmax(q, 0.0f);  // NaN -> 0
            q = std::fmin(q, top);
            c[a] = uint32_t(q);
        }
        keys[i] = c[0] + g->dims[0] * (c[1] + g->dims[1] * c[2]);
    }
}

static uint32_t cloth_key_bits(uint32_t c) {
    uint32_t b = 0;
    while ((1u << b) < c) ++b;
    return c <= 1 ? 0u : b;
}

void oracle_cell_keys_batch(const float* pos4, const uint32_t* offs, uint32_t nc, float cell_size, om_grid* grids,
                            uint32_t* keys) {
    const uint32_t cb = cloth_key_bits(nc);
    const uint32_t lb = 32 - cb;
    const uint64_t cap = (cb == 0 ? (1ull << 32) : (1ull << lb)) - 2;
    for (uint32_t c = 0; c < nc; ++c) {
        const size_t b = offs[c], e = offs[c + 1];
        cell_grid_capped(pos4, b, e, cell_size, cap, &grids[c]);
        const uint32_t base = cb == 0 ? 0u : (c << lb);
        oracle_cell_keys(pos4 + 4 * b, uint32_t(e - b), &grids[c], keys + b);
        for (size_t i = b; i < e; ++i) keys[i] += base;
    }
}

void oracle_sort_pairs(const uint32_t* keys, uint32_t n, uint32_t* sorted_keys, uint32_t* sorted_ids) {
    std::vector<uint32_t> ids(n);
    for (uint32_t i = 0; i < n; ++i) ids[i] = i;
    std::stable_sort(ids.begin(), ids.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    for (uint32_t k = 0; k < n; ++k) {
        sorted_ids[k] = ids[k];
        sorted_keys[k] = keys[ids[k]];
    }
}

double oracle_residual_f32(const om_solver_in* in, const float* pos) { return residual(in, pos); }
double oracle_residual_f64(const om_solver_in* in, const double* pos) { return residual(in, pos); }

void oracle_vertex_normals(const float* pos3, uint32_t n, const uint32_t* tris, uint32_t t, float* nrm3) {
    std::vector<V3<float>> acc(n, v3(0.f, 0.f, 0.f));
    for (uint32_t k = 0; k < t; ++k) {
        uint32_t i0 = tris[3 * k], i1 = tris[3 * k + 1], i2 = tris[3 * k + 2];
        V3<float> p0 = load3(pos3, i0), p1 = load3(pos3, i1), p2 = load3(pos3, i2);
        V3<float> c = cross(p1 - p0, p2 - p0);
        acc[i0] = acc[i0] + c;
        acc[i1] = acc[i1] + c;
        acc[i2] = acc[i2] + c;
    }
    for (uint32_t i = 0; i < n; ++i) {
        V3<float> v = acc[i];
        float ln = length(v);
        V3<float> r = ln > 1e-8f ? v / ln : v3(0.f, 1.f, 0.f);
        store3(nrm3, i, r);
    }
}

void oracle_set_threads(int t) {
#ifdef _OPENMP
    omp_set_num_threads(t > 0 ? t : 1);
#else
    (void)t;
#endif
}

}  // extern "C"
