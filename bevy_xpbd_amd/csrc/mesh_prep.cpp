// Topology preparation for the XPBD cloth path (reference: src/mesh_prep.rs).
//
// Float arithmetic follows glam's scalar Vec3 order (dot = (x*x + y*y) + z*z,
// length = sqrt(dot)) and the file is compiled with -ffp-contract=off, so rest
// lengths and inverse masses are bit-identical to the reference's.  The
// colouring keeps the reference's greedy order and tie-breaks exactly
// (mesh_prep.rs:153-221) but uses per-vertex colour bitmasks and a counting
// sort instead of HashSet + comparison sorts, so it runs in O(E log E) at
// millions of constraints.
#include <algorithm>
#include <array>
#include <cctype>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>

#include "mesh.hpp"

namespace xpbd {
namespace {

constexpr float BEND_COMPLIANCE = 3.8f;          // mesh_prep.rs:89
constexpr float STRETCH_COMPLIANCE = 2e-8f;      // mesh_prep.rs:93
constexpr float SKIP2_COMPLIANCE = 2.5e-8f;      // mesh_prep.rs:97
constexpr float CROSS_DIAG_COMPLIANCE = 6e-9f;   // mesh_prep.rs:101
constexpr uint32_t NO_NEIGHBOR = 0xFFFFFFFFu;

struct F3 {
    float x, y, z;
};
inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float len(F3 a) { return std::sqrt((a.x * a.x + a.y * a.y) + a.z * a.z); }
inline F3 crs(F3 a, F3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline F3 at(const std::vector<float>& p, size_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
inline float dist(const std::vector<float>& p, size_t a, size_t b) { return len(sub(at(p, a), at(p, b))); }
inline uint64_t pair_key(uint32_t a, uint32_t b) { return (uint64_t(a) << 32) | b; }

// Greedy colouring with the reference's processing order:
// (max endpoint degree descending, min endpoint, max endpoint, input index).
std::vector<uint32_t> colour_edges(size_t n, const std::vector<Constraint>& c, uint32_t& num_colours) {
    const size_t e = c.size();
    std::vector<uint32_t> deg(n, 0);
    for (const auto& k : c) {
        deg[k.i]++;
        deg[k.j]++;
    }
    struct Key {
        uint64_t hi;   // (~deg) << 32 | min
        uint64_t lo;   // max << 32 | idx
    };
    std::vector<Key> keys(e);
    for (size_t k = 0; k < e; ++k) {
        uint32_t a = std::min(c[k].i, c[k].j), b = std::max(c[k].i, c[k].j);
        uint32_t d = std::max(deg[c[k].i], deg[c[k].j]);
        keys[k] = {(uint64_t(~d) << 32) | a, (uint64_t(b) << 32) | uint64_t(k)};
    }
    std::sort(keys.begin(), keys.end(), [](const Key& x, const Key& y) {
        return x.hi != y.hi ? x.hi < y.hi : x.lo < y.lo;
    });
    std::vector<uint32_t> colour(e, 0);
    std::vector<uint64_t> used(n, 0);                 // colours 0..63
    std::vector<std::vector<uint32_t>> used_hi;      // colours >= 64 (rare)
    num_colours = 0;
    for (const auto& key : keys) {
        const size_t k = size_t(key.lo & 0xFFFFFFFFu);
        const uint32_t i = c[k].i, j = c[k].j;
        uint64_t taken = used[i] | used[j];
        uint32_t col;
        if (~taken) {
            col = uint32_t(__builtin_ctzll(~taken));
        } else {
            if (used_hi.empty()) used_hi.resize(n);
            std::vector<uint32_t> t = used_hi[i];
            t.insert(t.end(), used_hi[j].begin(), used_hi[j].end());
            std::sort(t.begin(), t.end());
            col = 64;
            for (uint32_t v : t) {
                if (v == col) col++;
                else if (v > col) break;
            }
        }
        colour[k] = col;
        if (col < 64) {
            used[i] |= (1ull << col);
            used[j] |= (1ull << col);
        } else {
            used_hi[i].push_back(col);
            used_hi[j].push_back(col);
        }
        num_colours = std::max(num_colours, col + 1);
    }
    return colour;
}

}  // namespace

void assemble_constraints(Mesh& m, size_t n, std::vector<Constraint>&& cons) {
    const size_t e = cons.size();
    m.constraint_i.resize(e);
    m.constraint_j.resize(e);
    m.constraint_rest.resize(e);
    m.constraint_comp.resize(e);
    if (e == 0) {
        m.batch_offsets = {0, 0};
    } else {
        uint32_t nc = 0;
        std::vector<uint32_t> colour = colour_edges(n, cons, nc);
        // Stable counting sort by colour == sort by (colour, index).
        std::vector<uint32_t> start(nc + 1, 0);
        for (uint32_t c : colour) start[c + 1]++;
        for (uint32_t c = 0; c < nc; ++c) start[c + 1] += start[c];
        m.batch_offsets.assign(start.begin(), start.end());
        std::vector<uint32_t> cur(start.begin(), start.end() - 1);
        for (size_t k = 0; k < e; ++k) {
            uint32_t dst = cur[colour[k]]++;
            m.constraint_i[dst] = cons[k].i;
            m.constraint_j[dst] = cons[k].j;
            m.constraint_rest[dst] = cons[k].rest;
            m.constraint_comp[dst] = cons[k].comp;
        }
    }
    std::vector<Constraint>().swap(cons);
    scatter_csr(m, n);
}

// CSR by cursor scatter in row order (mesh_prep.rs:588-634): row k is
// appended to the slice of its first endpoint, then of its second.
void scatter_csr(Mesh& m, size_t n) {
    const size_t e = m.constraint_i.size();
    m.neighbor_offsets.assign(n + 1, 0);
    for (size_t k = 0; k < e; ++k) {
        m.neighbor_offsets[m.constraint_i[k] + 1]++;
        m.neighbor_offsets[m.constraint_j[k] + 1]++;
    }
    for (size_t i = 0; i < n; ++i) m.neighbor_offsets[i + 1] += m.neighbor_offsets[i];
    const size_t nn = m.neighbor_offsets[n];
    m.neighbor_other.resize(nn);
    m.neighbor_rest_len.resize(nn);
    m.neighbor_compliance.resize(nn);
    m.neighbor_constraint_id.resize(nn);
    std::vector<uint32_t> cur(m.neighbor_offsets.begin(), m.neighbor_offsets.end() - 1);
    for (size_t k = 0; k < e; ++k) {
        const uint32_t a = m.constraint_i[k], b = m.constraint_j[k];
        uint32_t pa = cur[a]++;
        m.neighbor_other[pa] = b;
        m.neighbor_rest_len[pa] = m.constraint_rest[k];
        m.neighbor_compliance[pa] = m.constraint_comp[k];
        m.neighbor_constraint_id[pa] = uint32_t(k);
        uint32_t pb = cur[b]++;
        m.neighbor_other[pb] = a;
        m.neighbor_rest_len[pb] = m.constraint_rest[k];
        m.neighbor_compliance[pb] = m.constraint_comp[k];
        m.neighbor_constraint_id[pb] = uint32_t(k);
    }
}

namespace {

// Triangle-edge adjacency (mesh_prep.rs:661-695): edges sorted by (min, max)
// with ties in edge-number order; consecutive equal pairs link to each other.
std::vector<uint32_t> tri_neighbors(const std::vector<uint32_t>& idx) {
    const size_t t = idx.size() / 3;
    std::vector<std::array<uint32_t, 3>> edges(3 * t);
    for (size_t i = 0; i < t; ++i)
        for (uint32_t j = 0; j < 3; ++j) {
            uint32_t a = idx[3 * i + j], b = idx[3 * i + (j + 1) % 3];
            edges[3 * i + j] = {std::min(a, b), std::max(a, b), uint32_t(3 * i + j)};
        }
    std::sort(edges.begin(), edges.end());  // edge_nr last => stable order
    std::vector<uint32_t> nb(3 * t, NO_NEIGHBOR);
    for (size_t k = 0; k + 1 < edges.size(); ++k)
        if (edges[k][0] == edges[k + 1][0] && edges[k][1] == edges[k + 1][1]) {
            nb[edges[k][2]] = edges[k + 1][2];
            nb[edges[k + 1][2]] = edges[k][2];
        }
    return nb;
}

// finalize_cloth_mesh (mesh_prep.rs:532-659).
Mesh* finalize(std::vector<float>&& pos, std::vector<float>&& nrm, std::vector<float>&& uvs,
               std::vector<uint32_t>&& idx, std::vector<Constraint>&& extras) {
    auto* m = new Mesh();
    const size_t n = pos.size() / 3, t = idx.size() / 3;
    m->inv_mass.assign(n, 0.f);
    for (size_t k = 0; k < t; ++k) {
        const uint32_t a = idx[3 * k], b = idx[3 * k + 1], c = idx[3 * k + 2];
        const float area = 0.5f * len(crs(sub(at(pos, b), at(pos, a)), sub(at(pos, c), at(pos, a))));
        const float w = area > 0.f ? 1.0f / area / 3.0f : 0.f;
        m->inv_mass[a] += w;
        m->inv_mass[b] += w;
        m->inv_mass[c] += w;
    }
    m->tri_edge_neighbors = tri_neighbors(idx);
    float max_y = -INFINITY;
    for (size_t i = 0; i < n; ++i) max_y = std::fmax(max_y, pos[3 * i + 1]);
    for (size_t i = 0; i < n; ++i)
        if (pos[3 * i + 1] >= max_y - 1e-5f) m->inv_mass[i] = 0.f;

    std::vector<Constraint> cons;
    const auto& nb = m->tri_edge_neighbors;
    {   // unique triangle edges (mesh_prep.rs:697-722)
        std::unordered_set<uint64_t> seen;
        for (size_t k = 0; k < t; ++k)
            for (uint32_t j = 0; j < 3; ++j) {
                uint32_t a = idx[3 * k + j], b = idx[3 * k + (j + 1) % 3];
                if (nb[3 * k + j] != NO_NEIGHBOR && a >= b) continue;
                uint32_t lo = std::min(a, b), hi = std::max(a, b);
                if (seen.insert(pair_key(lo, hi)).second) cons.push_back({lo, hi, dist(pos, lo, hi), STRETCH_COMPLIANCE});
            }
    }
    {   // hinge bending pairs (mesh_prep.rs:724-747)
        std::unordered_set<uint64_t> seen;
        for (size_t k = 0; k < t; ++k)
            for (uint32_t j = 0; j < 3; ++j) {
                uint32_t o = nb[3 * k + j];
                if (o == NO_NEIGHBOR) continue;
                uint32_t a = idx[3 * k + (j + 2) % 3];
                uint32_t b = idx[3 * (o / 3) + ((o % 3) + 2) % 3];
                uint32_t lo = std::min(a, b), hi = std::max(a, b);
                if (seen.insert(pair_key(lo, hi)).second) cons.push_back({lo, hi, dist(pos, lo, hi), BEND_COMPLIANCE});
            }
    }
    cons.insert(cons.end(), extras.begin(), extras.end());
    assemble_constraints(*m, n, std::move(cons));
    m->rest_positions = pos;
    m->positions = std::move(pos);
    m->normals = std::move(nrm);
    m->uvs = std::move(uvs);
    m->indices = std::move(idx);
    m->ground.assign(n, 0);
    m->cloth_offsets = {0, uint32_t(n)};
    return m;
}

// Full-grid detection for welded OBJ meshes (mesh_prep.rs:428-501).
bool obj_grid_skip_two(const std::vector<float>& pos, std::vector<Constraint>& out) {
    const float EPS = 2e-4f;
    const size_t n = pos.size() / 3;
    if (n == 0) return false;
    auto uniq = [&](int axis) {
        std::vector<float> v(n);
        for (size_t i = 0; i < n; ++i) v[i] = pos[3 * i + axis];
        std::stable_sort(v.begin(), v.end());
        size_t w = 0;
        for (size_t r = 0; r < v.size(); ++r)
            if (w == 0 || std::fabs(v[r] - v[w - 1]) > EPS) v[w++] = v[r];
        v.resize(w);
        return v;
    };
    const std::vector<float> xs = uniq(0), ys = uniq(1);
    const size_t cols = xs.size(), rows = ys.size();
    if (cols < 3 || rows < 3 || cols * rows != n) return false;
    auto first_match = [&](const std::vector<float>& g, float v) -> long {
        for (size_t k = 0; k < g.size(); ++k)
            if (std::fabs(g[k] - v) <= EPS) return long(k);
        return -1;
    };
    std::vector<long> slot(cols * rows, -1);
    for (size_t p = 0; p < n; ++p) {
        long ix = first_match(xs, pos[3 * p]);
        if (ix < 0) return false;
        long iy = first_match(ys, pos[3 * p + 1]);
        if (iy < 0) return false;
        long& s = slot[size_t(iy) * cols + size_t(ix)];
        if (s >= 0) return false;
        s = long(p);
    }
    auto pid = [&](size_t ix, size_t iy) { return uint32_t(slot[iy * cols + ix]); };
    for (size_t iy = 0; iy < rows; ++iy)
        for (size_t ix = 0; ix + 2 < cols; ++ix) {
            uint32_t a = pid(ix, iy), b = pid(ix + 2, iy);
            out.push_back({std::min(a, b), std::max(a, b), dist(pos, a, b), SKIP2_COMPLIANCE});
        }
    for (size_t iy = 0; iy + 2 < rows; ++iy)
        for (size_t ix = 0; ix < cols; ++ix) {
            uint32_t a = pid(ix, iy), b = pid(ix, iy + 2);
            out.push_back({std::min(a, b), std::max(a, b), dist(pos, a, b), SKIP2_COMPLIANCE});
        }
    return true;
}

// str::parse::<f32> / ::<usize> accept the whole token or nothing.
bool tok_f32(const char* s, size_t len, float& v) {
    if (len == 0 || len > 63) return false;
    char buf[64];
    std::memcpy(buf, s, len);
    buf[len] = 0;
    char* end = nullptr;
    v = std::strtof(buf, &end);
    return end == buf + len;
}
bool tok_usize(const std::string& s, size_t& v) {
    size_t k = (!s.empty() && s[0] == '+') ? 1 : 0;
    if (k >= s.size()) return false;
    v = 0;
    for (; k < s.size(); ++k) {
        if (!std::isdigit((unsigned char)s[k])) return false;
        v = v * 10 + size_t(s[k] - '0');
    }
    return true;
}

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

}  // namespace

Mesh* grid_cloth_hanging(uint32_t qc, uint32_t qr, float cell) {
    const uint32_t vc = qc + 1, vr = qr + 1;
    const size_t n = size_t(vc) * vr;
    std::vector<float> pos(3 * n), nrm(3 * n), uvs(2 * n);
    const float x0 = -0.5f * (float(qc) * cell);
    for (uint32_t iy = 0; iy < vr; ++iy)
        for (uint32_t ix = 0; ix < vc; ++ix) {
            const size_t i = size_t(iy) * vc + ix;
            pos[3 * i] = x0 + float(ix) * cell;
            pos[3 * i + 1] = float(iy) * cell;
            pos[3 * i + 2] = 0.f;
            nrm[3 * i] = 0.f;
            nrm[3 * i + 1] = 0.f;
            nrm[3 * i + 2] = 1.f;
            uvs[2 * i] = float(ix) / float(qc);
            uvs[2 * i + 1] = float(iy) / float(qr);
        }
    std::vector<uint32_t> idx;
    idx.reserve(size_t(qc) * qr * 6);
    for (uint32_t iy = 0; iy < qr; ++iy)
        for (uint32_t ix = 0; ix < qc; ++ix) {
            uint32_t i0 = iy * vc + ix, i1 = i0 + 1, i2 = i0 + vc, i3 = i2 + 1;
            idx.insert(idx.end(), {i0, i1, i2, i1, i3, i2});
        }
    std::vector<Constraint> extras;   // skip-2 braces then cross diagonals (mesh_prep.rs:403-404)
    for (uint32_t iy = 0; iy < vr; ++iy)
        for (uint32_t ix = 0; ix + 2 < vc; ++ix) {
            uint32_t a = iy * vc + ix;
            extras.push_back({a, a + 2, dist(pos, a, a + 2), SKIP2_COMPLIANCE});
        }
    for (uint32_t iy = 0; iy + 2 < vr; ++iy)
        for (uint32_t ix = 0; ix < vc; ++ix) {
            uint32_t a = iy * vc + ix, b = (iy + 2) * vc + ix;
            extras.push_back({a, b, dist(pos, a, b), SKIP2_COMPLIANCE});
        }
    for (uint32_t iy = 0; iy < qr; ++iy)
        for (uint32_t ix = 0; ix < qc; ++ix) {
            uint32_t a = iy * vc + ix, b = (iy + 1) * vc + ix + 1;
            extras.push_back({a, b, dist(pos, a, b), CROSS_DIAG_COMPLIANCE});
        }
    Mesh* m = finalize(std::move(pos), std::move(nrm), std::move(uvs), std::move(idx), std::move(extras));
    if (m) m->cloth_cols = {vc};
    return m;
}

Mesh* parse_welded_obj(const std::string& text) {
    std::vector<std::array<float, 3>> vp, vn;
    std::vector<std::array<float, 2>> vt;
    std::vector<std::vector<std::string>> faces;
    size_t p = 0;
    while (p <= text.size()) {
        size_t e = text.find('\n', p);
        if (e == std::string::npos) e = text.size();
        // split_whitespace over the line
        std::vector<std::pair<size_t, size_t>> toks;
        size_t k = p;
        while (k < e) {
            while (k < e && std::isspace((unsigned char)text[k])) ++k;
            size_t s = k;
            while (k < e && !std::isspace((unsigned char)text[k])) ++k;
            if (k > s) toks.push_back({s, k - s});
        }
        p = e + 1;
        if (toks.empty() || text[toks[0].first] == '#') continue;
        const std::string head = text.substr(toks[0].first, toks[0].second);
        auto num = [&](size_t t) {
            float v = 0.f;
            if (t < toks.size() && tok_f32(text.data() + toks[t].first, toks[t].second, v)) return v;
            return 0.f;
        };
        if (head == "v") vp.push_back({num(1), num(2), num(3)});
        else if (head == "vt") vt.push_back({num(1), num(2)});
        else if (head == "vn") vn.push_back({num(1), num(2), num(3)});
        else if (head == "f") {
            std::vector<std::string> f;
            for (size_t t = 1; t < toks.size(); ++t) f.push_back(text.substr(toks[t].first, toks[t].second));
            faces.push_back(std::move(f));
        }
    }
    std::unordered_map<std::string, uint32_t> welded;
    std::vector<float> pos, nrm, uvs;
    std::vector<uint32_t> idx;
    std::vector<std::pair<uint32_t, uint32_t>> quad_cross;
    auto corner = [&](const std::string& s) -> uint32_t {
        auto it = welded.find(s);
        if (it != welded.end()) return it->second;
        const uint32_t id = uint32_t(pos.size() / 3);
        welded.emplace(s, id);
        std::string f[3];
        size_t nf = 0, a = 0;
        for (;;) {
            size_t b = s.find('/', a);
            if (nf < 3) f[nf] = s.substr(a, b == std::string::npos ? std::string::npos : b - a);
            ++nf;
            if (b == std::string::npos) break;
            a = b + 1;
        }
        size_t vi;
        if (!tok_usize(f[0], vi)) vi = 1;
        vi = vi ? vi - 1 : 0;
        const auto P = vi < vp.size() ? vp[vi] : std::array<float, 3>{0.f, 0.f, 0.f};
        pos.insert(pos.end(), P.begin(), P.end());
        size_t ti;
        if (nf > 1 && !f[1].empty() && tok_usize(f[1], ti)) {
            ti = ti ? ti - 1 : 0;
            const auto T = ti < vt.size() ? vt[ti] : std::array<float, 2>{0.f, 0.f};
            uvs.insert(uvs.end(), T.begin(), T.end());
        } else {
            uvs.insert(uvs.end(), {0.f, 0.f});
        }
        size_t ni;
        if (nf > 2 && tok_usize(f[2], ni)) {
            ni = ni ? ni - 1 : 0;
            const auto N = ni < vn.size() ? vn[ni] : std::array<float, 3>{0.f, 1.f, 0.f};
            nrm.insert(nrm.end(), N.begin(), N.end());
        } else {
            nrm.insert(nrm.end(), {0.f, 1.f, 0.f});
        }
        return id;
    };
    for (const auto& f : faces) {
        if (f.size() < 3) continue;
        if (f.size() == 4) {
            const uint32_t v0 = corner(f[0]), v1 = corner(f[1]), v2 = corner(f[2]), v3 = corner(f[3]);
            idx.insert(idx.end(), {v0, v1, v2, v0, v2, v3});
            quad_cross.push_back({std::min(v1, v3), std::max(v1, v3)});
            continue;
        }
        const uint32_t c0 = corner(f[0]);
        for (size_t i = 1; i + 1 < f.size(); ++i) {
            idx.push_back(c0);
            const uint32_t a = corner(f[i]);
            idx.push_back(a);
            const uint32_t b = corner(f[i + 1]);
            idx.push_back(b);
        }
    }
    std::vector<Constraint> extras;
    obj_grid_skip_two(pos, extras);
    for (auto [a, b] : quad_cross) extras.push_back({a, b, dist(pos, a, b), CROSS_DIAG_COMPLIANCE});
    return finalize(std::move(pos), std::move(nrm), std::move(uvs), std::move(idx), std::move(extras));
}

Mesh* sheet_batch(const xpbd_sheet_desc* descs, uint32_t count) {
    auto* m = new Mesh();
    size_t total_n = 0, total_e = 0;
    for (uint32_t c = 0; c < count; ++c) {
        const size_t C = descs[c].cols, R = descs[c].rows;
        total_n += C * R;
        total_e += (C - 1) * R + C * (R - 1) + 2 * (C - 1) * (R - 1) + (C > 2 ? (C - 2) * R : 0) +
                   (R > 2 ? C * (R - 2) : 0);
    }
    m->positions.reserve(3 * total_n);
    m->rest_positions.reserve(3 * total_n);
    m->normals.reserve(3 * total_n);
    m->uvs.reserve(2 * total_n);
    m->inv_mass.reserve(total_n);
    m->ground.reserve(total_n);
    std::vector<Constraint> cons;
    cons.reserve(total_e);
    m->cloth_offsets.push_back(0);
    for (uint32_t c = 0; c < count; ++c) {
        const xpbd_sheet_desc& d = descs[c];
        const uint32_t C = d.cols, R = d.rows;
        const uint32_t base = uint32_t(m->inv_mass.size());
        const float h = d.spacing;
        const float xa = -0.5f * (float(C - 1) * h), za = -0.5f * (float(R - 1) * h);
        const float cs = std::cos(d.angle), sn = std::sin(d.angle);
        const size_t rbase = m->rest_positions.size();
        for (uint32_t iy = 0; iy < R; ++iy)
            for (uint32_t ix = 0; ix < C; ++ix) {
                const float a = xa + float(ix) * h;
                if (d.layout == XPBD_LAYOUT_HANGING_XY) {
                    m->rest_positions.insert(m->rest_positions.end(), {a, d.height + float(iy) * h, 0.f});
                    m->normals.insert(m->normals.end(), {0.f, 0.f, 1.f});
                } else {
                    const float b = za + float(iy) * h;
                    m->rest_positions.insert(m->rest_positions.end(), {a * cs - b * sn, d.height, a * sn + b * cs});
                    m->normals.insert(m->normals.end(), {0.f, 1.f, 0.f});
                }
                m->uvs.insert(m->uvs.end(), {float(ix) / float(C - 1), float(iy) / float(R - 1)});
            }
        m->positions.insert(m->positions.end(), m->rest_positions.begin() + rbase, m->rest_positions.end());
        if (d.jitter > 0.f) {
            SplitMix64 rng{d.seed};
            const size_t axis = d.layout == XPBD_LAYOUT_HANGING_XY ? 2 : 1;
            for (size_t i = 0; i < size_t(C) * R; ++i) {
                float& v = m->positions[rbase + 3 * i + axis];
                v = v + (rng.unit() * 2.0f - 1.0f) * d.jitter;
            }
        }
        for (uint32_t iy = 0; iy + 1 < R; ++iy)
            for (uint32_t ix = 0; ix + 1 < C; ++ix) {
                const uint32_t i0 = base + iy * C + ix, i1 = i0 + 1, i2 = i0 + C, i3 = i2 + 1;
                m->indices.insert(m->indices.end(), {i0, i1, i2, i1, i3, i2});
            }
        m->inv_mass.insert(m->inv_mass.end(), size_t(C) * R, d.inv_mass);
        float* w = m->inv_mass.data() + base;
        if (d.pins == XPBD_PIN_TWO_CORNERS) {
            w[(R - 1) * C] = 0.f;
            w[(R - 1) * C + C - 1] = 0.f;
        } else if (d.pins == XPBD_PIN_TOP_ROW) {
            for (uint32_t ix = 0; ix < C; ++ix) w[(R - 1) * C + ix] = 0.f;
        } else if (d.pins == XPBD_PIN_CENTER) {
            w[(R / 2) * C + C / 2] = 0.f;
        }
        m->ground.insert(m->ground.end(), size_t(C) * R, d.ground ? 1u : 0u);
        const std::vector<float>& rp = m->rest_positions;
        auto add = [&](uint32_t a, uint32_t b, float comp) {
            a += base;
            b += base;
            cons.push_back({a, b, dist(rp, a, b), comp});
        };
        for (uint32_t iy = 0; iy < R; ++iy)
            for (uint32_t ix = 0; ix + 1 < C; ++ix) add(iy * C + ix, iy * C + ix + 1, d.stretch_compliance);
        for (uint32_t iy = 0; iy + 1 < R; ++iy)
            for (uint32_t ix = 0; ix < C; ++ix) add(iy * C + ix, (iy + 1) * C + ix, d.stretch_compliance);
        for (uint32_t iy = 0; iy + 1 < R; ++iy)
            for (uint32_t ix = 0; ix + 1 < C; ++ix) {
                const uint32_t i0 = iy * C + ix;
                add(i0, i0 + C + 1, d.shear_compliance);
                add(i0 + 1, i0 + C, d.shear_compliance);
            }
        for (uint32_t iy = 0; iy < R; ++iy)
            for (uint32_t ix = 0; ix + 2 < C; ++ix) add(iy * C + ix, iy * C + ix + 2, d.bend_compliance);
        for (uint32_t iy = 0; iy + 2 < R; ++iy)
            for (uint32_t ix = 0; ix < C; ++ix) add(iy * C + ix, (iy + 2) * C + ix, d.bend_compliance);
        m->cloth_offsets.push_back(uint32_t(m->inv_mass.size()));
        m->cloth_cols.push_back(C);
    }
    assemble_constraints(*m, m->inv_mass.size(), std::move(cons));
    return m;
}

// ClothSimConfig's topology arrays (cloth_compute.rs:142-179), as
// to_sim_config copies them out of a ClothMeshData (cloth_compute.rs:
// 1599-1700).  Rows keep the caller's order when batch offsets come with them
// (each batch checked vertex-disjoint); without, they are coloured and sorted
// as finalize_cloth_mesh does.  A given adjacency (pack_neighbor_gpu's vec4
// entries, cloth_jacobi.rs:95-111) is checked row by row and kept in its
// order; without one, the cursor scatter of mesh_prep.rs:588-635 builds it.
Mesh* mesh_from_arrays(const xpbd_mesh_desc& d, std::string& why, int& code) {
    const size_t n = d.num_particles, e = d.num_distance_constraints, t = d.num_tris;
    code = XPBD_E_INVALID;
    auto bad = [&](const std::string& w) -> Mesh* {
        why = w;
        return nullptr;
    };
    auto at = [](size_t k) { return std::to_string(k); };
    if (n == 0) return bad("num_particles is 0");
    if (n > (size_t(1) << 29)) return bad("num_particles above 2^29");
    if (e > (size_t(1) << 30)) return bad("num_distance_constraints above 2^30");
    if (t > (size_t(1) << 29)) return bad("num_tris above 2^29");
    if (!d.initial_pos || !d.inv_mass) return bad("initial_pos and inv_mass are required");
    if (e && (!d.constraint_i || !d.constraint_j || !d.constraint_rest_len || !d.constraint_compliance))
        return bad("constraint_i/j/rest_len/compliance are required when num_distance_constraints > 0");
    if (t && !d.tri_indices) return bad("tri_indices is required when num_tris > 0");
    if (!d.neighbor_offsets != !d.neighbor_packed) return bad("neighbor_offsets and neighbor_packed come together");
    if (d.neighbor_offsets && !d.constraint_batch_offsets)
        return bad("a given adjacency names constraint rows: it needs the constraint_batch_offsets of those rows");
    if (d.device_prepare && (d.constraint_batch_offsets || d.neighbor_offsets))
        return bad("device_prepare computes the colouring and the adjacency: give neither");

    auto m = std::make_unique<Mesh>();
    m->positions.resize(3 * n);
    m->rest_positions.resize(3 * n);
    m->normals.resize(3 * n);
    m->inv_mass.resize(n);
    const float* rest = d.rest_pos ? d.rest_pos : d.initial_pos;
    for (size_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            const float p = d.initial_pos[4 * i + a], r = rest[4 * i + a];
            if (!std::isfinite(p) || !std::isfinite(r)) return bad("position of particle " + at(i) + " is not finite");
            m->positions[3 * i + a] = p;
            m->rest_positions[3 * i + a] = r;
            m->normals[3 * i + a] = d.initial_normals ? d.initial_normals[4 * i + a] : (a == 1 ? 1.f : 0.f);
        }
        const float w = d.inv_mass[i];
        if (!std::isfinite(w) || w < 0.f) return bad("inv_mass of particle " + at(i) + " is negative or not finite");
        m->inv_mass[i] = w;
    }
    if (t) {
        m->indices.assign(d.tri_indices, d.tri_indices + 3 * t);
        for (size_t k = 0; k < 3 * t; ++k)
            if (m->indices[k] >= n) return bad("tri_indices[" + at(k) + "] out of range");
    }
    m->ground.assign(n, 0);
    if (d.ground)
        for (size_t i = 0; i < n; ++i) m->ground[i] = d.ground[i] ? 1u : 0u;
    m->cloth_offsets = {0, uint32_t(n)};
    std::vector<uint32_t> cloth_of;
    if (d.cloth_offsets && d.num_cloths > 1) {
        const uint32_t* co = d.cloth_offsets;
        if (co[0] != 0 || co[d.num_cloths] != n) return bad("cloth_offsets must start at 0 and end at N");
        cloth_of.resize(n);
        for (uint32_t c = 0; c < d.num_cloths; ++c) {
            if (co[c] >= co[c + 1]) return bad("cloth " + at(c) + " is empty or cloth_offsets decrease");
            for (uint32_t i = co[c]; i < co[c + 1]; ++i) cloth_of[i] = c;
        }
        m->cloth_offsets.assign(co, co + d.num_cloths + 1);
    }

    std::vector<Constraint> cons(e);
    for (size_t k = 0; k < e; ++k) {
        const Constraint c{d.constraint_i[k], d.constraint_j[k], d.constraint_rest_len[k], d.constraint_compliance[k]};
        if (c.i >= n || c.j >= n) return bad("constraint row " + at(k) + " has an endpoint out of range");
        if (c.i == c.j) return bad("constraint row " + at(k) + " joins a particle to itself");
        if (!cloth_of.empty() && cloth_of[c.i] != cloth_of[c.j])
            return bad("constraint row " + at(k) + " joins two cloths");
        if (!std::isfinite(c.rest) || c.rest < 0.f) return bad("rest length of row " + at(k) + " is negative or not finite");
        if (!std::isfinite(c.comp) || c.comp < 0.f) return bad("compliance of row " + at(k) + " is negative or not finite");
        cons[k] = c;
    }
    if (!d.constraint_batch_offsets) {
        if (d.device_prepare) {
            const int r = topology_on_device(n, m->cloth_offsets, cons, *m, why);
            if (r != XPBD_OK) {
                code = r;
                return nullptr;
            }
        } else {
            assemble_constraints(*m, n, std::move(cons));
        }
        return m.release();
    }

    // Rows in the caller's order; every batch vertex-disjoint.
    const size_t nb = d.constraint_batch_count;
    const uint32_t* bo = d.constraint_batch_offsets;
    if (e == 0 && nb == 0) {
        m->batch_offsets = {0, 0};   // partition_constraints_for_gs_batches on no rows
    } else {
        if (nb == 0) return bad("constraint_batch_count is 0 with rows present");
        if (bo[0] != 0 || bo[nb] != e) return bad("constraint_batch_offsets must start at 0 and end at E");
        std::vector<uint32_t> stamp(n, 0);
        for (size_t b = 0; b < nb; ++b) {
            if (bo[b] > bo[b + 1]) return bad("constraint_batch_offsets decrease at batch " + at(b));
            for (uint32_t k = bo[b]; k < bo[b + 1]; ++k) {
                const uint32_t i = cons[k].i, j = cons[k].j, s = uint32_t(b + 1);
                if (stamp[i] == s || stamp[j] == s)
                    return bad("batch " + at(b) + " is not vertex-disjoint (row " + at(k) + ")");
                stamp[i] = stamp[j] = s;
            }
        }
        m->batch_offsets.assign(bo, bo + nb + 1);
    }
    m->constraint_i.resize(e);
    m->constraint_j.resize(e);
    m->constraint_rest.resize(e);
    m->constraint_comp.resize(e);
    for (size_t k = 0; k < e; ++k) {
        m->constraint_i[k] = cons[k].i;
        m->constraint_j[k] = cons[k].j;
        m->constraint_rest[k] = cons[k].rest;
        m->constraint_comp[k] = cons[k].comp;
    }

    if (!d.neighbor_offsets) {   // cursor scatter over the rows (mesh_prep.rs:588-635)
        scatter_csr(*m, n);
        return m.release();
    }
    // The caller's adjacency: every row exactly once at each endpoint.
    const uint32_t* no = d.neighbor_offsets;
    if (no[0] != 0 || no[n] != 2 * e) return bad("neighbor_offsets must start at 0 and end at 2E");
    for (size_t i = 0; i < n; ++i)
        if (no[i] > no[i + 1]) return bad("neighbor_offsets decrease at particle " + at(i));
    if (e >= (size_t(1) << 24) || n >= (size_t(1) << 24))
        return bad("neighbor_packed holds indices as f32: N and E must stay below 2^24");
    m->neighbor_offsets.assign(no, no + n + 1);
    m->neighbor_other.resize(2 * e);
    m->neighbor_rest_len.resize(2 * e);
    m->neighbor_compliance.resize(2 * e);
    m->neighbor_constraint_id.resize(2 * e);
    std::vector<uint8_t> seen(e, 0);   // bit 0: seen from endpoint i, bit 1: from j
    auto as_index = [](float v, size_t lim, uint32_t& out) {
        if (!(v >= 0.f) || v != std::floor(v) || v >= float(lim)) return false;
        out = uint32_t(v);
        return true;
    };
    for (size_t i = 0; i < n; ++i)
        for (uint32_t k = no[i]; k < no[i + 1]; ++k) {
            const float* q = d.neighbor_packed + 4 * size_t(k);
            uint32_t other = 0, row = 0;
            if (!as_index(q[0], n, other) || !as_index(q[3], e, row))
                return bad("neighbor_packed entry " + at(k) + " has a bad particle or row index");
            const Constraint& c = cons[row];
            uint8_t side;
            if (c.i == i && c.j == other) side = 1;
            else if (c.j == i && c.i == other) side = 2;
            else return bad("neighbor_packed entry " + at(k) + " names row " + at(row) + ", which does not join " +
                            at(i) + " and " + at(other));
            if (seen[row] & side) return bad("row " + at(row) + " appears twice in the slice of particle " + at(i));
            seen[row] |= side;
            if (std::memcmp(&q[1], &c.rest, 4) != 0 || std::memcmp(&q[2], &c.comp, 4) != 0)
                return bad("neighbor_packed entry " + at(k) + " differs from row " + at(row) + " in rest or compliance");
            m->neighbor_other[k] = other;
            m->neighbor_rest_len[k] = q[1];
            m->neighbor_compliance[k] = q[2];
            m->neighbor_constraint_id[k] = row;
        }
    return m.release();   // 2E entries, each row seen at most once per side: all seen
}

// derive_collision_grid (mesh_prep.rs:16-85).
void derive_collision_grid(const float* rest, uint32_t n, float th, xpbd_coll_grid* g) {
    if (n == 0) {
        *g = xpbd_coll_grid{{0.f, 0.f, 0.f}, 1.f, {1, 1, 1}, 1, 1};
        return;
    }
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::fmin(mn[a], rest[4 * size_t(i) + a]);
            mx[a] = std::fmax(mx[a], rest[4 * size_t(i) + a]);
        }
    const float margin = std::fma(th, 6.0f, th * 0.5f);
    float ext[3];
    for (int a = 0; a < 3; ++a) {
        mn[a] = mn[a] - margin;
        mx[a] = mx[a] + margin;
        ext[a] = std::fmax(mx[a] - mn[a], 1e-4f);
    }
    const float target = std::min(std::max(std::sqrt(float(n)) * 6.25f, 256.f), 32768.f);
    const float vol = std::fmax(ext[0] * ext[1] * ext[2], 1e-9f);
    float side = std::fmax(std::fmax(std::pow(vol / target, 1.0f / 3.0f), th * 2.5f), 1e-5f);
    auto axis = [](float v) -> uint32_t {   // (v.ceil() as u32).clamp(1, 96)
        uint32_t u = !(v > 0.f) ? 0u : (v >= 4294967296.0f ? 0xFFFFFFFFu : uint32_t(v));
        return std::min(std::max(u, 1u), 96u);
    };
    uint32_t d[3];
    for (;;) {
        for (int a = 0; a < 3; ++a) d[a] = axis(std::ceil(ext[a] / side));
        const uint64_t cells = uint64_t(d[0]) * d[1] * d[2];
        if (cells <= 32768 && cells >= 1) break;
        side *= 1.25f;
        const float big = std::fmax(std::fmax(ext[0], ext[1]), ext[2]);
        if (side > big * 2.5f) {
            d[0] = d[1] = d[2] = 1;
            side = std::fmax(big, th * 2.5f);
            break;
        }
    }
    const uint32_t cells = d[0] * d[1] * d[2];
    const uint32_t m1 = cells - 1;
    uint32_t digits = 1;
    if (m1) digits = std::min(std::max((32u - uint32_t(__builtin_clz(m1)) + 7u) / 8u, 1u), 4u);
    *g = xpbd_coll_grid{{mn[0], mn[1], mn[2]}, 1.0f / side, {d[0], d[1], d[2]}, cells, digits};
}

}  // namespace xpbd
