// Topology preparation on the GPU: the greedy colouring, the colour sort and
// the CSR scatter of finalize_cloth_mesh (mesh_prep.rs:153-221, 580-635) for a
// caller's constraint rows, with the same results as the host path
// (assemble_constraints in mesh_prep.cpp).
//
//  1. degrees (integer atomics);
//  2. the greedy processing order -- rows by (max endpoint degree descending,
//     min endpoint, max endpoint, row index) -- as stable LSD radix sorts,
//     least significant key first, with the cloth as the most significant key
//     so each cloth's rows are contiguous;
//  3. the greedy itself.  A row's colour depends only on earlier rows sharing
//     an endpoint, so on rows of its own cloth: one wavefront per cloth walks
//     that cloth's rows in the processing order, 64 at a time.  The lanes load
//     their rows' endpoint colour masks, then are settled in order: lane t
//     takes the smallest colour free at both of its ends, and every lane
//     sharing an end with it marks that colour.  Lane by lane this is the
//     sequential greedy; the masks go back to memory before the next 64;
//  4. the colour sort -- rows by (colour, row index) -- as a stable radix sort;
//  5. the CSR: the directed entries (row, side) stably sorted by particle,
//     which is the cursor scatter's order; offsets from the run starts.
// Colours are limited to 64 (u64 masks); a mesh whose degrees allow more is
// refused before any launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "mesh.hpp"

namespace xpbd {
namespace {

// Rows as uint4: (i, j, rest bits, compliance bits) -- the layout of Constraint.
static_assert(sizeof(Constraint) == 16, "Constraint must be 16 bytes");

__global__ void k_topo_degree(const uint4* rows, uint32_t e, uint32_t* deg) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < e; k += gridDim.x * blockDim.x) {
        const uint4 r = rows[k];
        atomicAdd(deg + r.x, 1u);
        atomicAdd(deg + r.y, 1u);
    }
}

__global__ void k_topo_max(const uint32_t* v, uint32_t n, uint32_t* out) {
    uint32_t m = 0;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) m = max(m, v[k]);
    for (int off = 32; off > 0; off >>= 1) m = max(m, uint32_t(__shfl_xor(int(m), off, 64)));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void k_topo_cloth_of(const uint32_t* offsets, uint32_t* cloth_of) {
    const uint32_t c = blockIdx.x;
    for (uint32_t v = offsets[c] + threadIdx.x; v < offsets[c + 1]; v += blockDim.x) cloth_of[v] = c;
}

enum : uint32_t { KEY_MAX_END = 0, KEY_MIN_END = 1, KEY_DEGREE = 2, KEY_CLOTH = 3, KEY_COLOUR = 4 };

// keys[p] of row k = ids[p] (ids == nullptr: k = p, and ids_out[p] = p)
__global__ void k_topo_keys(const uint4* rows, const uint32_t* ids, uint32_t e, uint32_t mode, const uint32_t* deg,
                            uint32_t maxdeg, const uint32_t* aux, uint32_t* keys, uint32_t* ids_out) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < e; p += gridDim.x * blockDim.x) {
        const uint32_t k = ids ? ids[p] : p;
        if (!ids) ids_out[p] = p;
        const uint4 r = rows[k];
        uint32_t key;
        switch (mode) {
            case KEY_MAX_END: key = max(r.x, r.y); break;
            case KEY_MIN_END: key = min(r.x, r.y); break;
            case KEY_DEGREE: key = maxdeg - max(deg[r.x], deg[r.y]); break;
            case KEY_CLOTH: key = aux[r.x]; break;
            default: key = aux[k]; break;   // KEY_COLOUR: colour of row k
        }
        keys[p] = key;
    }
}

// The greedy walk; one 64-lane block per cloth.  M: colour mask word (u32 when
// at most 32 colours are possible, else u64).  The next 64 rows' endpoints
// are loaded while the current ones are settled.  stat[0] = 1 if some row
// found no colour below the mask width, stat[1] = largest colour used.
template <class M>
__global__ __launch_bounds__(64) void k_topo_greedy(const uint32_t* order, const uint32_t* start, const uint4* rows,
                                                    M* mask, uint32_t* colour, uint32_t* stat) {
    constexpr uint32_t WIDTH = 8 * sizeof(M);
    const uint32_t lane = threadIdx.x;
    const uint32_t r0 = start[blockIdx.x], r1 = start[blockIdx.x + 1];
    uint32_t cmax = 0;
    bool over = false;
    // software pipeline: the endpoints of this lane's row of the current 64
    // (k, i, j) and the row index of the next 64 (kn) are in registers; each
    // pass loads the next rows' endpoints, the masks (needed now) and the row
    // indices two batches ahead, unconditionally (positions clamped into the
    // cloth's range; lanes past its end compute on a copy of its last row and
    // store nothing), so that only the mask loads and the store drain wait
    if (r0 >= r1) return;
    const uint32_t last = r1 - 1;
    uint32_t k = order[min(r0 + lane, last)];
    uint4 q0 = rows[k];
    uint32_t i = q0.x, j = q0.y;
    uint32_t kn = order[min(r0 + 64 + lane, last)];
    for (uint32_t base = r0; base < r1; base += 64) {
        const bool valid = base + lane < r1;
        const uint4 qn = rows[kn];
        M mi = __hip_atomic_load(mask + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        M mj = __hip_atomic_load(mask + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t kn2 = order[min(base + 128 + lane, last)];
        const uint32_t cnt = min(64u, r1 - base);
        uint32_t col = 0;
        for (uint32_t t = 0; t < cnt; ++t) {
            const M freeb = M(~(mi | mj));
            const uint32_t mine = freeb ? uint32_t(sizeof(M) == 8 ? __builtin_ctzll(uint64_t(freeb))
                                                                  : __builtin_ctz(uint32_t(freeb)))
                                        : WIDTH;
            const uint32_t ct = __builtin_amdgcn_readlane(mine, t);
            const uint32_t it = __builtin_amdgcn_readlane(i, t);
            const uint32_t jt = __builtin_amdgcn_readlane(j, t);
            if (lane == t) col = ct;
            // no colour left (ct == WIDTH): marks nothing; reported below
            const M bit = ct < WIDTH ? M(M(1) << (ct & (WIDTH - 1))) : M(0);
            if (i == it || i == jt) mi |= bit;
            if (j == it || j == jt) mj |= bit;
        }
        if (valid) {
            over |= col >= WIDTH;
            colour[k] = col;
            cmax = max(cmax, col);
            // every lane sharing a particle holds the same final mask for it
            __hip_atomic_store(mask + i, mi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(mask + j, mj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // the masks are in L2 before the next 64 rows read them
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        k = kn;
        i = qn.x;
        j = qn.y;
        kn = kn2;
    }
    over = __ballot(over) != 0;
    for (int off = 32; off > 0; off >>= 1) cmax = max(cmax, uint32_t(__shfl_xor(int(cmax), off, 64)));
    if (lane == 0) {
        atomicMax(stat + 1, cmax);
        if (over) atomicExch(stat, 1u);
    }
}

__global__ void k_topo_gather_rows(const uint4* rows, const uint32_t* perm, uint32_t e, uint4* out) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < e; p += gridDim.x * blockDim.x) out[p] = rows[perm[p]];
}

// directed entry q = 2 row + side: key = its particle; ids = q
__global__ void k_topo_entry_keys(const uint4* rows, uint32_t e2, uint32_t* keys, uint32_t* ids) {
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < e2; q += gridDim.x * blockDim.x) {
        const uint4 r = rows[q >> 1];
        keys[q] = (q & 1u) ? r.y : r.x;
        ids[q] = q;
    }
}

// offsets[v] = first position whose sorted key is >= v, v = 0..n (the run
// starts of the sorted keys: CSR offsets, cloth ranges, colour batches)
__global__ void k_topo_offsets(const uint32_t* skeys, uint32_t e2, uint32_t n, uint32_t* offsets) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < e2; p += gridDim.x * blockDim.x) {
        const uint32_t v = skeys[p];
        const uint32_t from = p ? skeys[p - 1] + 1u : 0u;
        for (uint32_t u = from; u <= v; ++u) offsets[u] = p;
        if (p + 1 == e2)
            for (uint32_t u = v + 1; u <= n; ++u) offsets[u] = e2;
    }
}

__global__ void k_topo_entries(const uint4* rows, const uint32_t* sids, uint32_t e2, uint32_t* other, uint32_t* rest,
                               uint32_t* comp, uint32_t* eid) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < e2; p += gridDim.x * blockDim.x) {
        const uint32_t q = sids[p];
        const uint4 r = rows[q >> 1];
        other[p] = (q & 1u) ? r.x : r.y;
        rest[p] = r.z;
        comp[p] = r.w;
        eid[p] = q >> 1;
    }
}

uint32_t bits_for(uint64_t maxval) {
    uint32_t b = 0;
    while (b < 32 && (maxval >> b)) ++b;
    return b;
}

struct DeviceScratch {
    std::vector<void*> ptrs;
    ~DeviceScratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    bool alloc(T** p, size_t count) {
        void* q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(count * sizeof(T), 256)) != hipSuccess) return false;
        ptrs.push_back(q);
        *p = static_cast<T*>(q);
        return true;
    }
};

dim3 grid_for(size_t work) { return dim3(uint32_t(std::min<size_t>((work + 255) / 256, 8192)), 1, 1); }

// Stable LSD radix sort of (keys, ids) on the low `bits` bits of the keys,
// 8 bits per pass; the result is left in (*k, *i), the other pair is scratch.
void radix_bits(uint32_t** k, uint32_t** i, uint32_t** k2, uint32_t** i2, uint32_t n, uint32_t bits, uint32_t* hist,
                hipStream_t s) {
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        launch_radix_pass(*k, *i, *k2, *i2, n, shift, hist, s);
        std::swap(*k, *k2);
        std::swap(*i, *i2);
    }
}

}  // namespace

int topology_on_device(size_t n, const std::vector<uint32_t>& cloth_offsets, std::vector<Constraint>& rows, Mesh& m,
                       std::string& why) {
    auto bad = [&](const std::string& w, int code) {
        why = w;
        return code;
    };
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return bad("no HIP device for the topology", XPBD_E_NO_DEVICE);
    const size_t e = rows.size(), e2 = 2 * e;
    if (e2 >= (size_t(1) << 32) || n >= (size_t(1) << 31)) return bad("too many rows for the device path", XPBD_E_INVALID);
    const uint32_t nc = cloth_offsets.size() >= 2 ? uint32_t(cloth_offsets.size() - 1) : 1u;
    hipStream_t s = nullptr;
    if (hipStreamCreate(&s) != hipSuccess) return bad("hipStreamCreate failed", XPBD_E_HIP);
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    DeviceScratch d;
    uint4 *rows_d = nullptr, *out_d = nullptr;
    uint32_t *deg = nullptr, *cloth_of = nullptr, *offs_d = nullptr, *stat = nullptr, *colour = nullptr;
    uint32_t *ka = nullptr, *ia = nullptr, *kb = nullptr, *ib = nullptr, *hist = nullptr, *counts = nullptr;
    uint32_t *start = nullptr, *noff = nullptr, *nother = nullptr, *nrest = nullptr, *ncomp = nullptr, *neid = nullptr;
    uint64_t* mask = nullptr;
    const size_t big = std::max<size_t>(e2, 1);
    if (!d.alloc(&rows_d, std::max<size_t>(e, 1)) || !d.alloc(&out_d, std::max<size_t>(e, 1)) || !d.alloc(&deg, n) ||
        !d.alloc(&cloth_of, n) || !d.alloc(&offs_d, nc + 1) || !d.alloc(&stat, 4) || !d.alloc(&colour, std::max<size_t>(e, 1)) ||
        !d.alloc(&ka, big) || !d.alloc(&ia, big) || !d.alloc(&kb, big) || !d.alloc(&ib, big) ||
        !d.alloc(&hist, radix_hist_words(uint32_t(big))) || !d.alloc(&counts, std::max<uint32_t>(nc, 65)) ||
        !d.alloc(&start, nc + 1) || !d.alloc(&noff, n + 1) || !d.alloc(&nother, big) || !d.alloc(&nrest, big) ||
        !d.alloc(&ncomp, big) || !d.alloc(&neid, big) || !d.alloc(&mask, n))
        return bad("device allocation for the topology failed", XPBD_E_NOMEM);
#define TOPO_HIP(x)                                                               \
    do {                                                                          \
        hipError_t err_ = (x);                                                    \
        if (err_ != hipSuccess) return bad(hipGetErrorString(err_), XPBD_E_HIP); \
    } while (0)
    const uint32_t E = uint32_t(e), E2 = uint32_t(e2), N = uint32_t(n);
    TOPO_HIP(hipMemcpyAsync(rows_d, rows.data(), e * sizeof(Constraint), hipMemcpyHostToDevice, s));
    TOPO_HIP(hipMemsetAsync(deg, 0, n * sizeof(uint32_t), s));
    TOPO_HIP(hipMemsetAsync(stat, 0, 4 * sizeof(uint32_t), s));
    TOPO_HIP(hipMemsetAsync(mask, 0, n * sizeof(uint64_t), s));
    if (e) hipLaunchKernelGGL(k_topo_degree, grid_for(e), dim3(256), 0, s, rows_d, E, deg);
    hipLaunchKernelGGL(k_topo_max, grid_for(n), dim3(256), 0, s, deg, N, stat + 2);
    if (nc > 1) {
        TOPO_HIP(hipMemcpyAsync(offs_d, cloth_offsets.data(), (nc + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_topo_cloth_of, dim3(nc), dim3(256), 0, s, offs_d, cloth_of);
    }
    uint32_t st[4] = {0, 0, 0, 0};
    TOPO_HIP(hipMemcpyAsync(st, stat, sizeof(st), hipMemcpyDeviceToHost, s));
    TOPO_HIP(hipStreamSynchronize(s));
    const uint32_t maxdeg = st[2];
    // the greedy takes the smallest colour free at both ends: at most
    // (deg_i - 1) + (deg_j - 1) are taken, so 2 maxdeg - 1 colours suffice
    if (e && 2ull * maxdeg - 1 > 64) return bad("degrees allow more than 64 colours: use the host path", XPBD_E_UNSUPPORTED);

    // processing order: stable sorts, least significant key first
    uint32_t *k = ka, *i = ia, *k2 = kb, *i2 = ib;
    if (e) {
        const uint32_t vb = bits_for(n ? n - 1 : 0);
        hipLaunchKernelGGL(k_topo_keys, grid_for(e), dim3(256), 0, s, rows_d, nullptr, E, uint32_t(KEY_MAX_END), deg,
                           maxdeg, nullptr, k, i);
        radix_bits(&k, &i, &k2, &i2, E, vb, hist, s);
        hipLaunchKernelGGL(k_topo_keys, grid_for(e), dim3(256), 0, s, rows_d, i, E, uint32_t(KEY_MIN_END), deg, maxdeg,
                           nullptr, k, nullptr);
        radix_bits(&k, &i, &k2, &i2, E, vb, hist, s);
        hipLaunchKernelGGL(k_topo_keys, grid_for(e), dim3(256), 0, s, rows_d, i, E, uint32_t(KEY_DEGREE), deg, maxdeg,
                           nullptr, k, nullptr);
        radix_bits(&k, &i, &k2, &i2, E, bits_for(maxdeg), hist, s);
        if (nc > 1) {
            hipLaunchKernelGGL(k_topo_keys, grid_for(e), dim3(256), 0, s, rows_d, i, E, uint32_t(KEY_CLOTH), deg,
                               maxdeg, cloth_of, k, nullptr);
            radix_bits(&k, &i, &k2, &i2, E, bits_for(nc - 1), hist, s);
        }
    }
    uint32_t* order = i;
    // each cloth's range of the order: the run starts of the sorted cloth keys
    if (nc > 1 && e) {
        hipLaunchKernelGGL(k_topo_offsets, grid_for(e), dim3(256), 0, s, k, E, nc, start);
    } else {
        const uint32_t cstart[2] = {0, E};
        TOPO_HIP(hipMemcpyAsync(start, cstart, sizeof(cstart), hipMemcpyHostToDevice, s));
    }
    if (e) {
        if (2ull * maxdeg - 1 <= 32)
            hipLaunchKernelGGL(k_topo_greedy<uint32_t>, dim3(nc), dim3(64), 0, s, order, start, rows_d,
                               reinterpret_cast<uint32_t*>(mask), colour, stat);
        else
            hipLaunchKernelGGL(k_topo_greedy<uint64_t>, dim3(nc), dim3(64), 0, s, order, start, rows_d, mask, colour,
                               stat);
    }
    TOPO_HIP(hipGetLastError());
    TOPO_HIP(hipMemcpyAsync(st, stat, sizeof(st), hipMemcpyDeviceToHost, s));
    TOPO_HIP(hipStreamSynchronize(s));
    if (st[0]) return bad("a row found no colour below 64: use the host path", XPBD_E_UNSUPPORTED);
    const uint32_t ncol = e ? st[1] + 1 : 0;

    // colour sort: rows by (colour, row index); batch offsets from the counts
    std::vector<uint32_t> ccnt(ncol + 1, 0);   // colour batch starts
    uint32_t* perm = nullptr;
    if (e) {
        k = ka, i = ia, k2 = kb, i2 = ib;
        hipLaunchKernelGGL(k_topo_keys, grid_for(e), dim3(256), 0, s, rows_d, nullptr, E, uint32_t(KEY_COLOUR), deg,
                           maxdeg, colour, k, i);
        radix_bits(&k, &i, &k2, &i2, E, bits_for(ncol - 1), hist, s);
        perm = i;
        hipLaunchKernelGGL(k_topo_offsets, grid_for(e), dim3(256), 0, s, k, E, ncol, counts);
        TOPO_HIP(hipMemcpyAsync(ccnt.data(), counts, (ncol + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        hipLaunchKernelGGL(k_topo_gather_rows, grid_for(e), dim3(256), 0, s, rows_d, perm, E, out_d);
    }
    // CSR: directed entries stably by particle (the cursor scatter's order)
    if (e) {
        // the gather above has read perm before these kernels run (one stream)
        k = ka, i = ia, k2 = kb, i2 = ib;
        hipLaunchKernelGGL(k_topo_entry_keys, grid_for(e2), dim3(256), 0, s, out_d, E2, k, i);
        radix_bits(&k, &i, &k2, &i2, E2, bits_for(n ? n - 1 : 0), hist, s);
        hipLaunchKernelGGL(k_topo_offsets, grid_for(e2), dim3(256), 0, s, k, E2, N, noff);
        hipLaunchKernelGGL(k_topo_entries, grid_for(e2), dim3(256), 0, s, out_d, i, E2, nother, nrest, ncomp, neid);
    } else {
        TOPO_HIP(hipMemsetAsync(noff, 0, (n + 1) * sizeof(uint32_t), s));
    }
    TOPO_HIP(hipGetLastError());

    // back to the host mesh
    std::vector<Constraint> sorted(e);
    if (e) TOPO_HIP(hipMemcpyAsync(sorted.data(), out_d, e * sizeof(Constraint), hipMemcpyDeviceToHost, s));
    m.neighbor_offsets.resize(n + 1);
    m.neighbor_other.resize(e2);
    m.neighbor_rest_len.resize(e2);
    m.neighbor_compliance.resize(e2);
    m.neighbor_constraint_id.resize(e2);
    TOPO_HIP(hipMemcpyAsync(m.neighbor_offsets.data(), noff, (n + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    if (e) {
        TOPO_HIP(hipMemcpyAsync(m.neighbor_other.data(), nother, e2 * 4, hipMemcpyDeviceToHost, s));
        TOPO_HIP(hipMemcpyAsync(m.neighbor_rest_len.data(), nrest, e2 * 4, hipMemcpyDeviceToHost, s));
        TOPO_HIP(hipMemcpyAsync(m.neighbor_compliance.data(), ncomp, e2 * 4, hipMemcpyDeviceToHost, s));
        TOPO_HIP(hipMemcpyAsync(m.neighbor_constraint_id.data(), neid, e2 * 4, hipMemcpyDeviceToHost, s));
    }
    TOPO_HIP(hipStreamSynchronize(s));
#undef TOPO_HIP
    m.constraint_i.resize(e);
    m.constraint_j.resize(e);
    m.constraint_rest.resize(e);
    m.constraint_comp.resize(e);
    for (size_t q = 0; q < e; ++q) {
        m.constraint_i[q] = sorted[q].i;
        m.constraint_j[q] = sorted[q].j;
        m.constraint_rest[q] = sorted[q].rest;
        m.constraint_comp[q] = sorted[q].comp;
    }
    if (e == 0) {
        m.batch_offsets = {0, 0};
    } else {
        m.batch_offsets = ccnt;
    }
    std::vector<Constraint>().swap(rows);
    return XPBD_OK;
}

}  // namespace xpbd
