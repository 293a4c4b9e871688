// HIP kernels for the XPBD cloth hot path on MI355X (gfx950, wave64).
//
// Arithmetic restates the reference CPU solver (src/xpbd_cpu.rs) operation
// by operation: glam's scalar Vec3 order (dot = (x*x + y*y) + z*z,
// length = sqrtf(dot), component-wise '/' and '*'), correctly rounded sqrtf and
// division (HIP's default lowering), and no FMA contraction (this file is
// compiled with -ffp-contract=off).  With that, every kernel reproduces the
// f32 CPU solver bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "kernels.hpp"

#pragma clang fp contract(off)

namespace xpbd {

// Reference constants (cloth_compute.rs:82-90).
constexpr float JACOBI_CORRECTION_CAP = 0.28f;
constexpr float PREDICT_MAX_SPEED = 12.0f;
constexpr float GRAB_MAX_PULL = 0.065f;
constexpr float COLLISION_APPLY_CLAMP = 0.35f;
constexpr uint32_t EMPTY_KEY = 0xFFFFFFFFu;

__device__ __forceinline__ float sq3(float x, float y, float z) { return (x * x + y * y) + z * z; }
__device__ __forceinline__ float len3(float x, float y, float z) { return sqrtf(sq3(x, y, z)); }

// The correction cap's test without the square root: the correctly rounded
// square root is monotone, so sqrtf(s) > JACOBI_CORRECTION_CAP (which implies
// > 0) exactly when s > CAP_SQ_LIMIT, the largest float whose correctly
// rounded root is at most the cap (0x1.41205ep-4; NaN fails both tests).
// The root is then taken only for a capped correction.
constexpr float CAP_SQ_LIMIT = 0x1.41205ep-4f;

// (x, y, z) scaled to the cap where its length exceeds it (xpbd_cpu.rs's
// `if len > cap { v *= cap / len }`, same operations).
__device__ __forceinline__ void cap_correction(float& x, float& y, float& z) {
    const float s = sq3(x, y, z);
    if (s > CAP_SQ_LIMIT) {
        const float sc = JACOBI_CORRECTION_CAP / sqrtf(s);
        x = x * sc;
        y = y * sc;
        z = z * sc;
    }
}

// predict_particle (xpbd_cpu.rs:38-65); the inverse mass rides in .w.
__device__ __forceinline__ float4 predict(float4 x, float4 v, uint32_t i, const DevParams& p) {
    if (x.w <= 0.f) return x;
    float vx = v.x + p.gx * p.dt;
    float vy = v.y + p.gy * p.dt;
    float vz = v.z + p.gz * p.dt;
    const float speed = len3(vx, vy, vz);
    if (speed > PREDICT_MAX_SPEED && speed > 0.f) {
        const float s = PREDICT_MAX_SPEED / speed;
        vx = vx * s;
        vy = vy * s;
        vz = vz * s;
    }
    float px = x.x + vx * p.dt;
    float py = x.y + vy * p.dt;
    float pz = x.z + vz * p.dt;
    if (p.grab_active && int32_t(i) == p.grab_idx) {
        float ux = (p.grab_x - px) * p.grab_k;
        float uy = (p.grab_y - py) * p.grab_k;
        float uz = (p.grab_z - pz) * p.grab_k;
        const float pl = len3(ux, uy, uz);
        if (pl > GRAB_MAX_PULL && pl > 0.f) {
            const float s = GRAB_MAX_PULL / pl;
            ux = ux * s;
            uy = uy * s;
            uz = uz * s;
        }
        px = px + ux;
        py = py + uy;
        pz = pz + uz;
    }
    return make_float4(px, py, pz, x.w);
}

// Ground plane y = 0 then post_velocity (xpbd_cpu.rs:393-409) for one free
// particle: v = ((x - prev) * inv_dt) * damp.
__device__ __forceinline__ void finish_particle(float4 nx, float4 prev, uint32_t i, const uint8_t* ground,
                                                const DevParams& p, float4* X, float4* V) {
    if (ground && ground[i] && nx.y < 0.f) nx.y = 0.f;
    X[i] = nx;
    V[i] = make_float4(((nx.x - prev.x) * p.inv_dt) * p.damp, ((nx.y - prev.y) * p.inv_dt) * p.damp,
                       ((nx.z - prev.z) * p.inv_dt) * p.damp, 0.f);
}

// Block-wide bounds of the rows each thread holds, written as one partial
// (lo xyz, hi xyz) per block; the grid kernel reduces the partials.  -0 is
// folded to +0 and NaN ignored, as the oracle does.
__device__ __forceinline__ void aabb_fold(bool valid, float x, float y, float z, float lo[3], float hi[3]) {
    const float v[3] = {x + 0.f, y + 0.f, z + 0.f};
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (valid && v[a] == v[a]) {
            lo[a] = fminf(lo[a], v[a]);
            hi[a] = fmaxf(hi[a], v[a]);
        }
}

// v from lane (per DPP_CTRL) of the same wavefront; a lane whose source does
// not exist, or whose row ROW_MASK leaves out, gets its own v.
template <int DPP_CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_or_self(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), DPP_CTRL, ROW_MASK, 0xF, false));
}

// Wavefront bounds into lane 63 (every lane must be active): row_shr 1, 2,
// 4, 8 leave each row's extremum in its lane 15, row_bcast 15 and 31 carry
// them up to lane 63.  Each step is one DPP min / max (no LDS crossbar);
// min / max are exact, so the order does not matter.
__device__ __forceinline__ void wave_bounds_to_lane63(float lo[3], float hi[3]) {
#define XPBD_DPP_STEP(CTRL, ROWS)                                     \
    _Pragma("unroll") for (int a = 0; a < 3; ++a) {                   \
        lo[a] = fminf(lo[a], dpp_or_self<CTRL, ROWS>(lo[a]));         \
        hi[a] = fmaxf(hi[a], dpp_or_self<CTRL, ROWS>(hi[a]));         \
    }
    XPBD_DPP_STEP(0x111, 0xF)   // row_shr:1
    XPBD_DPP_STEP(0x112, 0xF)   // row_shr:2
    XPBD_DPP_STEP(0x114, 0xF)   // row_shr:4
    XPBD_DPP_STEP(0x118, 0xF)   // row_shr:8
    XPBD_DPP_STEP(0x142, 0xA)   // row_bcast:15 into rows 1 and 3
    XPBD_DPP_STEP(0x143, 0xC)   // row_bcast:31 into rows 2 and 3
#undef XPBD_DPP_STEP
}

__device__ __forceinline__ void block_aabb_store(float lo[3], float hi[3], float* part, uint32_t slot) {
    wave_bounds_to_lane63(lo, hi);
    __shared__ float red[16][6];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63)
        for (int a = 0; a < 3; ++a) {
            red[wave][a] = lo[a];
            red[wave][3 + a] = hi[a];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        const uint32_t nw = (blockDim.x + 63) >> 6;
        const int a = threadIdx.x;
        float r = red[0][a];
        for (uint32_t w = 1; w < nw; ++w) r = a < 3 ? fminf(r, red[w][a]) : fmaxf(r, red[w][a]);
        part[size_t(slot) * 8 + a] = r;
    }
    __syncthreads();
}

// Blocks are dealt round-robin to the 8 XCDs; this makes consecutive logical
// tiles land on one XCD so the +-2-row neighbour halo hits that XCD's L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb) {
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u, k = b >> 3;
    return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// ---------------------------------------------------------------------------
// Jacobi iteration: atomic-free per-particle gather.  The reference runs
// jacobi_edges (per constraint dlambda, cloth_sim_jacobi.wgsl:143-180) and then
// jacobi_gather (cloth_sim_jacobi.wgsl:182-218); since lambda is zero at
// every read (xpbd_cpu.rs:267-268), dlambda = (-C - alpha~*0) / (w_i + w_j + alpha~)
// depends only on the pair, and each endpoint recomputes it bit-identically
// (|a-b| == |b-a|, w_i + w_j == w_j + w_i).  One pass, no lambda buffers.
// ---------------------------------------------------------------------------
// Correctly rounded division without the general div_scale/div_fixup
// expansion, valid where no operand or result leaves the normal range.
// Fast correctly rounded division and square root for "tame" iterates.
//
// y = RN(1/b) from the hardware reciprocal and one FMA Newton step; then
// q = RN(a*y), r = a - b*q (exact with FMA), RN(q + r*y) = RN(a/b)
// (Markstein's theorem: y the correctly rounded reciprocal, q within one ulp,
// no intermediate under- or overflow).  sqrt_rn: the hardware square root is
// within one ulp; of its two neighbours, the one whose product with it lies on
// the other side of x is taken (Tuckerman's rounding test, the residuals
// x - s_prev*s and x - s_next*s exact with FMA).  Both are checked on the
// device by xpbd_selftest_division / xpbd_selftest_sqrt (reciprocal and
// square root exhaustively over the exponents used, the quotient on random
// operands across the operand ranges below).
//
// When they apply: every coordinate of the slices an entry reads is +0 or of
// magnitude in [2^-50, 2^20] (tracked per slice on the device; -0 is not
// tame), and the host has checked inverse masses <= 2^40, compliance / dt^2
// <= 2^40 and rest lengths 0 or in [2^-30, 2^30].  Then:
// - a difference component is +0 (never -0: x - y = -0 needs x = -0) or of
//   magnitude in [2^-73, 2^21], so its quotient by len needs no sign-of-zero
//   fix-up (div_rn_nz);
// - a squared length below 2^-100 has a root below 2^-50, which the hardware
//   root (within one ulp from about 2^-101 up, not exact below) keeps far
//   under 1e-8: the entry is skipped by the 1e-8 test on either path; above,
//   the root is exact;
// - past the 1e-8 test, len is in [1e-8, 2^22]; -C is 0 or in [2^-54, 2^30]
//   (a multiple of the smaller ulp of len and rest); the weight sum is in
//   [1e-8, 2^42];
// - every quotient is 0 or in [2^-96, 2^57] and every Markstein
//   intermediate normal.
__device__ __forceinline__ float rcp_rn(float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y0, 1.0f);
    return __builtin_fmaf(e, y0, y0);
}
__device__ __forceinline__ float div_rn(float a, float b, float y) {
    const float q = a * y;
    const float r = __builtin_fmaf(-b, q, a);
    const float v = __builtin_fmaf(r, y, q);
    return q == 0.f ? q : v;
}
// a != -0 (see above): a = +0 gives q = +0, r = +0 and +0; otherwise q != 0
__device__ __forceinline__ float div_rn_nz(float a, float b, float y) {
    const float q = a * y;
    const float r = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(r, y, q);
}
__device__ __forceinline__ float sqrt_rn(float x) {   // exact for x == 0 or x >= 2^-100
    float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    const float rd = __builtin_fmaf(-sd, s, x);
    const float ru = __builtin_fmaf(-su, s, x);
    s = rd <= 0.f ? sd : s;
    s = ru > 0.f ? su : s;
    return s;
}

// Slice-tameness byte of the wavefront's 64 outputs (every lane of the wave
// must reach this; `mine` = false only for a written value out of range).
__device__ __forceinline__ void write_slice_tame(uint8_t* out, uint32_t slice, bool in_range, bool mine) {
    const int all = __all(mine ? 1 : 0);
    if (out && in_range && (threadIdx.x & 63) == 0) out[slice] = all ? 1u : 0u;
}

// All slices holding neighbours of `slice` tame in the input iterate
// (wavefront-uniform).  Ranges wider than 128 slices never qualify.
__device__ __forceinline__ bool neighbourhood_tame(const uint8_t* tame, const uint2* srange, uint32_t slice,
                                                   uint32_t nslices) {
    if (!tame || slice >= nslices) return false;
    const uint2 rg = srange[slice];
    const uint32_t lane = threadIdx.x & 63;
    bool ok = rg.x <= rg.y && rg.y - rg.x < 128u;
    if (ok) {
        const uint32_t k0 = rg.x + lane, k1 = rg.x + lane + 64u;
        if (k0 <= rg.y && tame[k0] == 0) ok = false;
        if (k1 <= rg.y && tame[k1] == 0) ok = false;
    }
    return __all(ok ? 1 : 0) != 0;
}

// One constraint entry of the gather: accumulates (n * w_i) * dlambda in the
// reference's order of operations (xpbd_cpu.rs:159-182, 206-222).  Called only
// for w_i > 0, so the reference's both-pinned skip never applies.  pr is the
// entry's pair as staged: (rest, alpha~ = compliance / dt^2, alpha~ * 0).
// FAST: the input is tame (see above) and the shorter correctly rounded
// operations are exact; otherwise the compiler's general ones.
template <bool FAST>
__device__ __forceinline__ void jacobi_entry(float4 pi, float4 pj, float4 pr, float& ax, float& ay, float& az) {
    const float wi = pi.w, wj = pj.w;
    float gx = pi.x - pj.x, gy = pi.y - pj.y, gz = pi.z - pj.z;
    const float dot = (gx * gx + gy * gy) + gz * gz;
    const float len = FAST ? sqrt_rn(dot) : sqrtf(dot);
    if (len < 1e-8f) return;
    const float c = len - pr.x;
    const float wsum = wi + wj + pr.y;
    // fast slices only run on meshes whose inverse masses are 0 or >= 2^-26
    // and whose compliances are >= 0: there wsum >= w_i >= 2^-26 > 1e-8
    if (!FAST && wsum < 1e-8f) return;
    // -C - alpha~ * lambda with lambda = 0: pr.z = alpha~ * 0 is +0 whenever
    // alpha~ is finite (always on fast slices: alpha~ <= 2^40), and x - (+0)
    // is x for every x, so the fast form drops the subtraction
    const float num = FAST ? -c : -c - pr.z;
    float dlam;
    if (FAST) {
        const float yl = rcp_rn(len);
        gx = div_rn_nz(gx, len, yl);
        gy = div_rn_nz(gy, len, yl);
        gz = div_rn_nz(gz, len, yl);
        // -C = -0 (len == rest) gives +0 here where '/' gives -0.  dlam only
        // enters ax + (g * w) * dlam, with finite g * w on fast slices, so
        // the term is a zero either way; the accumulators start at +0 and a
        // round-to-nearest sum is -0 only when both addends are, so they are
        // never -0 and adding either zero leaves them unchanged.
        dlam = div_rn_nz(num, wsum, rcp_rn(wsum));
    } else {
        gx = gx / len;
        gy = gy / len;
        gz = gz / len;
        dlam = num / wsum;
    }
    ax = ax + (gx * wi) * dlam;
    ay = ay + (gy * wi) * dlam;
    az = az + (gz * wi) * dlam;
}

// (rest, compliance) -> the staged pair (rest, alpha~, alpha~ * 0): the same
// float operations the reference does per constraint, done once per pair.
__device__ __forceinline__ float4 stage_pair(float2 rc, const DevParams& p) {
    const float alpha_t = rc.y * p.inv_dt_sq;
    return make_float4(rc.x, alpha_t, alpha_t * 0.0f, 0.f);
}

// Entries are processed JU at a time: all adjacency loads of the chunk, then
// all neighbour gathers, then the arithmetic in entry order — several load
// chains in flight per wavefront.  Padding entries point at the particle
// itself (zero length: skipped by the reference's len < 1e-8 rule), so no
// load is conditional.  The slice width is wavefront-uniform.
constexpr uint32_t JU = 6;
// The tiled kernel's chunk: its neighbours come from LDS, so shorter chunks
// (fewer words held in registers) win there: 2 was fastest of 2..6 and 12
// on the 1024^2 sheet, while the per-slice kernel's global gathers want 6.
constexpr uint32_t JT_JU = 2;
constexpr uint32_t LDS_PAIRS_MAX = 2048;   // (rest, compliance) table staged in LDS up to this size (static)
constexpr uint32_t LDS_PAIRS_DYN_MAX = 4096;   // k_jacobi: table sized to the mesh's pairs (dynamic LDS, <= 32 KB)

// The staged pair of packed entry word w.  Two LDS table forms: (rest,
// compliance) as given (fewer registers: the per-slice kernels at the 64-VGPR
// cap), or already staged (two fewer operations per entry: the tiled kernel).
// LDSP: the table is in LDS (a compile-time choice: a run-time one made the
// compiler merge both into generic flat loads waited on per entry).
template <bool LDSP>
__device__ __forceinline__ float4 entry_pair(const JacobiArgs& a, const float2* lpairs, uint32_t w) {
    if constexpr (LDSP) return stage_pair(lpairs[w >> 16], a.p);
    else return stage_pair(a.sell.pairs[w >> 16], a.p);
}
template <bool LDSP>
__device__ __forceinline__ float4 entry_pair(const JacobiArgs& a, const float4* lpairs, uint32_t w) {
    if constexpr (LDSP) return lpairs[w >> 16];
    else return stage_pair(a.sell.pairs[w >> 16], a.p);
}
__device__ __forceinline__ void to_lds(float2& dst, float2 rc, const DevParams&) { dst = rc; }
__device__ __forceinline__ void to_lds(float4& dst, float2 rc, const DevParams& p) { dst = stage_pair(rc, p); }

template <bool PACKED, bool LDSP>
__device__ __forceinline__ void load_entry(const JacobiArgs& a, const float2* lpairs, uint32_t i, uint32_t e,
                                           uint32_t& j, float4& pr) {
    if (PACKED) {
        const uint32_t w = __builtin_nontemporal_load(a.sell.packed + e);
        j = uint32_t(int32_t(i) + int32_t(int16_t(w & 0xFFFFu)));
        pr = entry_pair<LDSP>(a, lpairs, w);
    } else {
        j = __builtin_nontemporal_load(a.sell.j + e);
        const float rest = __builtin_nontemporal_load(a.sell.rest + e);
        const float comp = __builtin_nontemporal_load(a.sell.comp + e);
        pr = stage_pair(make_float2(rest, comp), a.p);
    }
}

// One particle of a Jacobi iteration (the body shared by the per-iteration
// kernel and the persistent one).  Returns the new position; writes it (or,
// with JM_FINAL, the new state and velocity).
template <uint32_t MODE, bool PACKED, bool FASTDIV, bool LDSP>
__device__ __forceinline__ float4 jacobi_particle(const JacobiArgs& a, const float2* lpairs, uint32_t slice,
                                                  uint32_t lane, bool& valid) {
    const uint32_t i = slice * 64 + lane;
    const DevParams& p = a.p;
    valid = slice < p.nslices && i < p.n;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 out = pi;
    if (!valid) return out;
    if (MODE & JM_PREDICT) pi = predict(a.src[i], a.vel[i], i, p);
    else pi = a.src[i];
    out = pi;
    const float wi = pi.w;
    if (wi > 0.f) {
        float ax = 0.f, ay = 0.f, az = 0.f;
        const uint32_t base = a.sell.soff[slice] + lane;
        const uint32_t width = a.sell.swid[slice];
        uint32_t k0 = 0;
        if (PACKED && !(MODE & JM_PREDICT) && width >= JU) {
            // software-pipelined: the next chunk's packed words load while
            // this chunk gathers and computes (37.4 -> 33.8 us per iteration
            // on the 1024^2 sheet, where the entry loads' HBM latency showed)
            uint32_t wn[JU];
#pragma unroll
            for (uint32_t u = 0; u < JU; ++u) wn[u] = __builtin_nontemporal_load(a.sell.packed + base + u * 64);
            for (; k0 + JU <= width; k0 += JU) {
                uint32_t wc[JU];
#pragma unroll
                for (uint32_t u = 0; u < JU; ++u) wc[u] = wn[u];
                if (k0 + 2 * JU <= width) {
#pragma unroll
                    for (uint32_t u = 0; u < JU; ++u)
                        wn[u] = __builtin_nontemporal_load(a.sell.packed + base + (k0 + JU + u) * 64);
                }
                uint32_t jj[JU];
                float4 pp[JU];
#pragma unroll
                for (uint32_t u = 0; u < JU; ++u) {
                    jj[u] = uint32_t(int32_t(i) + int32_t(int16_t(wc[u] & 0xFFFFu)));
                    pp[u] = a.src[jj[u]];
                }
#pragma unroll
                for (uint32_t u = 0; u < JU; ++u) jacobi_entry<FASTDIV>(pi, pp[u], entry_pair<LDSP>(a, lpairs, wc[u]), ax, ay, az);
            }
        }
        for (; k0 + JU <= width; k0 += JU) {
            uint32_t jj[JU];
            float4 pr[JU];
            float4 pp[JU];
#pragma unroll
            for (uint32_t u = 0; u < JU; ++u) load_entry<PACKED, LDSP>(a, lpairs, i, base + (k0 + u) * 64, jj[u], pr[u]);
#pragma unroll
            for (uint32_t u = 0; u < JU; ++u) {
                if (MODE & JM_PREDICT) pp[u] = predict(a.src[jj[u]], a.vel[jj[u]], jj[u], p);
                else pp[u] = a.src[jj[u]];
            }
#pragma unroll
            for (uint32_t u = 0; u < JU; ++u) jacobi_entry<FASTDIV>(pi, pp[u], pr[u], ax, ay, az);
        }
        for (; k0 < width; ++k0) {
            uint32_t j;
            float4 pr;
            load_entry<PACKED, LDSP>(a, lpairs, i, base + k0 * 64, j, pr);
            const float4 pj = (MODE & JM_PREDICT) ? predict(a.src[j], a.vel[j], j, p) : a.src[j];
            jacobi_entry<FASTDIV>(pi, pj, pr, ax, ay, az);
        }
        float dx = p.omega * ax, dy = p.omega * ay, dz = p.omega * az;
        cap_correction(dx, dy, dz);
        out = make_float4(pi.x + dx, pi.y + dy, pi.z + dz, wi);
    }
    if (MODE & JM_FINAL) {
        if (pi.w > 0.f) finish_particle(out, a.dst[i], i, a.ground, p, a.dst, a.vout);
    } else {
        a.dst[i] = out;
    }
    return out;
}

template <bool PACKED, uint32_t MAXP, typename PT>
__device__ __forceinline__ bool stage_pairs(const JacobiArgs& a, PT* lpairs) {
    const bool lds = PACKED && a.num_pairs <= MAXP;
    if (lds) {
        for (uint32_t q = threadIdx.x; q < a.num_pairs; q += blockDim.x) to_lds(lpairs[q], a.sell.pairs[q], a.p);
        __syncthreads();
    }
    return lds;
}

// At least 8 wavefronts per SIMD (the compiler keeps it at <= 64 VGPRs, no
// spills): one more wavefront of gathers in flight than the unconstrained 70
// VGPRs allow; 39.6 -> 36.7 us per iteration on the 1024^2 sheet.
template <uint32_t MODE, bool PACKED, bool FASTDIV>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8))) void k_jacobi(JacobiArgs a) {
    if (a.gate && *a.gate == 0u) return;   // uniform over the grid
    extern __shared__ float2 jpairs_dyn[];   // num_pairs entries (launch-time size)
    float2* lpairs = jpairs_dyn;
    const bool lds = stage_pairs<PACKED, LDS_PAIRS_DYN_MAX>(a, lpairs);
    const uint32_t tile = a.remap ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t slice = tile * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    bool valid;
    // FASTDIV: the fast operations where the slices this wavefront reads are
    // proven tame (wavefront-uniform); lds: uniform over the launch
    const bool fast = FASTDIV && neighbourhood_tame(a.tame_in, a.srange, slice, a.p.nslices);
    float4 out;
    if (PACKED && lds)
        out = fast ? jacobi_particle<MODE, PACKED, true, true>(a, lpairs, slice, lane, valid)
                   : jacobi_particle<MODE, PACKED, false, true>(a, lpairs, slice, lane, valid);
    else
        out = fast ? jacobi_particle<MODE, PACKED, true, false>(a, lpairs, slice, lane, valid)
                   : jacobi_particle<MODE, PACKED, false, false>(a, lpairs, slice, lane, valid);
    if (!(MODE & JM_FINAL) && a.tame_out)
        write_slice_tame(a.tame_out, slice, slice < a.p.nslices, !valid || tame3(out));
    if (MODE & JM_AABB) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        aabb_fold(valid, out.x, out.y, out.z, lo, hi);
        block_aabb_store(lo, hi, a.aabb, tile);
    }
}

// ---------------------------------------------------------------------------
// Grid-tiled Jacobi iteration (one grid cloth, columns a multiple of 64).  A
// block owns a 64 x JT_ROWS tile of the grid (one wavefront per row = one
// 64-particle slice) and first stages the tile's positions with a 2-particle
// halo (bend constraints reach 2) in LDS with coalesced loads; every
// neighbour gather then reads LDS instead of L2.  The entries are the same
// CSR entries in the same order (their offsets re-encoded for the tile), so
// the sums, and the results, are those of k_jacobi.
// ---------------------------------------------------------------------------
// A tiled entry word: low 16 bits the neighbour's byte offset in the LDS
// tile from the particle's own slot (signed), high 16 bits the pair's byte
// offset in the staged pair table (pair index x 16), so that each address is
// one add or one shift.
__device__ __forceinline__ float4 tile_neighbour(const float4* self, uint32_t w) {
    return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(self) + int32_t(int16_t(w & 0xFFFFu)));
}
template <bool LDSP>
__device__ __forceinline__ float4 tile_pair(const JacobiArgs& a, const float4* lpairs, uint32_t w) {
    if constexpr (LDSP) return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(lpairs) + (w >> 16));
    else return stage_pair(a.sell.pairs[w >> 20], a.p);
}

// pre: the slice's first JT_JU entry words (those below width), loaded by
// the caller; base = the lane's first entry, width = the slice's width.
template <uint32_t MODE, bool FAST, bool LDSP>
__device__ __forceinline__ float4 jacobi_particle_tiled(const JacobiArgs& a, const float4* lpairs, const float4* tl, uint32_t slice, uint32_t lane,
                                                        uint32_t lidx, bool row_ok, bool& valid, const uint32_t* pre,
                                                        uint32_t base, uint32_t width) {
    const uint32_t i = slice * 64 + lane;
    const DevParams& p = a.p;
    valid = row_ok && slice < p.nslices && i < p.n;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 out = pi;
    if (!valid) return out;
    pi = tl[lidx];
    out = pi;
    const float wi = pi.w;
    if (wi > 0.f) {
        float ax = 0.f, ay = 0.f, az = 0.f;
        uint32_t k0 = 0;
        if (width >= JT_JU) {
            // the next chunk's entry words load while this chunk computes
            uint32_t wn[JT_JU];
#pragma unroll
            for (uint32_t u = 0; u < JT_JU; ++u) wn[u] = pre[u];
            for (; k0 + JT_JU <= width; k0 += JT_JU) {
                uint32_t wc[JT_JU];
#pragma unroll
                for (uint32_t u = 0; u < JT_JU; ++u) wc[u] = wn[u];
                if (k0 + 2 * JT_JU <= width) {
#pragma unroll
                    for (uint32_t u = 0; u < JT_JU; ++u)
                        wn[u] = __builtin_nontemporal_load(a.sell.tpacked + base + (k0 + JT_JU + u) * 64);
                }
#pragma unroll
                for (uint32_t u = 0; u < JT_JU; ++u) {
                    jacobi_entry<FAST>(pi, tile_neighbour(tl + lidx, wc[u]), tile_pair<LDSP>(a, lpairs, wc[u]), ax, ay,
                                       az);
                }
            }
        }
        for (; k0 < width; ++k0) {
            const uint32_t w = __builtin_nontemporal_load(a.sell.tpacked + base + k0 * 64);
            jacobi_entry<FAST>(pi, tile_neighbour(tl + lidx, w), tile_pair<LDSP>(a, lpairs, w), ax, ay, az);
        }
        float dx = p.omega * ax, dy = p.omega * ay, dz = p.omega * az;
        cap_correction(dx, dy, dz);
        out = make_float4(pi.x + dx, pi.y + dy, pi.z + dz, wi);
    }
    if (MODE & JM_KEEP) {
        // the caller keeps the result
    } else if (MODE & JM_FINAL) {
        if (pi.w > 0.f) finish_particle(out, a.dst[i], i, a.ground, p, a.dst, a.vout);
    } else {
        a.dst[i] = out;
    }
    return out;
}

template <uint32_t MODE, bool FASTDIV>
__global__ __launch_bounds__(64 * JT_ROWS) __attribute__((amdgpu_waves_per_eu(8))) void k_jacobi_tiled(JacobiArgs a) {
    if (a.gate && *a.gate == 0u) return;   // uniform over the grid
    __shared__ float4 lpairs[JT_PAIRS_MAX];
    __shared__ float4 tl[JT_H * JT_W];
    const uint32_t cols = a.sell.tcols, rows = a.sell.trows, ntx = cols / 64;
    const uint32_t b = a.remap ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t tx = b % ntx, ty = b / ntx;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t row = ty * JT_ROWS + wave;
    const bool row_ok = row < rows;
    const uint32_t slice = row_ok ? row * ntx + tx : a.p.nslices;
    uint32_t base = 0, width = 0;
    if (slice < a.p.nslices) {
        base = a.sell.soff[slice] + lane;
        width = a.sell.swid[slice];
    }
    const int32_t r0 = int32_t(ty * JT_ROWS) - 2, c0 = int32_t(tx * 64) - 2;
    bool tame = true;
    for (uint32_t q = threadIdx.x; q < JT_H * JT_W; q += blockDim.x) {
        const int32_t r = r0 + int32_t(q / JT_W), c = c0 + int32_t(q % JT_W);
        if (r >= 0 && r < int32_t(rows) && c >= 0 && c < int32_t(cols)) {
            const uint32_t id = uint32_t(r) * cols + uint32_t(c);
            const float4 v = (MODE & JM_PREDICT) ? predict(a.src[id], a.vel[id], id, a.p) : a.src[id];
            tl[q] = v;
            tame = tame && tame3(v);
        }
    }
    // the first chunk of entry words is requested before the barrier, so its
    // latency overlaps the other wavefronts' staging (-0.6 us per iteration
    // on the 1024^2 sheet against loading it after the barrier)
    uint32_t pre[JT_JU];
#pragma unroll
    for (uint32_t u = 0; u < JT_JU; ++u)
        pre[u] = u < width ? __builtin_nontemporal_load(a.sell.tpacked + base + u * 64) : 0u;
    const bool lds = stage_pairs<true, JT_PAIRS_MAX>(a, lpairs);
    // The block's entries read only staged positions, so the fast operations
    // need only these to be tame (predicted ones, in JM_PREDICT): one
    // block-wide AND, which is also the barrier that publishes the tile (so
    // it runs whatever FASTDIV is).
    const int all_tame = __syncthreads_and(tame ? 1 : 0);
    const bool fast = FASTDIV && all_tame != 0;
    const uint32_t lidx = (wave + 2) * JT_W + lane + 2;
    bool valid;
    float4 out;
    if (lds)
        out = fast ? jacobi_particle_tiled<MODE, true, true>(a, lpairs, tl, slice, lane, lidx, row_ok, valid, pre, base, width)
                   : jacobi_particle_tiled<MODE, false, true>(a, lpairs, tl, slice, lane, lidx, row_ok, valid, pre, base, width);
    else
        out = fast ? jacobi_particle_tiled<MODE, true, false>(a, lpairs, tl, slice, lane, lidx, row_ok, valid, pre, base, width)
                   : jacobi_particle_tiled<MODE, false, false>(a, lpairs, tl, slice, lane, lidx, row_ok, valid, pre, base, width);
    if (!(MODE & JM_FINAL) && a.tame_out)
        write_slice_tame(a.tame_out, slice, slice < a.p.nslices, !valid || tame3(out));
    if (MODE & JM_AABB) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        aabb_fold(valid, out.x, out.y, out.z, lo, hi);
        block_aabb_store(lo, hi, a.aabb, b);
    }
}

// ---------------------------------------------------------------------------
// Constraint-shared grid tiles: the tiled iteration with each constraint that
// touches a tile evaluated once per tile (the layout is described with
// SharedTiles).  Up to the last product, an entry's arithmetic is the same
// for both endpoints, bit for bit: the difference vector and the normal
// only change sign (a - b = -(b - a) and (-g) / len = -(g / len) exactly in
// round-to-nearest), and |g|, C, w_i + w_j + alpha~ and dlam are equal.
// So phase 1 stores (n, dlam) per constraint as its anchor a sees it, and
// phase 2 adds, in each particle's CSR order, (n * w) * dlam for the anchor
// and (n * (-w)) * dlam = ((-n) * w) * dlam for the other endpoint: the
// terms, and the sums, of k_jacobi_tiled.  A skipped constraint stores
// zeros; its term is a zero, and the accumulators (starting at +0, never -0
// in round-to-nearest) do not change when a zero is added.
// ---------------------------------------------------------------------------
// (n, dlam) of the constraint (a, b) as jacobi_entry computes it for a, or
// zeros where jacobi_entry adds nothing (and for `none`).  Branch-free (the
// skipped results are selected away), so several constraints can be in
// flight at once.
template <bool FAST>
__device__ __forceinline__ float4 constraint_result(float4 pa, float4 pb, float4 pr, bool none) {
    float gx = pa.x - pb.x, gy = pa.y - pb.y, gz = pa.z - pb.z;
    const float dot = (gx * gx + gy * gy) + gz * gz;
    const float len = FAST ? sqrt_rn(dot) : sqrtf(dot);
    const float c = len - pr.x;
    const float wsum = pa.w + pb.w + pr.y;
    const bool skip = none || len < 1e-8f || (!FAST && wsum < 1e-8f);
    const float num = FAST ? -c : -c - pr.z;
    float dlam;
    if (FAST) {
        const float yl = rcp_rn(len);
        gx = div_rn_nz(gx, len, yl);
        gy = div_rn_nz(gy, len, yl);
        gz = div_rn_nz(gz, len, yl);
        dlam = div_rn_nz(num, wsum, rcp_rn(wsum));   // see jacobi_entry for the zero's sign
    } else {
        gx = gx / len;
        gy = gy / len;
        gz = gz / len;
        dlam = num / wsum;
    }
    return skip ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(gx, gy, gz, dlam);
}

// Phase 1 of a block: every job's constraints into the result slots.  A
// wavefront's jobs are wave, wave + JT_ROWS, ...; the first JS_JOBS_PF of
// them are decoded, and their pair indices loaded, before the tile barrier
// (in k_jacobi_shared), so those loads overlap the staging.
constexpr uint32_t JS_JOBS_PF = 8;

struct SharedJob {
    uint32_t item, cl, pid;   // pid 0xFFFF: nothing to compute (the slot is stored as zeros)
};

__device__ __forceinline__ SharedJob shared_job(const JacobiArgs& a, uint32_t j, int32_t r0, int32_t c0) {
    const uint32_t lane = threadIdx.x & 63;
    SharedJob J;
    if (j < a.sh.njobs_main) {
        J.item = a.sh.jobs[j];
        J.cl = lane + 2;
    } else {
        J.item = a.sh.jobs[a.sh.njobs_main + (j - a.sh.njobs_main) * 64 + lane];
        J.cl = (J.item >> 12) & 0xFFu;
    }
    J.pid = 0xFFFFu;
    if (J.item == 0xFFFFFFFFu) return J;
    const int32_t rows = int32_t(a.sell.trows), cols = int32_t(a.sell.tcols);
    const int32_t dy = int32_t(J.item & 3u), dx = int32_t((J.item >> 2) & 7u) - 2;
    const uint32_t t = (J.item >> 5) & 7u, lr = (J.item >> 8) & 15u;
    const int32_t R = r0 + int32_t(lr), C = c0 + int32_t(J.cl);   // anchor (grid row, column)
    if (R >= 0 && R + dy < rows && C >= 0 && C < cols && C + dx >= 0 && C + dx < cols)
        J.pid = a.sh.jpair[size_t(t) * a.p.n + uint32_t(R) * uint32_t(cols) + uint32_t(C)];
    return J;
}

// Two jobs at once, branch-free until the stores: a job without
// a constraint reads tile slot 0 and pair 0 (both staged) and stores zeros.
template <bool FAST>
__device__ __forceinline__ void shared_compute(const SharedJob& J0, const SharedJob& J1, const float4* tl,
                                                const float4* lpairs, float4* res) {
    float4 pa[2], pb[2], pr[2];
    bool none[2];
    const SharedJob* J[2] = {&J0, &J1};
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const uint32_t item = J[u]->item;
        none[u] = item == 0xFFFFFFFFu || J[u]->pid == 0xFFFFu;
        const int32_t dy = int32_t(item & 3u), dx = int32_t((item >> 2) & 7u) - 2;
        const uint32_t lr = (item >> 8) & 15u;
        const uint32_t ia = none[u] ? 0u : lr * JT_W + J[u]->cl;
        const uint32_t ib = none[u] ? 0u : (lr + uint32_t(dy)) * JT_W + uint32_t(int32_t(J[u]->cl) + dx);
        pa[u] = tl[ia];
        pb[u] = tl[ib];
        pr[u] = lpairs[none[u] ? 0u : J[u]->pid];
    }
    float4 r[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) r[u] = constraint_result<FAST>(pa[u], pb[u], pr[u], none[u]);
#pragma unroll
    for (int u = 0; u < 2; ++u)
        if (J[u]->item != 0xFFFFFFFFu) res[(J[u]->item >> 20) * JT_W + J[u]->cl] = r[u];
}

template <bool FAST>
__device__ __forceinline__ void shared_phase1(const JacobiArgs& a, const float4* tl, const float4* lpairs, float4* res,
                                              int32_t r0, int32_t c0, const SharedJob (&pf)[JS_JOBS_PF]) {
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t njobs = a.sh.njobs_main + a.sh.njobs_edge;
#pragma unroll
    for (uint32_t k = 0; k < JS_JOBS_PF; k += 2) shared_compute<FAST>(pf[k], pf[k + 1], tl, lpairs, res);   // unused: item none
    SharedJob none;
    none.item = 0xFFFFFFFFu;
    none.cl = 0u;
    none.pid = 0xFFFFu;
    for (uint32_t j = wave + JS_JOBS_PF * JT_ROWS; j < njobs; j += JT_ROWS)
        shared_compute<FAST>(shared_job(a, j, r0, c0), none, tl, lpairs, res);
}

// pre: the lane's first JS_WORDS_PF entry words (those below ceil(width / 2)).
constexpr uint32_t JS_WORDS_PF = 8;

template <uint32_t MODE>
__device__ __forceinline__ float4 shared_particle(const JacobiArgs& a, const float4* tl, const float4* res,
                                                  uint32_t slice, uint32_t lane, uint32_t lidx, bool row_ok,
                                                  bool& valid, const uint32_t (&pre)[JS_WORDS_PF], uint32_t width) {
    const uint32_t i = slice * 64 + lane;
    const DevParams& p = a.p;
    valid = row_ok && slice < p.nslices && i < p.n;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 out = pi;
    if (!valid) return out;
    pi = tl[lidx];
    out = pi;
    const float wi = pi.w;
    if (wi > 0.f) {
        float ax = 0.f, ay = 0.f, az = 0.f;
        const float4* rb = res + lane;
        const uint32_t* ew = a.sh.entries + a.sh.eoff[slice] + lane;
        const uint32_t wsign = __float_as_uint(wi);
        auto add = [&](uint32_t code) {
            const float4 v = rb[code & 0x7FFFu];
            const float sw = __uint_as_float(wsign ^ ((code & 0x8000u) << 16));
            ax = ax + (v.x * sw) * v.w;
            ay = ay + (v.y * sw) * v.w;
            az = az + (v.z * sw) * v.w;
        };
        // the prefetched words (static indices keep them in registers), then
        // any further words from memory
#pragma unroll
        for (uint32_t q = 0; q < JS_WORDS_PF; ++q) {
            if (2 * q < width) add(pre[q] & 0xFFFFu);
            if (2 * q + 1 < width) add(pre[q] >> 16);
        }
        for (uint32_t k = 2 * JS_WORDS_PF; k < width; k += 2) {
            const uint32_t w = __builtin_nontemporal_load(ew + (k >> 1) * 64);
            add(w & 0xFFFFu);
            if (k + 1 < width) add(w >> 16);
        }
        float dx = p.omega * ax, dy = p.omega * ay, dz = p.omega * az;
        cap_correction(dx, dy, dz);
        out = make_float4(pi.x + dx, pi.y + dy, pi.z + dz, wi);
    }
    if (MODE & JM_FINAL) {
        if (pi.w > 0.f) finish_particle(out, a.dst[i], i, a.ground, p, a.dst, a.vout);
    } else {
        a.dst[i] = out;
    }
    return out;
}

template <uint32_t MODE, bool FASTDIV>
// LDS (position tile, result slots) allows 2 blocks per CU: 4 wavefronts per
// SIMD, so the compiler may use up to 128 VGPRs.
__global__ __launch_bounds__(64 * JT_ROWS) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_jacobi_shared(JacobiArgs a) {
    if (a.gate && *a.gate == 0u) return;   // uniform over the grid
    __shared__ float4 tl[JT_H * JT_W];
    extern __shared__ float4 js_dyn[];     // result slots, zero slot, then the staged pairs
    float4* res = js_dyn;
    float4* lpairs = js_dyn + a.sh.nslots + 1;
    const uint32_t cols = a.sell.tcols, rows = a.sell.trows, ntx = cols / 64;
    const uint32_t b = a.remap ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t tx = b % ntx, ty = b / ntx;
    const int32_t r0 = int32_t(ty * JT_ROWS) - 2, c0 = int32_t(tx * 64) - 2;
    bool tame = true;
    for (uint32_t q = threadIdx.x; q < JT_H * JT_W; q += blockDim.x) {
        const int32_t r = r0 + int32_t(q / JT_W), c = c0 + int32_t(q % JT_W);
        if (r >= 0 && r < int32_t(rows) && c >= 0 && c < int32_t(cols)) {
            const uint32_t id = uint32_t(r) * cols + uint32_t(c);
            const float4 v = (MODE & JM_PREDICT) ? predict(a.src[id], a.vel[id], id, a.p) : a.src[id];
            tl[q] = v;
            tame = tame && tame3(v);
        }
    }
    for (uint32_t q = threadIdx.x; q < a.num_pairs; q += blockDim.x) to_lds(lpairs[q], a.sell.pairs[q], a.p);
    if (threadIdx.x == 0) res[a.sh.nslots] = make_float4(0.f, 0.f, 0.f, 0.f);
    // the wavefront's first jobs (pair indices) and the lane's first entry
    // words are requested before the barrier: their latency overlaps the
    // staging
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t njobs = a.sh.njobs_main + a.sh.njobs_edge;
    SharedJob pf[JS_JOBS_PF];
#pragma unroll
    for (uint32_t k = 0; k < JS_JOBS_PF; ++k) {
        pf[k].item = 0xFFFFFFFFu;
        pf[k].cl = 0u;
        pf[k].pid = 0xFFFFu;
        if (wave + k * JT_ROWS < njobs) pf[k] = shared_job(a, wave + k * JT_ROWS, r0, c0);
    }
    const uint32_t row = ty * JT_ROWS + wave;
    const bool row_ok = row < rows;
    const uint32_t slice = row_ok ? row * ntx + tx : a.p.nslices;
    const uint32_t width = slice < a.p.nslices ? a.sell.swid[slice] : 0u;
    uint32_t pre[JS_WORDS_PF];
    {
        const uint32_t* ew = a.sh.entries + (slice < a.p.nslices ? a.sh.eoff[slice] : 0u) + lane;
#pragma unroll
        for (uint32_t k = 0; k < JS_WORDS_PF; ++k) pre[k] = 2 * k < width ? __builtin_nontemporal_load(ew + k * 64) : 0u;
    }
    // one block-wide AND, run whatever FASTDIV is: the tile and pairs are
    // published, and the fast operations apply where every staged position
    // is tame
    const int all_tame = __syncthreads_and(tame ? 1 : 0);
    const bool fast = FASTDIV && all_tame != 0;
    if (fast) shared_phase1<true>(a, tl, lpairs, res, r0, c0, pf);
    else shared_phase1<false>(a, tl, lpairs, res, r0, c0, pf);
    __syncthreads();
    const uint32_t lidx = (wave + 2) * JT_W + lane + 2;
    bool valid;
    const float4 out = shared_particle<MODE>(a, tl, res, slice, lane, lidx, row_ok, valid, pre, width);
    if (!(MODE & JM_FINAL) && a.tame_out)
        write_slice_tame(a.tame_out, slice, slice < a.p.nslices, !valid || tame3(out));
    if (MODE & JM_AABB) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        aabb_fold(valid, out.x, out.y, out.z, lo, hi);
        block_aabb_store(lo, hi, a.aabb, b);
    }
}

// Substep end without a collision pass when the last iterate is in F.
__global__ void k_finalize(const float4* F, float4* X, float4* V, const uint8_t* ground, DevParams p) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const float4 f = F[i];
    if (f.w <= 0.f) return;
    finish_particle(f, X[i], i, ground, p, X, V);
}

__global__ void k_copy4(const float4* __restrict__ a, float4* __restrict__ b, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

__global__ void k_aabb_of(const float4* P, uint32_t n, float* part) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const float4 v = valid ? P[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    aabb_fold(valid, v.x, v.y, v.z, lo, hi);
    block_aabb_store(lo, hi, part, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Collision broad phase.  Each cloth gets a grid over its current bounds: the
// cell side is cell_size * 65/64 (a guard so f32 rounding in the key can never
// put two particles within cell_size two cells apart), doubled until each axis
// has < 16383 cells and the cloth's cell count fits its key range
// (2^27 cells for one cloth; 2^(32-b) - 2 cells in a batch, b = bits
// numbering the cloths).
// Key = (cloth << (32-b)) + ix + gx * (iy + gy * iz): dense per cloth, so no
// two cells share a key and different cloths never meet.  Mirrors
// oracle_cell_grid / oracle_cell_keys / oracle_cell_keys_batch.
// ---------------------------------------------------------------------------
// One cloth's grid has at most 2^27 cells, so its keys sort in 3 passes of
// 9-bit digits (the cell doubles until the grid fits; larger cells only cost
// sweep time: the pair set is the same).
constexpr uint64_t ONE_CLOTH_MAX_CELLS = 1ull << 27;

__device__ void grid_from_bounds(float lo[3], float hi[3], float cell_size, uint64_t max_cells, uint32_t key_base,
                                 CollGrid* g) {
    for (int a = 0; a < 3; ++a)
        if (!(lo[a] <= hi[a])) lo[a] = hi[a] = 0.f;
    float cell = cell_size * 1.015625f;
    float inv = 1.0f / cell;
    uint32_t dims[3] = {1, 1, 1};
    bool ok = false;
    for (int it = 0; it < 64; ++it) {
        inv = 1.0f / cell;
        float e[3];
        bool fits = true;
        for (int a = 0; a < 3; ++a) {
            e[a] = (hi[a] - lo[a]) * inv;
            if (!(e[a] < 16383.0f)) fits = false;
        }
        if (fits) {
            for (int a = 0; a < 3; ++a) dims[a] = uint32_t(floorf(e[a])) + 1u;
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
    g->key_base = key_base;
}

// Many cloths: one block per cloth reduces its particle range.
// Per-cloth bounds of a batch in balanced pieces: block b reduces particles
// [span[b].x, span[b].y) of one cloth (at most CB_SPAN of them) into partial
// b; the grid kernel then folds each cloth's partials.  (One block per cloth
// let the largest cloths set the time; atomics on ~3000 words serialised.)
constexpr uint32_t CB_SPAN = 4096;

__global__ __launch_bounds__(256) void k_cloth_aabb(const float4* P, const uint2* span, float* part) {
    const uint2 r = span[blockIdx.x];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = r.x + threadIdx.x; i < r.y; i += blockDim.x) {
        const float4 v = P[i];
        aabb_fold(true, v.x, v.y, v.z, lo, hi);
    }
    block_aabb_store(lo, hi, part, blockIdx.x);
}

// Grids from the partials of each cloth (pieces [pfirst[c], pfirst[c+1])).
__global__ void k_coll_grids(const float* part, const uint32_t* pfirst, uint32_t nclothes, float cell_size,
                             uint32_t local_bits, CollGrid* grids) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nclothes) return;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t k = pfirst[c]; k < pfirst[c + 1]; ++k)
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], part[size_t(k) * 8 + a]);
            hi[a] = fmaxf(hi[a], part[size_t(k) * 8 + 3 + a]);
        }
    const uint64_t cap = (local_bits >= 32 ? (1ull << 32) : (1ull << local_bits)) - 2;
    grid_from_bounds(lo, hi, cell_size, cap, local_bits >= 32 ? 0u : (c << local_bits), &grids[c]);
}

__device__ __forceinline__ void cell_of(float4 p, const CollGrid& g, uint32_t c[3]) {
    const float v[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float q = (v[a] - g.lo[a]) * g.inv_cell;
        q = floorf(q);
        q = fmaxf(q, 0.0f);
        q = fminf(q, float(g.dims[a] - 1));
        c[a] = uint32_t(q);
    }
}

// ---- stable LSD radix sort of (key, id), DB-bit digits ----------------------
// A pass is three kernels: per-tile digit counts, an exclusive scan of the
// counts over tiles (per digit), and a stable scatter.  A tile is 4096 keys;
// each of its 8 wavefronts owns a contiguous eighth (512 keys, 8 rounds of
// 64), so ranks in input order are (wavefront, round, lane) and need no block
// barrier per round.  (2048-key tiles left 9-bit digit runs of about 4 keys
// in the scatter's output: poorly coalesced stores; 256-thread blocks left
// one wavefront per SIMD on a 1024^2 sheet.)  The count kernels (k_coll_keys' pass-0 counts and
// k_radix_hist) add each key to its wavefront's LDS histogram with an LDS
// atomic: exact integer counts, independent of the order of the adds.  The
// scatter ranks lanes stably: lanes with equal digits are found by matching
// the DB digit bits with ballots, and one leader per distinct digit updates
// the wavefront's own counters (distinct addresses, no atomics).
// Digit widths: 8 (4 passes over 32-bit keys; batches and topology sorts) or
// 9 (3 passes over the 27-bit keys of one cloth's grid: a pass is three
// launches of mostly fixed latency, so one pass fewer outweighs the larger
// per-tile count table).  Thread t of a block owns digits [t*DPT, t*DPT+DPT)
// (DPT = 1 and threads past the last digit own none when a block has more
// threads than digits).
constexpr uint32_t RADIX_BLOCK = 512;
constexpr uint32_t RADIX_WAVES = RADIX_BLOCK / 64;
constexpr uint32_t RADIX_ITEMS = 8;
constexpr uint32_t RADIX_TILE = RADIX_BLOCK * RADIX_ITEMS;
constexpr uint32_t RADIX_WAVE_SPAN = 64 * RADIX_ITEMS;   // keys owned by one wavefront
constexpr uint32_t RADIX_MAX_DIGITS = 512;

// Lanes (among `valid` ones) whose digit equals this lane's.
template <uint32_t DB>
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid) {
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (uint32_t b = 0; b < DB; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
    }
    return peers;
}

// Digits owned by one thread of a radix block.
constexpr uint32_t radix_dpt(uint32_t digits) { return digits >= RADIX_BLOCK ? digits / RADIX_BLOCK : 1u; }

// Block-wide digit counts of the per-wavefront LDS histograms -> hist[d][tile].
template <uint32_t DB>
__device__ __forceinline__ void store_tile_counts(uint32_t (*h)[1u << DB], uint32_t* hist, uint32_t ntiles) {
    constexpr uint32_t DIGITS = 1u << DB, DPT = radix_dpt(DIGITS);
#pragma unroll
    for (uint32_t k = 0; k < DPT; ++k) {
        const uint32_t d = threadIdx.x * DPT + k;
        if (d >= DIGITS) break;
        uint32_t c = 0;
#pragma unroll
        for (uint32_t w = 0; w < RADIX_WAVES; ++w) c += h[w][d];
        hist[size_t(d) * ntiles + blockIdx.x] = c;
    }
}

template <uint32_t DB>
__device__ __forceinline__ void clear_wave_hists(uint32_t (*h)[1u << DB]) {
    constexpr uint32_t DIGITS = 1u << DB, DPT = radix_dpt(DIGITS);
#pragma unroll
    for (uint32_t w = 0; w < RADIX_WAVES; ++w)
#pragma unroll
        for (uint32_t k = 0; k < DPT; ++k)
            if (threadIdx.x * DPT + k < DIGITS) h[w][threadIdx.x * DPT + k] = 0;
}

template <uint32_t DB>
__global__ __launch_bounds__(RADIX_BLOCK) void k_radix_hist(const uint32_t* keys, uint32_t n, uint32_t shift,
                                                          uint32_t* hist, uint32_t ntiles) {
    constexpr uint32_t MASK = (1u << DB) - 1u;
    __shared__ uint32_t h[RADIX_WAVES][1u << DB];
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    clear_wave_hists<DB>(h);
    __syncthreads();
    const uint32_t base = blockIdx.x * RADIX_TILE + wave * RADIX_WAVE_SPAN + lane;
    uint32_t kk[RADIX_ITEMS];
#pragma unroll
    for (uint32_t r = 0; r < RADIX_ITEMS; ++r) {
        const uint32_t idx = base + r * 64;
        kk[r] = idx < n ? keys[idx] : 0u;
    }
#pragma unroll
    for (uint32_t r = 0; r < RADIX_ITEMS; ++r) {
        const bool valid = base + r * 64 < n;
        const uint32_t d = (kk[r] >> shift) & MASK;
        if (valid) atomicAdd(&h[wave][d], 1u);   // LDS atomics: cheaper here than ranking the wave
    }
    __syncthreads();
    store_tile_counts<DB>(h, hist, ntiles);
}

// Cell keys of the current positions, laid out as the sort's tiles (one
// block per 4096 keys), with the sort's first per-tile digit counts (pass 0,
// the low DB bits) taken on the way, so that pass needs no histogram launch.
// Hashed group table (slot_rec != nullptr): also empties the slots the
// previous collision pass filled (slot_rec holds, per sorted index, the slot
// its group was inserted at, or EMPTY_KEY), so the table is empty again
// before this pass inserts.  Direct group table (epoch != nullptr): advances
// the pass tag instead; entries of earlier passes no longer match it.
// One cloth (part != nullptr): every block derives the grid itself from the
// nparts bounds partials of the producing kernel (min / max are exact and
// order-free, so all blocks agree), and block 0 stores it for the sweep;
// this replaces a one-block grid launch.  Batches read grids[] (k_coll_grids).
template <uint32_t DB>
__global__ __launch_bounds__(RADIX_BLOCK) void k_coll_keys(const float4* F, uint32_t n, const CollGrid* grids,
                                                         const uint32_t* cloth_of, uint32_t* keys, uint32_t* ids,
                                                         uint4* table, const uint32_t* slot_rec, uint32_t* hist,
                                                         uint32_t ntiles, const float* part, uint32_t nparts,
                                                         float cell_size, CollGrid* grid_out, uint32_t* epoch) {
    constexpr uint32_t MASK = (1u << DB) - 1u;
    __shared__ uint32_t h[RADIX_WAVES][1u << DB];
    __shared__ float red[RADIX_WAVES][6];
    __shared__ CollGrid sgrid;
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    clear_wave_hists<DB>(h);
    const uint32_t base = blockIdx.x * RADIX_TILE + wave * RADIX_WAVE_SPAN + lane;
    // every load of the thread's rounds is issued before the first store (the
    // table stores could otherwise alias them and serialise the rounds)
    const uint32_t last = n - 1;
    float4 f[RADIX_ITEMS];
    uint32_t sl[RADIX_ITEMS], co[RADIX_ITEMS];
#pragma unroll
    for (uint32_t r = 0; r < RADIX_ITEMS; ++r) {
        const uint32_t i = min(base + r * 64, last);
        f[r] = F[i];
        sl[r] = slot_rec && base + r * 64 < n ? slot_rec[i] : EMPTY_KEY;
        co[r] = cloth_of ? cloth_of[i] : 0u;
    }
    if (epoch && blockIdx.x == 0 && t == 0) epoch[0] = epoch[0] + 1u == 0u ? 1u : epoch[0] + 1u;   // 0: never written
    if (part) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        const float4* p4 = reinterpret_cast<const float4*>(part);   // (lo xyz, hi x), (hi yz, -, -) per partial
        for (uint32_t k = t; k < nparts; k += RADIX_BLOCK) {
            const float4 a = p4[2 * k], b = p4[2 * k + 1];
            lo[0] = fminf(lo[0], a.x);
            lo[1] = fminf(lo[1], a.y);
            lo[2] = fminf(lo[2], a.z);
            hi[0] = fmaxf(hi[0], a.w);
            hi[1] = fmaxf(hi[1], b.x);
            hi[2] = fmaxf(hi[2], b.y);
        }
        wave_bounds_to_lane63(lo, hi);
        if (lane == 63)
            for (int a = 0; a < 3; ++a) {
                red[wave][a] = lo[a];
                red[wave][3 + a] = hi[a];
            }
        __syncthreads();
        if (t == 0) {
            for (int a = 0; a < 3; ++a) {
                lo[a] = red[0][a];
                hi[a] = red[0][3 + a];
            }
            for (uint32_t w = 1; w < RADIX_WAVES; ++w)
                for (int a = 0; a < 3; ++a) {
                    lo[a] = fminf(lo[a], red[w][a]);
                    hi[a] = fmaxf(hi[a], red[w][3 + a]);
                }
            grid_from_bounds(lo, hi, cell_size, ONE_CLOTH_MAX_CELLS, 0u, &sgrid);
            if (blockIdx.x == 0) *grid_out = sgrid;
        }
    }
    __syncthreads();   // the cleared histograms (and the grid) are published
    const CollGrid g0 = part ? sgrid : grids[0];
#pragma unroll
    for (uint32_t r = 0; r < RADIX_ITEMS; ++r) {
        const uint32_t i = base + r * 64;
        if (sl[r] != EMPTY_KEY) table[sl[r]] = make_uint4(EMPTY_KEY, EMPTY_KEY, 0u, 0u);
        if (i < n) {
            const CollGrid g = cloth_of ? grids[co[r]] : g0;
            uint32_t c[3];
            cell_of(f[r], g, c);
            const uint32_t key = g.key_base + (c[0] + g.dims[0] * (c[1] + g.dims[1] * c[2]));
            keys[i] = key;
            ids[i] = i;
            atomicAdd(&h[wave][key & MASK], 1u);   // LDS atomics: cheaper here than ranking the wave
        }
    }
    __syncthreads();
    store_tile_counts<DB>(h, hist, ntiles);
}

// Exclusive scan of one value per thread over a block of WAVES wavefronts;
// `total` gets the block's sum.  lds holds WAVES words.
template <uint32_t WAVES>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off, 64);
        if (lane >= uint32_t(off)) incl += t;
    }
    if (lane == 63) lds[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
#pragma unroll
    for (uint32_t w = 0; w < WAVES; ++w) {
        before += w < wave ? lds[w] : 0u;
        total += lds[w];
    }
    __syncthreads();
    return before + incl - v;
}

// Per digit (one block each): exclusive prefix of the per-tile counts over
// tiles, in place; the digit's total goes to dtot[d].  Each thread scans a
// contiguous run of tiles serially, then one block scan joins the runs.
__global__ __launch_bounds__(256) void k_radix_scan_tiles(uint32_t* hist, uint32_t ntiles, uint32_t* dtot) {
    __shared__ uint32_t lds[4];
    uint32_t* row = hist + size_t(blockIdx.x) * ntiles;
    const uint32_t per = (ntiles + 255) / 256;
    const uint32_t b = threadIdx.x * per, e = min(b + per, ntiles);
    uint32_t sum = 0;
    for (uint32_t k = b; k < e; ++k) sum += row[k];
    uint32_t tot;
    uint32_t run = block_exclusive_scan<4>(sum, lds, tot);
    for (uint32_t k = b; k < e; ++k) {
        const uint32_t v = row[k];
        row[k] = run;
        run += v;
    }
    if (threadIdx.x == 0) dtot[blockIdx.x] = tot;
}

// Stable scatter: the tile is ranked in input order (see above), placed in
// LDS in digit order, and written out in that order, so each digit's run of
// the tile leaves as contiguous stores.  GATHER (the last pass of a
// collision sort): also the sweep's sorted copy sP[dst] = (F[id] xyz, key
// bits), read while the ids are at hand.
template <uint32_t DB, bool GATHER>
__global__ __launch_bounds__(RADIX_BLOCK) void k_radix_scatter(const uint32_t* kin, const uint32_t* iin,
                                                             uint32_t* kout, uint32_t* iout, uint32_t n,
                                                             uint32_t shift, const uint32_t* hist,
                                                             uint32_t ntiles, const uint32_t* dtot,
                                                             const float4* F, float4* sP) {
    constexpr uint32_t DIGITS = 1u << DB, MASK = DIGITS - 1u, DPT = radix_dpt(DIGITS);
    __shared__ uint32_t wcnt[RADIX_WAVES][DIGITS];   // per-wavefront digit counts, then its start in the tile
    __shared__ int32_t gbase[DIGITS];                 // global position of local sorted position 0 of digit d
    __shared__ uint32_t skey[RADIX_TILE];
    __shared__ uint32_t sid[RADIX_TILE];
    __shared__ uint32_t lds[RADIX_WAVES];
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t tile0 = blockIdx.x * RADIX_TILE;
    const uint32_t count = min(RADIX_TILE, n - tile0);
    // this thread's digits' tile offsets and totals, loaded first: their
    // latency then hides behind the key loads and the ranking
    uint32_t hv[DPT], dv[DPT];
#pragma unroll
    for (uint32_t k = 0; k < DPT; ++k) {
        const uint32_t d = t * DPT + k;
        hv[k] = d < DIGITS ? hist[size_t(d) * ntiles + blockIdx.x] : 0u;
        dv[k] = d < DIGITS ? dtot[d] : 0u;
    }
    clear_wave_hists<DB>(wcnt);
    __syncthreads();
    const uint32_t base = wave * RADIX_WAVE_SPAN + lane;   // tile-local index of round 0
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t kk[RADIX_ITEMS], ii[RADIX_ITEMS], rk[RADIX_ITEMS];
#pragma unroll
    for (uint32_t r = 0; r < RADIX_ITEMS; ++r) {
        const uint32_t q = base + r * 64;
        kk[r] = q < count ? kin[tile0 + q] : 0u;
        ii[r] = q < count ? iin[tile0 + q] : 0u;
    }
#pragma unroll
    for (uint32_t r = 0; r < RADIX_ITEMS; ++r) {
        const bool valid = base + r * 64 < count;
        const uint32_t d = (kk[r] >> shift) & MASK;
        const uint64_t peers = digit_peers<DB>(d, valid);
        const uint32_t before = wcnt[wave][d];   // read by every lane before the leader's update
        rk[r] = before + uint32_t(__popcll(peers & lt));
        if (valid && (peers & lt) == 0) wcnt[wave][d] = before + uint32_t(__popcll(peers));
    }
    __syncthreads();
    // this thread's digits: tile-local starts, then each wavefront's start
    // within them
    uint32_t c[DPT][RADIX_WAVES], tot_t = 0, dsum = 0;
#pragma unroll
    for (uint32_t k = 0; k < DPT; ++k) {
        const uint32_t d = t * DPT + k;
#pragma unroll
        for (uint32_t w = 0; w < RADIX_WAVES; ++w) {
            c[k][w] = d < DIGITS ? wcnt[w][d] : 0u;
            tot_t += c[k][w];
        }
        dsum += dv[k];
    }
    uint32_t tot;
    uint32_t lstart = block_exclusive_scan<RADIX_WAVES>(tot_t, lds, tot);
    uint32_t digit_base = block_exclusive_scan<RADIX_WAVES>(dsum, lds, tot);
#pragma unroll
    for (uint32_t k = 0; k < DPT; ++k) {
        const uint32_t d = t * DPT + k;
        if (d >= DIGITS) break;
        gbase[d] = int32_t(digit_base + hv[k]) - int32_t(lstart);
        digit_base += dv[k];
#pragma unroll
        for (uint32_t w = 0; w < RADIX_WAVES; ++w) {
            wcnt[w][d] = lstart;
            lstart += c[k][w];
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < RADIX_ITEMS; ++r) {
        if (base + r * 64 < count) {
            const uint32_t pos = wcnt[wave][(kk[r] >> shift) & MASK] + rk[r];
            skey[pos] = kk[r];
            sid[pos] = ii[r];
        }
    }
    __syncthreads();
    if (GATHER) {
        float4 f[RADIX_ITEMS];
#pragma unroll
        for (uint32_t r = 0; r < RADIX_ITEMS; ++r) {
            const uint32_t q = r * RADIX_BLOCK + t;
            f[r] = F[sid[q < count ? q : 0u]];
        }
#pragma unroll
        for (uint32_t r = 0; r < RADIX_ITEMS; ++r) {
            const uint32_t q = r * RADIX_BLOCK + t;
            if (q < count) {
                const uint32_t key = skey[q];
                const uint32_t dst = uint32_t(gbase[(key >> shift) & MASK] + int32_t(q));
                kout[dst] = key;
                iout[dst] = sid[q];
                sP[dst] = make_float4(f[r].x, f[r].y, f[r].z, __uint_as_float(key));
            }
        }
        return;
    }
#pragma unroll
    for (uint32_t r = 0; r < RADIX_ITEMS; ++r) {
        const uint32_t q = r * RADIX_BLOCK + t;
        if (q < count) {
            const uint32_t key = skey[q];
            const uint32_t dst = uint32_t(gbase[(key >> shift) & MASK] + int32_t(q));
            kout[dst] = key;
            iout[dst] = sid[q];
        }
    }
}

// Hash table: group of 8 consecutive x-cells (key >> 3) -> (group, first
// sorted index of the group, cell offsets).  The offsets are 8 bytes: byte
// k - 1 (k = 1..8) counts the group's particles in its cells below k, so the
// particles of cell k sit at [start + off[k], start + off[k + 1]) with
// off[0] = 0 and off[8] the group's count.  A group of more than 255
// particles stores no offsets (off[8] = 0; a stored group is never empty):
// rows through it are scanned from the group start with key tests.  A 3-cell
// row [kmin, kmax] spans one or two groups, so one probe (two where it
// crosses a group boundary) gives its exact sorted range; one insert per
// occupied group keeps the atomic traffic ~8x below one per cell.  The table
// has >= 4 slots per particle (>= 32 per occupied group on a dense sheet):
// a wave waits for its longest probe chain, and on a 1024x1024 sheet halving
// the slots (or the group) lengthened those chains enough to cost 1.5-2.5x in
// the sweep; the skipped cells before kmin cost less than the extra probes.
constexpr uint32_t GROUP_SHIFT = 3;
static_assert((ONE_CLOTH_MAX_CELLS >> GROUP_SHIFT) <= ONE_CLOTH_GROUPS, "direct group table too small");
// Slot of a group key: the low 3 bits stay (8 consecutive groups = 64 cells
// share one 64-byte line, so threads walking the sorted order probe the same
// lines); the block index g >> 3 is placed by Fibonacci hashing, taking the
// HIGH bits of (g >> 3) * 2^32/phi.  Those depend on every key bit (a batch
// keeps its cloth index in the top bits; the low bits of a product would
// ignore it and pile all cloths onto the same slots), and runs of consecutive
// blocks, which a dense cell grid produces, land nearly evenly spread.
__device__ __forceinline__ uint32_t slot_of(uint32_t g, uint32_t mask) {
    const uint32_t bits = __popc(mask >> 3);   // log2(slots / 8); the table has >= 64 slots
    const uint32_t blk = ((g >> 3) * 2654435761u) >> (32u - bits);
    return ((blk << 3) | (g & 7u)) & mask;
}

// Lookup whose first probe `e` (the entry at slot_of(g)) was loaded by the
// caller: most lookups end there; a collision continues at the next slot.
// Returns the group's entry, or one with .y = EMPTY_KEY if it is absent.
__device__ __forceinline__ uint4 table_resolve(const uint4* table, uint32_t mask, uint32_t g, uint4 e) {
    const uint4 none = make_uint4(EMPTY_KEY, EMPTY_KEY, 0u, 0u);
    if (e.x == g) return e;
    if (e.x == EMPTY_KEY) return none;
    uint32_t h = (slot_of(g, mask) + 1) & mask;
    for (uint32_t probe = 1; probe <= mask; ++probe) {
        const uint4 f = table[h];
        if (f.x == g) return f;
        if (f.x == EMPTY_KEY) return none;
        h = (h + 1) & mask;
    }
    return none;
}

// off[k] of a table entry (k = 0..8; see above).
__device__ __forceinline__ uint32_t group_off(const uint4& e, uint32_t k) {
    if (k == 0) return 0u;
    const uint32_t w = k <= 4 ? e.z : e.w;
    return (w >> (8u * ((k - 1u) & 3u))) & 0xFFu;
}

// Pair kernel of self_collision_resolve (xpbd_cpu.rs:84-121) for the pair
// (min(i,j), max(i,j)); returns the contribution to particle i.  fi/fj carry
// the positions (xyz), wi/wj the inverse masses, ri/rj the rest positions.
__device__ __forceinline__ bool collision_contrib(uint32_t i, uint32_t j, float4 fi, float wi, float4 fj, float wj,
                                                  float4 ri, float4 rj, const DevParams& p, float& cx, float& cy,
                                                  float& cz) {
    const bool i_lo = i < j;
    const float4 pa = i_lo ? fi : fj, pb = i_lo ? fj : fi;
    const float wa = i_lo ? wi : wj, wb = i_lo ? wj : wi;
    if (wa <= 0.f && wb <= 0.f) return false;
    const float th2 = p.thickness * p.thickness;
    float dx = pb.x - pa.x, dy = pb.y - pa.y, dz = pb.z - pa.z;
    const float dist2 = (dx * dx + dy * dy) + dz * dz;
    if (dist2 > th2 || dist2 < 1e-18f) return false;
    const float4 ra = i_lo ? ri : rj, rb = i_lo ? rj : ri;
    const float ex = rb.x - ra.x, ey = rb.y - ra.y, ez = rb.z - ra.z;
    const float rest2 = (ex * ex + ey * ey) + ez * ez;
    if (dist2 > rest2) return false;
    float min_d = p.thickness;
    if (rest2 < th2) min_d = sqrtf(rest2);
    const float dist = sqrtf(dist2);
    const float corr = (min_d - dist) * 0.5f * p.coll_scale;
    if (corr <= 0.f) return false;
    dx = (dx / dist) * corr;
    dy = (dy / dist) * corr;
    dz = (dz / dist) * corr;
    const float inv_w = 1.0f / fmaxf(wa + wb, 1e-8f);
    if (i_lo) {
        cx = (-dx * wa) * inv_w;
        cy = (-dy * wa) * inv_w;
        cz = (-dz * wa) * inv_w;
    } else {
        cx = (dx * wb) * inv_w;
        cy = (dy * wb) * inv_w;
        cz = (dz * wb) * inv_w;
    }
    return true;
}

// Narrow phase + apply + post_velocity, one particle per thread in sorted
// order.  Contributions must be added in ascending partner id — the order the
// reference's O(N^2) loop adds them (xpbd_cpu.rs:82-123).  One pass over the
// 3x3 neighbour rows (each a contiguous run of the sorted keys) collects up to
// LMAX candidates inside the contact distance; a particle with more goes to
// k_coll_sweep_overflow.  Computing each contribution inside the divergent
// scan made every wavefront walk the contact path whenever one lane had a
// contact (2.3x the VALU instructions on a folded sheet); collected first,
// the contributions run in one short convergent loop (frame 65 of the
// 1024^2 sheet: 146 -> 88 us).
constexpr int LMAX = 8;   // candidates inside the contact distance kept per particle
// candidates loaded per window of a row scan.  With the table's cell offsets
// a row starts at its first particle and most rows hold 1-3: on the 1024^2
// sheet 4 measured best (sweep 42.9 us; 3, 5, 6: 44.5, 44.8, 47.9 us).
constexpr uint32_t SW = 4;

// A neighbour row's sorted range.  The scan starts at lo, tests each key
// against [kmin, kmax] and stops past kmax or at hi.  With the table's
// offsets lo is the row's first particle, and hi its end when the row lies
// in one group (n otherwise: the scan's key test ends it); a group without
// offsets gives its start.  lo = EMPTY_KEY: no particles.
struct RowRange {
    uint32_t lo, hi, kmin, kmax;
};

// The 3 rows of the neighbourhood at z offset dz - 1 (rows dy = 0, 1, 2):
// their first groups are probed together, then the second groups of rows
// whose first group is empty and that cross into the next group.
__device__ __forceinline__ void sweep_rows3(const CollArgs& a, const CollGrid& g, const uint32_t c[3], int dz,
                                            uint32_t ep, RowRange rows[3]) {
    uint32_t g0[3], g1[3];
    const int nz = int(c[2]) + dz - 1;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        rows[r].lo = EMPTY_KEY;
        rows[r].hi = 0u;
        rows[r].kmin = 1u;
        rows[r].kmax = 0u;
        g0[r] = g1[r] = EMPTY_KEY;
        const int ny = int(c[1]) + r - 1;
        if (ny < 0 || nz < 0 || ny >= int(g.dims[1]) || nz >= int(g.dims[2])) continue;
        const uint32_t kb = g.key_base + g.dims[0] * (uint32_t(ny) + g.dims[1] * uint32_t(nz));
        rows[r].kmin = kb + (c[0] > 0 ? c[0] - 1 : 0);
        rows[r].kmax = kb + (c[0] + 1 < g.dims[0] ? c[0] + 1 : c[0]);
        g0[r] = rows[r].kmin >> GROUP_SHIFT;
        const uint32_t gk = rows[r].kmax >> GROUP_SHIFT;
        g1[r] = gk != g0[r] ? gk : EMPTY_KEY;
    }
    // unconditional loads (slot 0 for absent rows), so no branch splits them;
    // a direct table's entry is the group's own (current iff its tag is ep)
    const bool direct = a.epoch != nullptr;
    const uint4 none = make_uint4(EMPTY_KEY, EMPTY_KEY, 0u, 0u);
    uint4 e[3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
        e[r] = a.table[g0[r] == EMPTY_KEY ? 0u : direct ? g0[r] : slot_of(g0[r], a.table_mask)];
    const uint32_t n = a.p.n;
    bool second[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        second[r] = false;
        if (g0[r] == EMPTY_KEY) continue;
        const uint4 f = direct ? (e[r].x == ep ? e[r] : none) : table_resolve(a.table, a.table_mask, g0[r], e[r]);
        if (f.y == EMPTY_KEY) {
            second[r] = g1[r] != EMPTY_KEY;
            continue;
        }
        const uint32_t cnt = group_off(f, 8);
        if (cnt == 0u) {   // no offsets
            rows[r].lo = f.y;
            rows[r].hi = n;
        } else {
            rows[r].lo = f.y + group_off(f, rows[r].kmin & 7u);
            rows[r].hi = g1[r] == EMPTY_KEY ? f.y + group_off(f, (rows[r].kmax & 7u) + 1u) : n;
        }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) e[r] = a.table[!second[r] ? 0u : direct ? g1[r] : slot_of(g1[r], a.table_mask)];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        if (!second[r]) continue;
        const uint4 f = direct ? (e[r].x == ep ? e[r] : none) : table_resolve(a.table, a.table_mask, g1[r], e[r]);
        if (f.y == EMPTY_KEY) continue;
        const uint32_t cnt = group_off(f, 8);
        rows[r].lo = f.y;
        rows[r].hi = cnt == 0u ? n : f.y + group_off(f, (rows[r].kmax & 7u) + 1u);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
        if (rows[r].lo != EMPTY_KEY && rows[r].lo >= rows[r].hi) rows[r].lo = EMPTY_KEY;
}

// Squared distance of two particles; symmetric bit for bit (the components
// of a - b and b - a differ only in sign), so it can screen candidates before
// the canonical (min id, max id) orientation is known.
__device__ __forceinline__ float dist2_of(float4 a, float4 b) {
    const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    return (dx * dx + dy * dy) + dz * dz;
}

__device__ __forceinline__ float clamp_apply(float v) {
    if (v < -COLLISION_APPLY_CLAMP) v = -COLLISION_APPLY_CLAMP;
    if (v > COLLISION_APPLY_CLAMP) v = COLLISION_APPLY_CLAMP;
    return v;
}

// Sorted chunk swept by block b.  Blocks reach the XCDs round-robin (block b
// on XCD b % 8), so with chunk = b every XCD's L2 fetches nearly every row of
// the sorted array that a neighbourhood touches.  With seg > 0, XCD x sweeps
// segments x, x + 8, x + 16, ... of `seg` consecutive chunks: neighbouring
// rows stay in one L2, and contact clusters still spread over the XCDs.
// A bijection on [0, nb): blocks past the last whole round of 8 segments keep
// chunk = b.
__device__ __forceinline__ uint32_t sweep_chunk(uint32_t b, uint32_t nb, uint32_t seg) {
    if (seg == 0) return b;
    const uint32_t whole = nb / (8u * seg) * (8u * seg);
    if (b >= whole) return b;
    const uint32_t x = b & 7u, l = b >> 3;
    return ((l / seg) * 8u + x) * seg + l % seg;
}

__global__ __launch_bounds__(256) void k_coll_sweep(CollArgs a) {
    const uint32_t n = a.p.n;
    const uint32_t s = sweep_chunk(blockIdx.x, gridDim.x, a.xcd_seg) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float4 fi = a.sP[s];    // xyz, w = cell key bits
    const uint32_t i = a.sids[s];
    const float4 pv = a.prev[i];  // substep start (the velocity update's), w = inverse mass
    const float wi = pv.w;
    if (wi <= 0.f) {
        if (a.xout != a.prev) a.xout[i] = make_float4(fi.x, fi.y, fi.z, wi);   // timing runs write scratch
        return;
    }
    const CollGrid g = a.grids[a.cloth_of ? a.cloth_of[i] : 0u];
    const uint32_t ep = a.epoch ? a.epoch[0] : 0u;
    const float th2 = a.p.thickness * a.p.thickness;
    uint32_t c[3];
    cell_of(fi, g, c);

    // the scan only collects the sorted indices of candidates inside the
    // contact distance (cheap, convergent); their partner ids are loaded
    // together and sorted, and the contributions computed afterwards in
    // ascending partner id
    // each row is read in fixed windows of SW candidates, checked with
    // predication; hits go to a per-thread LDS list (slot LMAX absorbs the
    // misses, so every candidate does one unconditional store)
    __shared__ uint32_t cand[LMAX + 1][256];
    const uint32_t t = threadIdx.x;
    uint32_t cnt = 0;
    for (int dz = 0; dz < 3; ++dz) {
    RowRange rows[3];
    sweep_rows3(a, g, c, dz, ep, rows);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        if (rows[r].lo == EMPTY_KEY) continue;
        const uint32_t kmin = rows[r].kmin, kmax = rows[r].kmax, hi = rows[r].hi;
        uint32_t idx = rows[r].lo;
        for (;;) {
            float4 ff[SW];
#pragma unroll
            for (uint32_t u = 0; u < SW; ++u) ff[u] = a.sP[min(idx + u, hi - 1)];
            uint32_t last = 0;
#pragma unroll
            for (uint32_t u = 0; u < SW; ++u) {
                const uint32_t q = idx + u;
                const uint32_t key = q < hi ? __float_as_uint(ff[u].w) : EMPTY_KEY;
                const float d2 = dist2_of(fi, ff[u]);
                const bool hit = key >= kmin && key <= kmax && q != s && !(d2 > th2 || d2 < 1e-18f);
                cand[hit ? min(cnt, uint32_t(LMAX)) : uint32_t(LMAX)][t] = q;
                cnt += hit ? 1u : 0u;
                last = key;
            }
            if (last > kmax || idx + SW >= hi) break;   // the row ends inside this window
            idx += SW;
        }
    }
    }
    uint32_t lq[LMAX];
#pragma unroll
    for (int k = 0; k < LMAX; ++k) lq[k] = k < int(cnt) ? cand[k][t] : s;
    if (cnt > LMAX) {   // resolved by k_coll_sweep_overflow
        const uint32_t k = atomicAdd(&a.ovf[0], 1u);
        if (k < n) a.ovf[2 + k] = s;   // at most n entries per launch
        return;
    }
    float ax = 0.f, ay = 0.f, az = 0.f;
    if (cnt > 0) {
        const float4 ri = a.rest4[i];   // rest xyz (w = inverse mass): contacts only
        uint32_t lj[LMAX];
#pragma unroll
        for (int k = 0; k < LMAX; ++k) {
            const uint32_t j = a.sids[lq[k]];   // unused slots read this particle's own id
            lj[k] = k < cnt ? j : EMPTY_KEY;
        }
        // Batcher odd-even merge network on 8 (partner ids are distinct)
#define XPBD_CX(x, y)                                   \
        if (lj[y] < lj[x]) {                            \
            const uint32_t tj = lj[x]; lj[x] = lj[y]; lj[y] = tj; \
            const uint32_t tq = lq[x]; lq[x] = lq[y]; lq[y] = tq; \
        }
        XPBD_CX(0, 1) XPBD_CX(2, 3) XPBD_CX(4, 5) XPBD_CX(6, 7)
        XPBD_CX(0, 2) XPBD_CX(1, 3) XPBD_CX(4, 6) XPBD_CX(5, 7)
        XPBD_CX(1, 2) XPBD_CX(5, 6)
        XPBD_CX(0, 4) XPBD_CX(1, 5) XPBD_CX(2, 6) XPBD_CX(3, 7)
        XPBD_CX(2, 4) XPBD_CX(3, 5)
        XPBD_CX(1, 2) XPBD_CX(3, 4) XPBD_CX(5, 6)
#undef XPBD_CX
#pragma unroll
        for (int k = 0; k < LMAX; ++k) {
            if (k < cnt) {
                const float4 fj = a.sP[lq[k]];
                const float4 rj = a.rest4[lj[k]];
                float cx, cy, cz;
                if (collision_contrib(i, lj[k], fi, wi, fj, rj.w, ri, rj, a.p, cx, cy, cz)) {
                    ax = ax + cx;
                    ay = ay + cy;
                    az = az + cz;
                }
            }
        }
    }
    const float4 nx = make_float4(fi.x + clamp_apply(ax), fi.y + clamp_apply(ay), fi.z + clamp_apply(az), wi);
    finish_particle(nx, pv, i, a.ground, a.p, a.xout, a.vout);
}

// Particles with more than LMAX candidates.  One pass over the 27 cells
// collects up to OVF_MAX contacts into a list kept sorted by partner id in
// LDS (one list per thread); more than that falls back to repeated selection
// of the next-smallest partner.  Either way the sum runs in ascending partner
// id, the brute-force loop's order.  The list's count is cleared by
// k_coll_table, which runs before every sweep.
constexpr int OVF_MAX = 32;
constexpr uint32_t OVF_BLOCK = 64;

__global__ __launch_bounds__(OVF_BLOCK) void k_coll_sweep_overflow(CollArgs a) {
    __shared__ uint32_t lj[OVF_MAX][OVF_BLOCK];
    __shared__ float4 lc[OVF_MAX][OVF_BLOCK];
    const uint32_t n = a.p.n;
    const uint32_t total = min(__hip_atomic_load(&a.ovf[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), n);
    const float th2 = a.p.thickness * a.p.thickness;
    const uint32_t t = threadIdx.x;
    for (uint32_t k = blockIdx.x * blockDim.x + t; k < total; k += gridDim.x * blockDim.x) {
        const uint32_t s = a.ovf[2 + k];
        const float4 fi = a.sP[s];
        const uint32_t i = a.sids[s];
        const float4 pv = a.prev[i];
        const float4 ri = a.rest4[i];
        const float wi = pv.w;
        const CollGrid g = a.grids[a.cloth_of ? a.cloth_of[i] : 0u];
        const uint32_t ep = a.epoch ? a.epoch[0] : 0u;
        uint32_t c[3];
        cell_of(fi, g, c);
        RowRange rows[9];
#pragma unroll
        for (int dz = 0; dz < 3; ++dz) sweep_rows3(a, g, c, dz, ep, rows + 3 * dz);
        int cnt = 0;
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            if (rows[r].lo == EMPTY_KEY) continue;
            for (uint32_t idx = rows[r].lo; idx < rows[r].hi; ++idx) {
                const float4 fj = a.sP[idx];
                const uint32_t key = __float_as_uint(fj.w);
                if (key > rows[r].kmax) break;
                if (key < rows[r].kmin || idx == s) continue;
                const float d2 = dist2_of(fi, fj);
                if (d2 > th2 || d2 < 1e-18f) continue;
                const uint32_t j = a.sids[idx];
                const float4 rj = a.rest4[j];
                float cx, cy, cz;
                if (!collision_contrib(i, j, fi, wi, fj, rj.w, ri, rj, a.p, cx, cy, cz)) continue;
                if (cnt < OVF_MAX) {   // insertion into the sorted list
                    int w = cnt;
                    while (w > 0 && lj[w - 1][t] > j) {
                        lj[w][t] = lj[w - 1][t];
                        lc[w][t] = lc[w - 1][t];
                        --w;
                    }
                    lj[w][t] = j;
                    lc[w][t] = make_float4(cx, cy, cz, 0.f);
                }
                ++cnt;
            }
        }
        float ax = 0.f, ay = 0.f, az = 0.f;
        if (cnt <= OVF_MAX) {
            for (int q = 0; q < cnt; ++q) {
                const float4 v = lc[q][t];
                ax = ax + v.x;
                ay = ay + v.y;
                az = az + v.z;
            }
        } else {
            uint32_t last = 0;
            bool has_last = false;
            for (;;) {
                uint32_t best = EMPTY_KEY;
                float bx = 0.f, by = 0.f, bz = 0.f;
#pragma unroll
                for (int r = 0; r < 9; ++r) {
                    if (rows[r].lo == EMPTY_KEY) continue;
                    for (uint32_t idx = rows[r].lo; idx < rows[r].hi; ++idx) {
                        const float4 fj = a.sP[idx];
                        const uint32_t key = __float_as_uint(fj.w);
                        if (key > rows[r].kmax) break;
                        if (key < rows[r].kmin || idx == s) continue;
                        const float d2 = dist2_of(fi, fj);
                        if (d2 > th2 || d2 < 1e-18f) continue;
                        const uint32_t j = a.sids[idx];
                        if (j >= best || (has_last && j <= last)) continue;
                        const float4 rj = a.rest4[j];
                        float cx, cy, cz;
                        if (collision_contrib(i, j, fi, wi, fj, rj.w, ri, rj, a.p, cx, cy, cz)) {
                            best = j;
                            bx = cx;
                            by = cy;
                            bz = cz;
                        }
                    }
                }
                if (best == EMPTY_KEY) break;
                ax = ax + bx;
                ay = ay + by;
                az = az + bz;
                last = best;
                has_last = true;
            }
        }
        const float4 nx = make_float4(fi.x + clamp_apply(ax), fi.y + clamp_apply(ay), fi.z + clamp_apply(az), wi);
        finish_particle(nx, pv, i, a.ground, a.p, a.xout, a.vout);
    }
}

// Cell offsets of the group starting at sorted index s (see the table
// notes): the group's keys are read 8 at a time until one leaves the group;
// a group of more than 255 particles gets none (both words 0).  Keys below
// index lend come from the block's staged copy lk (index - l0), later ones
// from skeys.
constexpr uint32_t GATHER_BLOCK = 256;
constexpr uint32_t GATHER_STAGED = 2 * GATHER_BLOCK;   // the block's keys and the next block's

__device__ __forceinline__ uint2 group_offsets(const uint32_t* lk, uint32_t l0, uint32_t lend, const uint32_t* skeys,
                                               uint32_t n, uint32_t s, uint32_t g) {
    uint32_t lo = 0u, hi = 0u, cnt = 0u;
    for (uint32_t b = s; b < n; b += 8) {
        uint32_t kk[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
            const uint32_t q = b + u;
            kk[u] = q < lend ? lk[q - l0] : (q < n ? skeys[q] : EMPTY_KEY);
        }
        bool more = true;
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
            const bool in = more && (kk[u] >> GROUP_SHIFT) == g && kk[u] != EMPTY_KEY;
            more = in;
            if (in) {
                // byte k - 1 (k = 1..8) counts the cells below k: add 1 to
                // bytes k - 1 >= c of the particle's cell c
                const uint32_t c = kk[u] & 7u;
                lo += c < 4u ? (0x01010101u << (8u * c)) : 0u;
                hi += c < 4u ? 0x01010101u : (0x01010101u << (8u * (c - 4u)));
                ++cnt;
            }
        }
        if (!more || cnt > 255u) break;
    }
    return cnt > 255u ? make_uint2(0u, 0u) : make_uint2(lo, hi);
}

// Group-table entries over the sorted keys: the first sorted index of each
// group, with its cell offsets.  Direct table (one cloth, epoch != nullptr):
// entry g = (pass tag, start, offsets), one plain store per group.  Hashed
// table (batches): inserted by CAS probing, the slot recorded in slot_rec
// (EMPTY_KEY elsewhere) for k_coll_keys to empty next pass.  The block
// stages its keys and the next block's in LDS, so the offsets of groups
// shorter than that read no global memory.  Also clears the sweep's
// overflow count.
__global__ __launch_bounds__(GATHER_BLOCK) void k_coll_table(const uint32_t* skeys, uint32_t n, uint32_t* ovf,
                                                           uint4* table, uint32_t mask, uint32_t* slot_rec,
                                                           const uint32_t* epoch) {
    __shared__ uint32_t lk[GATHER_STAGED];
    const uint32_t l0 = blockIdx.x * GATHER_BLOCK;
    const uint32_t s = l0 + threadIdx.x;
    if (ovf && s == 0) ovf[0] = 0u;
#pragma unroll
    for (uint32_t u = 0; u < GATHER_STAGED / GATHER_BLOCK; ++u) {
        const uint32_t q = l0 + u * GATHER_BLOCK + threadIdx.x;
        lk[u * GATHER_BLOCK + threadIdx.x] = q < n ? skeys[q] : EMPTY_KEY;
    }
    __syncthreads();
    if (s >= n) return;
    const uint32_t key = lk[threadIdx.x];
    uint32_t rec = EMPTY_KEY;
    const uint32_t g = key >> GROUP_SHIFT;
    const uint32_t prev = threadIdx.x > 0 ? lk[threadIdx.x - 1] : (s > 0 ? skeys[s - 1] : EMPTY_KEY);
    const bool first = s == 0 || (prev >> GROUP_SHIFT) != g;
    if (epoch) {
        if (first) {
            const uint2 off = group_offsets(lk, l0, min(l0 + GATHER_STAGED, n), skeys, n, s, g);
            table[g] = make_uint4(epoch[0], s, off.x, off.y);
        }
        return;
    }
    if (first) {
        const uint2 off = group_offsets(lk, l0, min(l0 + GATHER_STAGED, n), skeys, n, s, g);
        uint32_t h = slot_of(g, mask);
        for (uint32_t probe = 0; probe <= mask; ++probe) {
            if (atomicCAS(&table[h].x, EMPTY_KEY, g) == EMPTY_KEY) {
                table[h].y = s;
                table[h].z = off.x;
                table[h].w = off.y;
                rec = h;
                break;
            }
            h = (h + 1) & mask;
        }
    }
    slot_rec[s] = rec;
}

// Standalone predict (xpbd_cpu.rs:251-264): P = predict(X, V).
__global__ void k_predict(const float4* X, const float4* V, float4* P, DevParams p, uint8_t* tame_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool t = true;
    if (i < p.n) {
        const float4 q = predict(X[i], V[i], i, p);
        P[i] = q;
        t = tame3(q);
    }
    if (tame_out) write_slice_tame(tame_out, i >> 6, i < p.n, t);   // blocks are whole slices
}

__global__ void k_predict_gated(const float4* X, const float4* V, float4* P, DevParams p, const uint32_t* gate) {
    if (*gate == 0u) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    P[i] = predict(X[i], V[i], i, p);
}

__global__ void k_copy4_if(const float4* a, float4* b, uint32_t n, const uint32_t* gate) {
    if (*gate == 0u) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// Warm-start acceptance in one launch.  The relative residual ||C|| / ||rest||
// of X over the constraints (ij, rest; ordered so that neighbouring threads
// load neighbouring vertices): C in f32 per constraint (hardware
// sqrt, about 1 ulp: a gate statistic, not the reported residual), summed in
// f32 per thread and in double across threads.  Every block stores its sum
// in its own slot, then arrives on one of WG_SLOTS counters (block b on
// counter b mod WG_SLOTS, each on its own 256-byte line, so that same-address
// atomics form short parallel chains); the last block of a counter arrives on
// the top counter, and the last of those adds the block slots in block order
// (a fixed-shape reduction: the total and the decision do not depend on
// arrival order), compares the total with the threshold -- a NaN residual
// falls back -- sets flag[0] (1: fall back), counts the substep (flag[1]) and
// the fallback (flag[2]), and clears the grid-barrier words of the fallback
// launch.  Each counter is reset by the block that reads it last, so the next
// launch finds zeros.
constexpr uint32_t WG_LINE = 64;           // words per counter line
constexpr uint32_t WG_SLOTS = 32;          // arrival counters (line 0 is the top counter)
constexpr uint32_t WG_MAX_BLOCKS = 2048;   // block sums (doubles) after the counter lines

__global__ __launch_bounds__(256) void k_warm_gate(const float4* X, const uint2* ij, const float* rest, uint32_t e,
                                                  double rest_sq, float threshold, uint32_t* acc, uint32_t* flag,
                                                  uint32_t* barrier) {
    float sf = 0.f;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < e; k += gridDim.x * blockDim.x) {
        const uint2 q = ij[k];
        const float4 a = X[q.x], b = X[q.y];
        const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
        const float c = __builtin_amdgcn_sqrtf((dx * dx + dy * dy) + dz * dz) - rest[k];
        sf += c * c;
    }
    double sd = sf;
    for (int off = 32; off > 0; off >>= 1) sd += __shfl_xor(sd, off, 64);
    __shared__ double red[4];
    __shared__ uint32_t last;
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sd;
    __syncthreads();
    double* part = reinterpret_cast<double*>(acc + WG_LINE * (WG_SLOTS + 1));
    if (threadIdx.x == 0) {
        const uint32_t slot = blockIdx.x % WG_SLOTS;
        const uint32_t in_slot = (gridDim.x - slot + WG_SLOTS - 1) / WG_SLOTS;
        const uint32_t slots = min(gridDim.x, WG_SLOTS);
        uint32_t* counter = acc + WG_LINE * (1 + slot);
        last = 0u;
        __hip_atomic_store(part + blockIdx.x, (red[0] + red[1]) + (red[2] + red[3]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        // the block sum is written before this block is counted
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        if (__hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 == in_slot) {
            __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__hip_atomic_fetch_add(acc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 == slots) {
                __hip_atomic_store(acc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                last = 1u;
            }
        }
    }
    __syncthreads();
    if (!last) return;
    // all of this thread's slots in flight at once, then added in slot order
    constexpr uint32_t PER_THREAD = WG_MAX_BLOCKS / 256;
    double v[PER_THREAD];
#pragma unroll
    for (uint32_t r = 0; r < PER_THREAD; ++r) {
        const uint32_t b = threadIdx.x + 256 * r;
        v[r] = b < gridDim.x ? __hip_atomic_load(part + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
    }
    double t = 0.0;
#pragma unroll
    for (uint32_t r = 0; r < PER_THREAD; ++r) t += v[r];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double tot = (red[0] + red[1]) + (red[2] + red[3]);
        const double rel = rest_sq > 0.0 ? sqrt(tot) / sqrt(rest_sq) : 0.0;
        const uint32_t f = rel <= double(threshold) ? 0u : 1u;
        flag[0] = f;
        flag[1] = flag[1] + 1u;
        flag[2] = flag[2] + f;
    }
    for (uint32_t w = threadIdx.x; w < GRID_SYNC_WORDS; w += blockDim.x) barrier[w] = 0u;
}

// The state update from the warm result F, or from alt when the gate chose
// the fallback (*gate != 0).
__global__ void k_finalize_sel(const float4* F, const float4* alt, const uint32_t* gate, float4* X, float4* V,
                               const uint8_t* ground, DevParams p) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const float4 f = (*gate != 0u ? alt : F)[i];
    if (f.w <= 0.f) return;
    finish_particle(f, X[i], i, ground, p, X, V);
}

// ---------------------------------------------------------------------------
// Coloured Gauss-Seidel: one persistent launch per substep.  Colours run in
// order with a grid barrier between them; a colour's constraints share no
// particle, so the parallel sweep equals the sequential one in colour-sorted
// order (xpbd_cpu.rs:344-386; gs_edges, cloth_sim.wgsl:154-210).
// ---------------------------------------------------------------------------
// Grid barrier for co-resident persistent launches.  Device memory from
// hipMalloc is cached in each XCD's L2 without coherence between XCDs, so data
// stored before a barrier has to be written back from the producing XCD's L2
// and invalidated in the consuming one.  Doing that in every block (an
// agent-scope release + acquire per block) costs ~95 us per barrier at 2048
// blocks; here the last block to arrive on each XCD does it once for the whole
// XCD (one L2 writeback per XCD) and the others only acquire (invalidate).  Blocks find their XCD
// from the XCC_ID hardware register.  Every wait is bounded: on expiry the
// error word is set and the kernel runs to its end instead of hanging.
//
// Words (each hot counter on its own 256-byte line, zeroed before the launch):
// [0] XCD leaders arrived, [L] generation, [2L] blocks counted at set-up,
// [(4+x)L] blocks of XCD x arrived, [(12+x)L] XCD x generation,
// [(20+x)L] blocks resident on XCD x.
constexpr uint32_t GS_LINE = 64;
static_assert(GRID_SYNC_WORDS >= 28 * GS_LINE, "grid barrier words");

__device__ __forceinline__ uint32_t gs_load(uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t gs_add(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Once any wait has expired (the error word is set, and it stays set until
// the host reads it), every later wait of every block returns at once, so a
// launch whose grid was not co-resident drains after one expiry instead of
// spending the full bound at each of its remaining barriers.
__device__ __forceinline__ void gs_wait(uint32_t* p, uint32_t v, uint32_t* err) {
    if (gs_load(err)) return;
    uint32_t spins = 0;
    while (gs_load(p) < v) {
        __builtin_amdgcn_s_sleep(1);
        ++spins;
        if ((spins & 1023u) == 0 && gs_load(err)) return;
        if (spins > (1u << 22)) {   // seconds at most: report instead of hanging
            atomicExch(err, 1u);
            return;
        }
    }
}

struct GridSync {
    uint32_t* w;
    uint32_t* err;
    uint32_t xcd, nx, nxcd, epoch;

    __device__ void init(uint32_t* words, uint32_t* error) {
        __shared__ uint32_t sh[2];
        w = words;
        err = error;
        epoch = 0;
        if (threadIdx.x == 0) {
            const uint32_t x = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 7u;   // HW_REG_XCC_ID[3:0]
            gs_add(w + GS_LINE * (20 + x), 1u);
            __hip_atomic_fetch_add(w + GS_LINE * 2, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            gs_wait(w + GS_LINE * 2, gridDim.x, err);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            uint32_t n = 0;
            for (uint32_t k = 0; k < 8; ++k) n += gs_load(w + GS_LINE * (20 + k)) ? 1u : 0u;
            sh[0] = x;
            sh[1] = gs_load(w + GS_LINE * (20 + x)) | (n << 16);
        }
        __syncthreads();
        xcd = sh[0];
        nx = sh[1] & 0xFFFFu;
        nxcd = sh[1] >> 16;
    }

    __device__ __forceinline__ void sync() {
        // __syncthreads does not wait for global stores (within a CU they are
        // ordered through the shared L1): each wave drains its own first, so
        // they are in L2 before the XCD leader writes L2 back.
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        epoch += 1;
        if (threadIdx.x == 0) {
            const uint32_t old = gs_add(w + GS_LINE * (4 + xcd), 1u);
            if (old + 1 == epoch * nx) {   // last block of this XCD
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");   // write back this XCD's L2
                if (gs_add(w, 1u) + 1 == epoch * nxcd)
                    __hip_atomic_store(w + GS_LINE, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                gs_wait(w + GS_LINE, epoch, err);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // invalidate this XCD's L2
                __hip_atomic_store(w + GS_LINE * (12 + xcd), epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                gs_wait(w + GS_LINE * (12 + xcd), epoch, err);
                // drop stale lines of this CU's L1 (gfx950 has no L1-only
                // invalidate that works here: the workgroup-scope one is a
                // no-op, so this is the agent-scope one, once per block)
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
        }
        __syncthreads();
    }
};

// The GS work iterate through ordinary loads and stores (ordered between
// blocks by launch boundaries or the grid barrier's L2 writeback).
struct PlainW {
    float4* W;
    __device__ __forceinline__ float4 ld(uint32_t i) const { return W[i]; }
    __device__ __forceinline__ void st(uint32_t i, float4 v) const { W[i] = v; }
};

// The GS work iterate at device scope, for blocks that hand particles to each
// other through flags without any L2 writeback or invalidate: buffer accesses
// with the SC1 cache policy (the policy of agent-scope atomics on this chip)
// miss the CU's L1 and the XCD's L2 copies on load and write through to the
// device's coherence point on store.  A store has reached that point once the
// wave's vmcnt has drained, so a flag stored after s_waitcnt vmcnt(0) and a
// barrier is seen after the data it announces.  The buffer range (n x 16
// bytes) bounds every access.
constexpr int CPOL_SC1 = 16;
typedef float v4f32 __attribute__((ext_vector_type(4)));
struct DeviceW {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ DeviceW(float4* W, uint32_t n)
        : r(__builtin_amdgcn_make_buffer_rsrc(W, 0, int(n * 16u), 0x00020000)) {}
    __device__ __forceinline__ float4 ld(uint32_t i) const {
        const v4f32 v = __builtin_amdgcn_raw_buffer_load_b128(r, int(i * 16u), 0, CPOL_SC1);
        return make_float4(v.x, v.y, v.z, v.w);
    }
    __device__ __forceinline__ void st(uint32_t i, float4 x) const {
        const v4f32 v = {x.x, x.y, x.z, x.w};
        __builtin_amdgcn_raw_buffer_store_b128(v, r, int(i * 16u), 0, CPOL_SC1);
    }
};

// One constraint of coloured Gauss-Seidel (xpbd_cpu.rs:312-391): project
// constraint e of the colour-sorted arrays on the work iterate W in place.
// Constraints of one colour share no vertex, so any order within a colour,
// and any split of a colour over launches, blocks or tiles, gives the same bits.
// gs_apply: the projection of constraint (ij, rc) from its endpoints' loaded
// iterates pi, pj; gs_project: the same with its own loads.
template <class Mem>
__device__ __forceinline__ void gs_apply(const GsArgs& a, uint2 ij, float2 rc, float4 pi, float4 pj, const Mem& m) {
    const DevParams& p = a.p;
    const float wi = pi.w, wj = pj.w;
    if (wj <= 0.f && wi <= 0.f) return;
    float gx = pi.x - pj.x, gy = pi.y - pj.y, gz = pi.z - pj.z;
    const float len = len3(gx, gy, gz);
    if (len < 1e-8f) return;
    gx = gx / len;
    gy = gy / len;
    gz = gz / len;
    const float c = len - rc.x;
    const float alpha_t = rc.y * p.inv_dt_sq;
    const float wsum = wi + wj + alpha_t;
    if (wsum < 1e-8f) return;
    const float dlam = (-c - alpha_t * 0.0f) / wsum;
    const float ox = p.omega * gx, oy = p.omega * gy, oz = p.omega * gz;
    float ix = (ox * wi) * dlam, iy = (oy * wi) * dlam, iz = (oz * wi) * dlam;
    float jx = -((ox * wj) * dlam), jy = -((oy * wj) * dlam), jz = -((oz * wj) * dlam);
    cap_correction(ix, iy, iz);
    cap_correction(jx, jy, jz);
    if (wi > 0.f) m.st(ij.x, make_float4(pi.x + ix, pi.y + iy, pi.z + iz, wi));
    if (wj > 0.f) m.st(ij.y, make_float4(pj.x + jx, pj.y + jy, pj.z + jz, wj));
}

template <class Mem>
__device__ __forceinline__ void gs_project(const GsArgs& a, uint32_t e, const Mem& m) {
    const uint2 ij = a.ij[e];
    const float4 pi = m.ld(ij.x), pj = m.ld(ij.y);
    gs_apply(a, ij, a.rc[e], pi, pj, m);
}

// One colour of one GS iteration as its own launch: constraints [e0, e1) of
// the colour-sorted arrays, one per thread.  The launch boundary orders the
// colours, in place of a grid barrier.
__global__ __launch_bounds__(256) void k_gs_colour(GsArgs a, uint32_t e0, uint32_t e1) {
    const uint32_t e = e0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (e < e1) gs_project(a, e, PlainW{a.W});
}

__global__ __launch_bounds__(256) void k_gs_substep(GsArgs a) {
    const DevParams& p = a.p;
    const uint32_t nthreads = gridDim.x * blockDim.x;
    const uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x;
    GridSync gsync;
    gsync.init(a.barrier, a.error);
    if (a.probe) {   // uniform: the co-residency check of this very kernel's grid
        gsync.sync();
        return;
    }
    for (uint32_t i = gt; i < p.n; i += nthreads) a.W[i] = predict(a.X[i], a.V[i], i, p);
    gsync.sync();
    for (uint32_t it = 0; it < a.iterations; ++it) {
        for (uint32_t b = 0; b < a.num_batches; ++b) {
            const uint32_t e1 = a.batch_offsets[b + 1];
            for (uint32_t e = a.batch_offsets[b] + gt; e < e1; e += nthreads) gs_project(a, e, PlainW{a.W});
            gsync.sync();
        }
    }
    if (a.final_write) {
        for (uint32_t i = gt; i < p.n; i += nthreads) {
            const float4 w = a.W[i];
            if (w.w > 0.f) finish_particle(w, a.X[i], i, a.ground, p, a.xout, a.vout);
        }
    } else {
        // bounds of the final iterate for the collision pass that follows
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = gt; i < p.n; i += nthreads) {
            const float4 w = a.W[i];
            aabb_fold(true, w.x, w.y, w.z, lo, hi);
        }
        block_aabb_store(lo, hi, a.aabb, blockIdx.x);
    }
}

// Wait until every tile in tile T's dependency list has finished step s - 1
// (its flag >= s).  One wavefront polls, a lane per dependency; the wait is
// bounded like gs_wait (on expiry the error word is set and this and every
// later wait of the launch return at once).
// Called by the lanes of wavefront 0: poll until each lane's flag (nullptr:
// none) is >= s, sleeping SLEEP x 64 clocks between polls.
template <int SLEEP = 1>
__device__ __forceinline__ void poll_flags(uint32_t* f, uint32_t s, uint32_t* err) {
    if (!__builtin_amdgcn_readfirstlane(gs_load(err))) {
        for (uint32_t spins = 1;; ++spins) {
            if (__all(f == nullptr || gs_load(f) >= s)) break;
            __builtin_amdgcn_s_sleep(SLEEP);
            if ((spins & 1023u) == 0 && __builtin_amdgcn_readfirstlane(gs_load(err))) break;
            if (spins > (1u << 22) / (1u + SLEEP / 32)) {   // seconds at most: report instead of hanging
                if (threadIdx.x == 0) atomicExch(err, 1u);
                break;
            }
        }
    }
}

__device__ __forceinline__ void gs_tile_wait(const GsTiles& t, uint32_t T, uint32_t s, uint32_t* err) {
    if (threadIdx.x < 64) {
        const uint32_t d0 = t.doff[T], nd = t.doff[T + 1] - d0;
        uint32_t* f = threadIdx.x < nd ? t.flag + t.dep[d0 + threadIdx.x] : nullptr;
        if (nd) poll_flags<16>(f, s, err);   // 16 x 64 clocks apart: 748 -> 766 at 1024^2, +4% at 256^2
    }
    __syncthreads();
}

// The end of a flagged step: every wave's stores drained, then the flag.
__device__ __forceinline__ void flag_step(uint32_t* flag, uint32_t v) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(flag, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One GS substep (xpbd_cpu.rs:312-391 inside the substep of :240-420) with
// the grid barriers replaced by per-tile flags.  Block b owns tiles b,
// b + gridDim.x, ...; for each step in order it runs each of its tiles'
// share: the predict of the tile's particles, then each colour's constraints
// of the tile, then the final write (or the bounds) of its particles.  A step
// waits only for the tiles whose particles the tile's constraints can share,
// so neighbouring tiles run several colours apart instead of every colour
// waiting for the slowest block.  Deadlock-free for a co-resident grid: the
// lowest unfinished step only waits for finished ones.  The iterate goes
// through DeviceW; a tile's flag is stored after its stores have drained.
// Constraints per thread whose loads go out together in a tile step.  2 on
// the 1024^2 sheet (1024-particle tiles): 771 substeps/s against 734 for 1
// and 654 for 4, whose registers leave too few blocks resident for one tile
// each.
constexpr uint32_t GS_TILE_K = 2;

__global__ __launch_bounds__(256) void k_gs_tiles(GsArgs a, GsTiles t) {
    if (a.probe) {   // uniform: the co-residency check of this kernel's grid
        GridSync gsync;
        gsync.init(a.barrier, a.error);
        return;
    }
    const DevParams& p = a.p;
    const DeviceW w(a.W, p.n);
    const uint32_t nb = a.num_batches;
    const uint32_t last = 1 + a.iterations * nb;   // the final step
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t s = 0; s <= last; ++s) {
        const bool colour = s > 0 && s < last;
        const uint32_t* to = colour ? t.toff + ((s - 1) % nb) * (t.ntiles + 1) : nullptr;
        for (uint32_t T = blockIdx.x; T < t.ntiles; T += gridDim.x) {
            // the tile's constraint rows of this step do not depend on other
            // tiles: fetched before the wait, GS_TILE_K per thread
            uint32_t e0 = 0, e1 = 0;
            uint2 ij[GS_TILE_K];
            float2 rc[GS_TILE_K];
            if (colour) {
                e0 = to[T] + threadIdx.x;
                e1 = to[T + 1];
#pragma unroll
                for (uint32_t k = 0; k < GS_TILE_K; ++k) {
                    const uint32_t e = e0 + k * blockDim.x;
                    ij[k] = e < e1 ? a.ij[e] : make_uint2(0u, 0u);
                    rc[k] = e < e1 ? a.rc[e] : make_float2(0.f, 0.f);
                }
            }
            if (s > 0) gs_tile_wait(t, T, s, a.error);
            const uint32_t i0 = T * t.tile, i1 = min(i0 + t.tile, p.n);
            if (s == 0) {
                for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) w.st(i, predict(a.X[i], a.V[i], i, p));
            } else if (colour) {
                // the constraints of one colour share no particle: all loads
                // first, then the updates
                float4 pi[GS_TILE_K], pj[GS_TILE_K];
#pragma unroll
                for (uint32_t k = 0; k < GS_TILE_K; ++k) {
                    if (e0 + k * blockDim.x < e1) {
                        pi[k] = w.ld(ij[k].x);
                        pj[k] = w.ld(ij[k].y);
                    }
                }
#pragma unroll
                for (uint32_t k = 0; k < GS_TILE_K; ++k)
                    if (e0 + k * blockDim.x < e1) gs_apply(a, ij[k], rc[k], pi[k], pj[k], w);
                for (uint32_t e = e0 + GS_TILE_K * blockDim.x; e < e1; e += blockDim.x) gs_project(a, e, w);
            } else {
                for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
                    const float4 x = w.ld(i);
                    if (a.final_write) {
                        if (x.w > 0.f) finish_particle(x, a.X[i], i, a.ground, p, a.xout, a.vout);
                    } else {
                        aabb_fold(true, x.x, x.y, x.z, lo, hi);
                    }
                }
                continue;   // nothing waits for the final step
            }
            flag_step(t.flag + T, s + 1);
        }
    }
    if (!a.final_write) block_aabb_store(lo, hi, a.aabb, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Coloured GS on grid tiles resident in LDS, ordered by neighbour flags.  One
// grid cloth (row-major, cols a multiple of 64) is cut into tiles of 64
// columns x th rows, one block each, all resident at once.  A block keeps its
// tile and a 2-particle ring around it (the window, (th + 4) x 68 particles)
// in LDS for the whole substep.  Its constraints (those whose lower endpoint
// is in the tile) reach at most 2 rows down and 2 columns either way, so they
// stay inside the window; a particle more than 2 from the tile's edge is
// touched by no other tile's constraints and never leaves LDS between the
// predict and the final write.  The shared particles (the ring inside and
// outside the tile's edge) go through the iterate at device scope: each step
// re-reads them after the 8 neighbour tiles have finished the previous step,
// and writes back the ones its constraints change.  A neighbour may change a
// shared particle during the same step; a colour's constraints share no
// vertex, so this tile neither uses nor writes that particle in that step and
// reads it again in the next one.
constexpr uint32_t GSG_W = 64;              // tile columns
constexpr uint32_t GSG_WW = GSG_W + 4;      // window columns
constexpr uint32_t GSG_SHARED = 0x8000u;    // window position flag: written back to the iterate

__device__ __forceinline__ bool gsg_shared(uint32_t lr, uint32_t lc, uint32_t th) {
    return lr < 4 || lr >= th || lc < 4 || lc >= GSG_W;
}

// Window accessor for gs_apply: positions are window indices with
// GSG_SHARED set where the particle is written back.
struct TileWindow {
    float4* win;
    DeviceW w;
    int gbase;   // global index of window position 0: (r0 - 2) * cols + c0 - 2
    uint32_t cols;
    __device__ __forceinline__ float4 ld(uint32_t x) const { return win[x & (GSG_SHARED - 1u)]; }
    __device__ __forceinline__ void st(uint32_t x, float4 v) const {
        const uint32_t k = x & (GSG_SHARED - 1u);
        win[k] = v;
        if (x & GSG_SHARED) w.st(uint32_t(gbase + int(k / GSG_WW) * int(cols) + int(k % GSG_WW)), v);
    }
};

// Re-read the window's shared particles that a constraint of this tile or of
// a neighbour can touch (window rows 2..3 and th..th+3 whole, columns 0..3 and
// 64..67 of rows 4..th-1; rows 0..1 lie above every endpoint of this tile).
// Threads [0, nthreads) of the block take part, thread `tid` from k = tid.
__device__ __forceinline__ void gsg_ring_load(const TileWindow& m, int r0, int c0, uint32_t th, uint32_t rows,
                                              uint32_t tid, uint32_t nthreads) {
    const uint32_t nfull = 6 * GSG_WW, nring = nfull + (th - 4) * 8;
    for (uint32_t k = tid; k < nring; k += nthreads) {
        uint32_t lr, lc;
        if (k < nfull) {
            const uint32_t q = k / GSG_WW;
            lr = q < 2 ? 2 + q : th + q - 2;
            lc = k % GSG_WW;
        } else {
            const uint32_t q = k - nfull, c8 = q % 8;
            lr = 4 + q / 8;
            lc = c8 < 4 ? c8 : GSG_W + c8 - 4;
        }
        const int r = r0 - 2 + int(lr), c = c0 - 2 + int(lc);
        if (r >= 0 && r < int(rows) && c >= 0 && c < int(m.cols))
            m.win[lr * GSG_WW + lc] = m.w.ld(uint32_t(r) * m.cols + uint32_t(c));
    }
}

// One grid-tile step's constraint rows held in registers: this thread's first
// GSG_PRE rounds of the interior group [e0, em) and its first round of the
// shared group [em, e1) (emt: its row).
constexpr uint32_t GSG_PRE = 2;
struct GsgRows {
    uint32_t e0, em, e1, emt;
    uint32_t c[GSG_PRE + 1];
    float2 r[GSG_PRE + 1];
};

__device__ __forceinline__ GsgRows gsg_rows(const GsGrid& t, const uint32_t* co, uint32_t colour) {
    GsgRows g;
    g.e0 = co[2 * colour];
    g.em = co[2 * colour + 1];
    g.e1 = co[2 * colour + 2];
    g.emt = g.em + threadIdx.x;
#pragma unroll
    for (uint32_t k = 0; k <= GSG_PRE; ++k) {
        const uint32_t q = k < GSG_PRE ? g.e0 + threadIdx.x + k * blockDim.x : g.emt;
        const bool ok = k < GSG_PRE ? q < g.em : q < g.e1;
        g.c[k] = ok ? t.con[q] : 0u;
        g.r[k] = ok ? t.crc[q] : make_float2(0.f, 0.f);
    }
    return g;
}

// Wavefront 0 polls the 8 neighbour tiles' flags (lanes 0..8), then the block
// synchronises.  Polls 64 x 64 clocks apart: the polls' own loads at the
// coherence point slowed the other tiles' ring traffic (sleeps of 1, 4, 16,
// 32, 64, 127: 1 172, 1 187-1 207, 1 234-1 242, 1 259-1 261, 1 275-1 280,
// 1 143 substeps/s on the 1024^2 sheet).
__device__ __forceinline__ void gsg_wait(const GsGrid& t, uint32_t tx, uint32_t ty, uint32_t s, uint32_t* err) {
    if (threadIdx.x < 64) {
        const int l = int(threadIdx.x), nx = int(tx) + l % 3 - 1, ny = int(ty) + l / 3 - 1;
        const bool nb = l < 9 && l != 4 && nx >= 0 && ny >= 0 && nx < int(t.tx) && ny < int(t.ty);
        poll_flags<64>(nb ? t.flag + ny * int(t.tx) + nx : nullptr, s, err);
    }
    __syncthreads();
}

// One GS substep on LDS-resident grid tiles (block = tile, grid = all tiles,
// co-resident): the predict of the tile (step 0), I x colours constraint
// steps, then the final write of the tile's particles to the iterate and the
// state update (or the bounds).  Same results as k_gs_substep.
__global__ __launch_bounds__(1024) void k_gs_grid(GsArgs a, GsGrid t) {
    if (a.probe) {   // uniform: the co-residency check of this kernel's grid
        GridSync gsync;
        gsync.init(a.barrier, a.error);
        return;
    }
    extern __shared__ float4 win[];
    const DevParams& p = a.p;
    const uint32_t T = blockIdx.x, tx = T % t.tx, ty = T / t.tx, th = t.th;
    const int r0 = int(ty * th), c0 = int(tx * GSG_W);
    const uint32_t ntr = min(th, t.rows - ty * th);
    const TileWindow m{win, DeviceW(a.W, p.n), (r0 - 2) * int(t.cols) + c0 - 2, t.cols};
    for (uint32_t k = threadIdx.x; k < ntr * GSG_W; k += blockDim.x) {
        const uint32_t lr = k / GSG_W + 2, lc = k % GSG_W + 2;
        const uint32_t i = uint32_t(r0 + int(lr) - 2) * t.cols + uint32_t(c0 + int(lc) - 2);
        const float4 v = predict(a.X[i], a.V[i], i, p);
        win[lr * GSG_WW + lc] = v;
        if (gsg_shared(lr, lc, th)) m.w.st(i, v);
    }
    flag_step(t.flag + T, 1);
    const uint32_t last = 1 + a.iterations * t.colours;
    const uint32_t* co = t.coff + size_t(T) * t.colours * 2;
    // A step's constraint rows do not depend on any other tile: each step's
    // first rounds (GSG_PRE of the interior, one of the rest) are fetched a
    // step ahead, so their latency overlaps the previous step.
    GsgRows cur = gsg_rows(t, co, 0u), nxt = cur;
    for (uint32_t s = 1; s < last; ++s) {
        if (s + 1 < last) nxt = gsg_rows(t, co, s % t.colours);
        // the colour's constraints inside the tile's interior touch no
        // particle of another tile: they run before the wait
#pragma unroll
        for (uint32_t k = 0; k < GSG_PRE; ++k) {
            if (cur.e0 + threadIdx.x + k * blockDim.x < cur.em) {
                const uint2 ij = make_uint2(cur.c[k] & 0xFFFFu, cur.c[k] >> 16);
                gs_apply(a, ij, cur.r[k], m.ld(ij.x), m.ld(ij.y), m);
            }
        }
        for (uint32_t q = cur.e0 + threadIdx.x + GSG_PRE * blockDim.x; q < cur.em; q += blockDim.x) {
            const uint32_t w2 = t.con[q];
            const uint2 ij = make_uint2(w2 & 0xFFFFu, w2 >> 16);
            gs_apply(a, ij, t.crc[q], m.ld(ij.x), m.ld(ij.y), m);
        }
        gsg_wait(t, tx, ty, s, a.error);
        gsg_ring_load(m, r0, c0, th, t.rows, threadIdx.x, blockDim.x);
        __syncthreads();
        const uint32_t e = cur.emt;
        if (e < cur.e1) {
            const uint2 ij = make_uint2(cur.c[GSG_PRE] & 0xFFFFu, cur.c[GSG_PRE] >> 16);
            gs_apply(a, ij, cur.r[GSG_PRE], m.ld(ij.x), m.ld(ij.y), m);
        }
        for (uint32_t q = e + blockDim.x; q < cur.e1; q += blockDim.x) {
            const uint32_t w2 = t.con[q];
            const uint2 ij = make_uint2(w2 & 0xFFFFu, w2 >> 16);
            gs_apply(a, ij, t.crc[q], m.ld(ij.x), m.ld(ij.y), m);
        }
        flag_step(t.flag + T, s + 1);
        cur = nxt;
    }
    gsg_wait(t, tx, ty, last, a.error);
    gsg_ring_load(m, r0, c0, th, t.rows, threadIdx.x, blockDim.x);
    __syncthreads();
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t k = threadIdx.x; k < ntr * GSG_W; k += blockDim.x) {
        const uint32_t lr = k / GSG_W + 2, lc = k % GSG_W + 2;
        const uint32_t i = uint32_t(r0 + int(lr) - 2) * t.cols + uint32_t(c0 + int(lc) - 2);
        const float4 v = win[lr * GSG_WW + lc];
        a.W[i] = v;
        if (a.final_write) {
            if (v.w > 0.f) finish_particle(v, a.X[i], i, a.ground, p, a.xout, a.vout);
        } else {
            aabb_fold(true, v.x, v.y, v.z, lo, hi);
        }
    }
    if (!a.final_write) block_aabb_store(lo, hi, a.aabb, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Persistent Jacobi substep: predict and all I iterations in one launch, a
// grid barrier between iterations (the batch layout's "one launch per
// substep").  Blocks walk the slices in a grid-stride loop; the result is
// identical to the per-iteration kernels.
// ---------------------------------------------------------------------------
template <bool PACKED, bool FASTDIV>
__global__ __launch_bounds__(256) void k_jacobi_persistent(JacobiArgs a, PersistArgs q) {
    if (q.gate && *q.gate == 0u) return;   // uniform: every block leaves before any barrier
    __shared__ float2 lpairs[LDS_PAIRS_MAX];
    const bool lds = stage_pairs<PACKED, LDS_PAIRS_MAX>(a, lpairs);
    const DevParams& p = a.p;
    GridSync gsync;
    gsync.init(q.barrier, q.error);
    if (q.probe) {   // uniform: the co-residency check of this very kernel's grid
        gsync.sync();
        return;
    }
    const uint32_t nthreads = gridDim.x * blockDim.x;
    // predict (whole wavefronts per 64 particles: the slice tame bytes of A)
    const bool track = FASTDIV && q.tameA && q.tameB;
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < p.n; i0 += nthreads) {
        const uint32_t i = i0 + threadIdx.x;
        const bool valid = i < p.n;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (valid) {
            v = predict(q.X[i], q.V[i], i, p);
            q.A[i] = v;
        }
        if (track) write_slice_tame(q.tameA, i >> 6, valid, !valid || tame3(v));
    }
    gsync.sync();
    const uint32_t spb = blockDim.x >> 6;
    const uint32_t groups = (p.nslices + spb - 1) / spb;
    JacobiArgs b = a;
    float4* cur = q.A;
    float4* nxt = q.B;
    uint8_t* tcur = q.tameA;
    uint8_t* tnxt = q.tameB;
    for (uint32_t k = 0; k < q.iterations; ++k) {
        const bool last = k + 1 == q.iterations;
        b.src = cur;
        for (uint32_t gi = blockIdx.x; gi < groups; gi += gridDim.x) {
            const uint32_t slice = gi * spb + (threadIdx.x >> 6);
            bool valid;
            const uint32_t lane = threadIdx.x & 63;
            // the short exact operations where every slice this one reads is tame
            const bool fast = track && neighbourhood_tame(tcur, a.srange, slice, p.nslices);
            if (last && q.final_write) {
                b.dst = q.X;
                b.vout = q.V;
                if (PACKED && lds) {
                    if (fast) jacobi_particle<JM_FINAL, PACKED, true, true>(b, lpairs, slice, lane, valid);
                    else jacobi_particle<JM_FINAL, PACKED, false, true>(b, lpairs, slice, lane, valid);
                } else {
                    if (fast) jacobi_particle<JM_FINAL, PACKED, true, false>(b, lpairs, slice, lane, valid);
                    else jacobi_particle<JM_FINAL, PACKED, false, false>(b, lpairs, slice, lane, valid);
                }
            } else {
                b.dst = nxt;
                float4 out;
                if (PACKED && lds) {
                    out = fast ? jacobi_particle<0, PACKED, true, true>(b, lpairs, slice, lane, valid)
                               : jacobi_particle<0, PACKED, false, true>(b, lpairs, slice, lane, valid);
                } else {
                    out = fast ? jacobi_particle<0, PACKED, true, false>(b, lpairs, slice, lane, valid)
                               : jacobi_particle<0, PACKED, false, false>(b, lpairs, slice, lane, valid);
                }
                if (track && !last) write_slice_tame(tnxt, slice, slice < p.nslices, !valid || tame3(out));
            }
        }
        if (!last) gsync.sync();
        float4* t = cur;
        cur = nxt;
        nxt = t;
        uint8_t* tt = tcur;
        tcur = tnxt;
        tnxt = tt;
    }
}

// ---------------------------------------------------------------------------
// Jacobi substep on LDS-resident grid tiles: predict and all I iterations in
// one launch, without grid barriers and without restaging the tile every
// iteration.  One grid cloth whose column count is a multiple of 64 is cut
// into tiles of 64 columns x th rows, one block each (all resident at once).
// A block keeps its tile's iterate and a 2-particle ring around it (the
// window, (th + 4) x 68, the layout of the per-iteration grid tiles, so the
// same tile-offset entry words apply) in LDS, twice: the iterate read and the
// one written.  Between iterations only the ring crosses tiles: a tile
// publishes its own particles within 2 of its edge (at device scope, DeviceW)
// into the iterate buffer of that iteration (A for the predict, then B, A,
// ...), raises its flag, and before the next iteration reads the 2-particle
// ring outside its edge that its 8 neighbours published.  Double buffering:
// iteration k reads ring k from buffer k % 2 and publishes ring k + 1 into the
// other, which no neighbour still reads (each waited for this tile's ring k
// before it published its own ring k + 1).  Per particle the gather is that
// of k_jacobi_tiled (same entries, same order, same arithmetic), so the
// results are those of the per-iteration launches.  The fast division is
// used in an iteration where the whole window is tame (a block-wide AND).
// The final iterate goes to the state (final_write) or, with its bounds (one
// partial per tile), to buffer I % 2 for the collision pass.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool jr_ring(uint32_t lr, uint32_t lc, uint32_t ntr) {
    return lr < 4 || lr >= ntr || lc < 4 || lc >= 64;   // own positions: window rows/cols 2..ntr+1 / 2..65
}

// Load the 2-particle ring outside the tile (window rows 0, 1, ntr + 2,
// ntr + 3 whole; columns 0, 1, 66, 67 of rows 2..ntr+1) from the published
// iterate; tame: AND over the loaded particles.
__device__ __forceinline__ bool jr_halo(float4* win, const DeviceW& src, int r0, int c0, uint32_t ntr, uint32_t rows,
                                        uint32_t cols) {
    bool tame = true;
    const uint32_t nfull = 4 * JT_W, n = nfull + ntr * 4;
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        uint32_t lr, lc;
        if (k < nfull) {
            const uint32_t q = k / JT_W;
            lr = q < 2 ? q : ntr + q;
            lc = k % JT_W;
        } else {
            const uint32_t q = k - nfull, c4 = q % 4;
            lr = 2 + q / 4;
            lc = c4 < 2 ? c4 : 64 + c4;
        }
        const int r = r0 - 2 + int(lr), c = c0 - 2 + int(lc);
        if (r >= 0 && r < int(rows) && c >= 0 && c < int(cols)) {
            const float4 v = src.ld(uint32_t(r) * cols + uint32_t(c));
            win[lr * JT_W + lc] = v;
            tame = tame && tame3(v);
        }
    }
    return tame;
}

template <bool FASTDIV>
__global__ __launch_bounds__(1024) void k_jacobi_resident(JacobiArgs a, PersistArgs q, JResident t) {
    if (q.probe) {   // uniform: the co-residency check of this kernel's grid
        GridSync gsync;
        gsync.init(q.barrier, q.error);
        return;
    }
    extern __shared__ float4 jr_smem[];
    const DevParams& p = a.p;
    const uint32_t cols = a.sell.tcols, rows = a.sell.trows, ntx = cols / 64, th = t.th;
    const uint32_t T = blockIdx.x, tx = T % ntx, ty = T / ntx;
    const int r0 = int(ty * th), c0 = int(tx * 64);
    const uint32_t ntr = min(th, rows - ty * th);
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    float4* lpairs = jr_smem;
    float4* cur = jr_smem + t.npairs;
    float4* nxt = cur + (th + 4) * JT_W;
    if (t.npairs)
        for (uint32_t k = threadIdx.x; k < a.num_pairs; k += blockDim.x) lpairs[k] = stage_pair(a.sell.pairs[k], p);
    // step 0: predict the tile, publish its ring into A
    bool tame = true;
    {
        const DeviceW pub(q.A, p.n);
        for (uint32_t k = threadIdx.x; k < ntr * 64; k += blockDim.x) {
            const uint32_t lr = k / 64 + 2, lc = k % 64 + 2;
            const uint32_t i = uint32_t(r0 + int(lr) - 2) * cols + uint32_t(c0 + int(lc) - 2);
            const float4 v = predict(q.X[i], q.V[i], i, p);
            cur[lr * JT_W + lc] = v;
            tame = tame && tame3(v);
            if (jr_ring(lr, lc, ntr)) pub.st(i, v);
        }
    }
    flag_step(t.flag + T, 1);
    for (uint32_t k = 0; k < q.iterations; ++k) {
        const bool last = k + 1 == q.iterations;
        const DeviceW src((k & 1) ? q.B : q.A, p.n), pub((k & 1) ? q.A : q.B, p.n);
        if (threadIdx.x < 64) {
            const int l = int(threadIdx.x), nx = int(tx) + l % 3 - 1, ny = int(ty) + l / 3 - 1;
            const bool nb = l < 9 && l != 4 && nx >= 0 && ny >= 0 && nx < int(ntx) && ny < int(t.ty);
            poll_flags(nb ? t.flag + ny * int(ntx) + nx : nullptr, k + 1, q.error);
        }
        __syncthreads();
        tame = jr_halo(cur, src, r0, c0, ntr, rows, cols) && tame;
        const int all_tame = __syncthreads_and(tame ? 1 : 0);
        const bool fast = FASTDIV && all_tame != 0;
        tame = true;
        for (uint32_t row = wave; row < ntr; row += nw) {
            const uint32_t slice = (uint32_t(r0) + row) * ntx + tx;
            const uint32_t base = a.sell.soff[slice] + lane, width = a.sell.swid[slice];
            uint32_t pre[JT_JU];
#pragma unroll
            for (uint32_t u = 0; u < JT_JU; ++u)
                pre[u] = u < width ? __builtin_nontemporal_load(a.sell.tpacked + base + u * 64) : 0u;
            const uint32_t lidx = (row + 2) * JT_W + lane + 2;
            bool valid;
            float4 out;
            if (t.npairs)
                out = fast ? jacobi_particle_tiled<JM_KEEP, true, true>(a, lpairs, cur, slice, lane, lidx, true, valid,
                                                                         pre, base, width)
                           : jacobi_particle_tiled<JM_KEEP, false, true>(a, lpairs, cur, slice, lane, lidx, true, valid,
                                                                          pre, base, width);
            else
                out = fast ? jacobi_particle_tiled<JM_KEEP, true, false>(a, lpairs, cur, slice, lane, lidx, true, valid,
                                                                          pre, base, width)
                           : jacobi_particle_tiled<JM_KEEP, false, false>(a, lpairs, cur, slice, lane, lidx, true,
                                                                           valid, pre, base, width);
            nxt[lidx] = out;
            tame = tame && tame3(out);
            if (!last && jr_ring(row + 2, lane + 2, ntr)) pub.st(slice * 64 + lane, out);
        }
        if (!last) flag_step(t.flag + T, k + 2);
        else __syncthreads();
        float4* tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    // the final iterate of the tile
    float4* F = (q.iterations & 1) ? q.B : q.A;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t k = threadIdx.x; k < ntr * 64; k += blockDim.x) {
        const uint32_t lr = k / 64 + 2, lc = k % 64 + 2;
        const uint32_t i = uint32_t(r0 + int(lr) - 2) * cols + uint32_t(c0 + int(lc) - 2);
        const float4 v = cur[lr * JT_W + lc];
        if (q.final_write) {
            if (v.w > 0.f) finish_particle(v, q.X[i], i, a.ground, p, q.X, q.V);
        } else {
            F[i] = v;
            aabb_fold(true, v.x, v.y, v.z, lo, hi);
        }
    }
    if (!q.final_write) block_aabb_store(lo, hi, t.aabb, T);
}

// ---------------------------------------------------------------------------
// Render publish: finalize_normals (cloth_sim_jacobi.wgsl:527-547) with the
// face normals gathered per vertex over its incident triangles in ascending
// order (no fixed-point atomics), fallback +Y; positions copied with w = 0.
// ---------------------------------------------------------------------------
__global__ void k_publish(const float4* X, uint32_t n, const uint32_t* tris, const uint32_t* vt_off,
                          const uint32_t* vt_tri, float4* dpos, float4* dnrm) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 x = X[i];
    if (dpos) dpos[i] = make_float4(x.x, x.y, x.z, 0.f);
    if (!dnrm) return;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (uint32_t k = vt_off[i]; k < vt_off[i + 1]; ++k) {
        const uint32_t t = vt_tri[k];
        const float4 p0 = X[tris[3 * t]], p1 = X[tris[3 * t + 1]], p2 = X[tris[3 * t + 2]];
        const float ax = p1.x - p0.x, ay = p1.y - p0.y, az = p1.z - p0.z;
        const float bx = p2.x - p0.x, by = p2.y - p0.y, bz = p2.z - p0.z;
        sx = sx + (ay * bz - by * az);
        sy = sy + (az * bx - bz * ax);
        sz = sz + (ax * by - bx * ay);
    }
    const float ln = len3(sx, sy, sz);
    if (ln > 1e-8f) dnrm[i] = make_float4(sx / ln, sy / ln, sz / ln, 0.f);
    else dnrm[i] = make_float4(0.f, 1.f, 0.f, 0.f);
}

// Sum over constraints of C^2 (C = |x_i - x_j| - rest), in double: one
// partial per block (no atomics), summed on the host in block order, so the
// figure is the same on every run.
__global__ __launch_bounds__(256) void k_residual(const float4* X, const uint2* ij, const float* rest, uint32_t e,
                                                 double* part) {
    double s = 0.0;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < e; k += gridDim.x * blockDim.x) {
        const uint2 q = ij[k];
        const float4 a = X[q.x], b = X[q.y];
        const double dx = double(a.x) - double(b.x), dy = double(a.y) - double(b.y), dz = double(a.z) - double(b.z);
        const double c = sqrt((dx * dx + dy * dy) + dz * dz) - double(rest[k]);
        s += c * c;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// ---------------------------------------------------------------------------
// Allocation-free relative residual: the sums of k_residual plus the sum of
// rest^2, over caller-owned device arrays, with the per-block partials in a
// fixed file-scope device array instead of handle memory.  Each block owns
// two slots (C^2 at [b], rest^2 at [blocks + b]) and rewrites them on every
// call, so nothing has to start at zero; the finishing block sums them in
// block order.  A constraint index >= n is not loaded: its block's C^2
// partial becomes NaN, and so does the result.
// ---------------------------------------------------------------------------
constexpr uint32_t NOALLOC_RESIDUAL_BLOCKS = 2048;
static __device__ __align__(256) double scratch[2 * NOALLOC_RESIDUAL_BLOCKS];

__global__ __launch_bounds__(256) void k_residual_parts(const float4* X, uint32_t n, const uint32_t* ci,
                                                        const uint32_t* cj, const float* rest, uint32_t e,
                                                        double* part) {
    double s = 0.0, r = 0.0;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < e; k += gridDim.x * blockDim.x) {
        const uint32_t i = ci[k], j = cj[k];
        const double rk = double(rest[k]);
        r += rk * rk;
        if (i >= n || j >= n) {
            s = __builtin_nan("");
            continue;
        }
        const float4 a = X[i], b = X[j];
        const double dx = double(a.x) - double(b.x), dy = double(a.y) - double(b.y), dz = double(a.z) - double(b.z);
        const double c = sqrt((dx * dx + dy * dy) + dz * dz) - rk;
        s += c * c;
    }
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off, 64);
        r += __shfl_xor(r, off, 64);
    }
    __shared__ double red[8];
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = s;
        red[4 + (threadIdx.x >> 6)] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
        part[gridDim.x + blockIdx.x] = (red[4] + red[5]) + (red[6] + red[7]);
    }
}

__global__ __launch_bounds__(64) void k_residual_finish(const double* part, uint32_t blocks, double* out) {
    if (threadIdx.x != 0) return;
    double c2 = 0.0, r2 = 0.0;
    for (uint32_t b = 0; b < blocks; ++b) c2 += part[b];
    for (uint32_t b = 0; b < blocks; ++b) r2 += part[blocks + b];
    out[0] = c2 != c2 ? c2 : (r2 > 0.0 ? sqrt(c2) / sqrt(r2) : 0.0);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

uint32_t jacobi_blocks(uint32_t nslices, uint32_t block) { return cdiv(nslices, block / 64); }

template <bool PACKED, bool FASTDIV>
static void launch_jacobi_t(uint32_t mode, const JacobiArgs& a, uint32_t blocks, uint32_t block, hipStream_t s) {
    // the pair table's LDS, as k_jacobi decides it (stage_pairs)
    const size_t sh = (PACKED && a.num_pairs <= LDS_PAIRS_DYN_MAX) ? size_t(a.num_pairs) * sizeof(float2) : 0;
    const dim3 g(blocks), b(block);
    switch (mode) {
        case 0: hipLaunchKernelGGL((k_jacobi<0, PACKED, FASTDIV>), g, b, sh, s, a); break;
        case JM_PREDICT: hipLaunchKernelGGL((k_jacobi<JM_PREDICT, PACKED, FASTDIV>), g, b, sh, s, a); break;
        case JM_FINAL: hipLaunchKernelGGL((k_jacobi<JM_FINAL, PACKED, FASTDIV>), g, b, sh, s, a); break;
        case JM_AABB: hipLaunchKernelGGL((k_jacobi<JM_AABB, PACKED, FASTDIV>), g, b, sh, s, a); break;
        case JM_PREDICT | JM_AABB:
            hipLaunchKernelGGL((k_jacobi<JM_PREDICT | JM_AABB, PACKED, FASTDIV>), g, b, sh, s, a);
            break;
        default: break;
    }
}

template <bool FASTDIV>
static void launch_jacobi_tiled_t(uint32_t mode, const JacobiArgs& a, uint32_t blocks, hipStream_t s) {
    const dim3 g(blocks), b(64 * JT_ROWS);
    switch (mode) {
        case 0: hipLaunchKernelGGL((k_jacobi_tiled<0, FASTDIV>), g, b, 0, s, a); break;
        case JM_PREDICT: hipLaunchKernelGGL((k_jacobi_tiled<JM_PREDICT, FASTDIV>), g, b, 0, s, a); break;
        case JM_FINAL: hipLaunchKernelGGL((k_jacobi_tiled<JM_FINAL, FASTDIV>), g, b, 0, s, a); break;
        case JM_AABB: hipLaunchKernelGGL((k_jacobi_tiled<JM_AABB, FASTDIV>), g, b, 0, s, a); break;
        case JM_PREDICT | JM_AABB:
            hipLaunchKernelGGL((k_jacobi_tiled<JM_PREDICT | JM_AABB, FASTDIV>), g, b, 0, s, a);
            break;
        default: break;
    }
}

template <bool FASTDIV>
static void launch_jacobi_shared_t(uint32_t mode, const JacobiArgs& a, uint32_t blocks, hipStream_t s) {
    const dim3 g(blocks), b(64 * JT_ROWS);
    const size_t lds = (size_t(a.sh.nslots) + 1 + std::max(a.num_pairs, 1u)) * sizeof(float4);
    switch (mode) {
        case 0: hipLaunchKernelGGL((k_jacobi_shared<0, FASTDIV>), g, b, lds, s, a); break;
        case JM_PREDICT: hipLaunchKernelGGL((k_jacobi_shared<JM_PREDICT, FASTDIV>), g, b, lds, s, a); break;
        case JM_FINAL: hipLaunchKernelGGL((k_jacobi_shared<JM_FINAL, FASTDIV>), g, b, lds, s, a); break;
        case JM_AABB: hipLaunchKernelGGL((k_jacobi_shared<JM_AABB, FASTDIV>), g, b, lds, s, a); break;
        case JM_PREDICT | JM_AABB:
            hipLaunchKernelGGL((k_jacobi_shared<JM_PREDICT | JM_AABB, FASTDIV>), g, b, lds, s, a);
            break;
        default: break;
    }
}

size_t jacobi_shared_lds_bytes(uint32_t nslots, uint32_t num_pairs) {
    return (size_t(nslots) + 1 + std::max(num_pairs, 1u)) * sizeof(float4) + sizeof(float4) * JT_H * JT_W;
}

uint32_t launch_jacobi(uint32_t mode, const JacobiArgs& a, uint32_t block, hipStream_t s) {
    if (a.sh.jobs && a.sell.tcols >= 64 && a.sell.tcols % 64 == 0) {   // constraint-shared grid tiles
        const uint32_t blocks = (a.sell.tcols / 64) * cdiv(a.sell.trows, JT_ROWS);
        if (blocks == 0) return 0;
        if (a.fastdiv) launch_jacobi_shared_t<true>(mode, a, blocks, s);
        else launch_jacobi_shared_t<false>(mode, a, blocks, s);
        return blocks;
    }
    if (a.sell.tpacked && a.sell.tcols >= 64 && a.sell.tcols % 64 == 0) {   // grid tiles: one 64 x 8 tile per block
        const uint32_t blocks = (a.sell.tcols / 64) * cdiv(a.sell.trows, JT_ROWS);
        if (blocks == 0) return 0;
        if (a.fastdiv) launch_jacobi_tiled_t<true>(mode, a, blocks, s);
        else launch_jacobi_tiled_t<false>(mode, a, blocks, s);
        return blocks;
    }
    const uint32_t blocks = jacobi_blocks(a.p.nslices, block);
    if (blocks == 0) return 0;
    if (a.sell.packed) {
        if (a.fastdiv) launch_jacobi_t<true, true>(mode, a, blocks, block, s);
        else launch_jacobi_t<true, false>(mode, a, blocks, block, s);
    } else {
        if (a.fastdiv) launch_jacobi_t<false, true>(mode, a, blocks, block, s);
        else launch_jacobi_t<false, false>(mode, a, blocks, block, s);
    }
    return blocks;
}

// ---------------------------------------------------------------------------
// Device self-test of div_rn / rcp_rn: the refined reciprocal against the
// correctly rounded 1.0f / b for every significand at the given exponents,
// and div_rn against '/' on pseudo-random pairs.  Counts mismatches.
// ---------------------------------------------------------------------------
__global__ void k_selftest_rcp(int exp_lo, int exp_hi, uint32_t* bad, uint32_t* first) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (1u << 23)) return;
    for (int e = exp_lo; e <= exp_hi; ++e)
        for (int sgn = 0; sgn < 2; ++sgn) {
            const float b = __uint_as_float((uint32_t(sgn) << 31) | (uint32_t(e + 127) << 23) | m);
            const float y = rcp_rn(b);
            const float ref = 1.0f / b;
            if (__float_as_uint(y) != __float_as_uint(ref)) {
                if (atomicAdd(bad, 1u) == 0) *first = __float_as_uint(b);
            }
        }
}

__device__ __forceinline__ uint32_t mix32(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return uint32_t(x);
}

__global__ void k_selftest_div(uint64_t n, uint64_t seed, uint32_t* bad, uint32_t* first) {
    for (uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t ua = mix32(seed * 0x9E3779B97F4A7C15ull + 2 * k);
        const uint32_t ub = mix32(seed * 0x9E3779B97F4A7C15ull + 2 * k + 1);
        // numerator exponents [-74, 32], denominator exponents [-28, 43] (the
        // operand ranges of the gather's divisions on tame iterates), random
        // significands and signs
        const float a = __uint_as_float((ua & 0x807FFFFFu) | (uint32_t(53 + (ua >> 24) % 107) << 23));
        const float b = __uint_as_float((ub & 0x807FFFFFu) | (uint32_t(99 + (ub >> 24) % 72) << 23));
        const float q = div_rn(a, b, rcp_rn(b));
        const float ref = a / b;
        if (__float_as_uint(q) != __float_as_uint(ref)) {
            if (atomicAdd(bad, 1u) == 0) {
                first[0] = __float_as_uint(a);
                first[1] = __float_as_uint(b);
            }
        }
    }
}

__global__ void k_selftest_sqrt(int exp_lo, int exp_hi, uint32_t* bad, uint32_t* first) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (1u << 23)) return;
    for (int e = exp_lo; e <= exp_hi; ++e) {
        const float x = __uint_as_float((uint32_t(e + 127) << 23) | m);
        const float y = sqrt_rn(x);
        const float ref = sqrtf(x);
        if (__float_as_uint(y) != __float_as_uint(ref)) {
            if (atomicAdd(bad, 1u) == 0) *first = __float_as_uint(x);
        }
    }
    if (m == 0 && __float_as_uint(sqrt_rn(0.0f)) != 0u) atomicAdd(bad, 1u);
}

void launch_selftest_sqrt(int exp_lo, int exp_hi, uint32_t* bad, uint32_t* first, hipStream_t s) {
    hipLaunchKernelGGL(k_selftest_sqrt, dim3((1u << 23) / 256), dim3(256), 0, s, exp_lo, exp_hi, bad, first);
}

void launch_selftest_rcp(int exp_lo, int exp_hi, uint32_t* bad, uint32_t* first, hipStream_t s) {
    hipLaunchKernelGGL(k_selftest_rcp, dim3((1u << 23) / 256), dim3(256), 0, s, exp_lo, exp_hi, bad, first);
}

void launch_selftest_div(uint64_t n, uint64_t seed, uint32_t* bad, uint32_t* first, hipStream_t s) {
    hipLaunchKernelGGL(k_selftest_div, dim3(8192), dim3(256), 0, s, n, seed, bad, first);
}

void launch_predict(const float4* X, const float4* V, float4* P, DevParams p, hipStream_t s, uint8_t* tame_out) {
    if (p.n == 0) return;
    hipLaunchKernelGGL(k_predict, dim3(cdiv(p.n, 256)), dim3(256), 0, s, X, V, P, p, tame_out);
}

__global__ void k_tame_scan(const float4* P, uint32_t n, uint8_t* slice_tame) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool t = i >= n || tame3(P[i]);
    write_slice_tame(slice_tame, i >> 6, i < n, t);
}

void launch_tame_scan(const float4* P, uint32_t n, uint8_t* slice_tame, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_tame_scan, dim3(cdiv(n, 256)), dim3(256), 0, s, P, n, slice_tame);
}

void launch_predict_gated(const float4* X, const float4* V, float4* P, DevParams p, const uint32_t* gate,
                          hipStream_t s) {
    if (p.n == 0) return;
    hipLaunchKernelGGL(k_predict_gated, dim3(cdiv(p.n, 256)), dim3(256), 0, s, X, V, P, p, gate);
}

void launch_copy4_if(const float4* a, float4* b, uint32_t n, const uint32_t* gate, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_copy4_if, dim3(cdiv(n, 256)), dim3(256), 0, s, a, b, n, gate);
}

uint32_t warm_gate_words() { return WG_LINE * (WG_SLOTS + 1) + 2 * WG_MAX_BLOCKS; }

void launch_warm_gate(const float4* X, const uint2* ij, const float* rest, uint32_t e, double rest_sq,
                      float threshold, uint32_t* acc, uint32_t* flag, uint32_t* barrier, hipStream_t s) {
    const uint32_t blocks = std::max(1u, std::min(cdiv(e, 256 * 4), WG_MAX_BLOCKS));
    hipLaunchKernelGGL(k_warm_gate, dim3(blocks), dim3(256), 0, s, X, ij, rest, e, rest_sq, threshold, acc, flag,
                       barrier);
}

void launch_finalize_sel(const float4* F, const float4* alt, const uint32_t* gate, float4* X, float4* V,
                         const uint8_t* ground, DevParams p, hipStream_t s) {
    if (p.n == 0) return;
    hipLaunchKernelGGL(k_finalize_sel, dim3(cdiv(p.n, 256)), dim3(256), 0, s, F, alt, gate, X, V, ground, p);
}

void launch_finalize(const float4* F, float4* X, float4* V, const uint8_t* ground, DevParams p, hipStream_t s) {
    if (p.n == 0) return;
    hipLaunchKernelGGL(k_finalize, dim3(cdiv(p.n, 256)), dim3(256), 0, s, F, X, V, ground, p);
}

void launch_copy4(const float4* a, float4* b, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_copy4, dim3(std::max(1u, cdiv(n, 256))), dim3(256), 0, s, a, b, n);
}

void launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u32, dim3(std::max(1u, std::min(cdiv(n, 256), 1024u))), dim3(256), 0, s, p, n);
}

uint32_t launch_aabb_of(const float4* P, uint32_t n, float* part, hipStream_t s) {
    const uint32_t blocks = std::max(1u, cdiv(n, 256));
    hipLaunchKernelGGL(k_aabb_of, dim3(blocks), dim3(256), 0, s, P, n, part);
    return blocks;
}

void launch_coll_grids(const float4* P, const uint2* span, uint32_t nspans, const uint32_t* pfirst,
                       uint32_t nclothes, float* part, float cell_size, uint32_t local_bits, CollGrid* grids,
                       hipStream_t s) {
    if (nspans) hipLaunchKernelGGL(k_cloth_aabb, dim3(nspans), dim3(256), 0, s, P, span, part);
    hipLaunchKernelGGL(k_coll_grids, dim3(cdiv(nclothes, 64)), dim3(64), 0, s, part, pfirst, nclothes, cell_size,
                       local_bits, grids);
}

uint32_t cloth_span_max() { return CB_SPAN; }

void launch_coll_keys(const float4* F, uint32_t n, CollGrid* grids, const uint32_t* cloth_of, uint32_t* keys,
                      uint32_t* ids, uint4* table, const uint32_t* slot_rec, uint32_t* hist, uint32_t digit_bits,
                      const float* part, uint32_t nparts, float cell_size, uint32_t* epoch, hipStream_t s) {
    if (n == 0) return;
    const uint32_t ntiles = std::max(1u, cdiv(n, RADIX_TILE));
    if (digit_bits == 9)
        hipLaunchKernelGGL(k_coll_keys<9>, dim3(ntiles), dim3(RADIX_BLOCK), 0, s, F, n, grids, cloth_of, keys, ids,
                           table, slot_rec, hist, ntiles, part, nparts, cell_size, grids, epoch);
    else
        hipLaunchKernelGGL(k_coll_keys<8>, dim3(ntiles), dim3(RADIX_BLOCK), 0, s, F, n, grids, cloth_of, keys, ids,
                           table, slot_rec, hist, ntiles, part, nparts, cell_size, grids, epoch);
}

uint32_t radix_hist_words(uint32_t n) {
    return RADIX_MAX_DIGITS * std::max(1u, cdiv(n, RADIX_TILE)) + RADIX_MAX_DIGITS;
}

void launch_radix_pass(const uint32_t* kin, const uint32_t* iin, uint32_t* kout, uint32_t* iout, uint32_t n,
                       uint32_t shift, uint32_t* hist, hipStream_t s, bool hist_done, uint32_t digit_bits,
                       const float4* F, float4* sP) {
    const uint32_t ntiles = std::max(1u, cdiv(n, RADIX_TILE));
    const uint32_t digits = 1u << digit_bits;
    uint32_t* dtot = hist + size_t(digits) * ntiles;
    const dim3 g(ntiles), b(RADIX_BLOCK);
    if (digit_bits == 9) {
        if (!hist_done) hipLaunchKernelGGL(k_radix_hist<9>, g, b, 0, s, kin, n, shift, hist, ntiles);
        hipLaunchKernelGGL(k_radix_scan_tiles, dim3(digits), dim3(256), 0, s, hist, ntiles, dtot);
        if (sP)
            hipLaunchKernelGGL((k_radix_scatter<9, true>), g, b, 0, s, kin, iin, kout, iout, n, shift, hist, ntiles,
                               dtot, F, sP);
        else
            hipLaunchKernelGGL((k_radix_scatter<9, false>), g, b, 0, s, kin, iin, kout, iout, n, shift, hist, ntiles,
                               dtot, F, sP);
    } else {
        if (!hist_done) hipLaunchKernelGGL(k_radix_hist<8>, g, b, 0, s, kin, n, shift, hist, ntiles);
        hipLaunchKernelGGL(k_radix_scan_tiles, dim3(digits), dim3(256), 0, s, hist, ntiles, dtot);
        if (sP)
            hipLaunchKernelGGL((k_radix_scatter<8, true>), g, b, 0, s, kin, iin, kout, iout, n, shift, hist, ntiles,
                               dtot, F, sP);
        else
            hipLaunchKernelGGL((k_radix_scatter<8, false>), g, b, 0, s, kin, iin, kout, iout, n, shift, hist, ntiles,
                               dtot, F, sP);
    }
}

bool launch_radix_sort(uint32_t* keys, uint32_t* ids, uint32_t* keys_alt, uint32_t* ids_alt, uint32_t n,
                       uint32_t* hist, hipStream_t s, bool first_hist_done, uint32_t digit_bits, const float4* F,
                       float4* sP) {
    const uint32_t passes = digit_bits == 9 ? 3u : 4u;
    if (n == 0) return false;
    for (uint32_t p = 0; p < passes; ++p) {
        const bool odd = p & 1u, last = p + 1 == passes;
        launch_radix_pass(odd ? keys_alt : keys, odd ? ids_alt : ids, odd ? keys : keys_alt, odd ? ids : ids_alt, n,
                          p * digit_bits, hist, s, p == 0 && first_hist_done, digit_bits, last ? F : nullptr,
                          last ? sP : nullptr);
    }
    return passes & 1u;
}

void launch_coll_table(const uint32_t* skeys, uint32_t n, uint32_t* ovf, uint4* table, uint32_t mask,
                       uint32_t* slot_rec, const uint32_t* epoch, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_coll_table, dim3(cdiv(n, GATHER_BLOCK)), dim3(GATHER_BLOCK), 0, s, skeys, n, ovf, table, mask,
                       slot_rec, epoch);
}

void launch_coll_sweep(const CollArgs& a, hipStream_t s) {
    if (a.p.n == 0) return;
    hipLaunchKernelGGL(k_coll_sweep, dim3(cdiv(a.p.n, 256)), dim3(256), 0, s, a);
    // overflow is rare: a small grid (it strides over the list) keeps the
    // usually empty launch short
    hipLaunchKernelGGL(k_coll_sweep_overflow, dim3(std::min(cdiv(a.p.n, OVF_BLOCK), 256u)), dim3(OVF_BLOCK), 0, s, a);
}

// Blocks per CU a grid-barrier launch may use: the occupancy figure, but at
// most 7 wavefronts per SIMD.  A grid filling all 8 wavefront slots of every
// SIMD was not fully co-resident on MI355X (its barrier wait expired), so
// one slot per SIMD stays free.
int coresident_per_cu(int per_cu, uint32_t block_threads) {
    const int waves = int((block_threads + 63) / 64);
    const int cap = std::max(1, (7 * 4) / std::max(1, waves));
    int use = std::min(per_cu, cap);
    // a kernel whose own limit is below 8 wavefronts per SIMD also keeps a
    // block's worth of slots free rather than filling that limit
    if (use == per_cu && per_cu * waves < 8 * 4) use = per_cu - 1;
    return std::max(1, use);
}

uint32_t persistent_max_blocks(bool packed, bool fastdiv) {
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    hipError_t e;
    if (packed) e = fastdiv ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_jacobi_persistent<true, true>, 256, 0)
                            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_jacobi_persistent<true, false>, 256, 0);
    else e = fastdiv ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_jacobi_persistent<false, true>, 256, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_jacobi_persistent<false, false>, 256, 0);
    if (e != hipSuccess) return 0;
    return uint32_t(coresident_per_cu(per_cu, 256) * std::max(0, cus));
}

void launch_jacobi_persistent(const JacobiArgs& a, const PersistArgs& q, uint32_t blocks, hipStream_t s) {
    if (a.sell.packed) {
        if (a.fastdiv) hipLaunchKernelGGL((k_jacobi_persistent<true, true>), dim3(blocks), dim3(256), 0, s, a, q);
        else hipLaunchKernelGGL((k_jacobi_persistent<true, false>), dim3(blocks), dim3(256), 0, s, a, q);
    } else {
        if (a.fastdiv) hipLaunchKernelGGL((k_jacobi_persistent<false, true>), dim3(blocks), dim3(256), 0, s, a, q);
        else hipLaunchKernelGGL((k_jacobi_persistent<false, false>), dim3(blocks), dim3(256), 0, s, a, q);
    }
}

void launch_gs_colours(const GsArgs& a, const std::vector<uint32_t>& off, uint32_t iterations, hipStream_t s) {
    for (uint32_t it = 0; it < iterations; ++it)
        for (size_t b = 0; b + 1 < off.size(); ++b) {
            const uint32_t e0 = off[b], e1 = off[b + 1];
            if (e1 > e0) hipLaunchKernelGGL(k_gs_colour, dim3(cdiv(e1 - e0, 256)), dim3(256), 0, s, a, e0, e1);
        }
}

uint32_t gs_max_blocks(uint32_t block) {
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gs_substep, int(block), 0) != hipSuccess) return 0;
    return uint32_t(coresident_per_cu(per_cu, block) * std::max(0, cus));
}

void launch_gs(const GsArgs& a, uint32_t blocks, uint32_t block, hipStream_t s) {
    hipLaunchKernelGGL(k_gs_substep, dim3(blocks), dim3(block), 0, s, a);
}

uint32_t gs_tiles_max_blocks() {
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gs_tiles, 256, 0) != hipSuccess) return 0;
    return uint32_t(coresident_per_cu(per_cu, 256) * std::max(0, cus));
}

void launch_gs_tiles(const GsArgs& a, const GsTiles& t, uint32_t blocks, hipStream_t s) {
    launch_zero_u32(t.flag, t.ntiles, s);
    hipLaunchKernelGGL(k_gs_tiles, dim3(blocks), dim3(256), 0, s, a, t);
}

uint32_t gs_grid_lds_bytes(uint32_t th) { return (th + 4) * GSG_WW * 16u; }

uint32_t gs_grid_block(uint32_t th) { return th >= 64 ? 1024u : th >= 32 ? 512u : 256u; }

hipError_t gs_grid_prepare(uint32_t th) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(k_gs_grid), hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(gs_grid_lds_bytes(th)));
}

uint32_t gs_grid_max_blocks(uint32_t th) {
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gs_grid, int(gs_grid_block(th)),
                                                     gs_grid_lds_bytes(th)) != hipSuccess)
        return 0;
    return uint32_t(std::max(0, per_cu) * std::max(0, cus));
}

uint32_t jacobi_resident_lds_bytes(uint32_t th, uint32_t npairs) { return (npairs + 2 * (th + 4) * JT_W) * 16u; }

hipError_t jacobi_resident_prepare(uint32_t th, uint32_t npairs) {
    const int bytes = int(jacobi_resident_lds_bytes(th, npairs));
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_jacobi_resident<true>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(k_jacobi_resident<false>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

uint32_t jacobi_resident_max_blocks(uint32_t th, uint32_t npairs) {
    int dev = 0, per_cu = 0, per_cu2 = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    const size_t bytes = jacobi_resident_lds_bytes(th, npairs);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_jacobi_resident<true>, int(th * 16), bytes) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu2, k_jacobi_resident<false>, int(th * 16), bytes) !=
            hipSuccess)
        return 0;
    return uint32_t(std::max(0, std::min(per_cu, per_cu2)) * std::max(0, cus));
}

void launch_jacobi_resident(const JacobiArgs& a, const PersistArgs& q, const JResident& t, hipStream_t s) {
    const uint32_t tiles = (a.sell.tcols / 64) * t.ty;
    launch_zero_u32(t.flag, tiles, s);
    const size_t bytes = jacobi_resident_lds_bytes(t.th, t.npairs);
    if (a.fastdiv) hipLaunchKernelGGL(k_jacobi_resident<true>, dim3(tiles), dim3(t.th * 16), bytes, s, a, q, t);
    else hipLaunchKernelGGL(k_jacobi_resident<false>, dim3(tiles), dim3(t.th * 16), bytes, s, a, q, t);
}

void launch_gs_grid(const GsArgs& a, const GsGrid& t, hipStream_t s) {
    launch_zero_u32(t.flag, t.tx * t.ty, s);
    hipLaunchKernelGGL(k_gs_grid, dim3(t.tx * t.ty), dim3(gs_grid_block(t.th)), gs_grid_lds_bytes(t.th), s, a, t);
}

void launch_publish(const float4* X, uint32_t n, const uint32_t* tris, const uint32_t* vt_off, const uint32_t* vt_tri,
                    float4* dpos, float4* dnrm, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_publish, dim3(cdiv(n, 256)), dim3(256), 0, s, X, n, tris, vt_off, vt_tri, dpos, dnrm);
}

uint32_t residual_noalloc_max_blocks() { return NOALLOC_RESIDUAL_BLOCKS; }

namespace {
// Device address of `scratch` on each device, looked up on the first call
// made with that device current (each device holds its own copy).
std::mutex g_scratch_mu;
double* g_scratch_ptr[64] = {};
}  // namespace

hipError_t launch_residual_noalloc(const float4* X, uint32_t n, const uint32_t* ci, const uint32_t* cj,
                                   const float* rest, uint32_t e, hipStream_t s, double* out) {
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return err;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    double* part = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void* p = nullptr;
            err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (err != hipSuccess) return err;
            g_scratch_ptr[dev] = static_cast<double*>(p);
        }
        part = g_scratch_ptr[dev];
    }
    const uint32_t blocks = e ? std::min(cdiv(e, 256 * 4), NOALLOC_RESIDUAL_BLOCKS) : 0u;
    if (blocks) hipLaunchKernelGGL(k_residual_parts, dim3(blocks), dim3(256), 0, s, X, n, ci, cj, rest, e, part);
    hipLaunchKernelGGL(k_residual_finish, dim3(1), dim3(64), 0, s, part, blocks, out);
    return hipGetLastError();
}

uint32_t residual_max_blocks() { return 2048u; }

uint32_t launch_residual(const float4* X, const uint2* ij, const float* rest, uint32_t e, double* part,
                         hipStream_t s) {
    if (e == 0) return 0;
    const uint32_t blocks = std::min(cdiv(e, 256 * 4), residual_max_blocks());
    hipLaunchKernelGGL(k_residual, dim3(blocks), dim3(256), 0, s, X, ij, rest, e, part);
    return blocks;
}

}  // namespace xpbd
