// Learned warm start kernels for gfx950: features, BF16 MFMA forward /
// backward (FP32 accumulate), Adam, and the warm-start apply step.
//
// MFMA: v_mfma_f32_16x16x16_bf16 (wave64).  Lane l = 16 g + r holds
//   A[m = r][k = 4g + i], B[k = 4g + i][n = r], D[m = 4g + i][n = r], i = 0..3.
// The forward runs transposed, Z^T = W^T H^T (M = units, N = 16 rows), so the
// D fragments of one layer are exactly the B fragments of the next and no
// data moves between layers.  Backward: dH^T = W dZ^T keeps that layout; the
// weight gradients (K = rows) read both operands transposed through a small
// per-wave LDS tile.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>

#include "mlp.hpp"

namespace xpbd {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// f32 -> bf16, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f)); }

typedef __bf16 bf16v2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
// two v_cvt_pk_bf16_f32, one per pair
__device__ __forceinline__ bf16x4 pack4(f32x4 v) {
    const bf16v2 lo = __builtin_convertvector((f32x2){v[0], v[1]}, bf16v2);
    const bf16v2 hi = __builtin_convertvector((f32x2){v[2], v[3]}, bf16v2);
    return __builtin_bit_cast(bf16x4, (u32x2){__builtin_bit_cast(uint32_t, lo), __builtin_bit_cast(uint32_t, hi)});
}

__device__ __forceinline__ bf16x4 ld4(const uint16_t* p) { return *reinterpret_cast<const bf16x4*>(p); }

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// GELU, tanh form, and its derivative, written with s = sigmoid(2u) =
// (1 + tanh u) / 2, u = C (x + A x^3): GELU(x) = x s, and since
// 1 - tanh^2 u = 4 s (1 - s), GELU'(x) = s + x s (1 - s) (2C + 6CA x^2).
// exp(-2u) = 2^(x (K1 + K2 x^2)) with K1 = -2C log2(e), K2 = A K1, through
// the hardware exp2 and reciprocal (about 1e-6 relative, far inside the bf16
// rounding the activations get next).  Large |x|: exp2 -> inf or 0, so
// s -> 0 or 1 and GELU -> 0 or x.
constexpr float GELU_C = 0.7978845608028654f;   // sqrt(2 / pi)
constexpr float GELU_A = 0.044715f;
constexpr float GELU_K1 = float(-2.0 * 0.7978845608028654 * 1.4426950408889634);
constexpr float GELU_K2 = float(-2.0 * 0.7978845608028654 * 1.4426950408889634 * 0.044715);
__device__ __forceinline__ float sigmoid_2u(float x, float x2) {
    const float t = __builtin_fmaf(GELU_K2, x2, GELU_K1);
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(t * x));
}
__device__ __forceinline__ float gelu(float x) { return x * sigmoid_2u(x, x * x); }
__device__ __forceinline__ float gelu_grad(float x) {
    const float x2 = x * x;
    const float s = sigmoid_2u(x, x2);
    const float q = __builtin_fmaf(6.0f * GELU_C * GELU_A, x2, 2.0f * GELU_C);
    const float d = __builtin_fmaf(-s, s, s);   // s (1 - s)
    return __builtin_fmaf(x * d, q, s);
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float sum16(float v) {   // over the 16 lanes sharing l / 16
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    return v;
}

// LDS copy of the biases: b1 [0, 64), b2 [64, 128), b3 [128, 144)
constexpr int LB_B1 = 0, LB_B2 = MLP_HID, LB_B3 = 2 * MLP_HID, LB_WORDS = 2 * MLP_HID + MLP_OUTP;

__device__ __forceinline__ void stage_weights(const uint16_t* bf, const float* P, uint16_t* W, float* B) {
    const uint4* src = reinterpret_cast<const uint4*>(bf);
    uint4* dst = reinterpret_cast<uint4*>(W);
    for (uint32_t q = threadIdx.x; q < uint32_t(MLP_BF / 8); q += blockDim.x) dst[q] = src[q];
    for (uint32_t q = threadIdx.x; q < uint32_t(LB_WORDS); q += blockDim.x) {
        B[q] = q < uint32_t(LB_B2) ? P[OFF_B1 + q] : q < uint32_t(LB_B3) ? P[OFF_B2 + q - LB_B2] : P[OFF_B3 + q - LB_B3];
    }
    __syncthreads();
}

__device__ __forceinline__ f32x4 ld_b4(const float* B) { return *reinterpret_cast<const f32x4*>(B); }

struct TileAct {
    f32x4 z1[4], z2[4], o;
    bf16x4 h1[4], h2[4];
};

// Forward of one 16-row tile.  x[s]: B fragments of X^T (features 16s..16s+15).
// The bias initialises the accumulator (z = b + W^T h).
__device__ __forceinline__ void tile_forward(const uint16_t* W, const float* B, const bf16x4 x[2], TileAct& t,
                                             int r, int g) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        f32x4 acc = ld_b4(B + LB_B1 + 16 * m + 4 * g);
#pragma unroll
        for (int s = 0; s < 2; ++s) acc = mfma(ld4(W + BF_W1T + (16 * m + r) * MLP_IN + 16 * s + 4 * g), x[s], acc);
        f32x4 h;
#pragma unroll
        for (int i = 0; i < 4; ++i) h[i] = gelu(acc[i]);
        t.z1[m] = acc;
        t.h1[m] = pack4(h);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        f32x4 acc = ld_b4(B + LB_B2 + 16 * m + 4 * g);
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mfma(ld4(W + BF_W2T + (16 * m + r) * MLP_HID + 16 * s + 4 * g), t.h1[s], acc);
        f32x4 h;
#pragma unroll
        for (int i = 0; i < 4; ++i) h[i] = gelu(acc[i]);
        t.z2[m] = acc;
        t.h2[m] = pack4(h);
    }
    f32x4 acc = ld_b4(B + LB_B3 + 4 * g);
#pragma unroll
    for (int s = 0; s < 4; ++s) acc = mfma(ld4(W + BF_W3T + r * MLP_HID + 16 * s + 4 * g), t.h2[s], acc);
    t.o = acc;   // lanes g = 0: outputs 0..3 of row r
}

__device__ __forceinline__ void load_x(const uint16_t* feat, uint32_t row, bool valid, int g, bf16x4 x[2]) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if (valid) {
            x[s] = ld4(feat + size_t(row) * MLP_IN + 16 * s + 4 * g);
        } else {
            x[s] = bf16x4{0, 0, 0, 0};
        }
    }
}

// ---------------------------------------------------------------- features
// Grid-neighbour feature of vertex i: (x_j - x_i) - (r_j - r_i), over the
// spacing; +0 where the neighbour is absent (j < 0).  An absent neighbour
// loads vertex i itself and its feature bits are masked to +0 (not a select,
// which the compiler turns into a branch around the loads), so that every
// gather of a vertex is in flight at once.
__device__ __forceinline__ float keep_if(float v, uint32_t mask) {
    return __uint_as_float(__float_as_uint(v) & mask);
}

__device__ __forceinline__ void nbr_delta(const WarmFeatArgs& a, uint32_t i, int32_t j, float4 pi, float4 ri,
                                          float ih, float d[3]) {
    const uint32_t mask = uint32_t(~(j >> 31));   // j >= 0: all ones
    const uint32_t jj = j >= 0 ? uint32_t(j) : i;
    const float4 pj = a.xp[jj], rj = a.rest4[jj];
    d[0] = keep_if(((pj.x - pi.x) - (rj.x - ri.x)) * ih, mask);
    d[1] = keep_if(((pj.y - pi.y) - (rj.y - ri.y)) * ih, mask);
    d[2] = keep_if(((pj.z - pi.z) - (rj.z - ri.z)) * ih, mask);
}

// The 32 bf16 features of vertex i, packed in pairs (feature 2q in the low
// half of packed[q]).
__device__ __forceinline__ void row_features(const WarmFeatArgs& a, uint32_t i, uint32_t packed[MLP_IN / 2]) {
    const float4 pi = a.xp[i], ri = a.rest4[i], ax = a.vaux[i], vi = a.vel[i];
    const float ih = ax.w;
    const int4* nb = reinterpret_cast<const int4*>(a.nbr8 + size_t(i) * 8);
    const int4 n0 = nb[0], n1 = nb[1];
    const int32_t js[8] = {n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, n1.z, n1.w};
    float f[MLP_IN];
#pragma unroll
    for (int k = 0; k < 8; ++k) nbr_delta(a, i, js[k], pi, ri, ih, f + 3 * k);
    const float sv = a.dt * ih;
    f[24] = vi.x * sv;
    f[25] = vi.y * sv;
    f[26] = vi.z * sv;
    f[27] = ax.x;
    f[28] = ax.y;
    f[29] = ax.z;
    f[30] = 0.f;
    f[31] = 0.f;
#pragma unroll
    for (int q = 0; q < MLP_IN / 2; ++q) packed[q] = uint32_t(f2bf(f[2 * q])) | (uint32_t(f2bf(f[2 * q + 1])) << 16);
}

__global__ __launch_bounds__(256) void k_warm_features(WarmFeatArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    uint32_t packed[MLP_IN / 2];
    row_features(a, i, packed);
    uint4* dst = reinterpret_cast<uint4*>(a.feat + size_t(i) * MLP_IN);
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[q] = make_uint4(packed[4 * q], packed[4 * q + 1], packed[4 * q + 2], packed[4 * q + 3]);
}

__global__ void k_warm_targets(const float4* xp, const float4* xc, const float4* vaux, uint32_t n, float* t) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = xp[i], c = xc[i];
    const float ih = vaux[i].w;
    t[3 * size_t(i)] = (c.x - p.x) * ih;
    t[3 * size_t(i) + 1] = (c.y - p.y) * ih;
    t[3 * size_t(i) + 2] = (c.z - p.z) * ih;
}

// ---------------------------------------------------------------- forward
__global__ __launch_bounds__(256) void k_mlp_forward(const uint16_t* feat, uint32_t rows, const float* P,
                                                     const uint16_t* bf, float* out) {
    __shared__ __align__(16) uint16_t W[MLP_BF];
    __shared__ __align__(16) float Bs[LB_WORDS];
    stage_weights(bf, P, W, Bs);
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t tile = wave; tile * 16 < rows; tile += nwaves) {
        const uint32_t row = tile * 16 + r;
        const bool valid = row < rows;
        bf16x4 x[2];
        load_x(feat, row, valid, g, x);
        TileAct t;
        tile_forward(W, Bs, x, t, r, g);
        if (g == 0 && valid) {
            out[3 * size_t(row)] = t.o[0];
            out[3 * size_t(row) + 1] = t.o[1];
            out[3 * size_t(row) + 2] = t.o[2];
        }
    }
}

// A fragments of the three layers for lane 16 g + r, held in registers by
// an inference wave for all its tiles (56 VGPRs): no LDS traffic (and no
// bank conflicts of the 64 B / 128 B weight rows) per tile.
struct WFrag {
    bf16x4 w1[4][2], w2[4][4], w3[4];
};

__device__ __forceinline__ void load_wfrag(const uint16_t* bf, int r, int g, WFrag& f) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
#pragma unroll
        for (int s = 0; s < 2; ++s) f.w1[m][s] = ld4(bf + BF_W1T + (16 * m + r) * MLP_IN + 16 * s + 4 * g);
#pragma unroll
        for (int s = 0; s < 4; ++s) f.w2[m][s] = ld4(bf + BF_W2T + (16 * m + r) * MLP_HID + 16 * s + 4 * g);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) f.w3[s] = ld4(bf + BF_W3T + r * MLP_HID + 16 * s + 4 * g);
}

// Forward of NT tiles from register-held weights (the same products, in the
// same order, as tile_forward; the tiles' chains interleave to hide the MFMA
// and LDS latencies); o[t]: lanes g = 0 hold outputs 0..3 of row r of tile t.
template <int NT>
__device__ __forceinline__ void tile_infer(const WFrag& wf, const float* B, const bf16x4 (&x)[NT][2], int g,
                                           f32x4 (&o)[NT]) {
    bf16x4 h1[NT][4], h2[NT][4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const f32x4 b = ld_b4(B + LB_B1 + 16 * m + 4 * g);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            f32x4 acc = b;
#pragma unroll
            for (int s = 0; s < 2; ++s) acc = mfma(wf.w1[m][s], x[t][s], acc);
            f32x4 h;
#pragma unroll
            for (int i = 0; i < 4; ++i) h[i] = gelu(acc[i]);
            h1[t][m] = pack4(h);
        }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const f32x4 b = ld_b4(B + LB_B2 + 16 * m + 4 * g);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            f32x4 acc = b;
#pragma unroll
            for (int s = 0; s < 4; ++s) acc = mfma(wf.w2[m][s], h1[t][s], acc);
            f32x4 h;
#pragma unroll
            for (int i = 0; i < 4; ++i) h[i] = gelu(acc[i]);
            h2[t][m] = pack4(h);
        }
    }
    const f32x4 b = ld_b4(B + LB_B3 + 4 * g);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        f32x4 acc = b;
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mfma(wf.w3[s], h2[t][s], acc);
        o[t] = acc;
    }
}

// Warm start of one substep: features of the predicted state, the forward,
// and x0 = xp + spacing * MLP / out_scale (pinned vertices keep xp), with no
// feature buffer in HBM.  Each pass of a block takes 256 vertices: thread t
// builds the features of vertex base + t (row_features, as k_warm_features
// does) into an LDS row, then each wave runs the forward of 4 of the 16
// tiles straight from LDS.
constexpr uint32_t WI_ROWS = 256;            // vertices per block pass (= threads)
constexpr uint32_t WI_STRIDE = MLP_IN + 4;   // bf16 per LDS row: 72 B, spreads rows over the banks
#ifndef WI_NT
#define WI_NT 1u   // tiles per forward step of a wave (1, 2 or 4)
#endif

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_warm_infer(WarmFeatArgs a, const float* P, const uint16_t* bf,
                                                    float inv_out_scale, float4* x0, uint32_t* count) {
    if (count && blockIdx.x == 0 && threadIdx.x == 0) {   // ungated: accepted, one more warm substep
        count[0] = 0u;
        count[1] = count[1] + 1u;
    }
    __shared__ __align__(16) float Bs[LB_WORDS];
    __shared__ __align__(16) uint16_t F[WI_ROWS * WI_STRIDE];
    for (uint32_t q = threadIdx.x; q < uint32_t(LB_WORDS); q += blockDim.x) {
        Bs[q] = q < uint32_t(LB_B2) ? P[OFF_B1 + q] : q < uint32_t(LB_B3) ? P[OFF_B2 + q - LB_B2] : P[OFF_B3 + q - LB_B3];
    }
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4, wv = threadIdx.x >> 6;
    WFrag wf;
    load_wfrag(bf, r, g, wf);
    for (uint32_t base = blockIdx.x * WI_ROWS; base < a.n; base += gridDim.x * WI_ROWS) {
        const uint32_t i = base + threadIdx.x;
        uint32_t packed[MLP_IN / 2];
        if (i < a.n) {
            row_features(a, i, packed);
        } else {
#pragma unroll
            for (int q = 0; q < MLP_IN / 2; ++q) packed[q] = 0u;
        }
        uint2* frow = reinterpret_cast<uint2*>(F + threadIdx.x * WI_STRIDE);
#pragma unroll
        for (int q = 0; q < MLP_IN / 4; ++q) frow[q] = make_uint2(packed[2 * q], packed[2 * q + 1]);
        __syncthreads();   // (the first pass: also the bias copy)
        bool tame_rows = true;   // the rows this lane wrote (64 per wave: one slice)
#pragma unroll 1
        for (uint32_t tt = 0; tt < 4; tt += WI_NT) {
            const uint32_t lt = 4 * wv + tt;   // first tile of the step
            if (base + 16 * lt >= a.n) break;
            bf16x4 x[WI_NT][2];
            f32x4 o[WI_NT];
#pragma unroll
            for (int t = 0; t < int(WI_NT); ++t)
#pragma unroll
                for (int s2 = 0; s2 < 2; ++s2) x[t][s2] = ld4(F + (16 * (lt + t) + r) * WI_STRIDE + 16 * s2 + 4 * g);
            tile_infer<WI_NT>(wf, Bs, x, g, o);
#pragma unroll
            for (int t = 0; t < int(WI_NT); ++t) {
                const uint32_t row = base + 16 * (lt + t) + r;
                if (g == 0 && row < a.n) {
                    const float4 p = a.xp[row];
                    const float ih = a.vaux[row].w;
                    float4 q = p;
                    if (p.w > 0.f && ih > 0.f) {
                        const float sc = inv_out_scale / ih;   // spacing / out_scale
                        q.x = p.x + o[t][0] * sc;
                        q.y = p.y + o[t][1] * sc;
                        q.z = p.z + o[t][2] * sc;
                    }
                    x0[row] = q;
                    tame_rows = tame_rows && tame3(q);
                }
            }
        }
        // wave wv wrote rows base + 64 wv .. + 63: the slice's tame byte
        const uint32_t first = base + 64 * wv;
        const int all_tame = __all(tame_rows ? 1 : 0);
        if (a.tame_out && first < a.n && lane == 0) a.tame_out[first >> 6] = all_tame ? 1u : 0u;
        __syncthreads();
    }
}

// ---------------------------------------------------------------- training
constexpr int TR_X = 0;                       // [32][16]
constexpr int TR_H1 = TR_X + MLP_IN * 16;      // [64][16]
constexpr int TR_H2 = TR_H1 + MLP_HID * 16;
constexpr int TR_D1 = TR_H2 + MLP_HID * 16;
constexpr int TR_D2 = TR_D1 + MLP_HID * 16;
constexpr int TR_DO = TR_D2 + MLP_HID * 16;   // [16][16]
constexpr int TR_WORDS = TR_DO + MLP_OUTP * 16;

__device__ __forceinline__ void st_t(uint16_t* base, int m, bf16x4 v, int r, int g) {
#pragma unroll
    for (int i = 0; i < 4; ++i) base[(16 * m + 4 * g + i) * 16 + r] = uint16_t(v[i]);
}
__device__ __forceinline__ bf16x4 ld_t(const uint16_t* base, int m, int r, int g) {
    return ld4(base + (16 * m + r) * 16 + 4 * g);
}

__global__ __launch_bounds__(256) void k_mlp_train(const uint16_t* feat, const float* targets, uint32_t rows,
                                                   float tscale, float inv_count, const float* P,
                                                   const uint16_t* bf, float* partial) {
    __shared__ __align__(16) uint16_t W[MLP_BF];
    __shared__ __align__(16) uint16_t tr_all[4][TR_WORDS];
    __shared__ __align__(16) float Bs[LB_WORDS];
    stage_weights(bf, P, W, Bs);
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4, wv = threadIdx.x >> 6;
    const uint32_t wave = blockIdx.x * 4 + wv, nwaves = gridDim.x * 4;
    uint16_t* tr = tr_all[wv];

    const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};
    f32x4 gw1[2][4], gw2[4][4], gw3[4], gb1[4], gb2[4], gb3 = z4;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gw3[a] = gb1[a] = gb2[a] = z4;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            gw2[a][b] = z4;
            if (a < 2) gw1[a][b] = z4;
        }
    }
    float lsum = 0.f;

    for (uint32_t tile = wave; tile * 16 < rows; tile += nwaves) {
        const uint32_t row = tile * 16 + r;
        const bool valid = row < rows;
        bf16x4 x[2];
        load_x(feat, row, valid, g, x);
        TileAct t;
        tile_forward(W, Bs, x, t, r, g);

        f32x4 dO = z4;
        if (g == 0 && valid) {
#pragma unroll
            for (int i = 0; i < MLP_OUT; ++i) {
                const float d = t.o[i] - targets[3 * size_t(row) + i] * tscale;
                lsum += d * d;
                dO[i] = 2.0f * d * inv_count;
            }
        }
        const bf16x4 dOb = pack4(dO);
        gb3 += dO;

        f32x4 dz2[4];
        bf16x4 dz2b[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            f32x4 acc = mfma(ld4(W + BF_W3N + (16 * m + r) * MLP_OUTP + 4 * g), dOb, z4);
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[i] *= gelu_grad(t.z2[m][i]);
            dz2[m] = acc;
            dz2b[m] = pack4(acc);
            gb2[m] += acc;
        }
        bf16x4 dz1b[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            f32x4 acc = z4;
#pragma unroll
            for (int s = 0; s < 4; ++s) acc = mfma(ld4(W + BF_W2N + (16 * m + r) * MLP_HID + 16 * s + 4 * g), dz2b[s], acc);
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[i] *= gelu_grad(t.z1[m][i]);
            dz1b[m] = pack4(acc);
            gb1[m] += acc;
        }
        (void)dz2;

        // operands of the weight gradients, transposed through LDS
#pragma unroll
        for (int s = 0; s < 2; ++s) st_t(tr + TR_X, s, x[s], r, g);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            st_t(tr + TR_H1, m, t.h1[m], r, g);
            st_t(tr + TR_H2, m, t.h2[m], r, g);
            st_t(tr + TR_D1, m, dz1b[m], r, g);
            st_t(tr + TR_D2, m, dz2b[m], r, g);
        }
        st_t(tr + TR_DO, 0, dOb, r, g);
        wave_sync();
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const bf16x4 d1 = ld_t(tr + TR_D1, b, r, g);
            const bf16x4 d2 = ld_t(tr + TR_D2, b, r, g);
#pragma unroll
            for (int a = 0; a < 2; ++a) gw1[a][b] = mfma(ld_t(tr + TR_X, a, r, g), d1, gw1[a][b]);
#pragma unroll
            for (int a = 0; a < 4; ++a) gw2[a][b] = mfma(ld_t(tr + TR_H1, a, r, g), d2, gw2[a][b]);
        }
        const bf16x4 dob = ld_t(tr + TR_DO, 0, r, g);
#pragma unroll
        for (int a = 0; a < 4; ++a) gw3[a] = mfma(ld_t(tr + TR_H2, a, r, g), dob, gw3[a]);
        wave_sync();
    }

    float* out = partial + size_t(wave) * MLP_PSTRIDE;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (a < 2) out[OFF_W1 + (16 * a + 4 * g + i) * MLP_HID + 16 * b + r] = gw1[a][b][i];
                out[OFF_W2 + (16 * a + 4 * g + i) * MLP_HID + 16 * b + r] = gw2[a][b][i];
            }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 4; ++i) out[OFF_W3 + (16 * a + 4 * g + i) * MLP_OUTP + r] = gw3[a][i];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float s1 = sum16(gb1[m][i]);
            const float s2 = sum16(gb2[m][i]);
            if (r == 0) {
                out[OFF_B1 + 16 * m + 4 * g + i] = s1;
                out[OFF_B2 + 16 * m + 4 * g + i] = s2;
            }
        }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float s3 = sum16(gb3[i]);
        if (r == 0) out[OFF_B3 + 4 * g + i] = s3;
    }
    float l = lsum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) l += __shfl_xor(l, o);
    if (lane == 0) out[MLP_PARAMS] = l;
}

__device__ __forceinline__ bool fixed_zero(int p) {
    if (p >= OFF_W1 + MLP_FEAT * MLP_HID && p < OFF_B1) return true;           // padded inputs
    if (p >= OFF_W3 && p < OFF_B3) return (p - OFF_W3) % MLP_OUTP >= MLP_OUT;   // padded outputs
    if (p >= OFF_B3) return p - OFF_B3 >= MLP_OUT;
    return false;
}

__device__ __forceinline__ void pack_param(int p, float val, uint16_t* bf) {
    const uint16_t b = f2bf(val);
    if (p < OFF_B1) {
        const int f = p / MLP_HID, h = p % MLP_HID;
        bf[BF_W1T + h * MLP_IN + f] = b;
    } else if (p >= OFF_W2 && p < OFF_B2) {
        const int q = p - OFF_W2, h1 = q / MLP_HID, h2 = q % MLP_HID;
        bf[BF_W2T + h2 * MLP_HID + h1] = b;
        bf[BF_W2N + h1 * MLP_HID + h2] = b;
    } else if (p >= OFF_W3 && p < OFF_B3) {
        const int q = p - OFF_W3, h2 = q / MLP_OUTP, o = q % MLP_OUTP;
        bf[BF_W3T + o * MLP_HID + h2] = b;
        bf[BF_W3N + h2 * MLP_OUTP + o] = b;
    }
}

// 64 parameters per block (one per lane), the partials split over 16
// waves; each wave sums its share with four independent accumulators, then
// the waves' sums are added in a fixed order: deterministic, and ~16x the
// parallelism of one thread per parameter.
constexpr uint32_t RED_WAVES = 16;

__global__ __launch_bounds__(64 * RED_WAVES) void k_mlp_reduce_adam(const float* partial, uint32_t nwaves, float* P,
                                                                   float* m, float* v, uint16_t* bf, float* grad_out,
                                                                   float* loss_out, float lr, float bc1, float bc2) {
    __shared__ float part[RED_WAVES][64];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const int p = int(blockIdx.x * 64u + lane);
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    if (p <= MLP_PARAMS) {
        uint32_t w = wv;
        for (; w + 3u * RED_WAVES < nwaves; w += 4u * RED_WAVES) {
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) acc[u] += partial[size_t(w + u * RED_WAVES) * MLP_PSTRIDE + p];
        }
        for (; w < nwaves; w += RED_WAVES) acc[0] += partial[size_t(w) * MLP_PSTRIDE + p];
    }
    part[wv][lane] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    __syncthreads();
    if (wv != 0 || p > MLP_PARAMS) return;
    float s = 0.f;
#pragma unroll
    for (uint32_t k = 0; k < RED_WAVES; ++k) s += part[k][lane];
    if (p == MLP_PARAMS) {
        if (loss_out) *loss_out = s;
        return;
    }
    if (fixed_zero(p)) s = 0.f;
    if (grad_out) grad_out[p] = s;
    if (lr > 0.f) {
        constexpr float B1 = 0.9f, B2 = 0.999f, EPS = 1e-8f;
        const float mm = B1 * m[p] + (1.f - B1) * s;
        const float vv = B2 * v[p] + (1.f - B2) * s * s;
        m[p] = mm;
        v[p] = vv;
        float w = P[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + EPS);
        if (fixed_zero(p)) w = 0.f;
        P[p] = w;
        pack_param(p, w, bf);
    }
}

__global__ void k_mlp_pack(const float* P, uint16_t* bf) {
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p < MLP_PARAMS) pack_param(p, P[p], bf);
}

}  // namespace

void launch_warm_features(const WarmFeatArgs& a, hipStream_t s) {
    if (a.n == 0) return;
    hipLaunchKernelGGL(k_warm_features, dim3(cdiv(a.n, 256)), dim3(256), 0, s, a);
}

void launch_warm_targets(const float4* xp, const float4* xconv, const float4* vaux, uint32_t n, float* t,
                         hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_warm_targets, dim3(cdiv(n, 256)), dim3(256), 0, s, xp, xconv, vaux, n, t);
}

void launch_mlp_forward(const uint16_t* feat, uint32_t rows, const float* P, const uint16_t* bf, float* out,
                        hipStream_t s) {
    if (rows == 0) return;
    const uint32_t blocks = std::min<uint32_t>(cdiv(cdiv(rows, 16), 4), 2048);
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(256), 0, s, feat, rows, P, bf, out);
}

void launch_mlp_train(const uint16_t* feat, const float* targets, uint32_t rows, float target_scale,
                      const float* P, const uint16_t* bf, float* partial, hipStream_t s) {
    const float inv_count = rows ? 1.0f / (3.0f * float(rows)) : 0.f;
    hipLaunchKernelGGL(k_mlp_train, dim3(MLP_TRAIN_BLOCKS), dim3(256), 0, s, feat, targets, rows, target_scale,
                       inv_count, P, bf, partial);
}

void launch_mlp_reduce_adam(const float* partial, float* P, float* m, float* v, uint16_t* bf, float* grad_out,
                            float* loss_out, float lr, float beta1_t, float beta2_t, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3(cdiv(MLP_PARAMS + 1, 64)), dim3(64 * RED_WAVES), 0, s, partial,
                       MLP_TRAIN_WAVES, P, m, v, bf, grad_out, loss_out, lr, 1.0f - beta1_t, 1.0f - beta2_t);
}

void launch_mlp_pack(const float* P, uint16_t* bf, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(cdiv(MLP_PARAMS, 256)), dim3(256), 0, s, P, bf);
}

void launch_warm_infer(const WarmFeatArgs& a, const float* P, const uint16_t* bf, float inv_out_scale, float4* x0,
                       uint32_t* count, hipStream_t s) {
    if (a.n == 0) return;
    // one co-resident wave of blocks: each loads the weights once and walks
    // its passes (more blocks only reload the weights)
    static std::atomic<uint32_t> resident[64];   // per device, 0 until first queried
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    uint32_t cap = resident[dev].load(std::memory_order_relaxed);
    if (cap == 0) {
        int cus = 0, per_cu = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_warm_infer, int(WI_ROWS), 0) == hipSuccess)
            cap = uint32_t(std::max(1, cus * per_cu));
        else
            cap = 2048;
        resident[dev].store(cap, std::memory_order_relaxed);
    }
    const uint32_t blocks = std::min<uint32_t>(cdiv(a.n, WI_ROWS), cap);
    hipLaunchKernelGGL(k_warm_infer, dim3(blocks), dim3(256), 0, s, a, P, bf, inv_out_scale, x0, count);
}

}  // namespace xpbd
