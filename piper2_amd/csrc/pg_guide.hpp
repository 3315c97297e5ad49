// Neural path guiding on gfx950 matrix cores: network layout, feature encoding, BF16 MFMA layers
// and the 8-lobe von Mises-Fisher mixture used as a one-sample-MIS proposal next to the BSDF.
//
// Network: 144 input features -> 4 hidden layers of 64 (ReLU) -> 40 outputs in 48 rows (8 zero
// rows). Weights are row-major [out][in] FP32 masters with a BF16 copy; matrix products run with
// FP32 accumulation on v_mfma_f32_16x16x32_bf16 (inference) and v_mfma_f32_16x16x16_bf16
// (training), one instruction shape per accumulator chain. Inputs and outputs are ordered so that each lane
// computes the features it feeds and decodes the lobes it holds (guideFeatureTiles, output layout).
//
// MFMA operand layout used throughout (lane l, g = l / 16, s = l % 16, j = 0..3):
//   A (16 x 16, rows m, cols k): lane holds A[s][4g + j]
//   B (16 x 16, rows k, cols n): lane holds B[4g + j][s]
//   C/D (16 x 16):              lane holds D[4g + j][s]
// A layer computes Z^T (neurons x samples) = W (neurons x in) . X^T (in x samples): weights are the
// A operand (read from LDS), activations the B operand, so a layer's D tile (neurons 4g+j of sample
// s) is, after ReLU and rounding, exactly the next layer's B operand for the same 16 neurons.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pg_shared.hpp"

namespace pg {

constexpr int kGIn = 144, kGHid = 64, kGOut = 48, kGLobes = 8;
// Output layout: lane group g holds rows 16 mt + 4 g + j (mt = 0..2, j = 0..3) of its sample; slot
// u = 4 mt + j is lobe 2 g + u / 5, parameter u % 5 (mu x, mu y, mu z, log kappa, weight logit) for
// u < 10. Slot 10 of lane group 0 (row 34) is the logit of the learned guide-vs-BSDF selection
// probability (trained only when a guide learns it, zero otherwise); the other slots 10 and 11 are
// the zero rows 35, 38, 39, 42, 43, 46, 47.
constexpr int kGFreq = 12, kGBlob = 32;
// FP32 parameter offsets (floats), layer l: W then b
constexpr int kGW0 = 0, kGB0 = kGW0 + kGHid * kGIn;
constexpr int kGW1 = kGB0 + kGHid, kGB1 = kGW1 + kGHid * kGHid;
constexpr int kGW2 = kGB1 + kGHid, kGB2 = kGW2 + kGHid * kGHid;
constexpr int kGW3 = kGB2 + kGHid, kGB3 = kGW3 + kGHid * kGHid;
constexpr int kGW4 = kGB3 + kGHid, kGB4 = kGW4 + kGOut * kGHid;
constexpr int kGParams = kGB4 + kGOut;  // 24880
static_assert(kGParams == 24880, "guide parameter count");
constexpr float kGKappaLo = -4.0f, kGKappaHi = 9.0f;  // log-kappa clamp
// Mixture record per sample: 8 x (mu xyz, kappa), 8 weights pi, 8 x pi * vMF normalisation, then
// the learned selection probability of the guide and three zeros.
constexpr int kGMixFloats = 52;
constexpr int kGAlphaSlot = 10;                     // output slot of lane group 0 holding its logit
constexpr float kGAlphaLo = 0.05f, kGAlphaHi = 0.95f;  // a learned selection probability stays inside
// Selection probability from its logit: kGAlphaLo + (kGAlphaHi - kGAlphaLo) sigmoid(z), so that both
// the guide and the BSDF keep a share of the samples wherever the guide is used.
__device__ __forceinline__ float alphaFromLogit(float z) {
    return kGAlphaLo + (kGAlphaHi - kGAlphaLo) * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

// LDS row strides (bf16 elements): 4 extra elements keep the 16 rows of one A-operand read on
// distinct banks.
constexpr int kGLd0 = kGIn + 4, kGLdH = kGHid + 4, kGLdOT = kGOut + 4;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16v8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16v2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));

// Round to nearest even: gfx950's v_cvt_pk_bf16_f32 (what a __bf16 cast lowers to).
__device__ __forceinline__ uint16_t toBf16(float f) {
    const __bf16 b = static_cast<__bf16>(f);
    uint16_t u;
    __builtin_memcpy(&u, &b, 2);
    return u;
}
__device__ __forceinline__ float fromBf16(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// gfx950's K = 32 BF16 product (twice the K of mfma16 per instruction). Operand layout, lane l
// (s = l % 16, g = l / 16): A holds A[s][8g + j], B holds B[8g + j][s] (j = 0..7); D as mfma16.
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16v8, a), __builtin_bit_cast(bf16v8, b), c, 0,
                                                  0, 0);
}
__device__ __forceinline__ bf16x8 cat(bf16x4 lo, bf16x4 hi) {
    return bf16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

// Two floats to two BF16 (round to nearest even) in one v_cvt_pk_bf16_f32.
__device__ __forceinline__ i16x2 toBf16x2(float a, float b) {
    return __builtin_bit_cast(i16x2, __builtin_convertvector(f32x2{a, b}, bf16v2));
}
// ReLU on BF16 bit patterns: as signed 16-bit integers every negative value (and -0) is below 0,
// positive values keep their order, so max(x, 0) in v_pk_max_i16 is the float ReLU.
__device__ __forceinline__ i16x2 reluBf16x2(i16x2 v) { return __builtin_elementwise_max(v, i16x2{0, 0}); }

__device__ __forceinline__ bf16x4 packRelu(f32x4 v) {
    const i16x2 lo = reluBf16x2(toBf16x2(v[0], v[1])), hi = reluBf16x2(toBf16x2(v[2], v[3]));
    return bf16x4{lo[0], lo[1], hi[0], hi[1]};
}
__device__ __forceinline__ bf16x4 pack(f32x4 v) {
    const i16x2 lo = toBf16x2(v[0], v[1]), hi = toBf16x2(v[2], v[3]);
    return bf16x4{lo[0], lo[1], hi[0], hi[1]};
}

// Weights of one block in LDS (BF16, padded rows) and biases (FP32).
struct GuideLds {
    uint16_t* w0;  // [64][kGLd0]
    uint16_t* wh;  // 3 x [64][kGLdH]
    uint16_t* w4;  // [48][kGLdH]
    float* b;      // 64 * 4 + 48
};
constexpr int kGLdsWeightBytes = (kGHid * kGLd0 + 3 * kGHid * kGLdH + kGOut * kGLdH) * 2 + (4 * kGHid + kGOut) * 4;

__device__ __forceinline__ GuideLds guideLdsLayout(unsigned char* base) {
    GuideLds L;
    L.w0 = reinterpret_cast<uint16_t*>(base);
    L.wh = L.w0 + kGHid * kGLd0;
    L.w4 = L.wh + 3 * kGHid * kGLdH;
    L.b = reinterpret_cast<float*>(L.w4 + kGOut * kGLdH);
    return L;
}

// Rows of BF16 weights from global (unpadded, 16-byte aligned) into padded LDS rows: 8 weights per
// 16-byte load; the padded rows are 8-byte aligned, so each load becomes two 8-byte stores.
template <int ROWS, int COLS, int LD>
__device__ __forceinline__ void guideCopyRows(uint16_t* dst, const uint16_t* __restrict__ src) {
    static_assert(COLS % 8 == 0 && LD % 4 == 0, "vector copy");
    constexpr int kVec = COLS / 8;
    for (int i = threadIdx.x; i < ROWS * kVec; i += blockDim.x) {
        const int r = i / kVec, q = i - r * kVec;
        const uint4 w = reinterpret_cast<const uint4*>(src + r * COLS)[q];
        uint2* d = reinterpret_cast<uint2*>(dst + r * LD + 8 * q);
        d[0] = make_uint2(w.x, w.y);
        d[1] = make_uint2(w.z, w.w);
    }
}
static_assert(kGW1 % 8 == 0 && kGW2 % 8 == 0 && kGW3 % 8 == 0 && kGW4 % 8 == 0, "16-byte aligned weight rows");

// Cooperative copy of the BF16 weights (global, unpadded [out][in]) and FP32 biases into LDS.
__device__ __forceinline__ void guideLoadLds(const GuideLds& L, const uint16_t* __restrict__ wb,
                                             const float* __restrict__ theta) {
    guideCopyRows<kGHid, kGIn, kGLd0>(L.w0, wb + kGW0);
    guideCopyRows<kGHid, kGHid, kGLdH>(L.wh, wb + kGW1);
    guideCopyRows<kGHid, kGHid, kGLdH>(L.wh + kGHid * kGLdH, wb + kGW2);
    guideCopyRows<kGHid, kGHid, kGLdH>(L.wh + 2 * kGHid * kGLdH, wb + kGW3);
    guideCopyRows<kGOut, kGHid, kGLdH>(L.w4, wb + kGW4);
    for (int i = threadIdx.x; i < kGHid; i += blockDim.x) {
        L.b[i] = theta[kGB0 + i];
        L.b[kGHid + i] = theta[kGB1 + i];
        L.b[2 * kGHid + i] = theta[kGB2 + i];
        L.b[3 * kGHid + i] = theta[kGB3 + i];
    }
    for (int i = threadIdx.x; i < kGOut; i += blockDim.x) L.b[4 * kGHid + i] = theta[kGB4 + i];
}

__device__ __forceinline__ bf16x4 ldsA(const uint16_t* w, int ld, int row, int col) {
    return *reinterpret_cast<const bf16x4*>(w + row * ld + col);
}

// Inputs of one sample: position in [0,1]^3, unit shading normal, unit outgoing direction.
struct GuideInput {
    float px, py, pz, nx, ny, nz, wx, wy, wz;
};
__device__ __forceinline__ GuideInput guideInput(float3 p01, float3 n, float3 wo) {
    return GuideInput{p01.x, p01.y, p01.z, n.x, n.y, n.z, wo.x, wo.y, wo.z};
}

// sin(pi x) and cos(pi x): x reduced to r = x - n / 2 with n = rint(2x) (exact for |x| < 2^22),
// Taylor polynomials of sin and cos at t = pi r in [-pi/4, pi/4] (truncation below 3e-8), then the
// quarter turn n applied. Cheaper than a library sincospi, accurate far beyond the BF16 features.
__device__ __forceinline__ void sinCosPi(float x, float& s, float& c) {
    const float n = rintf(2.0f * x);
    const float r = fmaf(n, -0.5f, x);
    const float t = r * 3.14159265358979323846f, t2 = t * t;
    const float sp = t * fmaf(t2, fmaf(t2, fmaf(t2, fmaf(t2, 2.75573192e-6f, -1.98412698e-4f), 8.33333333e-3f),
                                       -1.66666667e-1f), 1.0f);
    const float cp = fmaf(t2, fmaf(t2, fmaf(t2, fmaf(t2, 2.48015873e-5f, -1.38888889e-3f), 4.16666667e-2f), -0.5f), 1.0f);
    const int q = static_cast<int>(n);
    const float a = (q & 1) ? cp : sp, b = (q & 1) ? sp : cp;
    s = (q & 2) ? -a : a;
    c = ((q + 1) & 2) ? -b : b;
}
// atan2(y, x) in [-pi, pi]: atan of min / max on [0, 1] from an odd polynomial of degree 11 fitted
// (reweighted least squares toward minimax) to atan, maximum error 1.8e-6 rad, then the octant.
__device__ __forceinline__ float atan2Guide(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float a = mx > 0.0f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f, a2 = a * a;
    float p = a * fmaf(a2, fmaf(a2, fmaf(a2, fmaf(a2, fmaf(a2, -0.011719105f, 0.052647319f), -0.11642652f), 0.19354044f),
                                -0.33262286f), 0.99997723f);
    if (ay > ax) p = 1.57079632679489662f - p;
    if (x < 0.0f) p = 3.14159265358979323846f - p;
    return copysignf(p, y);
}

// Layer-0 input layout. Lane group g (lanes 16g..16g+15, one sample each) feeds the B-operand
// slots k = 16 kt + 4 g + j (kt = 0..8, j = 0..3); slot t = 4 kt + j of group g holds
//   t = 0..17  : position item P = 9 g + t / 2, coordinate P / 12 at frequency f = P % 12:
//                sin(2^f pi p) for even t, cos(2^f pi p) for odd t (72 in all);
//   t = 18..33 : one-blob bin b = 16 (g % 2) + t - 18 of theta (g < 2) or phi (g >= 2):
//                exp(-(32 a - b - 0.5)^2 / 2), a = theta / pi or phi / 2 pi in [0, 1) (64 in all),
//                from two anchors (bins 0 and 8 of the lane's run) and the product recurrence
//                g(d - 1) = g(d) exp(d - 1/2); where an anchor underflows its 8 bins are below 2e-12;
//   t = 34, 35 : n.x, n.y (g = 0); n.z, 0 (g = 1); 0, 0 (g = 2, 3).
// Every lane thus computes exactly the 36 features it feeds, with the same instructions as every
// other lane (no divergence, nothing computed twice). Zero slots: features 135, 138, 139, 142, 143.
// Frequencies come from one sinCosPi at the start of each coordinate run and the double-angle
// recurrence (sin 2x = 2 sin x cos x, cos 2x = 1 - 2 sin^2 x), at most 8 steps.
__device__ __forceinline__ void guideFeatureTiles(const GuideInput& g, int grp, bf16x4 (&x0)[9]) {
    float v[36];
    // position items: a run of coordinate c0 from frequency f0, then coordinate c0 + 1 from 0
    const int p0 = 9 * grp, c0 = p0 / 12, f0 = p0 % 12, brk = 12 - f0;
    const float va = c0 == 0 ? g.px : (c0 == 1 ? g.py : g.pz);
    const float vb = c0 == 0 ? g.py : g.pz;
    float s, c, s1, c1;
    sinCosPi(ldexpf(va, f0), s, c);  // 2^f0 va is exact
    sinCosPi(vb, s1, c1);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        if (i == brk) s = s1, c = c1;
        v[2 * i] = s;
        v[2 * i + 1] = c;
        const float t = s + s;  // sin 2x = 2 sin x cos x, cos 2x = 1 - 2 sin^2 x
        const float ns = t * c;
        c = fmaf(-t, s, 1.0f);
        s = ns;
    }
    // one-blob of theta (groups 0, 1) or phi (groups 2, 3): both from one atan2
    const bool th = grp < 2;
    const float ay = th ? __builtin_amdgcn_sqrtf(g.wx * g.wx + g.wy * g.wy) : g.wy, ax = th ? g.wz : g.wx;
    float a = atan2Guide(ay, ax) * (th ? 0.31830988618379067154f : 0.15915494309189533577f);
    if (!th && a < 0.0f) a += 1.0f;
    const float u = a * kGBlob - static_cast<float>(16 * (grp & 1)) - 0.5f;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const float d = u - static_cast<float>(8 * h);  // distance of the anchor bin from the blob's centre
        float gk = __builtin_amdgcn_exp2f(d * d * -0.72134752044448170368f);      // exp(-d^2 / 2)
        float rk = __builtin_amdgcn_exp2f((d - 0.5f) * 1.44269504088896340736f);  // exp(d - 1/2)
        v[18 + 8 * h] = gk;
#pragma unroll
        for (int k = 1; k < 8; ++k) {
            gk = gk * rk;
            rk = rk * 0.36787944117144233f;  // exp(-1)
            v[18 + 8 * h + k] = gk;
        }
    }
    v[34] = grp == 0 ? g.nx : (grp == 1 ? g.nz : 0.0f);
    v[35] = grp == 0 ? g.ny : 0.0f;
#pragma unroll
    for (int kt = 0; kt < 9; ++kt) x0[kt] = pack(f32x4{v[4 * kt], v[4 * kt + 1], v[4 * kt + 2], v[4 * kt + 3]});
}

// Forward pass of one 16-sample tile. x0: layer-0 B operands (9 K tiles). out: 3 D tiles of the
// output layer. If `save` is set, the B operands of layers 1..4 (post-ReLU activations) are kept.
__device__ __forceinline__ void guideForwardTile(const GuideLds& L, const bf16x4 (&x0)[9], f32x4 (&out)[3],
                                                 bf16x4 (*save)[4] = nullptr) {
    const int lane = __lane_id();
    const int s = lane & 15, grp = lane >> 4;
    f32x4 acc[4];
    bf16x4 h[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        acc[mt] = f32x4{L.b[mt * 16 + 4 * grp], L.b[mt * 16 + 4 * grp + 1], L.b[mt * 16 + 4 * grp + 2],
                        L.b[mt * 16 + 4 * grp + 3]};
#pragma unroll
        for (int kt = 0; kt < 9; ++kt) acc[mt] = mfma16(ldsA(L.w0, kGLd0, mt * 16 + s, kt * 16 + 4 * grp), x0[kt], acc[mt]);
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) h[mt] = packRelu(acc[mt]);
    if (save)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) save[0][mt] = h[mt];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        const uint16_t* w = L.wh + l * kGHid * kGLdH;
        const float* b = L.b + (l + 1) * kGHid;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            acc[mt] = f32x4{b[mt * 16 + 4 * grp], b[mt * 16 + 4 * grp + 1], b[mt * 16 + 4 * grp + 2],
                            b[mt * 16 + 4 * grp + 3]};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) acc[mt] = mfma16(ldsA(w, kGLdH, mt * 16 + s, kt * 16 + 4 * grp), h[kt], acc[mt]);
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) h[mt] = packRelu(acc[mt]);
        if (save)
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) save[l + 1][mt] = h[mt];
    }
    const float* b4 = L.b + 4 * kGHid;
#pragma unroll
    for (int mt = 0; mt < 3; ++mt) {
        out[mt] = f32x4{b4[mt * 16 + 4 * grp], b4[mt * 16 + 4 * grp + 1], b4[mt * 16 + 4 * grp + 2],
                        b4[mt * 16 + 4 * grp + 3]};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) out[mt] = mfma16(ldsA(L.w4, kGLdH, mt * 16 + s, kt * 16 + 4 * grp), h[kt], out[mt]);
    }
}

// ---- inference on K = 32 products --------------------------------------------------------------
// A K = 32 tile's B operand is two K = 16 tiles of the layout above side by side: lane group g holds
// rows 4g..4g+3 of the first and of the second, which are K positions 8g..8g+7 of the K = 32 tile.
// So a layer's D tiles mt = 2kt, 2kt + 1 are the next layer's K = 32 B operand kt as they stand, and
// the layer-0 feature tiles x0[2kt], x0[2kt + 1] are layer 0's K = 32 operand kt, provided the A
// operand (the weights) holds, at K position 32kt + 8g + j, the input 32kt + 4g + j (j < 4) or
// 32kt + 16 + 4g + (j - 4) (j >= 4). The inference LDS copy stores every 32-column block of every
// weight row in that order; layer 0's last 16 columns (features 128..143) stay in a K = 16 product.
// Parameters, features and outputs are unchanged.
constexpr int kGLd0X = kGIn + 8, kGLdHX = kGHid + 8;  // 16-byte rows; 16 rows of a read hit distinct banks
struct GuideLdsX {
    uint16_t* w0;  // [64][kGLd0X]
    uint16_t* wh;  // 3 x [64][kGLdHX]
    uint16_t* w4;  // [48][kGLdHX]
    float* b;      // 64 * 4 + 48
};
constexpr int kGLdsXWeightBytes = (kGHid * kGLd0X + 3 * kGHid * kGLdHX + kGOut * kGLdHX) * 2 + (4 * kGHid + kGOut) * 4;
static_assert(((kGHid * kGLd0X + 3 * kGHid * kGLdHX + kGOut * kGLdHX) * 2) % 16 == 0, "16-byte aligned biases");

__device__ __forceinline__ GuideLdsX guideLdsXLayout(unsigned char* base) {
    GuideLdsX L;
    L.w0 = reinterpret_cast<uint16_t*>(base);
    L.wh = L.w0 + kGHid * kGLd0X;
    L.w4 = L.wh + 3 * kGHid * kGLdHX;
    L.b = reinterpret_cast<float*>(L.w4 + kGOut * kGLdHX);
    return L;
}

// Rows of COLS BF16 weights into LDS rows of stride LD, each 32-column block below PERM in the K = 32
// operand order (8 consecutive LDS elements 32b + 8q.. = inputs 32b + 4q..+3 and 32b + 16 + 4q..+3),
// the columns from PERM on copied as they are.
template <int ROWS, int COLS, int LD, int PERM>
__device__ __forceinline__ void guideCopyRowsX(uint16_t* dst, const uint16_t* __restrict__ src) {
    static_assert(COLS % 8 == 0 && PERM % 32 == 0 && LD % 8 == 0, "16-byte groups");
    constexpr int kVec = COLS / 8;
    for (int i = threadIdx.x; i < ROWS * kVec; i += blockDim.x) {
        const int r = i / kVec, q8 = i - r * kVec, c = 8 * q8;
        const uint16_t* row = src + r * COLS;
        uint4 w;
        if (c < PERM) {
            const int b = c & ~31, q = (c & 31) >> 3;
            const uint2 lo = *reinterpret_cast<const uint2*>(row + b + 4 * q);
            const uint2 hi = *reinterpret_cast<const uint2*>(row + b + 16 + 4 * q);
            w = make_uint4(lo.x, lo.y, hi.x, hi.y);
        } else {
            w = *reinterpret_cast<const uint4*>(row + c);
        }
        *reinterpret_cast<uint4*>(dst + r * LD + c) = w;
    }
}

__device__ __forceinline__ void guideLoadLdsX(const GuideLdsX& L, const uint16_t* __restrict__ wb,
                                              const float* __restrict__ theta) {
    guideCopyRowsX<kGHid, kGIn, kGLd0X, 128>(L.w0, wb + kGW0);
    guideCopyRowsX<kGHid, kGHid, kGLdHX, kGHid>(L.wh, wb + kGW1);
    guideCopyRowsX<kGHid, kGHid, kGLdHX, kGHid>(L.wh + kGHid * kGLdHX, wb + kGW2);
    guideCopyRowsX<kGHid, kGHid, kGLdHX, kGHid>(L.wh + 2 * kGHid * kGLdHX, wb + kGW3);
    guideCopyRowsX<kGOut, kGHid, kGLdHX, kGHid>(L.w4, wb + kGW4);
    for (int i = threadIdx.x; i < kGHid; i += blockDim.x) {
        L.b[i] = theta[kGB0 + i];
        L.b[kGHid + i] = theta[kGB1 + i];
        L.b[2 * kGHid + i] = theta[kGB2 + i];
        L.b[3 * kGHid + i] = theta[kGB3 + i];
    }
    for (int i = threadIdx.x; i < kGOut; i += blockDim.x) L.b[4 * kGHid + i] = theta[kGB4 + i];
}

__device__ __forceinline__ bf16x8 ldsA8(const uint16_t* w, int ld, int row, int col) {
    return *reinterpret_cast<const bf16x8*>(w + row * ld + col);
}
__device__ __forceinline__ f32x4 biasTile(const float* b, int mt, int grp) {
    return *reinterpret_cast<const f32x4*>(b + mt * 16 + 4 * grp);
}

// Forward pass on K = 32 products: layer 0 in 4 K = 32 steps and one K = 16 tail step per output
// tile, the other layers in 2 K = 32 steps (50 matrix instructions per 16-sample tile).
// Layer 0 of one tile: the pre-activations acc (4 output tiles) from its feature tiles x0.
__device__ __forceinline__ void guideLayer0X32(const GuideLdsX& L, const bf16x4 (&x0)[9], f32x4 (&acc)[4]) {
    const int lane = __lane_id();
    const int s = lane & 15, grp = lane >> 4;
    const bf16x4 z = {};
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        acc[mt] = biasTile(L.b, mt, grp);
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
            acc[mt] = mfma32(ldsA8(L.w0, kGLd0X, mt * 16 + s, 32 * kt + 8 * grp), cat(x0[2 * kt], x0[2 * kt + 1]), acc[mt]);
        // K tail 128..143 as a K = 32 product whose upper 16 columns are zero in both operands:
        // an accumulator chain keeps one MFMA shape. A 16x16x16 product that reads as C the result
        // of a 16x16x32 product issued one or two instructions earlier has read it unfinished on
        // gfx950 (results then depend on timing), and the compiler pads only same-shape chains.
        acc[mt] = mfma32(cat(ldsA(L.w0, kGLd0X, mt * 16 + s, 128 + 4 * grp), z), cat(x0[8], z), acc[mt]);
    }
}
// Layers 1..4 of T tiles side by side from their layer-0 pre-activations: every weight operand read
// from LDS feeds the T tiles' products, and the T independent accumulator chains of an output tile
// keep the matrix pipe busy within one wave.
template <int T>
__device__ __forceinline__ void guideHiddenX32(const GuideLdsX& L, f32x4 (&acc)[T][4], f32x4 (&out)[T][3]) {
    const int lane = __lane_id();
    const int s = lane & 15, grp = lane >> 4;
    bf16x4 h[T][4];
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) h[t][mt] = packRelu(acc[t][mt]);
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        const uint16_t* w = L.wh + l * kGHid * kGLdHX;
        const float* b = L.b + (l + 1) * kGHid;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const f32x4 bias = biasTile(b, mt, grp);
#pragma unroll
            for (int t = 0; t < T; ++t) acc[t][mt] = bias;
#pragma unroll
            for (int kt = 0; kt < 2; ++kt) {
                const bf16x8 a = ldsA8(w, kGLdHX, mt * 16 + s, 32 * kt + 8 * grp);
#pragma unroll
                for (int t = 0; t < T; ++t) acc[t][mt] = mfma32(a, cat(h[t][2 * kt], h[t][2 * kt + 1]), acc[t][mt]);
            }
        }
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) h[t][mt] = packRelu(acc[t][mt]);
    }
    const float* b4 = L.b + 4 * kGHid;
#pragma unroll
    for (int mt = 0; mt < 3; ++mt) {
        const f32x4 bias = biasTile(b4, mt, grp);
#pragma unroll
        for (int t = 0; t < T; ++t) out[t][mt] = bias;
#pragma unroll
        for (int kt = 0; kt < 2; ++kt) {
            const bf16x8 a = ldsA8(L.w4, kGLdHX, mt * 16 + s, 32 * kt + 8 * grp);
#pragma unroll
            for (int t = 0; t < T; ++t) out[t][mt] = mfma32(a, cat(h[t][2 * kt], h[t][2 * kt + 1]), out[t][mt]);
        }
    }
}

// ---- von Mises-Fisher mixture ---------------------------------------------------------------
// vMF density on the sphere: kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1)).
__device__ __forceinline__ float vmfNorm(float kappa) {
    return kappa / (2.0f * 3.14159265358979f * (-expm1f(-2.0f * kappa)));
}
__device__ __forceinline__ float vmfPdf(float kappa, float cosang) {
    return vmfNorm(kappa) * __expf(kappa * (cosang - 1.0f));
}
// The normalisation of the density mixtureSample draws from: it inverts the CDF
// (exp(kappa (t - 1)) - E) / (1 - E) with E = __expf(-2 kappa), so its density is
// kappa exp(kappa (cos - 1)) / (2 pi (1 - E)) with that same E.
__device__ __forceinline__ float vmfNormSampled(float kappa) {
    return kappa * __builtin_amdgcn_rcpf(2.0f * 3.14159265358979f * (1.0f - __expf(-2.0f * kappa)));
}

// Mixture record of one sample: lobe[i] = (mu xyz, kappa) for i < 8 at float4 i, weights pi at
// 32..39, pi_i times lobe i's normalisation at 40..47 (so the density costs one exp per lobe).
__device__ __forceinline__ float mixturePdf(const float* __restrict__ mix, float3 w) {
    float p = 0.0f;
#pragma unroll
    for (int i = 0; i < kGLobes; ++i) {
        const float4 l = reinterpret_cast<const float4*>(mix)[i];
        p += mix[40 + i] * __expf(l.w * (l.x * w.x + l.y * w.y + l.z * w.z - 1.0f));
    }
    return p;
}

// Sample the mixture with u1 (lobe choice, reused after remapping) and u2 (Jakob 2012 vMF method).
__device__ __forceinline__ float3 mixtureSample(const float* __restrict__ mix, float u1, float u2) {
    // lobe by the CDF of the weights; the last lobe with weight absorbs rounding at the top
    int sel = -1, last = 0;
    float acc = 0.0f, lo = 0.0f, wsel = 1.0f, loLast = 0.0f;
    for (int i = 0; i < kGLobes; ++i) {
        const float wi = mix[32 + i];
        if (wi > 0.0f) {
            if (sel < 0 && u1 < acc + wi) sel = i, lo = acc, wsel = wi;
            last = i, loLast = acc;
        }
        acc += wi;
    }
    if (sel < 0) sel = last, lo = loLast, wsel = mix[32 + last];
    u1 = fminf(fmaxf((u1 - lo) * __builtin_amdgcn_rcpf(wsel), 0.0f), 1.0f);
    const float4 l = reinterpret_cast<const float4*>(mix)[sel];
    const float k = l.w;
    const float wz = fmaxf(-1.0f, 1.0f + logf(fmaxf(u1 + (1.0f - u1) * __expf(-2.0f * k), 1e-30f)) * __builtin_amdgcn_rcpf(k));
    const float r = __builtin_amdgcn_sqrtf(fmaxf(0.0f, 1.0f - wz * wz));
    float sp, cp;
    sincosf(2.0f * 3.14159265358979f * u2, &sp, &cp);
    // frame around mu (Duff et al. 2017 orthonormal basis)
    const float sign = copysignf(1.0f, l.z);
    const float a = -__builtin_amdgcn_rcpf(sign + l.z), bb = l.x * l.y * a;
    const float3 t = make_float3(1.0f + sign * l.x * l.x * a, sign * bb, -sign * l.x);
    const float3 b = make_float3(bb, sign + l.y * l.y * a, -l.y);
    const float x = r * cp, y = r * sp;
    return make_float3(t.x * x + b.x * y + l.x * wz, t.y * x + b.y * y + l.y * wz, t.z * x + b.z * y + l.z * wz);
}

}  // namespace pg
