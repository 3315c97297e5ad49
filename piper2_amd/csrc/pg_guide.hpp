// Neural path guiding on gfx950 matrix cores: network layout, feature encoding, BF16 MFMA layers
// and the 8-lobe von Mises-Fisher mixture used as a one-sample-MIS proposal next to the BSDF.
//
// Network: 144 input features -> 4 hidden layers of 64 (ReLU) -> 40 outputs (48 rows, the last 8
// padding). Weights are row-major [out][in] FP32 masters with a BF16 copy; matrix products run on
// v_mfma_f32_16x16x16_bf16 with FP32 accumulation.
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

constexpr int kGIn = 144, kGHid = 64, kGOut = 48, kGOutUsed = 40, kGLobes = 8;
constexpr int kGFreq = 12, kGBlob = 32, kGFeat = 139;
// FP32 parameter offsets (floats), layer l: W then b
constexpr int kGW0 = 0, kGB0 = kGW0 + kGHid * kGIn;
constexpr int kGW1 = kGB0 + kGHid, kGB1 = kGW1 + kGHid * kGHid;
constexpr int kGW2 = kGB1 + kGHid, kGB2 = kGW2 + kGHid * kGHid;
constexpr int kGW3 = kGB2 + kGHid, kGB3 = kGW3 + kGHid * kGHid;
constexpr int kGW4 = kGB3 + kGHid, kGB4 = kGW4 + kGOut * kGHid;
constexpr int kGParams = kGB4 + kGOut;  // 24880
static_assert(kGParams == 24880, "guide parameter count");
constexpr float kGKappaLo = -4.0f, kGKappaHi = 9.0f;  // log-kappa clamp
constexpr int kGMixFloats = 48;  // per sample: 8 x (mu xyz, kappa), 8 x weight, 8 padding

// LDS row strides (bf16 elements): 4 extra elements keep the 16 rows of one A-operand read on
// distinct banks.
constexpr int kGLd0 = kGIn + 4, kGLdH = kGHid + 4, kGLdOT = kGOut + 4;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

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

__device__ __forceinline__ bf16x4 packRelu(f32x4 v) {
    bf16x4 r;
    r[0] = static_cast<short>(toBf16(fmaxf(v[0], 0.0f)));
    r[1] = static_cast<short>(toBf16(fmaxf(v[1], 0.0f)));
    r[2] = static_cast<short>(toBf16(fmaxf(v[2], 0.0f)));
    r[3] = static_cast<short>(toBf16(fmaxf(v[3], 0.0f)));
    return r;
}
__device__ __forceinline__ bf16x4 pack(f32x4 v) {
    bf16x4 r;
    r[0] = static_cast<short>(toBf16(v[0]));
    r[1] = static_cast<short>(toBf16(v[1]));
    r[2] = static_cast<short>(toBf16(v[2]));
    r[3] = static_cast<short>(toBf16(v[3]));
    return r;
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

// Cooperative copy of the BF16 weights (global, unpadded [out][in]) and FP32 biases into LDS.
__device__ __forceinline__ void guideLoadLds(const GuideLds& L, const uint16_t* __restrict__ wb,
                                             const float* __restrict__ theta) {
    for (int i = threadIdx.x; i < kGHid * kGIn; i += blockDim.x) L.w0[(i / kGIn) * kGLd0 + i % kGIn] = wb[kGW0 + i];
    for (int l = 0; l < 3; ++l) {
        const int off = l == 0 ? kGW1 : (l == 1 ? kGW2 : kGW3);
        for (int i = threadIdx.x; i < kGHid * kGHid; i += blockDim.x)
            L.wh[l * kGHid * kGLdH + (i / kGHid) * kGLdH + i % kGHid] = wb[off + i];
    }
    for (int i = threadIdx.x; i < kGOut * kGHid; i += blockDim.x) L.w4[(i / kGHid) * kGLdH + i % kGHid] = wb[kGW4 + i];
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

// Feature k of a sample (positions in [0,1]^3, unit normal, unit outgoing direction).
struct GuideInput {
    float px, py, pz, nx, ny, nz, theta01, phi01;
};
__device__ __forceinline__ GuideInput guideInput(float3 p01, float3 n, float3 wo) {
    GuideInput g;
    g.px = p01.x, g.py = p01.y, g.pz = p01.z;
    g.nx = n.x, g.ny = n.y, g.nz = n.z;
    g.theta01 = acosf(fminf(1.0f, fmaxf(-1.0f, wo.z))) * 0.31830988618379067154f;
    float phi = atan2f(wo.y, wo.x) * 0.15915494309189533577f;
    g.phi01 = phi < 0.0f ? phi + 1.0f : phi;
    return g;
}
__device__ __forceinline__ float guideFeature(const GuideInput& g, int k) {
    if (k < 6 * kGFreq) {
        const int f = k / 6, j = k % 6;
        const float p = (j % 3) == 0 ? g.px : ((j % 3) == 1 ? g.py : g.pz);
        const float x = ldexpf(p, f);  // 2^f p, exact
        return j < 3 ? sinpif(x) : cospif(x);
    }
    if (k < 75) return k == 72 ? g.nx : (k == 73 ? g.ny : g.nz);
    if (k < kGFeat) {
        const bool th = k < 75 + kGBlob;
        const int bin = th ? k - 75 : k - 75 - kGBlob;
        const float d = ((th ? g.theta01 : g.phi01) - (bin + 0.5f) * (1.0f / kGBlob)) * kGBlob;
        return __expf(-0.5f * d * d);
    }
    return 0.0f;
}
// B operands of layer 0 (all 9 K tiles): features kt*16 + 4*grp + j of this lane's sample.
// The 72 frequency features sin/cos(2^f pi p) come from one sincospi per coordinate and the
// double-angle recurrence (sin 2a = 2 sin a cos a, cos 2a = (cos a - sin a)(cos a + sin a)); the
// phase error doubles per step, about 2^11 ulp at f = 11, far below BF16 rounding. Every lane
// forms all 72 (the same instructions across the wave) and keeps the ones its slots hold.
__device__ __forceinline__ void guideFeatureTiles(const GuideInput& g, int grp, bf16x4 (&x0)[9]) {
    float v[9][4];
#pragma unroll
    for (int kt = 0; kt < 9; ++kt)
#pragma unroll
        for (int j = 0; j < 4; ++j) v[kt][j] = 0.0f;
    float s[3], c[3];
    sincospif(g.px, &s[0], &c[0]);
    sincospif(g.py, &s[1], &c[1]);
    sincospif(g.pz, &s[2], &c[2]);
#pragma unroll
    for (int f = 0; f < kGFreq; ++f) {
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const int k = 6 * f + j;
            if (((k >> 2) & 3) == grp) v[k >> 4][k & 3] = j < 3 ? s[j] : c[j - 3];
        }
        if (f + 1 < kGFreq)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const float ns = 2.0f * s[i] * c[i];
                c[i] = (c[i] - s[i]) * (c[i] + s[i]);
                s[i] = ns;
            }
    }
    // normal (72..74) and the one-blob encodings (75..138) of this lane's slots
#pragma unroll
    for (int kt = 4; kt < 9; ++kt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = kt * 16 + 4 * grp + j;
            if (k >= 6 * kGFreq) v[kt][j] = guideFeature(g, k);
        }
#pragma unroll
    for (int kt = 0; kt < 9; ++kt)
#pragma unroll
        for (int j = 0; j < 4; ++j) x0[kt][j] = static_cast<short>(toBf16(v[kt][j]));
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

// ---- von Mises-Fisher mixture ---------------------------------------------------------------
__device__ __forceinline__ float vmfPdf(float kappa, float cosang) {
    return kappa / (2.0f * 3.14159265358979f * (-expm1f(-2.0f * kappa))) * __expf(kappa * (cosang - 1.0f));
}

// Mixture record of one sample: lobe[i] = (mu xyz, kappa) for i < 8 at float4 i, weights at 32..39.
__device__ __forceinline__ float mixturePdf(const float* __restrict__ mix, float3 w) {
    float p = 0.0f;
#pragma unroll
    for (int i = 0; i < kGLobes; ++i) {
        const float4 l = reinterpret_cast<const float4*>(mix)[i];
        p += mix[32 + i] * vmfPdf(l.w, l.x * w.x + l.y * w.y + l.z * w.z);
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
    u1 = fminf(fmaxf((u1 - lo) / wsel, 0.0f), 1.0f);
    const float4 l = reinterpret_cast<const float4*>(mix)[sel];
    const float k = l.w;
    const float wz = fmaxf(-1.0f, 1.0f + logf(fmaxf(u1 + (1.0f - u1) * __expf(-2.0f * k), 1e-30f)) / k);
    const float r = sqrtf(fmaxf(0.0f, 1.0f - wz * wz));
    float sp, cp;
    sincosf(2.0f * 3.14159265358979f * u2, &sp, &cp);
    // frame around mu (Duff et al. 2017 orthonormal basis)
    const float sign = copysignf(1.0f, l.z);
    const float a = -1.0f / (sign + l.z), bb = l.x * l.y * a;
    const float3 t = make_float3(1.0f + sign * l.x * l.x * a, sign * bb, -sign * l.x);
    const float3 b = make_float3(bb, sign + l.y * l.y * a, -l.y);
    const float x = r * cp, y = r * sp;
    return make_float3(t.x * x + b.x * y + l.x * wz, t.y * x + b.y * y + l.y * wz, t.z * x + b.z * y + l.z * wz);
}

}  // namespace pg
