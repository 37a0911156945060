// rsx_sdf.h — the learned distance field of the arena mesh and the exact
// distance it is trained on and judged by.
//
// The network is an MLP 3 -> 128 -> 128 -> 128 -> 1 with SiLU activations,
// BF16 weights and FP32 accumulation.  Its input is a point in Unreal units,
// scaled into [-1, 1] per axis over the mesh bounds; its output is the
// unsigned distance to the nearest mesh triangle in units of 100 uu (the four
// static planes stay analytic and are not part of the field).
//
// In the tick it is only a conservative broadphase cull: a body is dropped
// from the mesh queries (ball and car narrowphase, wheel rays) when the
// predicted distance minus a safety margin exceeds the body's reach, the
// radius around its query point inside which a mesh triangle could take part
// in any of those queries.  The exact triangle test stays the fallback.
//
// Layers 2 and 3 run on the matrix cores: one wavefront evaluates 64 points
// (one per lane) with v_mfma_f32_16x16x16_bf16 over activations staged in
// LDS.  Operand layout of that instruction (lane l of 64, j = 0..3):
//   A (16x16, rows x k):  A[l % 16][4 * (l / 16) + j]
//   B (16x16, k x cols):  B[4 * (l / 16) + j][l % 16]
//   D (16x16):            D[4 * (l / 16) + j][l % 16]
// B operands (the weights) are stored pre-swizzled in that lane order, one
// 8-byte load per lane per tile.
#pragma once

#include "rsx_math.h"

namespace rsx {
namespace sdf {

constexpr int kHidden = 128;
constexpr int kTiles = kHidden / 16;                  // 16-wide tiles per layer side
constexpr int kSwz = kTiles * kTiles * 64 * 4;        // bf16 values of one swizzled 128x128 operand
constexpr int kActStride = kHidden + 8;               // LDS row stride (bf16), padded against bank conflicts
constexpr int kLdsActBytes = 64 * kActStride * 2;     // one activation buffer of 64 rows (layers run in place)
constexpr float kOutScale = 100.f;                    // network output unit (uu)

// FP32 parameter vector layout (the master copy the optimiser updates).
constexpr int kOffW1 = 0;                                 // [128][3]
constexpr int kOffB1 = kOffW1 + kHidden * 3;
constexpr int kOffW2 = kOffB1 + kHidden;                  // [out 128][in 128]
constexpr int kOffB2 = kOffW2 + kHidden * kHidden;
constexpr int kOffW3 = kOffB2 + kHidden;
constexpr int kOffB3 = kOffW3 + kHidden * kHidden;
constexpr int kOffW4 = kOffB3 + kHidden;                  // [128]
constexpr int kOffB4 = kOffW4 + kHidden;
constexpr int kNumParams = kOffB4 + 1;                    // 33665

// Mesh bounds used to scale inputs (uu); a point maps to (p - centre) / half.
struct InputScale {
    float cx, cy, cz, hx, hy, hz;
};

// Device view of one network.  w2s / w3s: forward B operands (B[k=in][n=out]
// = W[out][in]), pre-swizzled; w2t / w3t: backward-data B operands
// (B[k=out][n=in] = W[out][in]).
struct Net {
    const float* w1;       // [128][3]
    const float* b1;       // [128]
    const uint16_t* w2s;   // kSwz bf16
    const float* b2;
    const uint16_t* w3s;
    const float* b3;
    const float* w4;       // [128]
    const float* b4;       // [1]
    InputScale in;
};

// Cull configuration handed to the tick kernels.  The network runs every
// `every` ticks; a prediction made k ticks ago still culls when it exceeds
// the body's margin + reach + k * (the most it can move in one tick).
struct Cull {
    Net net;
    int enabled;
    int audit;              // also run the exact check on every culled body
    int every;              // ticks between network evaluations (>= 1)
    float margin_ball_uu;   // safety margins subtracted from the prediction
    float margin_car_uu;
    float reach_ball_uu;    // ball: radius + gather pad + breaking threshold
    float reach_car_uu;     // car: box + wheel rays around the box centre
    float step_ball_uu;     // bound on the ball's motion in one tick
    float step_car_uu;      // bound on a car's motion in one tick
    unsigned long long* stats;  // [0] network queries, [1] culled body-ticks, [2] audit defects
};

RSX_HD float bf16_to_f32(uint16_t h) {
    union {
        uint32_t u;
        float f;
    } v;
    v.u = (uint32_t)h << 16;
    return v.f;
}
// Round to nearest even (finite inputs).
RSX_HD uint16_t f32_to_bf16(float f) {
    union {
        uint32_t u;
        float f;
    } v;
    v.f = f;
    uint32_t r = v.u + 0x7FFFu + ((v.u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}

// SiLU and its derivative.  On the device the exponential and the
// reciprocal use the hardware approximations (v_exp_f32, v_rcp_f32); the
// network is trained and validated with the same functions it is culled
// with, so its safety margin covers them.
#if defined(__HIP_DEVICE_COMPILE__)
RSX_HD float sig(float z) { return __frcp_rn(1.f + __expf(-z)); }
#else
RSX_HD float sig(float z) { return 1.f / (1.f + expf(-z)); }
#endif
RSX_HD float silu(float z) { return z * sig(z); }
RSX_HD float silu_grad(float z) {
    float s = sig(z);
    return s * (1.f + z * (1.f - s));
}

// Index of B operand element (k, n) of a 128x128 matrix in the swizzled
// layout: tile (k / 16, n / 16), lane (n % 16) + 16 * ((k % 16) / 4), slot k % 4.
RSX_HD int swz_index(int k, int n) {
    int kt = k >> 4, nt = n >> 4, kk = k & 15, nn = n & 15;
    int lane = nn + 16 * (kk >> 2);
    return (((kt * kTiles + nt) * 64) + lane) * 4 + (kk & 3);
}

// Squared distance from p to triangle (a, b, c): closest point by Voronoi
// region (Ericson, Real-Time Collision Detection, 5.1.5), in FP32.
RSX_HD float point_tri_dist2(V3 p, V3 a, V3 b, V3 c) {
    V3 ab = b - a, ac = c - a, ap = p - a;
    float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) return len2(p - a);
    V3 bp = p - b;
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.f && d4 <= d3) return len2(p - b);
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
        float v = d1 / (d1 - d3);
        return len2(p - (a + ab * v));
    }
    V3 cp = p - c;
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.f && d5 <= d6) return len2(p - c);
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
        float w = d2 / (d2 - d6);
        return len2(p - (a + ac * w));
    }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return len2(p - (b + (c - b) * w));
    }
    float denom = 1.f / (va + vb + vc);
    float v = vb * denom, w = vc * denom;
    return len2(p - (a + ab * v + ac * w));
}

#if defined(__HIPCC__)
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Orders the LDS accesses of one wavefront: writes issued before it are seen
// by every lane's reads after it.  Synchronises nothing beyond the wavefront,
// so wavefronts of one workgroup may take different branches around it.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One 128 -> 128 layer for the 64 rows in `in` (LDS, bf16, row stride
// kActStride): out = SiLU(in x B + bias) in bf16 (and, if z_out is given,
// the pre-activation for row r, column n at z_out[r * 128 + n]).  out may be
// in: each 16-row tile is read whole (into the matrix operands) before any
// of its rows is written, and no tile reads another tile's rows.
// Every lane of the wavefront must call it.
__device__ inline void mfma_layer(const uint16_t* in, uint16_t* out, const uint16_t* bswz, const float* bias,
                                  float* z_out) {
    const int l = threadIdx.x & 63;
    const int r16 = l & 15, q = l >> 4;
    for (int mt = 0; mt < 4; mt++) {
        v4f acc[kTiles];
        for (int nt = 0; nt < kTiles; nt++) acc[nt] = v4f{0.f, 0.f, 0.f, 0.f};
        for (int kt = 0; kt < kTiles; kt++) {
            const uint16_t* ap = in + (mt * 16 + r16) * kActStride + kt * 16 + 4 * q;
            v4s a = *(const v4s*)ap;
            for (int nt = 0; nt < kTiles; nt++) {
                v4s b = *(const v4s*)(bswz + (((kt * kTiles + nt) * 64) + l) * 4);
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
            }
        }
        wave_sync();
        for (int nt = 0; nt < kTiles; nt++) {
            const int n = nt * 16 + r16;
            const float bn = bias[n];
            for (int j = 0; j < 4; j++) {
                const int row = mt * 16 + 4 * q + j;
                float z = acc[nt][j] + bn;
                if (z_out) z_out[row * kHidden + n] = z;
                out[row * kActStride + n] = f32_to_bf16(silu(z));
            }
        }
    }
}

// Network output (distance, uu) for this lane's point p (uu).  The
// wavefront's 64 lanes each bring one point; lds holds kLdsActBytes bytes of
// this wavefront's own.  Every lane of the wavefront must call it; other
// wavefronts of the workgroup are not waited for.
__device__ inline float eval_wave(const Net& net, V3 p, uint16_t* lds) {
    const int l = threadIdx.x & 63;
    const float x = (p.x - net.in.cx) / net.in.hx, y = (p.y - net.in.cy) / net.in.hy,
                z = (p.z - net.in.cz) / net.in.hz;
    wave_sync();
    for (int o = 0; o < kHidden; o++) {
        float zz = net.w1[o * 3 + 0] * x + net.w1[o * 3 + 1] * y + net.w1[o * 3 + 2] * z + net.b1[o];
        lds[l * kActStride + o] = f32_to_bf16(silu(zz));
    }
    wave_sync();
    mfma_layer(lds, lds, net.w2s, net.b2, nullptr);
    wave_sync();
    mfma_layer(lds, lds, net.w3s, net.b3, nullptr);
    wave_sync();
    float s = *net.b4;
    for (int k = 0; k < kHidden; k++) s += net.w4[k] * bf16_to_f32(lds[l * kActStride + k]);
    wave_sync();
    return s * kOutScale;
}
#endif

}  // namespace sdf
}  // namespace rsx
