// rsx_sdf.hip — the learned cull's GPU side: exact mesh distance, the MLP's
// training kernels (forward with saved activations, backward, Adam) and the
// C-ABI of include/rsx_sdf.h.
//
// Training layout: a batch of n points is processed in rows of 64 (one
// wavefront per 64 points).  Activations a1..a3 are kept in BF16 (the matrix
// core operands), pre-activations z1..z3 in FP32 for the SiLU derivative,
// layer deltas in BF16.  Weight gradients of the 128x128 layers are
// delta^T x activation products on the matrix cores, summed over the batch
// with FP32 atomics; the Adam step updates the FP32 master parameters and
// re-packs the BF16 operand copies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rsx_sdf.h"
#include "rsx_internal.h"

using namespace rsx;
using namespace rsx::sdf;

struct rsx_sdf {
    rsx_engine* eng = nullptr;
    int device = 0;
    float* d_params = nullptr;
    float* d_grads = nullptr;
    float* d_m = nullptr;
    float* d_v = nullptr;
    uint16_t* d_w2s = nullptr;  // forward B operands
    uint16_t* d_w3s = nullptr;
    uint16_t* d_w2t = nullptr;  // backward-data B operands
    uint16_t* d_w3t = nullptr;
    float* d_tris = nullptr;    // uu
    int ntri = 0;
    std::vector<float> h_tris;
    InputScale in{};
    int adam_t = 0;
    int cap = 0;                // training scratch capacity (points)
    float *d_z1 = nullptr, *d_z2 = nullptr, *d_z3 = nullptr;
    uint16_t *d_a1 = nullptr, *d_a2 = nullptr, *d_a3 = nullptr;
    uint16_t *d_dl1 = nullptr, *d_dl2 = nullptr, *d_dl3 = nullptr;
    float* d_loss = nullptr;
    unsigned long long* d_stats = nullptr;
    float reach_ball = 0.f, reach_car = 0.f;
    bool attached = false;
};

namespace {

#define SDF_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess)                                                                          \
            return record_failure(RSX_HIP_ERROR, (std::string(#expr) + ": " + hipGetErrorString(_e)).c_str()); \
    } while (0)

Net net_view(const rsx_sdf* s) {
    Net n;
    n.w1 = s->d_params + kOffW1;
    n.b1 = s->d_params + kOffB1;
    n.w2s = s->d_w2s;
    n.b2 = s->d_params + kOffB2;
    n.w3s = s->d_w3s;
    n.b3 = s->d_params + kOffB3;
    n.w4 = s->d_params + kOffW4;
    n.b4 = s->d_params + kOffB4;
    n.in = s->in;
    return n;
}

// ---- exact distance: every point against every triangle, tiles in LDS ----
constexpr int kTriTile = 256;

__global__ void __launch_bounds__(256) k_exact(const float* tris, int ntri, const float* pts, int n, float* out) {
    __shared__ float st[kTriTile * 9];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    V3 p = live ? V3{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]} : V3{0, 0, 0};
    float best = 3.0e38f;
    for (int base = 0; base < ntri; base += kTriTile) {
        const int cnt = min(kTriTile, ntri - base);
        for (int k = threadIdx.x; k < cnt * 9; k += blockDim.x) st[k] = tris[(size_t)base * 9 + k];
        __syncthreads();
        for (int t = 0; t < cnt; t++) {
            const float* v = st + t * 9;
            float d2 = point_tri_dist2(p, V3{v[0], v[1], v[2]}, V3{v[3], v[4], v[5]}, V3{v[6], v[7], v[8]});
            best = fminf(best, d2);
        }
        __syncthreads();
    }
    if (live) out[i] = sqrtf(best);
}

// ---- inference for arbitrary points (the tick's code path) ----
__global__ void __launch_bounds__(64) k_eval(const float* pts, int n, Net net, float* out) {
    __shared__ __align__(16) uint16_t lds[kLdsActBytes / 2];
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = i < n ? i : n - 1;
    V3 p = V3{pts[3 * j], pts[3 * j + 1], pts[3 * j + 2]};
    float y = eval_wave(net, p, lds);
    if (i < n) out[i] = y;
}

// ---- training forward ----
__global__ void k_fwd1(const float* pts, int n, const float* P, InputScale in, float* z1, uint16_t* a1) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * kHidden) return;
    const int p = idx / kHidden, o = idx % kHidden;
    const float x = (pts[3 * p] - in.cx) / in.hx, y = (pts[3 * p + 1] - in.cy) / in.hy,
                z = (pts[3 * p + 2] - in.cz) / in.hz;
    const float* w1 = P + kOffW1;
    float zz = w1[o * 3 + 0] * x + w1[o * 3 + 1] * y + w1[o * 3 + 2] * z + P[kOffB1 + o];
    z1[idx] = zz;
    a1[idx] = f32_to_bf16(silu(zz));
}

// Loads 64 rows of a [n][128] bf16 matrix into LDS (row stride kActStride),
// zero past n.
__device__ void load_rows(const uint16_t* src, int n, int base, uint16_t* dst) {
    for (int k = threadIdx.x; k < 64 * kHidden; k += blockDim.x) {
        const int r = k / kHidden, c = k % kHidden;
        dst[r * kActStride + c] = (base + r < n) ? src[(size_t)(base + r) * kHidden + c] : (uint16_t)0;
    }
}

// mode 0 (forward): zout = acc + bias, out = bf16(SiLU(zout));
// mode 1 (backward data): out = bf16(acc * SiLU'(zin)).
template <int MODE>
__global__ void __launch_bounds__(64) k_layer(const uint16_t* src, int n, const uint16_t* bswz, const float* bias,
                                             const float* zin, float* zout, uint16_t* out) {
    __shared__ __align__(16) uint16_t lds[64 * kActStride];
    const int base = blockIdx.x * 64;
    load_rows(src, n, base, lds);
    __syncthreads();
    const int l = threadIdx.x;
    const int r16 = l & 15, q = l >> 4;
    for (int mt = 0; mt < 4; mt++) {
        v4f acc[kTiles];
        for (int nt = 0; nt < kTiles; nt++) acc[nt] = v4f{0.f, 0.f, 0.f, 0.f};
        for (int kt = 0; kt < kTiles; kt++) {
            v4s a = *(const v4s*)(lds + (mt * 16 + r16) * kActStride + kt * 16 + 4 * q);
            for (int nt = 0; nt < kTiles; nt++) {
                v4s b = *(const v4s*)(bswz + (((kt * kTiles + nt) * 64) + l) * 4);
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
            }
        }
        for (int nt = 0; nt < kTiles; nt++) {
            const int c = nt * 16 + r16;
            for (int j = 0; j < 4; j++) {
                const int row = base + mt * 16 + 4 * q + j;
                if (row >= n) continue;
                const size_t o = (size_t)row * kHidden + c;
                if (MODE == 0) {
                    float z = acc[nt][j] + bias[c];
                    zout[o] = z;
                    out[o] = f32_to_bf16(silu(z));
                } else {
                    out[o] = f32_to_bf16(acc[nt][j] * silu_grad(zin[o]));
                }
            }
        }
    }
}

// Output layer, L1 loss and the first backward step, 64 points per block.
__global__ void __launch_bounds__(64) k_head(const uint16_t* a3, const float* z3, int n, const float* P,
                                            const float* labels, float clamp_uu, float over_w, float inv_n, float* G,
                                            float* loss, uint16_t* d3) {
    const int p = blockIdx.x * 64 + threadIdx.x;
    const bool live = p < n;
    const int pp = live ? p : n - 1;
    const float* w4 = P + kOffW4;
    float y = P[kOffB4];
    for (int k = 0; k < kHidden; k++) y += w4[k] * bf16_to_f32(a3[(size_t)pp * kHidden + k]);
    const float t = fminf(labels[pp], clamp_uu) / kOutScale;
    const float e = y - t;
    // over-estimates weigh over_w: the cull's margin is set by them
    const float w = e > 0.f ? over_w : 1.f;
    const float gy = live ? ((e > 0.f ? 1.f : (e < 0.f ? -1.f : 0.f)) * w * inv_n) : 0.f;
    float lv = live ? fabsf(e) * w * inv_n : 0.f;
    for (int o = 32; o > 0; o >>= 1) lv += __shfl_xor(lv, o);
    if (threadIdx.x == 0) atomicAdd(loss, lv);
    float gb = gy;
    for (int o = 32; o > 0; o >>= 1) gb += __shfl_xor(gb, o);
    if (threadIdx.x == 0) atomicAdd(&G[kOffB4], gb);
    for (int k = 0; k < kHidden; k++) {
        const size_t o = (size_t)pp * kHidden + k;
        float g = gy * bf16_to_f32(a3[o]);
        for (int s = 32; s > 0; s >>= 1) g += __shfl_xor(g, s);
        if (threadIdx.x == 0) atomicAdd(&G[kOffW4 + k], g);
        if (live) d3[o] = f32_to_bf16(gy * w4[k] * silu_grad(z3[o]));
    }
}

// Weight gradients of a 128x128 layer: gW[o][i] += sum_p d[p][o] a[p][i],
// gb[o] += sum_p d[p][o].  Block (chunk of 64 points, quarter of the rows).
__global__ void __launch_bounds__(64) k_wgrad(const uint16_t* d, const uint16_t* a, int n, float* gW, float* gb) {
    __shared__ __align__(16) uint16_t sd[64 * kActStride];
    __shared__ __align__(16) uint16_t sa[64 * kActStride];
    const int base = blockIdx.x * 64, qo = blockIdx.y;
    load_rows(d, n, base, sd);
    load_rows(a, n, base, sa);
    __syncthreads();
    const int l = threadIdx.x;
    const int r16 = l & 15, q = l >> 4;
    for (int mt = 0; mt < 2; mt++) {
        const int orow = qo * 32 + mt * 16;
        v4f acc[kTiles];
        for (int nt = 0; nt < kTiles; nt++) acc[nt] = v4f{0.f, 0.f, 0.f, 0.f};
        for (int kt = 0; kt < 4; kt++) {
            v4s av;
            for (int j = 0; j < 4; j++) av[j] = (short)sd[(kt * 16 + 4 * q + j) * kActStride + orow + r16];
            for (int nt = 0; nt < kTiles; nt++) {
                v4s bv;
                for (int j = 0; j < 4; j++) bv[j] = (short)sa[(kt * 16 + 4 * q + j) * kActStride + nt * 16 + r16];
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[nt], 0, 0, 0);
            }
        }
        for (int nt = 0; nt < kTiles; nt++)
            for (int j = 0; j < 4; j++) atomicAdd(&gW[(orow + 4 * q + j) * kHidden + nt * 16 + r16], acc[nt][j]);
    }
    if (qo == 0) {
        for (int o = threadIdx.x; o < kHidden; o += 64) {
            float s = 0.f;
            for (int r = 0; r < 64; r++) s += bf16_to_f32(sd[r * kActStride + o]);
            atomicAdd(&gb[o], s);
        }
    }
}

// First layer gradients: gW1[o][j] += sum_p d1[p][o] x_p[j], gb1[o].
__global__ void __launch_bounds__(128) k_wgrad1(const uint16_t* d1, const float* pts, int n, InputScale in, float* G) {
    const int o = threadIdx.x;
    const int base = blockIdx.x * 64;
    float g0 = 0.f, g1 = 0.f, g2 = 0.f, gb = 0.f;
    for (int r = 0; r < 64 && base + r < n; r++) {
        const int p = base + r;
        const float dv = bf16_to_f32(d1[(size_t)p * kHidden + o]);
        const float x = (pts[3 * p] - in.cx) / in.hx, y = (pts[3 * p + 1] - in.cy) / in.hy,
                    z = (pts[3 * p + 2] - in.cz) / in.hz;
        g0 += dv * x;
        g1 += dv * y;
        g2 += dv * z;
        gb += dv;
    }
    atomicAdd(&G[kOffW1 + o * 3 + 0], g0);
    atomicAdd(&G[kOffW1 + o * 3 + 1], g1);
    atomicAdd(&G[kOffW1 + o * 3 + 2], g2);
    atomicAdd(&G[kOffB1 + o], gb);
}

__global__ void k_adam(float* P, float* G, float* M, float* V, int np, float lr, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float g = G[i];
    const float m = b1 * M[i] + (1.f - b1) * g;
    const float v = b2 * V[i] + (1.f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    P[i] -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
    G[i] = 0.f;
}

__global__ void k_zero(float* x, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.f;
}

// BF16 operand copies of W2 and W3 in both swizzled layouts.
__global__ void k_pack(const float* P, uint16_t* w2s, uint16_t* w3s, uint16_t* w2t, uint16_t* w3t) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= kHidden * kHidden) return;
    const int o = idx / kHidden, i = idx % kHidden;
    const uint16_t v2 = f32_to_bf16(P[kOffW2 + idx]), v3 = f32_to_bf16(P[kOffW3 + idx]);
    w2s[swz_index(i, o)] = v2;
    w3s[swz_index(i, o)] = v3;
    w2t[swz_index(o, i)] = v2;
    w3t[swz_index(o, i)] = v3;
}

// splitmix64 stream for the weight initialisation
uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
float uniform(uint64_t& s, float lo, float hi) {
    return lo + (hi - lo) * (float)((splitmix(s) >> 40) * (1.0 / 16777216.0));
}

rsx_status ensure_scratch(rsx_sdf* s, int n) {
    if (n <= s->cap) return RSX_OK;
    float** fz[] = {&s->d_z1, &s->d_z2, &s->d_z3};
    uint16_t** ha[] = {&s->d_a1, &s->d_a2, &s->d_a3, &s->d_dl1, &s->d_dl2, &s->d_dl3};
    for (auto p : fz) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    for (auto p : ha) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    const size_t m = (size_t)n * kHidden;
    for (auto p : fz) SDF_TRY(hipMalloc((void**)p, m * sizeof(float)));
    for (auto p : ha) SDF_TRY(hipMalloc((void**)p, m * sizeof(uint16_t)));
    s->cap = n;
    return RSX_OK;
}

// Forward with saved activations, loss and backward into d_grads (not zeroed
// here).  The batch loss is added to d_loss.
rsx_status forward_backward(rsx_sdf* s, const float* pts, const float* labels, int n, float clamp_uu,
                            float over_w, float* loss, hipStream_t st) {
    rsx_status r = ensure_scratch(s, n);
    if (r != RSX_OK) return r;
    const int rows = (n + 63) / 64;
    const float* P = s->d_params;
    float* G = s->d_grads;
    hipLaunchKernelGGL(k_fwd1, dim3((n * kHidden + 255) / 256), dim3(256), 0, st, pts, n, P, s->in, s->d_z1, s->d_a1);
    hipLaunchKernelGGL(k_layer<0>, dim3(rows), dim3(64), 0, st, s->d_a1, n, s->d_w2s, P + kOffB2, nullptr, s->d_z2,
                       s->d_a2);
    hipLaunchKernelGGL(k_layer<0>, dim3(rows), dim3(64), 0, st, s->d_a2, n, s->d_w3s, P + kOffB3, nullptr, s->d_z3,
                       s->d_a3);
    hipLaunchKernelGGL(k_head, dim3(rows), dim3(64), 0, st, s->d_a3, s->d_z3, n, P, labels, clamp_uu, over_w, 1.f / n,
                       G, loss, s->d_dl3);
    hipLaunchKernelGGL(k_wgrad, dim3(rows, 4), dim3(64), 0, st, s->d_dl3, s->d_a2, n, G + kOffW3, G + kOffB3);
    hipLaunchKernelGGL(k_layer<1>, dim3(rows), dim3(64), 0, st, s->d_dl3, n, s->d_w3t, nullptr, s->d_z2, nullptr,
                       s->d_dl2);
    hipLaunchKernelGGL(k_wgrad, dim3(rows, 4), dim3(64), 0, st, s->d_dl2, s->d_a1, n, G + kOffW2, G + kOffB2);
    hipLaunchKernelGGL(k_layer<1>, dim3(rows), dim3(64), 0, st, s->d_dl2, n, s->d_w2t, nullptr, s->d_z1, nullptr,
                       s->d_dl1);
    hipLaunchKernelGGL(k_wgrad1, dim3(rows), dim3(128), 0, st, s->d_dl1, pts, n, s->in, G);
    SDF_TRY(hipGetLastError());
    return RSX_OK;
}

rsx_status repack(rsx_sdf* s, hipStream_t st) {
    hipLaunchKernelGGL(k_pack, dim3((kHidden * kHidden + 255) / 256), dim3(256), 0, st, s->d_params, s->d_w2s,
                       s->d_w3s, s->d_w2t, s->d_w3t);
    SDF_TRY(hipGetLastError());
    return RSX_OK;
}

}  // namespace

extern "C" {

int32_t rsx_sdf_num_params(void) { return kNumParams; }

rsx_status rsx_sdf_create(rsx_engine* e, uint64_t seed, rsx_sdf** out) {
    if (!e || !out) return record_failure(RSX_INVALID_ARG, "null engine or output pointer");
    rsx_sdf* s = new rsx_sdf();
    s->eng = e;
    s->device = engine_device(e);
    const HostWorld& hw = engine_host_world(e);
    s->ntri = (int)hw.tri_mesh.size();
    s->h_tris.resize((size_t)s->ntri * 9);
    float mn[3] = {3e38f, 3e38f, 3e38f}, mx[3] = {-3e38f, -3e38f, -3e38f};
    for (size_t k = 0; k < s->h_tris.size(); k++) {
        float v = hw.tri[k] * kBT2UU;
        s->h_tris[k] = v;
        mn[k % 3] = std::min(mn[k % 3], v);
        mx[k % 3] = std::max(mx[k % 3], v);
    }
    s->in = InputScale{0.5f * (mn[0] + mx[0]), 0.5f * (mn[1] + mx[1]), 0.5f * (mn[2] + mx[2]),
                       0.5f * (mx[0] - mn[0]), 0.5f * (mx[1] - mn[1]), 0.5f * (mx[2] - mn[2])};
    // reaches (uu): the ball's sphere-triangle query pads its radius by 0.08
    // and accepts contacts within the breaking threshold; a car's box (with
    // margin) reaches its half diagonal plus the threshold around the box
    // centre, and each wheel ray runs from its hardpoint down the full ray
    // length.
    const BallGeom& bg = hw.bg;
    const CarGeom& cg = hw.cg;
    s->reach_ball = (bg.radius + 0.08f + bg.breaking) * kBT2UU + 1.f;
    V3 h = cg.box_implicit + V3{cg.box_margin, cg.box_margin, cg.box_margin};
    float box_reach = len(h) + cg.breaking;
    float ray_reach = 0.f;
    for (int i = 0; i < 4; i++) {
        float real_len = cg.rest_len[i] + kMaxSuspensionTravel * kUU2BT + cg.wheel_radius[i] - kSuspensionSubtraction;
        ray_reach = std::max(ray_reach, len(cg.wheel_cs[i] - cg.box_offset) + real_len);
    }
    s->reach_car = std::max(box_reach, ray_reach) * kBT2UU + 1.f;

    rsx_status r = RSX_OK;
    auto bail = [&](rsx_status st) {
        rsx_sdf_destroy(s);
        return st;
    };
    if (hipSetDevice(s->device) != hipSuccess) return bail(record_failure(RSX_HIP_ERROR, "hipSetDevice failed"));
    float** fp[] = {&s->d_params, &s->d_grads, &s->d_m, &s->d_v};
    for (auto p : fp)
        if (hipMalloc((void**)p, kNumParams * sizeof(float)) != hipSuccess)
            return bail(record_failure(RSX_OOM, "parameter buffers"));
    uint16_t** hp[] = {&s->d_w2s, &s->d_w3s, &s->d_w2t, &s->d_w3t};
    for (auto p : hp)
        if (hipMalloc((void**)p, kSwz * sizeof(uint16_t)) != hipSuccess)
            return bail(record_failure(RSX_OOM, "operand buffers"));
    if (hipMalloc((void**)&s->d_tris, std::max<size_t>(1, s->h_tris.size()) * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&s->d_loss, sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&s->d_stats, 4 * sizeof(unsigned long long)) != hipSuccess)
        return bail(record_failure(RSX_OOM, "mesh / loss / stats buffers"));
    if (hipMemcpy(s->d_tris, s->h_tris.data(), s->h_tris.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(s->d_stats, 0, 4 * sizeof(unsigned long long)) != hipSuccess)
        return bail(record_failure(RSX_HIP_ERROR, "upload"));
    // initial weights: U(-1/sqrt(fan_in), 1/sqrt(fan_in)); the output bias
    // starts at 5 (500 uu), the middle of the clamped label range
    std::vector<float> P(kNumParams);
    uint64_t st = seed ^ 0x5DF5DF5DF5ull;
    const float a1 = 1.f / std::sqrt(3.f), ah = 1.f / std::sqrt((float)kHidden);
    for (int i = 0; i < kHidden * 3; i++) P[kOffW1 + i] = uniform(st, -a1, a1);
    for (int i = 0; i < kHidden; i++) P[kOffB1 + i] = uniform(st, -a1, a1);
    for (int i = 0; i < kHidden * kHidden; i++) P[kOffW2 + i] = uniform(st, -ah, ah);
    for (int i = 0; i < kHidden; i++) P[kOffB2 + i] = uniform(st, -ah, ah);
    for (int i = 0; i < kHidden * kHidden; i++) P[kOffW3 + i] = uniform(st, -ah, ah);
    for (int i = 0; i < kHidden; i++) P[kOffB3 + i] = uniform(st, -ah, ah);
    for (int i = 0; i < kHidden; i++) P[kOffW4 + i] = uniform(st, -ah, ah);
    P[kOffB4] = 5.f;
    r = rsx_sdf_set_params(s, P.data());
    if (r != RSX_OK) return bail(r);
    *out = s;
    return RSX_OK;
}

rsx_status rsx_sdf_destroy(rsx_sdf* s) {
    if (!s) return RSX_OK;
    (void)hipSetDevice(s->device);
    if (s->attached && s->eng) {
        Cull off{};
        engine_set_cull(s->eng, off);
    }
    void* ptrs[] = {s->d_params, s->d_grads, s->d_m,  s->d_v,   s->d_w2s, s->d_w3s, s->d_w2t, s->d_w3t,
                    s->d_tris,   s->d_z1,    s->d_z2, s->d_z3,  s->d_a1,  s->d_a2,  s->d_a3,  s->d_dl1,
                    s->d_dl2,    s->d_dl3,   s->d_loss, s->d_stats};
    for (void* p : ptrs) (void)hipFree(p);
    delete s;
    return RSX_OK;
}

rsx_status rsx_sdf_mesh_triangles(rsx_sdf* s, float* out, int64_t cap, int32_t* n) {
    if (!s || !n) return record_failure(RSX_INVALID_ARG, "null argument");
    *n = s->ntri;
    if (out) {
        if (cap < (int64_t)s->h_tris.size()) return record_failure(RSX_INVALID_ARG, "output too small");
        memcpy(out, s->h_tris.data(), s->h_tris.size() * sizeof(float));
    }
    return RSX_OK;
}

rsx_status rsx_sdf_exact_distance(rsx_sdf* s, const float* d_points, int32_t n, float* d_dist, void* stream) {
    if (!s || (!d_points && n > 0) || (!d_dist && n > 0) || n < 0) return record_failure(RSX_INVALID_ARG, "bad arguments");
    if (n == 0) return RSX_OK;
    SDF_TRY(hipSetDevice(s->device));
    hipLaunchKernelGGL(k_exact, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, s->d_tris, s->ntri, d_points,
                       n, d_dist);
    SDF_TRY(hipGetLastError());
    return RSX_OK;
}

rsx_status rsx_sdf_eval(rsx_sdf* s, const float* d_points, int32_t n, float* d_dist, void* stream) {
    if (!s || (!d_points && n > 0) || (!d_dist && n > 0) || n < 0) return record_failure(RSX_INVALID_ARG, "bad arguments");
    if (n == 0) return RSX_OK;
    SDF_TRY(hipSetDevice(s->device));
    hipLaunchKernelGGL(k_eval, dim3((n + 63) / 64), dim3(64), 0, (hipStream_t)stream, d_points, n, net_view(s), d_dist);
    SDF_TRY(hipGetLastError());
    return RSX_OK;
}

rsx_status rsx_sdf_train_step(rsx_sdf* s, const float* d_points, const float* d_labels, int32_t n, float lr,
                              float clamp_uu, float over_weight, float* d_loss, void* stream) {
    if (!s || !d_points || !d_labels || n <= 0 || !(lr > 0.f) || !(clamp_uu > 0.f) || !(over_weight > 0.f))
        return record_failure(RSX_INVALID_ARG, "bad arguments");
    SDF_TRY(hipSetDevice(s->device));
    hipStream_t st = (hipStream_t)stream;
    float* loss = d_loss ? d_loss : s->d_loss;
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, loss, 1);
    rsx_status r = forward_backward(s, d_points, d_labels, n, clamp_uu, over_weight, loss, st);
    if (r != RSX_OK) return r;
    s->adam_t++;
    const float bc1 = 1.f - std::pow(0.9f, (float)s->adam_t), bc2 = 1.f - std::pow(0.999f, (float)s->adam_t);
    hipLaunchKernelGGL(k_adam, dim3((kNumParams + 255) / 256), dim3(256), 0, st, s->d_params, s->d_grads, s->d_m,
                       s->d_v, kNumParams, lr, bc1, bc2);
    return repack(s, st);
}

rsx_status rsx_sdf_gradients(rsx_sdf* s, const float* d_points, const float* d_labels, int32_t n, float clamp_uu,
                             float over_weight, float* h_grads, float* h_loss) {
    if (!s || !d_points || !d_labels || n <= 0 || !h_grads || !(clamp_uu > 0.f) || !(over_weight > 0.f))
        return record_failure(RSX_INVALID_ARG, "bad arguments");
    SDF_TRY(hipSetDevice(s->device));
    hipLaunchKernelGGL(k_zero, dim3((kNumParams + 255) / 256), dim3(256), 0, nullptr, s->d_grads, kNumParams);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, nullptr, s->d_loss, 1);
    rsx_status r = forward_backward(s, d_points, d_labels, n, clamp_uu, over_weight, s->d_loss, nullptr);
    if (r != RSX_OK) return r;
    SDF_TRY(hipMemcpy(h_grads, s->d_grads, kNumParams * sizeof(float), hipMemcpyDeviceToHost));
    if (h_loss) SDF_TRY(hipMemcpy(h_loss, s->d_loss, sizeof(float), hipMemcpyDeviceToHost));
    hipLaunchKernelGGL(k_zero, dim3((kNumParams + 255) / 256), dim3(256), 0, nullptr, s->d_grads, kNumParams);
    SDF_TRY(hipDeviceSynchronize());
    return RSX_OK;
}

rsx_status rsx_sdf_get_params(rsx_sdf* s, float* h_params) {
    if (!s || !h_params) return record_failure(RSX_INVALID_ARG, "null argument");
    SDF_TRY(hipSetDevice(s->device));
    SDF_TRY(hipMemcpy(h_params, s->d_params, kNumParams * sizeof(float), hipMemcpyDeviceToHost));
    return RSX_OK;
}

rsx_status rsx_sdf_set_params(rsx_sdf* s, const float* h_params) {
    if (!s || !h_params) return record_failure(RSX_INVALID_ARG, "null argument");
    SDF_TRY(hipSetDevice(s->device));
    SDF_TRY(hipMemcpy(s->d_params, h_params, kNumParams * sizeof(float), hipMemcpyHostToDevice));
    SDF_TRY(hipMemset(s->d_grads, 0, kNumParams * sizeof(float)));
    SDF_TRY(hipMemset(s->d_m, 0, kNumParams * sizeof(float)));
    SDF_TRY(hipMemset(s->d_v, 0, kNumParams * sizeof(float)));
    s->adam_t = 0;
    rsx_status r = repack(s, nullptr);
    if (r != RSX_OK) return r;
    SDF_TRY(hipDeviceSynchronize());
    return RSX_OK;
}

rsx_status rsx_attach_cull(rsx_engine* e, rsx_sdf* s, float margin_ball_uu, float margin_car_uu, int32_t every_ticks,
                           int32_t audit) {
    if (!e) return record_failure(RSX_INVALID_ARG, "null engine");
    Cull c{};
    if (s) {
        if (s->eng != e) return record_failure(RSX_INVALID_ARG, "the network was created for another engine");
        if (!(margin_ball_uu >= 0.f) || !(margin_car_uu >= 0.f))
            return record_failure(RSX_INVALID_ARG, "margins must be >= 0");
        if (every_ticks < 1) return record_failure(RSX_INVALID_ARG, "every_ticks must be >= 1");
        c.net = net_view(s);
        c.enabled = 1;
        c.audit = audit ? 1 : 0;
        c.every = every_ticks;
        c.margin_ball_uu = margin_ball_uu;
        c.margin_car_uu = margin_car_uu;
        c.reach_ball_uu = s->reach_ball;
        c.reach_car_uu = s->reach_car;
        // motion bounds per tick: the ball's speed is clamped to 6000 uu/s
        // (50 uu per tick) after each tick and a car's to 2300 uu/s; contact
        // impulses can exceed the clamp inside the tick that integrates them,
        // so twice the ball's and more than twice the car's figure
        c.step_ball_uu = 100.f;
        c.step_car_uu = 50.f;
        c.stats = s->d_stats;
        s->attached = true;
    }
    engine_set_cull(e, c);
    return RSX_OK;
}

rsx_status rsx_get_cull_stats(rsx_sdf* s, rsx_cull_stats* out) {
    if (!s || !out) return record_failure(RSX_INVALID_ARG, "null argument");
    unsigned long long h[4];
    SDF_TRY(hipSetDevice(s->device));
    SDF_TRY(hipMemcpy(h, s->d_stats, sizeof(h), hipMemcpyDeviceToHost));
    out->queries = h[0];
    out->culled = h[1];
    out->defects = h[2];
    return RSX_OK;
}

rsx_status rsx_reset_cull_stats(rsx_sdf* s) {
    if (!s) return record_failure(RSX_INVALID_ARG, "null argument");
    SDF_TRY(hipSetDevice(s->device));
    SDF_TRY(hipMemset(s->d_stats, 0, 4 * sizeof(unsigned long long)));
    return RSX_OK;
}

rsx_status rsx_cull_reach(rsx_sdf* s, float* ball_uu, float* car_uu) {
    if (!s) return record_failure(RSX_INVALID_ARG, "null argument");
    if (ball_uu) *ball_uu = s->reach_ball;
    if (car_uu) *car_uu = s->reach_car;
    return RSX_OK;
}

}  // extern "C"
