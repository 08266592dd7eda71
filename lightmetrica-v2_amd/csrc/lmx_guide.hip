// lmx_guide.hip -- the learned directional sampler on gfx950.
//
// No counterpart in the reference: this is the third sampler beside the
// plain (renderer_pt.cpp) and bidirectional ones.  At every diffuse or glossy
// vertex a small MLP maps (position, shading normal, direction to the
// previous vertex) to a mixture of 8 von Mises-Fisher lobes and a BSDF
// selection probability alpha; the next direction is drawn from
// alpha BSDF + (1 - alpha) mixture (one-sample MIS), so the estimator stays
// unbiased whatever the network says.  The network is trained online on the
// radiance the path tracer itself carries back along its paths.
//
// Network: 7 inputs in [0, 1] -> 12-frequency sin/cos encoding (168) ->
// 4 x (64, ReLU) -> 41 outputs.  BF16 weights and activations, FP32
// accumulation, on v_mfma_f32_16x16x32_bf16:
//   A (16 x 32)  lane l: row l%16, k = 8 (l/16) .. +7
//   B (32 x 16)  lane l: col l%16, k = 8 (l/16) .. +7
//   D (16 x 16)  lane l: col l%16, rows 4 (l/16) .. +3
// Layers are computed transposed (H_out^T = W^T H_in^T, queries as columns),
// so the D tiles of one layer are, lane for lane, the B operand of the next:
// the two D tiles of units [32c, 32c+16) and [32c+16, 32c+32) give lane group
// g the units 32c + 4g + {0..3} and 32c + 16 + 4g + {0..3}, and the next
// layer's weights are packed with k slot 8g + j bound to exactly those units.
// The encoding is computed in registers directly in operand layout (group g
// holds input dimensions g and g + 4, 4 frequencies per 32-slot chunk), so
// no activation goes through LDS in the forward pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "lmx.h"
#include "lmx_internal.h"
#include "rmath.h"
#include "rshade.h"
#include "wbvh.h"

using namespace lmx;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define MFMA16(a, b, c) __builtin_amdgcn_mfma_f32_16x16x32_bf16((a), (b), (c), 0, 0, 0)

namespace {

// canonical parameter layout (W[k][n]: input k -> output n)
constexpr int GP_W1 = 0;
constexpr int GP_W2 = GP_W1 + 169 * 64;
constexpr int GP_B2 = GP_W2 + 64 * 64;
constexpr int GP_W3 = GP_B2 + 64;
constexpr int GP_B3 = GP_W3 + 64 * 64;
constexpr int GP_W4 = GP_B3 + 64;
constexpr int GP_B4 = GP_W4 + 64 * 64;
constexpr int GP_W5 = GP_B4 + 64;
constexpr int GP_B5 = GP_W5 + 64 * 41;
constexpr int GP_N = GP_B5 + 41;
static_assert(GP_N == LMX_GUIDE_NPARAM, "parameter layout");

// packed operand records (16 B = 8 bf16 = one lane's operand of one MFMA)
constexpr int N_A1 = 4 * 6 * 64;      // forward layer 1: [nt 4][kc 6][lane]
constexpr int N_AH = 4 * 2 * 64;      // forward layers 2-4: [nt 4][kc 2][lane]
constexpr int N_A5 = 3 * 2 * 64;      // forward output: [nt 3][kc 2][lane]
constexpr int OFF_A1 = 0;
constexpr int OFF_A2 = OFF_A1 + N_A1;
constexpr int OFF_A5 = OFF_A2 + 3 * N_AH;
constexpr int N_FWD = OFF_A5 + N_A5;  // 3456 records = 55,296 B
constexpr int OFF_B5 = N_FWD;         // backward output layer: W5 as [hidden][output]
constexpr int OFF_B4 = OFF_B5 + N_AH; // backward layers 4, 3, 2: W as [input][output]
constexpr int N_PACK = OFF_B4 + 3 * N_AH;
constexpr int N_BIAS = 3 * 64 + 48;   // b2, b3, b4, b5 (zero-padded to 48)
constexpr int YS = 41;                // LDS row stride of the 41 outputs (9 banks apart: conflict-free rows)
constexpr int XS = 7;                 // LDS row stride of the 7 inputs

constexpr float LOGK_MIN = -4.0f, LOGK_MAX = 9.0f;
// BSDF selection probability alpha in [0.5, 0.9]: a direction the BSDF draws
// is weighted at most 1 / alpha = 2 times its plain weight, so a vertex where
// the lobes miss the integrand cannot start a firefly; lower bounds (0.1,
// 0.25) measured more variance than they saved
constexpr float ALPHA_MIN = 0.5f, ALPHA_SPAN = 0.4f;

// hidden unit bound to k slot 8g + j of 32-unit chunk kc
__host__ __device__ inline int unit_of_slot(int kc, int g, int j)
{
    return 32 * kc + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}

// layer-1 slot -> canonical W1 row: feature d*24 + 2f (+1 cos), 168 = the
// constant-1 input carrying the bias, -1 = zero padding
__host__ __device__ inline int feat_of_slot(int kc, int g, int j)
{
    const int L = kc * 8 + j;
    const int d = L < 24 ? g : g + 4;
    if (d < 7) return d * 24 + (L % 24);
    return L == 24 ? 168 : -1;
}

struct GuideDev {
    const bf16x8* pack;        // N_PACK records
    const float* bias;         // N_BIAS
    f3 lo, inv_ext;            // position normalisation
};

// ---------------------------------------------------------------------------
// packing and optimiser

__global__ void k_pack(const float* __restrict__ P, __bf16* __restrict__ out, float* __restrict__ bias)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N_BIAS) {
        float b;
        if (t < 64) b = P[GP_B2 + t];
        else if (t < 128) b = P[GP_B3 + t - 64];
        else if (t < 192) b = P[GP_B4 + t - 128];
        else b = (t - 192) < 41 ? P[GP_B5 + t - 192] : 0.0f;
        bias[t] = b;
    }
    if (t >= N_PACK * 8) return;
    const int rec = t >> 3, j = t & 7;
    float v = 0.0f;
    if (rec < OFF_A2) {                                   // W1^T
        const int lane = rec & 63, kc = (rec >> 6) % 6, nt = (rec >> 6) / 6;
        const int row = feat_of_slot(kc, lane >> 4, j);
        if (row >= 0) v = P[GP_W1 + row * 64 + nt * 16 + (lane & 15)];
    } else if (rec < OFF_A5) {                            // W2..W4 ^T
        const int r = rec - OFF_A2, layer = r / N_AH, q = r % N_AH;
        const int lane = q & 63, kc = (q >> 6) & 1, nt = q >> 7;
        const int W = layer == 0 ? GP_W2 : (layer == 1 ? GP_W3 : GP_W4);
        v = P[W + unit_of_slot(kc, lane >> 4, j) * 64 + nt * 16 + (lane & 15)];
    } else if (rec < OFF_B5) {                            // W5^T
        const int q = rec - OFF_A5, lane = q & 63, kc = (q >> 6) & 1, nt = q >> 7;
        const int n = nt * 16 + (lane & 15);
        if (n < 41) v = P[GP_W5 + unit_of_slot(kc, lane >> 4, j) * 41 + n];
    } else if (rec < OFF_B4) {                            // W5 as [hidden][output], natural output order
        const int q = rec - OFF_B5, lane = q & 63, kc = (q >> 6) & 1, nt = q >> 7;
        const int o = 32 * kc + 8 * (lane >> 4) + j;
        if (o < 41) v = P[GP_W5 + (nt * 16 + (lane & 15)) * 41 + o];
    } else {                                              // W4, W3, W2 as [input][output]
        const int r = rec - OFF_B4, layer = r / N_AH, q = r % N_AH;
        const int lane = q & 63, kc = (q >> 6) & 1, nt = q >> 7;
        const int W = layer == 0 ? GP_W4 : (layer == 1 ? GP_W3 : GP_W2);
        v = P[W + (nt * 16 + (lane & 15)) * 64 + unit_of_slot(kc, lane >> 4, j)];
    }
    out[t] = (__bf16)v;
}

// Sums the per-workgroup gradient partials, then one Adam step
// (bias-corrected).  A workgroup of 16 waves owns 64 consecutive
// parameters: wave q sums the partials q, q + 16, q + 32, ... with eight
// loads in flight per lane (a 128-partial sum is one round of loads), and
// the first wave adds the 16 sums.  The reduction is latency-bound: with 4
// waves and 4 loads in flight a 2^15-record step spent 31 us here.
constexpr int ADAM_PARAMS = 64;
constexpr int ADAM_WAVES = 16;
__global__ __launch_bounds__(ADAM_WAVES * 64) void k_adam(float* __restrict__ P, float* __restrict__ m,
                                                          float* __restrict__ v, const float* __restrict__ partial,
                                                          int nparts, float scale, float lr, float b1, float b2,
                                                          float c1, float c2, float eps, float* __restrict__ grad_out)
{
    __shared__ float red[ADAM_WAVES][ADAM_PARAMS];
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int i = blockIdx.x * ADAM_PARAMS + lane;
    float g[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (i < GP_N) {
        int k = q;
        for (; k + 7 * ADAM_WAVES < nparts; k += 8 * ADAM_WAVES) {
#pragma unroll
            for (int u = 0; u < 8; u++) g[u] += partial[(size_t)(k + u * ADAM_WAVES) * GP_N + i];
        }
        for (; k < nparts; k += ADAM_WAVES) g[0] += partial[(size_t)k * GP_N + i];
    }
    red[q][lane] = ((g[0] + g[1]) + (g[2] + g[3])) + ((g[4] + g[5]) + (g[6] + g[7]));
    __syncthreads();
    if (q != 0 || i >= GP_N) return;
    float h[4];
#pragma unroll
    for (int w = 0; w < 4; w++)
        h[w] = (red[4 * w][lane] + red[4 * w + 1][lane]) + (red[4 * w + 2][lane] + red[4 * w + 3][lane]);
    float gs = ((h[0] + h[1]) + (h[2] + h[3])) * scale;
    if (grad_out) grad_out[i] = gs;
    if (lr == 0.0f) return;
    const float mi = b1 * m[i] + (1.0f - b1) * gs;
    const float vi = b2 * v[i] + (1.0f - b2) * gs * gs;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------------
// the network

// Orders the LDS accesses of one wave across its lanes (the per-wave LDS
// regions need no workgroup barrier, and the loops around them are not
// workgroup-uniform).
RD void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef float f32x2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));
typedef short s16x2v __attribute__((ext_vector_type(2)));

// two values rounded to bf16 as one packed pair, then the ReLU on the pair
// as two 16-bit integers: a set sign bit (negative, or -0) is a negative
// integer and becomes +0, the rest keep their bits
RD bf16x2v relu2(float a, float b)
{
    const bf16x2v p = __builtin_convertvector((f32x2v){a, b}, bf16x2v);
    const s16x2v v = __builtin_elementwise_max(__builtin_bit_cast(s16x2v, p), (s16x2v){0, 0});
    return __builtin_bit_cast(bf16x2v, v);
}

RD bf16x8 relu_pack(f32x4 lo, f32x4 hi)
{
    const bf16x4v l = __builtin_shufflevector(relu2(lo[0], lo[1]), relu2(lo[2], lo[3]), 0, 1, 2, 3);
    const bf16x4v h = __builtin_shufflevector(relu2(hi[0], hi[1]), relu2(hi[2], hi[3]), 0, 1, 2, 3);
    return __builtin_shufflevector(l, h, 0, 1, 2, 3, 4, 5, 6, 7);
}

// Operand of the encoding for k chunk kc, for this lane's query of tile mt.
// sx, cx: running sin/cos of 2^f pi x of the lane's current dimension.
RD bf16x8 encode_chunk(const float* xs, int kc, int g, int r, int mt, float& sx, float& cx)
{
    bf16x8 b;
    const int d = kc < 3 ? g : g + 4;
    if (d < 7) {
        // sin / cos of pi x by the hardware's sine and cosine (argument in
        // revolutions: pi x = 2 pi (x / 2)), ~1e-6 absolute for x in [0, 1],
        // far inside the bf16 rounding of every operand
        if (kc == 0 || kc == 3) {
            const float h = 0.5f * xs[(mt * 16 + r) * XS + d];
            sx = __builtin_amdgcn_sinf(h);
            cx = __builtin_amdgcn_cosf(h);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            b[2 * q] = (__bf16)sx;
            b[2 * q + 1] = (__bf16)cx;
            const float sc = sx * cx;
            const float s2 = sc + sc, c2 = fmaf(cx, cx, -(sx * sx));
            sx = s2;
            cx = c2;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 8; j++) b[j] = (__bf16)((kc == 3 && j == 0) ? 1.0f : 0.0f);
    }
    return b;
}

// Observer of the forward pass: sees every operand of the encoding (chunk kc
// of tile mt) and every hidden activation (layer L = 0..3, tile mt, chunk c),
// in operand layout.  Inference passes NoSink.
struct NoSink {
    RD void enc(int, int, const bf16x8&) {}
    RD void hidden(int, int, int, const bf16x8&) {}
};

// Forward pass of MT tiles of 16 queries (inputs xs[MT*16][8]) for one wave;
// outputs to ys[MT*16][YS].
template <int MT, class Sink>
RD void mlp_forward(const bf16x8* __restrict__ W, const float* __restrict__ bias, const float* xs, float* ys,
                    Sink& sink)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    f32x4 acc[4][MT];
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
        for (int mt = 0; mt < MT; mt++) acc[nt][mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float sx[MT], cx[MT];
#pragma unroll
    for (int mt = 0; mt < MT; mt++) { sx[mt] = 0.0f; cx[mt] = 0.0f; }
    for (int kc = 0; kc < 6; kc++) {
        bf16x8 b[MT];
#pragma unroll
        for (int mt = 0; mt < MT; mt++) {
            b[mt] = encode_chunk(xs, kc, g, r, mt, sx[mt], cx[mt]);
            sink.enc(kc, mt, b[mt]);
        }
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const bf16x8 a = W[OFF_A1 + (nt * 6 + kc) * 64 + lane];
#pragma unroll
            for (int mt = 0; mt < MT; mt++) acc[nt][mt] = MFMA16(a, b[mt], acc[nt][mt]);
        }
    }
    bf16x8 h[MT][2];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int c = 0; c < 2; c++) {
            h[mt][c] = relu_pack(acc[2 * c][mt], acc[2 * c + 1][mt]);
            sink.hidden(0, mt, c, h[mt][c]);
        }
#pragma unroll
    for (int L = 0; L < 3; L++) {
        const bf16x8* A = W + OFF_A2 + L * N_AH;
        const float* bl = bias + 64 * L;
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const f32x4 bv = {bl[nt * 16 + 4 * g], bl[nt * 16 + 4 * g + 1], bl[nt * 16 + 4 * g + 2],
                              bl[nt * 16 + 4 * g + 3]};
#pragma unroll
            for (int mt = 0; mt < MT; mt++) acc[nt][mt] = bv;
        }
#pragma unroll
        for (int kc = 0; kc < 2; kc++)
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                const bf16x8 a = A[(nt * 2 + kc) * 64 + lane];
#pragma unroll
                for (int mt = 0; mt < MT; mt++) acc[nt][mt] = MFMA16(a, h[mt][kc], acc[nt][mt]);
            }
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
            for (int c = 0; c < 2; c++) {
                h[mt][c] = relu_pack(acc[2 * c][mt], acc[2 * c + 1][mt]);
                sink.hidden(L + 1, mt, c, h[mt][c]);
            }
    }
    f32x4 o[3][MT];
    const float* b5 = bias + 192;
#pragma unroll
    for (int nt = 0; nt < 3; nt++) {
        const f32x4 bv = {b5[nt * 16 + 4 * g], b5[nt * 16 + 4 * g + 1], b5[nt * 16 + 4 * g + 2], b5[nt * 16 + 4 * g + 3]};
#pragma unroll
        for (int mt = 0; mt < MT; mt++) o[nt][mt] = bv;
    }
#pragma unroll
    for (int kc = 0; kc < 2; kc++)
#pragma unroll
        for (int nt = 0; nt < 3; nt++) {
            const bf16x8 a = W[OFF_A5 + (nt * 2 + kc) * 64 + lane];
#pragma unroll
            for (int mt = 0; mt < MT; mt++) o[nt][mt] = MFMA16(a, h[mt][kc], o[nt][mt]);
        }
#pragma unroll
    for (int nt = 0; nt < 3; nt++)
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                if (nt * 16 + 4 * g + i < 41) ys[(mt * 16 + r) * YS + nt * 16 + 4 * g + i] = o[nt][mt][i];
}

// Network inputs of a shading point: position in the scene bounds, shading
// normal and direction to the previous vertex as (theta / pi, phi / 2 pi).
RD void guide_inputs(const GuideDev& G, f3 p, f3 n, f3 wi, float* x)
{
    x[0] = fminf(fmaxf((p.x - G.lo.x) * G.inv_ext.x, 0.0f), 1.0f);
    x[1] = fminf(fmaxf((p.y - G.lo.y) * G.inv_ext.y, 0.0f), 1.0f);
    x[2] = fminf(fmaxf((p.z - G.lo.z) * G.inv_ext.z, 0.0f), 1.0f);
    const float inv_pi = LMX_INV_PI_F, inv_2pi = 0.5f * LMX_INV_PI_F;
    x[3] = acosf(fminf(fmaxf(n.z, -1.0f), 1.0f)) * inv_pi;
    float ph = atan2f(n.y, n.x);
    x[4] = (ph < 0.0f ? ph + 2.0f * LMX_PI_F : ph) * inv_2pi;
    x[5] = acosf(fminf(fmaxf(wi.z, -1.0f), 1.0f)) * inv_pi;
    ph = atan2f(wi.y, wi.x);
    x[6] = (ph < 0.0f ? ph + 2.0f * LMX_PI_F : ph) * inv_2pi;
}

// ---------------------------------------------------------------------------
// the mixture

struct Mix {
    f3 mu[8];
    float mlen[8], kap[8], pi[8], alpha, sig;
    bool kclamp[8];
};

// The mixture's arithmetic uses the hardware's exp / log / reciprocal /
// square root (a few ulp, against the correctly rounded library calls):
// it runs for every guided vertex, and its densities only need to agree
// with the directions actually drawn to well within the Monte Carlo noise
// (relative 1e-5 at most here; the tests bound pdfs at 1e-4).
RD float mx_exp(float x) { return __expf(x); }
RD float mx_log(float x) { return __logf(x); }
RD float mx_div(float a, float b) { return __fdividef(a, b); }
RD float mx_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

RD void mix_decode(const float* y, Mix& m)
{
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const f3 r = mk3(y[3 * k], y[3 * k + 1], y[3 * k + 2]);
        const float l = mx_sqrt(dot(r, r));
        m.mlen[k] = l;
        const float il = mx_div(1.0f, l);
        m.mu[k] = l > 0.0f ? mk3(r.x * il, r.y * il, r.z * il) : mk3(0.0f, 0.0f, 1.0f);
        float lk = y[24 + k];
        m.kclamp[k] = lk < LOGK_MIN || lk > LOGK_MAX;
        lk = fminf(fmaxf(lk, LOGK_MIN), LOGK_MAX);
        m.kap[k] = mx_exp(lk);
    }
    float mx = y[32];
#pragma unroll
    for (int k = 1; k < 8; k++) mx = fmaxf(mx, y[32 + k]);
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) { m.pi[k] = mx_exp(y[32 + k] - mx); s += m.pi[k]; }
    const float is = mx_div(1.0f, s);
#pragma unroll
    for (int k = 0; k < 8; k++) m.pi[k] *= is;
    m.sig = mx_div(1.0f, 1.0f + mx_exp(-y[40]));
    m.alpha = ALPHA_MIN + ALPHA_SPAN * m.sig;
}

// kappa >= e^-4, so 1 - e^{-2 kappa} >= 0.036 keeps its relative error
// below 1e-5 without expm1
RD float vmf_pdf(f3 mu, float k, f3 w)
{
    return mx_div(k, 2.0f * LMX_PI_F * (1.0f - mx_exp(-2.0f * k))) * mx_exp(k * (dot(mu, w) - 1.0f));
}

RD float mix_pdf(const Mix& m, f3 w, float* v = nullptr)
{
    float p = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float vk = vmf_pdf(m.mu[k], m.kap[k], w);
        if (v) v[k] = vk;
        p += m.pi[k] * vk;
    }
    return p;
}

// Jakob 2012: cos theta = 1 + log(u1 + (1 - u1) e^{-2 kappa}) / kappa
RD f3 vmf_sample(f3 mu, float k, float u1, float u2)
{
    const float W = fmaxf(1.0f + mx_div(mx_log(u1 + (1.0f - u1) * mx_exp(-2.0f * k)), k), -1.0f);
    const float rr = mx_sqrt(fmaxf(0.0f, 1.0f - W * W));
    float sn, cs;
    sincospif(2.0f * u2, &sn, &cs);
    f3 b, c;
    onb(mu, b, c);
    return add(mul(mu, W), add(mul(b, rr * cs), mul(c, rr * sn)));
}

// the lobe of the cumulative weights that u falls in (the last one when
// rounding leaves u above the sum); static indices only, so the mixture
// stays in registers
RD int mix_lobe(const Mix& m, float u, f3& mu, float& kap)
{
    int lobe = 7;
    mu = m.mu[7];
    kap = m.kap[7];
    float c = 0.0f;
    bool found = false;
#pragma unroll
    for (int k = 0; k < 7; k++) {
        c += m.pi[k];
        if (!found && u < c) { found = true; lobe = k; mu = m.mu[k]; kap = m.kap[k]; }
    }
    return lobe;
}

// d(-w log q)/dy of one record (q = alpha pb + (1 - alpha) p_vmf(dir))
RD float out_grad(const float* y, f3 dir, float pb, float w, float* g)
{
    Mix m;
    mix_decode(y, m);
    float v[8];
    const float pv = mix_pdf(m, dir, v);
    const float q = m.alpha * pb + (1.0f - m.alpha) * pv;
    const float gq = q > 0.0f ? -w / q : 0.0f;
    g[40] = gq * (pb - pv) * ALPHA_SPAN * m.sig * (1.0f - m.sig);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float base = gq * (1.0f - m.alpha) * m.pi[k];
        g[32 + k] = base * (v[k] - pv);
        const float kap = m.kap[k], c = dot(m.mu[k], dir);
        g[24 + k] = m.kclamp[k] ? 0.0f : base * v[k] * (1.0f - 2.0f * kap / expm1f(2.0f * kap) + kap * (c - 1.0f));
        const float s = m.mlen[k] > 0.0f ? base * v[k] * kap / m.mlen[k] : 0.0f;
        g[3 * k] = s * (dir.x - c * m.mu[k].x);
        g[3 * k + 1] = s * (dir.y - c * m.mu[k].y);
        g[3 * k + 2] = s * (dir.z - c * m.mu[k].z);
    }
    return q > 0.0f ? -w * logf(q) : 0.0f;
}

// ---------------------------------------------------------------------------
// kernels: evaluation and test hooks

// Inference workgroups: EVAL_WAVES waves sharing the weights in LDS
// (55 KB); each wave's inputs (64 x 7 floats) live at the start of its
// output rows (64 x 41 floats), which layer 1 has read before layer 5
// writes them.  140 KB of LDS: one workgroup, 8 waves, per CU.
#ifndef LMX_EVAL_WAVES
#define LMX_EVAL_WAVES 8
#endif
constexpr int EVAL_WAVES = LMX_EVAL_WAVES;
static_assert(N_FWD * 16 + N_BIAS * 4 + EVAL_WAVES * 64 * YS * 4 <= 163840, "inference LDS");

__global__ __launch_bounds__(EVAL_WAVES * 64) void k_eval(GuideDev G, const float* __restrict__ x, uint32_t n,
                                                          float* __restrict__ y, int vec)
{
    __shared__ bf16x8 wl[N_FWD];
    __shared__ float bl[N_BIAS];
    __shared__ __attribute__((aligned(16))) float ys[EVAL_WAVES][64 * YS];
    for (int t = threadIdx.x; t < N_FWD; t += blockDim.x) wl[t] = G.pack[t];
    for (int t = threadIdx.x; t < N_BIAS; t += blockDim.x) bl[t] = G.bias[t];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float* yw = ys[wave];
    NoSink dummy;
    for (uint32_t base = (blockIdx.x * EVAL_WAVES + wave) * 64u; base < n; base += gridDim.x * EVAL_WAVES * 64u) {
        // the wave's 64 x 7 inputs and 64 x 41 outputs are contiguous blocks
        // in both memories: 16-byte copies for a full block
        const uint32_t nq = min(64u, n - base);
        if (nq == 64u && vec) {
            const float4* src = reinterpret_cast<const float4*>(x + (size_t)base * 7);
            for (uint32_t k = (uint32_t)lane; k < 112u; k += 64u) reinterpret_cast<float4*>(yw)[k] = src[k];
        } else {
            for (uint32_t k = (uint32_t)lane; k < 64u * 7u; k += 64u) yw[k] = k < nq * 7u ? x[(size_t)base * 7 + k] : 0.0f;
        }
        wave_sync();
        mlp_forward<4>(wl, bl, yw, yw, dummy);
        wave_sync();
        if (nq == 64u && vec) {
            float4* dst = reinterpret_cast<float4*>(y + (size_t)base * 41);
            for (uint32_t k = (uint32_t)lane; k < 656u; k += 64u) dst[k] = reinterpret_cast<const float4*>(yw)[k];
        } else {
            for (uint32_t k = (uint32_t)lane; k < nq * 41u; k += 64u) y[(size_t)base * 41 + k] = yw[k];
        }
        wave_sync();
    }
}

// Test hook: decode outputs y and draw the mixture direction with (u, u1, u2);
// out: direction xyz, p_vmf of it, alpha, lobe, 0, 0
__global__ void k_mix_sample(const float* __restrict__ y, const float* __restrict__ u, uint32_t n, float* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Mix m;
    mix_decode(y + (size_t)i * 41, m);
    f3 mu;
    float kap;
    const int k = mix_lobe(m, u[3 * i], mu, kap);
    const f3 d = vmf_sample(mu, kap, u[3 * i + 1], u[3 * i + 2]);
    float* o = out + (size_t)i * 8;
    o[0] = d.x; o[1] = d.y; o[2] = d.z;
    o[3] = mix_pdf(m, d);
    o[4] = m.alpha;
    o[5] = (float)k;
    o[6] = 0.0f; o[7] = 0.0f;
}

// ---------------------------------------------------------------------------
// training
//
// Record (5 float4): [x0 x1 x2 x3] [x4 x5 x6 pb] [dir.xyz, sample] [T.rgb, w]
// [L.rgb, 0] -- inputs, direction, its BSDF density (solid angle), the
// throughput at the vertex, the loss weight, and the path's contribution so
// far (the weight is finished once the path has ended).

// Two kernels per step.  k_train_fb: forward (keeping the layer inputs),
// the loss gradient and the backward pass, 32 records per wave pass; it
// writes every layer input A_l and every layer-output delta D_l as bf16 rows
// [unit][record] to a scratch array.  k_wgrad: dW_l^T = D_l A_{l-1}^T as a
// split-K GEMM over records, each workgroup staging 32-record slices of all
// rows in LDS and its waves accumulating all 108 output tiles in registers;
// one gradient partial per workgroup, summed by k_adam.
//
// Scratch layout: one block of NROWS x 32 bf16 per 32-record pass,
// [pass][row][record], so that a pass's rows are one contiguous 47 KB span:
constexpr int ROW_A0 = 0;              // encoding, 192 slots (layer-1 operand order)
constexpr int ROW_A1 = 192;            // h1..h4, 64 units each, natural order
constexpr int ROW_D1 = ROW_A1 + 256;   // deltas of layers 1..4 (pre-activation), 64 each
constexpr int ROW_D5 = ROW_D1 + 256;   // output gradient, 48 (41 used)
constexpr int NROWS = ROW_D5 + 48;     // 752
constexpr int FB_WAVES = 4;
constexpr int WG_WAVES = 4;
constexpr int SROW = 40;               // bf16 per staged row (32 records + bank padding)

// one operand-layout chunk (tile mt, chunk kc: units unit_of_slot(kc, g, j))
// -> rows [row0 + unit][record] of a pass block
RD void store_chunk(__bf16* blk, int row0, int mt, int kc, const bf16x8& v, int lane)
{
    const int g = lane >> 4, r = lane & 15;
    __bf16* base = blk + mt * 16 + r;
#pragma unroll
    for (int j = 0; j < 8; j++) base[(row0 + unit_of_slot(kc, g, j)) * 32] = v[j];
}

// ReLU mask of an operand-layout activation: bit 16 mt + 8 c + j
RD uint32_t relu_bits(int mt, int c, const bf16x8& h)
{
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) m |= ((float)h[j] > 0.0f ? 1u : 0u) << (16 * mt + 8 * c + j);
    return m;
}

// forward observer of training: the encoding (layer-1 slot order) and
// h1..h4 go to the scratch rows; the ReLU masks stay in registers
struct TrainSink {
    __bf16* blk;
    int lane;
    uint32_t mask[4];
    RD void enc(int kc, int mt, const bf16x8& b)
    {
        const int g = lane >> 4, r = lane & 15;
        __bf16* base = blk + mt * 16 + r;
#pragma unroll
        for (int j = 0; j < 8; j++) base[(ROW_A0 + kc * 32 + 8 * g + j) * 32] = b[j];
    }
    RD void hidden(int L, int mt, int c, const bf16x8& h)
    {
        store_chunk(blk, ROW_A1 + 64 * L, mt, c, h, lane);
        const uint32_t m = relu_bits(mt, c, h);
        mask[L] = (mt == 0 && c == 0) ? m : (mask[L] | m);
    }
};

// one backward layer: delta of the layer input = W (natural rows) * delta,
// masked by the ReLU of that input (bits of relu_bits)
RD void back_layer(const bf16x8* __restrict__ Wb, const bf16x8 (&din)[2][2], uint32_t mask, bf16x8 (&dout)[2][2],
                   int lane)
{
    f32x4 acc[4][2];
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
        for (int mt = 0; mt < 2; mt++) acc[nt][mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kc = 0; kc < 2; kc++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const bf16x8 a = Wb[(nt * 2 + kc) * 64 + lane];
#pragma unroll
            for (int mt = 0; mt < 2; mt++) acc[nt][mt] = MFMA16(a, din[mt][kc], acc[nt][mt]);
        }
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int bit = 16 * mt + 8 * c;
                dout[mt][c][i] = (mask >> (bit + i)) & 1u ? (__bf16)acc[2 * c][mt][i] : (__bf16)0.0f;
                dout[mt][c][4 + i] = (mask >> (bit + 4 + i)) & 1u ? (__bf16)acc[2 * c + 1][mt][i] : (__bf16)0.0f;
            }
}

// Forward, loss gradient and backward over `n` records (indices idx, or
// 0..n-1), padded to whole 32-record passes (padding records have weight 0);
// writes the scratch blocks of all passes.
__global__ __launch_bounds__(FB_WAVES * 64) void k_train_fb(GuideDev G, const float4* __restrict__ recs,
                                                           const uint32_t* __restrict__ idx, uint32_t n,
                                                           float wnorm, float wclamp, float inv_n,
                                                           __bf16* __restrict__ act, float* __restrict__ loss)
{
    __shared__ float xs_all[FB_WAVES][32 * 8];
    __shared__ float ys_all[FB_WAVES][32 * YS];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    float* xs = xs_all[wave];
    float* ys = ys_all[wave];
    float lsum = 0.0f;
    const uint32_t npass = (n + 31) / 32;
    for (uint32_t pass = blockIdx.x * FB_WAVES + wave; pass < npass; pass += gridDim.x * FB_WAVES) {
        __bf16* blk = act + (size_t)pass * NROWS * 32;
        const uint32_t j = pass * 32 + (uint32_t)lane;
        const bool mine = lane < 32 && j < n;
        f3 dir = mk3(0, 0, 0);
        float pb = 0.0f, w = 0.0f;
        if (lane < 32) {
            float x[7] = {0, 0, 0, 0, 0, 0, 0};
            if (mine) {
                const size_t q = idx ? idx[j] : j;
                const float4 a = recs[5 * q], b = recs[5 * q + 1], c = recs[5 * q + 2], d = recs[5 * q + 3];
                x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w; x[4] = b.x; x[5] = b.y; x[6] = b.z;
                pb = b.w;
                dir = mk3(c.x, c.y, c.z);
                w = fminf(d.w * wnorm, wclamp) * inv_n;
            }
            for (int k = 0; k < 7; k++) xs[lane * XS + k] = x[k];
        }
        wave_sync();
        TrainSink sink{blk, lane, {0u, 0u, 0u, 0u}};
        mlp_forward<2>(G.pack, G.bias, xs, ys, sink);
        wave_sync();
        // output gradient (rows of ys, in place; pre-scaled by 1/n) and D5
        if (lane < 32) {
            float gy[41];
            if (mine) lsum += out_grad(ys + lane * YS, dir, pb, w, gy);
            else for (int o = 0; o < 41; o++) gy[o] = 0.0f;
            for (int o = 0; o < 41; o++) {
                ys[lane * YS + o] = gy[o];
                blk[(ROW_D5 + o) * 32 + lane] = (__bf16)gy[o];
            }
            for (int o = 41; o < 48; o++) blk[(ROW_D5 + o) * 32 + lane] = (__bf16)0.0f;
        }
        wave_sync();
        // backward: W5 (natural output order) * dY, then W4, W3, W2
        bf16x8 dy[2][2], d[2][2], dn[2][2];
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int kc = 0; kc < 2; kc++)
#pragma unroll
                for (int jj = 0; jj < 8; jj++) {
                    const int o = 32 * kc + 8 * g + jj;
                    dy[mt][kc][jj] = o < 41 ? (__bf16)ys[(mt * 16 + r) * YS + o] : (__bf16)0.0f;
                }
        back_layer(G.pack + OFF_B5, dy, sink.mask[3], d, lane);     // D4
#pragma unroll
        for (int L = 3; L >= 0; L--) {                              // store D_L+1, make D_L
#pragma unroll
            for (int mt = 0; mt < 2; mt++)
#pragma unroll
                for (int kc = 0; kc < 2; kc++) store_chunk(blk, ROW_D1 + 64 * L, mt, kc, d[mt][kc], lane);
            if (L == 0) break;
            back_layer(G.pack + OFF_B4 + (3 - L) * N_AH, d, sink.mask[L - 1], dn, lane);
#pragma unroll
            for (int mt = 0; mt < 2; mt++)
#pragma unroll
                for (int kc = 0; kc < 2; kc++) d[mt][kc] = dn[mt][kc];
        }
        wave_sync();
    }
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (lane == 0 && loss) atomicAdd(loss, lsum);
}

// The 108 weight-gradient tiles of the network (16 x 16 each, rows n of a
// layer's outputs, columns k of its inputs): layer 1 4 x 12, layers 2-4
// 4 x 4 each, layer 5 3 x 4.  Tile t: its layer, the first scratch row of its
// delta rows and of its input rows.
struct WTile {
    int layer, nt, kt, drow, arow;
};
RD WTile tile_of(int t)
{
    WTile T;
    if (t < 48) {
        T.layer = 1; T.nt = t / 12; T.kt = t % 12;
        T.drow = ROW_D1 + 16 * T.nt; T.arow = ROW_A0 + 16 * T.kt;
    } else if (t < 96) {
        const int u = t - 48;
        T.layer = 2 + u / 16; T.nt = (u % 16) / 4; T.kt = u % 4;
        T.drow = ROW_D1 + 64 * (T.layer - 1) + 16 * T.nt; T.arow = ROW_A1 + 64 * (T.layer - 2) + 16 * T.kt;
    } else {
        const int u = t - 96;
        T.layer = 5; T.nt = u / 4; T.kt = u % 4;
        T.drow = ROW_D5 + 16 * T.nt; T.arow = ROW_A1 + 192 + 16 * T.kt;
    }
    return T;
}
constexpr int TILES_PER_WAVE = 27;
static_assert(TILES_PER_WAVE * WG_WAVES == 108, "every tile has one wave");

// Split-K weight gradients: workgroup b sums the passes [b*chunk, (b+1)*chunk)
// of the scratch into partial[b] (every parameter written).
__global__ __launch_bounds__(WG_WAVES * 64) void k_wgrad(const __bf16* __restrict__ act, uint32_t npass,
                                                         uint32_t chunk, float* __restrict__ partial)
{
    __shared__ alignas(16) __bf16 st[NROWS * SROW];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    f32x4 acc[TILES_PER_WAVE];
#pragma unroll
    for (int t = 0; t < TILES_PER_WAVE; t++) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;                                   // thread t < 240: bias row t of D2..D5
    const uint32_t p_begin = blockIdx.x * chunk;
    const uint32_t p_end = min(npass, p_begin + chunk);
    // software pipeline: the next pass block is loaded into registers while
    // the MFMAs of the current one run (752 rows x 64 B = 12 x 16 B per thread)
    constexpr int NPRE = (NROWS * 4 + WG_WAVES * 64 - 1) / (WG_WAVES * 64);
    uint4 pre[NPRE];
    auto fetch = [&](uint32_t pass) {
        const uint4* src = reinterpret_cast<const uint4*>(act + (size_t)pass * NROWS * 32);
#pragma unroll
        for (int u = 0; u < NPRE; u++) {
            const int t = (int)threadIdx.x + u * WG_WAVES * 64;
            if (t < NROWS * 4) pre[u] = src[t];
        }
    };
    if (p_begin < p_end) fetch(p_begin);
    for (uint32_t pass = p_begin; pass < p_end; pass++) {
        __syncthreads();
        // stage one pass block: 752 rows x 64 B, 16 B per thread per step
#pragma unroll
        for (int u = 0; u < NPRE; u++) {
            const int t = (int)threadIdx.x + u * WG_WAVES * 64;
            if (t < NROWS * 4) *reinterpret_cast<uint4*>(&st[(t >> 2) * SROW + (t & 3) * 8]) = pre[u];
        }
        __syncthreads();
        if (pass + 1 < p_end) fetch(pass + 1);
#pragma unroll
        for (int i = 0; i < TILES_PER_WAVE; i++) {
            const WTile T = tile_of(wave * TILES_PER_WAVE + i);
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&st[(T.drow + r) * SROW + 8 * g]);
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(&st[(T.arow + r) * SROW + 8 * g]);
            acc[i] = MFMA16(a, b, acc[i]);
        }
        if (threadIdx.x < 240) {
            const int row = ROW_D1 + 64 + threadIdx.x;      // D2, D3, D4 (64 each), D5 (48)
            for (int q = 0; q < 32; q++) bsum += (float)st[row * SROW + q];
        }
    }
    float* P = partial + (size_t)blockIdx.x * GP_N;
#pragma unroll
    for (int i = 0; i < TILES_PER_WAVE; i++) {
        const WTile T = tile_of(wave * TILES_PER_WAVE + i);
        const int k = T.kt * 16 + r;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int nn = T.nt * 16 + 4 * g + e;
            if (T.layer == 1) {
                const int row = feat_of_slot(k >> 5, (k & 31) >> 3, k & 7);
                if (row >= 0) P[GP_W1 + row * 64 + nn] = acc[i][e];
            } else if (T.layer == 5) {
                if (nn < 41) P[GP_W5 + k * 41 + nn] = acc[i][e];
            } else {
                const int W = T.layer == 2 ? GP_W2 : (T.layer == 3 ? GP_W3 : GP_W4);
                P[W + k * 64 + nn] = acc[i][e];
            }
        }
    }
    if (threadIdx.x < 240) {
        const int t = threadIdx.x;
        if (t < 64) P[GP_B2 + t] = bsum;
        else if (t < 128) P[GP_B3 + t - 64] = bsum;
        else if (t < 192) P[GP_B4 + t - 128] = bsum;
        else if (t - 192 < 41) P[GP_B5 + t - 192] = bsum;
    }
}

// Loss weights of the records of one pass: w = mean over channels of
// (contribution after the vertex) / (throughput at the vertex); their sum.
__global__ __launch_bounds__(256) void k_finalize(float4* __restrict__ recs, const uint32_t* __restrict__ count,
                                                 uint32_t cap, const float* __restrict__ pathL,
                                                 float* __restrict__ wsum, uint32_t* __restrict__ nz,
                                                 uint32_t* __restrict__ nz_count)
{
    __shared__ uint32_t lds[5];
    const uint32_t n = min(*count, cap);
    float s = 0.0f;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        float w = 0.0f;
        if (i < n) {
        const float4 c = recs[5 * i + 2], t = recs[5 * i + 3], l = recs[5 * i + 4];
        const uint32_t sample = __float_as_uint(c.w);
        const float* Lf = pathL + 3 * (size_t)sample;
        const float Sr = fmaxf(Lf[0] - l.x, 0.0f), Sg = fmaxf(Lf[1] - l.y, 0.0f), Sb = fmaxf(Lf[2] - l.z, 0.0f);
        if (t.x > 0.0f) w += Sr / t.x;
        if (t.y > 0.0f) w += Sg / t.y;
        if (t.z > 0.0f) w += Sb / t.z;
        w *= (1.0f / 3.0f);
        if (!(w < 3.0e38f)) w = 0.0f;
        recs[5 * i + 3] = make_float4(t.x, t.y, t.z, w);
        s += w;
        }
        // records that carry radiance: the only ones with a loss gradient
        const uint32_t slot = block_reserve<4>(nz_count, w > 0.0f ? 1u : 0u, lds);
        if (w > 0.0f) nz[slot] = i;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(wsum, s);
}

// minibatch indices: uniform with replacement from the records of the pass
__global__ void k_pick(uint32_t* __restrict__ idx, uint32_t m, const uint32_t* __restrict__ nz,
                       const uint32_t* __restrict__ nz_count, uint32_t cap, uint2 key, uint32_t step)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t n = min(*nz_count, cap);
    const uint4 o = philox4x32_10(make_uint4(i, step, 0x67756964u, LMX_PHILOX_TAG), key);
    idx[i] = n ? nz[((uint64_t)o.x * n) >> 32] : 0u;
}

__global__ void k_fill(float* a, size_t n, float v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}

__global__ void k_scale_f(float* a, size_t n, float w)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] *= w;
}

}  // namespace

// ---------------------------------------------------------------------------
// guided shading

namespace lmx {
struct GuideLaunch {
    GuideDev dev;
    float4* recs;              // training records or null
    uint32_t* rec_count;
    uint32_t rec_cap;
    float* pathL;              // per-sample contribution so far (rgb), with records
};
}  // namespace lmx

namespace {

// GS_WAVES waves per workgroup share the weights in LDS, as k_eval's do
// (224 VGPRs: two waves per SIMD)
constexpr int GS_WAVES = 8;
__global__ __launch_bounds__(GS_WAVES * 64) void k_guided_shade(RScene s, RParams p, GuideLaunch gl,
                                                                  Ctl* __restrict__ ctl, Queue in,
                                                                  const float4* __restrict__ hits, Queue out)
{
    __shared__ bf16x8 wl[N_FWD];
    __shared__ float bl[N_BIAS];
    __shared__ float ys[GS_WAVES][64 * YS];       // inputs at the start of each wave's rows
    __shared__ uint32_t lds[GS_WAVES + 1];
    for (int t = threadIdx.x; t < N_FWD; t += blockDim.x) wl[t] = gl.dev.pack[t];
    for (int t = threadIdx.x; t < N_BIAS; t += blockDim.x) bl[t] = gl.dev.bias[t];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    NoSink dummy;
    const uint32_t n = ctl->n_trace;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool guided = false, cont = false, rec = false;
        float rec_pb = 0.0f;
        f3 rec_L = mk3(0, 0, 0);
        float4 o_ray0, o_ray1, o_thr;
        uint4 o_meta;
        f3 thr = mk3(0, 0, 0), d = mk3(0, 0, 0), wi = mk3(0, 0, 0);
        Frame fr;
        RMat m;
        uint32_t sample = 0, seg = 0, pixel = 0;
        float x[7] = {0, 0, 0, 0, 0, 0, 0};
        if (i < n) {
            const uint4 meta = in.meta[i];
            const float4 h = hits[i];
            const uint32_t id = __float_as_uint(h.w);
            if (!(meta.w & META_DEAD) && id != WBVH_MISS) {
                const float4 r0 = in.ray[2 * i], r1 = in.ray[2 * i + 1];
                const float4 t4 = in.thr[i];
                thr = mk3(t4.x, t4.y, t4.z);
                const f3 o = mk3(r0.x, r0.y, r0.z);
                d = mk3(r1.x, r1.y, r1.z);
                sample = meta.x; seg = meta.y; pixel = meta.z;
                surface_at(s, o, d, h.x, h.y, h.z, id, fr);
                m = s.mats[s.tri_mat[id]];
                int nv = (int)seg + 1;
                // emission (renderer_pt.cpp:183-194)
                if (m.Le[0] > 0.0f || m.Le[1] > 0.0f || m.Le[2] > 0.0f) {
                    if (nv + 1 >= p.min_nv && to_local(fr, neg(d)).z > 0.0f) {
                        const f3 C = mk3(thr.x * m.Le[0], thr.y * m.Le[1], thr.z * m.Le[2]);
                        float* fp = p.film + 3 * (size_t)pixel;
                        atomicAdd(fp, C.x); atomicAdd(fp + 1, C.y); atomicAdd(fp + 2, C.z);
                        if (gl.pathL) {
                            float* L = gl.pathL + 3 * (size_t)sample;
                            L[0] += C.x; L[1] += C.y; L[2] += C.z;
                        }
                        if (p.sample_out) {
                            float* so = p.sample_out + 8 * (size_t)sample;
                            so[2] += C.x; so[3] += C.y; so[4] += C.z; so[5] += 1.0f;
                        }
                    }
                }
                const Uniforms<RParams> U{p, sample};
                bool alive = true;
                if ((int)seg + 1 >= p.rr_depth) {
                    const float q = rr_continue(p, thr);
                    if (!(q > 0.0f) || U.one(LMX_PT_SLOT_SEG0 + 4u * seg + 3u) > q) alive = false;
                    else thr = mk3(thr.x / q, thr.y / q, thr.z / q);
                }
                nv++;
                if (alive && p.max_nv != -1 && nv >= p.max_nv) alive = false;
                if (alive && is_scatterer(m)) {
                    guided = true;
                    wi = neg(d);
                    guide_inputs(gl.dev, fr.p, fr.sn, wi, x);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 7; k++) ys[wave][lane * XS + k] = x[k];
        __syncthreads();
        if (__ballot(guided)) mlp_forward<4>(wl, bl, ys[wave], ys[wave], dummy);   // wave-uniform
        __syncthreads();
        if (guided) {
            Mix mx;
            mix_decode(ys[wave] + lane * YS, mx);
            const Uniforms<RParams> U{p, sample};
            const uint32_t seg2 = seg + 1u;
            const float4 u = U.block(LMX_PT_SLOT_SEG0 + 4u * seg2);
            f3 wo;
            bool ok;
            if (u.z < mx.alpha) {
                ok = bsdf_sample(m, fr, wi, u.x, u.y, wo);
            } else {
                f3 mu;
                float kap;
                (void)mix_lobe(mx, (u.z - mx.alpha) / (1.0f - mx.alpha), mu, kap);
                wo = vmf_sample(mu, kap, u.x, u.y);
                ok = true;
            }
            if (ok) {
                const float co = fabsf(to_local(fr, wo).z);
                const float pb = bsdf_pdf(m, fr, wi, wo) * co;          // solid angle
                const float q = mx.alpha * pb + (1.0f - mx.alpha) * mix_pdf(mx, wo);
                const f3 fs = bsdf_eval(m, fr, wi, wo);
                if (q > 0.0f && !(fs.x == 0.0f && fs.y == 0.0f && fs.z == 0.0f)) {
                    const float sc = co / q;
                    const f3 t2 = mk3(thr.x * fs.x * sc, thr.y * fs.y * sc, thr.z * fs.z * sc);
                    o_ray0 = make_float4(fr.p.x, fr.p.y, fr.p.z, LMX_EPS_ISECT);
                    o_ray1 = make_float4(wo.x, wo.y, wo.z, LMX_FLT_MAX);
                    o_thr = make_float4(t2.x, t2.y, t2.z, 0.0f);
                    o_meta = make_uint4(sample, seg2, pixel, 0u);
                    cont = true;
                    if (gl.recs) {
                        rec = true;
                        // record: inputs, direction, pb, throughput at the vertex, contribution so far
                        const float* L = gl.pathL + 3 * (size_t)sample;
                        rec_pb = pb;
                        rec_L = mk3(L[0], L[1], L[2]);
                    }
                }
            }
        }
        const uint32_t slot = block_reserve<GS_WAVES>(&ctl->n_next, cont ? 1u : 0u, lds);
        if (cont) {
            out.ray[2 * slot] = o_ray0;
            out.ray[2 * slot + 1] = o_ray1;
            out.thr[slot] = o_thr;
            out.meta[slot] = o_meta;
        }
        if (gl.recs) {
            const uint32_t r = block_reserve<GS_WAVES>(gl.rec_count, rec ? 1u : 0u, lds);
            if (rec && r < gl.rec_cap) {
                float4* R = gl.recs + 5 * (size_t)r;
                R[0] = make_float4(x[0], x[1], x[2], x[3]);
                R[1] = make_float4(x[4], x[5], x[6], rec_pb);
                R[2] = make_float4(o_ray1.x, o_ray1.y, o_ray1.z, __uint_as_float(sample));
                R[3] = make_float4(thr.x, thr.y, thr.z, 0.0f);
                R[4] = make_float4(rec_L.x, rec_L.y, rec_L.z, 0.0f);
            }
        }
    }
}

}  // namespace

int launch_guided_shade(lmx_scene* sc, const RParams& rp, Queue qa, Queue qb, hipStream_t st)
{
    const int grid = std::max(1, std::min((int)((sc->cap + GS_WAVES * 64 - 1) / (GS_WAVES * 64)), sc->cus));
    k_guided_shade<<<grid, GS_WAVES * 64, 0, st>>>(sc->rs, rp, *rp.guide, sc->ctl.p, qa, sc->hits.p, qb);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? LMX_OK : hip_error(e, "guided shade launch");
}

// ---------------------------------------------------------------------------
// the object behind the C ABI

struct lmx_guide {
    lmx_scene* scene = nullptr;
    lmx_guide_opts opts;
    DevBuf<float> P, m, v, grad, partial, loss, wsum;
    DevBuf<__bf16> pack, act;
    DevBuf<float> bias;
    DevBuf<float4> recs;
    DevBuf<uint32_t> rec_count, idx, nz, nz_count;
    DevBuf<float> pathL;
    uint32_t rec_cap = 0, path_cap = 0;
    int step = 0;
    int nparts = 0;
    GuideDev dev() const
    {
        GuideDev G;
        G.pack = reinterpret_cast<const bf16x8*>(pack.p);
        G.bias = bias.p;
        G.lo = lo;
        G.inv_ext = inv_ext;
        return G;
    }
    f3 lo, inv_ext;
};

namespace {

int guide_pack(lmx_guide* g, hipStream_t st)
{
    k_pack<<<(N_PACK * 8 + 255) / 256, 256, 0, st>>>(g->P.p, g->pack.p, g->bias.p);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? LMX_OK : hip_error(e, "guide pack");
}

// gradient of the mean loss over n records (device) with an optional Adam step
int guide_train_step(lmx_guide* g, const float4* recs, const uint32_t* idx, uint32_t n, float wnorm, float wclamp,
                     bool adam, float* grad_out, hipStream_t st)
{
    hipError_t e;
    const uint32_t npass = std::max<uint32_t>((n + 31) / 32, 1);
    if (g->act.n < (size_t)NROWS * 32 * npass) {
        if ((e = hipStreamSynchronize(st)) != hipSuccess || (e = g->act.alloc((size_t)NROWS * 32 * npass)) != hipSuccess)
            return hip_error(e, "guide training scratch");
    }
    // forward/backward: one wave per 32-record pass, up to 8 waves per CU
    const int fb_wg = std::max(1, std::min((int)((npass + FB_WAVES - 1) / FB_WAVES), 2 * g->nparts));
    // weight gradients: at least 4 passes per workgroup (a 2^15-record step
    // fills the CUs with one workgroup each)
    const int nwg = std::max(1, std::min((int)((npass + 3) / 4), g->nparts));
    const uint32_t chunk = (npass + nwg - 1) / nwg;
    k_fill<<<1, 64, 0, st>>>(g->loss.p, 1, 0.0f);
    k_train_fb<<<fb_wg, FB_WAVES * 64, 0, st>>>(g->dev(), recs, idx, n, wnorm, wclamp, 1.0f / (float)n, g->act.p,
                                                g->loss.p);
    k_wgrad<<<nwg, WG_WAVES * 64, 0, st>>>(g->act.p, npass, chunk, g->partial.p);
    float lr = 0.0f, c1 = 1.0f, c2 = 1.0f;
    if (adam) {
        g->step++;
        lr = g->opts.lr;
        c1 = 1.0f - powf(g->opts.beta1, (float)g->step);
        c2 = 1.0f - powf(g->opts.beta2, (float)g->step);
    }
    k_adam<<<(GP_N + ADAM_PARAMS - 1) / ADAM_PARAMS, ADAM_WAVES * 64, 0, st>>>(g->P.p, g->m.p, g->v.p, g->partial.p, nwg, 1.0f, lr, g->opts.beta1,
                                               g->opts.beta2, c1, c2, g->opts.eps, grad_out);
    if ((e = hipGetLastError()) != hipSuccess) return hip_error(e, "guide train launch");
    return adam ? guide_pack(g, st) : LMX_OK;
}

void init_params(std::vector<float>& P, uint64_t seed)
{
    // He-normal weights from a counter hash, zero biases; lobe concentrations
    // start near kappa = e^1 and the BSDF fraction near 0.84, so that an
    // untrained network samples mostly as the plain path tracer does
    P.assign(GP_N, 0.0f);
    auto normal = [&](uint32_t i) {
        uint32_t a = (uint32_t)(seed * 0x9E3779B97F4A7C15ull >> 32) ^ (i * 0x85EBCA6Bu);
        auto nx = [&]() { a ^= a >> 16; a *= 0x7feb352du; a ^= a >> 15; a *= 0x846ca68bu; a ^= a >> 16; return a; };
        const double u1 = ((nx() >> 8) + 0.5) / 16777216.0, u2 = ((nx() >> 8) + 0.5) / 16777216.0;
        return (float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    };
    for (int k = 0; k < 169 * 64; k++) P[GP_W1 + k] = k < 168 * 64 ? normal(k) * std::sqrt(2.0f / 168.0f) : 0.0f;
    for (int L = 0; L < 3; L++) {
        const int W = L == 0 ? GP_W2 : (L == 1 ? GP_W3 : GP_W4);
        for (int k = 0; k < 64 * 64; k++) P[W + k] = normal(100000 * (L + 1) + k) * std::sqrt(2.0f / 64.0f);
    }
    for (int k = 0; k < 64 * 41; k++) P[GP_W5 + k] = normal(900000 + k) * std::sqrt(1.0f / 64.0f) * 0.5f;
    for (int k = 0; k < 8; k++) P[GP_B5 + 24 + k] = 1.0f;
    P[GP_B5 + 40] = 2.5f;
}

}  // namespace

extern "C" {

int lmx_guide_create(lmx_scene* sc, const lmx_guide_opts* o, lmx_guide** out)
{
    if (!out) return set_error(LMX_EINVAL, "lmx_guide_create: out is NULL");
    *out = nullptr;
    if (!sc) return set_error(LMX_EINVAL, "lmx_guide_create: scene is NULL");
    std::unique_ptr<lmx_guide> g(new (std::nothrow) lmx_guide());
    if (!g) return set_error(LMX_ENOMEM, "lmx_guide_create: out of host memory");
    g->scene = sc;
    if (o) g->opts = *o;
    else lmx_guide_default_opts(&g->opts);
    if (!(g->opts.lr > 0.0f) || g->opts.batch == 0 || g->opts.batch > (1u << 24) || !(g->opts.wclamp > 0.0f))
        return set_error(LMX_EINVAL, "lmx_guide_create: bad options (lr > 0, batch in [1, 2^24], wclamp > 0)");
    hipError_t e;
    g->nparts = std::max(1, sc->cus);
    if ((e = g->P.alloc(GP_N)) != hipSuccess || (e = g->m.alloc(GP_N)) != hipSuccess ||
        (e = g->v.alloc(GP_N)) != hipSuccess || (e = g->grad.alloc(GP_N)) != hipSuccess ||
        (e = g->partial.alloc((size_t)g->nparts * GP_N)) != hipSuccess || (e = g->loss.alloc(1)) != hipSuccess ||
        (e = g->wsum.alloc(1)) != hipSuccess || (e = g->pack.alloc((size_t)N_PACK * 8)) != hipSuccess ||
        (e = g->bias.alloc(N_BIAS)) != hipSuccess || (e = g->rec_count.alloc(1)) != hipSuccess ||
        (e = g->idx.alloc(g->opts.batch)) != hipSuccess || (e = g->nz_count.alloc(1)) != hipSuccess ||
        (e = g->act.alloc((size_t)NROWS * ((g->opts.batch + 31) / 32 * 32))) != hipSuccess)
        return hip_error(e, "lmx_guide_create: device memory");
    std::vector<float> P;
    init_params(P, g->opts.seed);
    if ((e = hipMemcpy(g->P.p, P.data(), GP_N * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(g->m.p, 0, GP_N * sizeof(float))) != hipSuccess ||
        (e = hipMemset(g->v.p, 0, GP_N * sizeof(float))) != hipSuccess)
        return hip_error(e, "lmx_guide_create: upload");
    // position normalisation: the scene's bounds
    const float* bb = sc->bounds;
    g->lo = mk3(bb[0], bb[1], bb[2]);
    g->inv_ext = mk3(1.0f / std::max(bb[3] - bb[0], 1e-6f), 1.0f / std::max(bb[4] - bb[1], 1e-6f),
                     1.0f / std::max(bb[5] - bb[2], 1e-6f));
    int rc = guide_pack(g.get(), nullptr);
    if (rc == LMX_OK && (e = hipDeviceSynchronize()) != hipSuccess) rc = hip_error(e, "lmx_guide_create: sync");
    if (rc != LMX_OK) return rc;
    *out = g.release();
    return LMX_OK;
}

void lmx_guide_free(lmx_guide* g) { delete g; }

void lmx_guide_default_opts(lmx_guide_opts* o)
{
    if (!o) return;
    o->lr = 1e-3f;
    o->beta1 = 0.9f;
    o->beta2 = 0.999f;
    o->eps = 1e-8f;
    o->batch = 1u << 15;
    o->steps_per_pass = 16;
    o->wclamp = 50.0f;
    o->seed = 7;
}

int lmx_guide_params(lmx_guide* g, float* host, int set)
{
    if (!g || !host) return set_error(LMX_EINVAL, "lmx_guide_params: NULL argument");
    hipError_t e;
    if (set) {
        if ((e = hipMemcpy(g->P.p, host, GP_N * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
            return hip_error(e, "lmx_guide_params");
        int rc = guide_pack(g, nullptr);
        if (rc != LMX_OK) return rc;
        if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_error(e, "lmx_guide_params: sync");
        return LMX_OK;
    }
    if ((e = hipMemcpy(host, g->P.p, GP_N * sizeof(float), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_error(e, "lmx_guide_params");
    return LMX_OK;
}

int lmx_guide_eval(lmx_guide* g, const float* x, uint32_t n, float* y, void* stream)
{
    if (!g || (n && (!x || !y))) return set_error(LMX_EINVAL, "lmx_guide_eval: NULL argument");
    if (n == 0) return LMX_OK;
    hipStream_t st = (hipStream_t)stream;
    const int grid = std::max(1, std::min((int)((n + EVAL_WAVES * 64 - 1) / (EVAL_WAVES * 64)), g->scene->cus));
    const int vec = ((uintptr_t)x % 16 == 0 && (uintptr_t)y % 16 == 0) ? 1 : 0;    // 16-byte block copies
    k_eval<<<grid, EVAL_WAVES * 64, 0, st>>>(g->dev(), x, n, y, vec);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? LMX_OK : hip_error(e, "lmx_guide_eval");
}

int lmx_guide_mix_sample(const float* y, const float* u, uint32_t n, float* out, void* stream)
{
    if (n && (!y || !u || !out)) return set_error(LMX_EINVAL, "lmx_guide_mix_sample: NULL argument");
    if (n == 0) return LMX_OK;
    k_mix_sample<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(y, u, n, out);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? LMX_OK : hip_error(e, "lmx_guide_mix_sample");
}

int lmx_guide_grad(lmx_guide* g, const float* recs, uint32_t n, float* grad, float* loss, int adam, void* stream)
{
    if (!g || !recs || n == 0) return set_error(LMX_EINVAL, "lmx_guide_grad: NULL argument or n == 0");
    hipStream_t st = (hipStream_t)stream;
    int rc = guide_train_step(g, reinterpret_cast<const float4*>(recs), nullptr, n, 1.0f, 3.0e38f, adam != 0,
                              grad ? grad : g->grad.p, st);
    hipError_t e;
    if (rc == LMX_OK && loss &&
        (e = hipMemcpyAsync(loss, g->loss.p, sizeof(float), hipMemcpyDeviceToHost, st)) != hipSuccess)
        rc = hip_error(e, "lmx_guide_grad: loss");
    if (rc == LMX_OK && (e = hipStreamSynchronize(st)) != hipSuccess) rc = hip_error(e, "lmx_guide_grad: sync");
    return rc;
}

int lmx_render_guided(lmx_scene* sc, lmx_guide* g, const lmx_render_params* p, int train, float* film_rgb,
                      void* stream, lmx_guide_stats* gs, lmx_render_stats* stats)
{
    if (!sc || !g || !p || !film_rgb) return set_error(LMX_EINVAL, "lmx_render_guided: NULL argument");
    if (g->scene != sc) return set_error(LMX_EINVAL, "lmx_render_guided: guide belongs to another scene");
    if (p->render_time > 0.0) return set_error(LMX_EINVAL, "lmx_render_guided: render_time is not supported");
    if (p->num_samples == 0 || p->num_samples > 0xffffffffull)
        return set_error(LMX_EINVAL, "lmx_render_guided: num_samples must be in [1, 2^32)");
    if (p->max_num_vertices != -1 && p->max_num_vertices < 2)
        return set_error(LMX_EINVAL, "lmx_render_guided: max_num_vertices must be -1 or >= 2");
    if (p->uniform_table && (p->max_num_vertices == -1 ||
                             p->table_width < (uint32_t)(LMX_PT_SLOT_SEG0 + 4 * (p->max_num_vertices - 1))))
        return set_error(LMX_EINVAL, "lmx_render_guided: uniform table narrower than 8 + 4 (max_num_vertices - 1)");
    if (!(p->rr_prob > 0.0f && p->rr_prob <= 1.0f)) return set_error(LMX_EINVAL, "lmx_render_guided: rr_prob must be in (0, 1]");
    if (p->rr_mode != LMX_RR_FIXED && p->rr_mode != LMX_RR_THROUGHPUT)
        return set_error(LMX_EINVAL, "lmx_render_guided: unknown rr_mode");
    if (p->raster_mode != LMX_RASTER_RANDOM && p->raster_mode != LMX_RASTER_STRATIFIED)
        return set_error(LMX_EINVAL, "lmx_render_guided: unknown raster_mode");
    hipStream_t st = (hipStream_t)stream;
    hipStream_t own = nullptr;
    hipError_t e;
    if (!st) {
        if ((e = hipStreamCreateWithFlags(&own, hipStreamNonBlocking)) != hipSuccess) return hip_error(e, "stream create");
        st = own;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const size_t npix = (size_t)sc->width * (size_t)sc->height;
    const int zgrid = std::max(1, std::min((int)((3 * npix + 255) / 256), sc->cus * 8));
    int rc = LMX_OK;
    GuideLaunch gl;
    gl.dev = g->dev();
    gl.recs = nullptr; gl.rec_count = nullptr; gl.rec_cap = 0; gl.pathL = nullptr;
    const int maxv = p->max_num_vertices == -1 ? 64 : p->max_num_vertices;
    if (train) {
        // records: at most one per vertex before the last
        const uint64_t want = std::min<uint64_t>((uint64_t)p->num_samples * (uint64_t)(maxv - 1), 1ull << 27);
        if (g->rec_cap < want) {
            if ((e = g->recs.alloc(5 * (size_t)want)) != hipSuccess) rc = hip_error(e, "guide records");
            else g->rec_cap = (uint32_t)want;
        }
        if (rc == LMX_OK && g->path_cap < p->num_samples) {
            if ((e = g->pathL.alloc(3 * (size_t)p->num_samples)) != hipSuccess) rc = hip_error(e, "guide path buffer");
            else g->path_cap = (uint32_t)p->num_samples;
        }
        gl.recs = g->recs.p; gl.rec_count = g->rec_count.p; gl.rec_cap = g->rec_cap; gl.pathL = g->pathL.p;
        if (rc == LMX_OK) {
            k_fill<<<std::max(1, std::min((int)((3 * p->num_samples + 255) / 256), sc->cus * 8)), 256, 0, st>>>(
                g->pathL.p, 3 * (size_t)p->num_samples, 0.0f);
            k_fill<<<1, 64, 0, st>>>(reinterpret_cast<float*>(g->rec_count.p), 1, 0.0f);
        }
    }
    RParams rp;
    rp.max_nv = p->max_num_vertices; rp.min_nv = p->min_num_vertices; rp.rr_depth = p->rr_depth;
    rp.rr_prob = p->rr_prob;
    rp.key = make_uint2((uint32_t)p->seed, (uint32_t)(p->seed >> 32));
    rp.total = (uint32_t)p->num_samples;
    rp.raster_mode = p->raster_mode;
    rp.table = p->uniform_table; rp.table_width = p->table_width;
    rp.sample_out = p->sample_out;
    rp.film = film_rgb;
    rp.guide = &gl;
    rp.rr_mode = p->rr_mode;
    rp.nee = NEE_NONE;
    rp.seg_stride = 4;
    rp.so_stride = 8;
    uint32_t cap = p->queue_capacity ? p->queue_capacity : (sc->tuned_cap ? sc->tuned_cap : (1u << 22));
    cap = (uint32_t)std::min<uint64_t>(cap, std::max<uint64_t>(p->num_samples, 64));
    uint64_t rays = 0;
    uint32_t iters = 0;
    if (rc == LMX_OK) {
        k_fill<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix, 0.0f);
        rc = wavefront_render(sc, rp, cap, st, &rays, &iters);
    }
    if (rc == LMX_OK)
        k_scale_f<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix, (float)((double)npix / (double)p->num_samples));
    const auto t1 = std::chrono::steady_clock::now();
    uint32_t nrec = 0, nrec_nz = 0;
    float loss_sum = 0.0f;
    if (rc == LMX_OK && train) {
        // loss weights, normalised to mean one over the records that carry
        // radiance (the others have no gradient), then the steps on those
        if (g->nz.n < g->rec_cap && (e = g->nz.alloc(g->rec_cap)) != hipSuccess) rc = hip_error(e, "guide index");
        uint32_t nnz = 0;
        float hsum = 0.0f;
        if (rc == LMX_OK) {
            k_fill<<<1, 64, 0, st>>>(g->wsum.p, 1, 0.0f);
            k_fill<<<1, 64, 0, st>>>(reinterpret_cast<float*>(g->nz_count.p), 1, 0.0f);
            k_finalize<<<sc->cus * 4, 256, 0, st>>>(g->recs.p, g->rec_count.p, g->rec_cap, g->pathL.p, g->wsum.p,
                                                     g->nz.p, g->nz_count.p);
            if ((e = hipMemcpyAsync(&nrec, g->rec_count.p, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
                (e = hipMemcpyAsync(&nnz, g->nz_count.p, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
                (e = hipMemcpyAsync(&hsum, g->wsum.p, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
                (e = hipStreamSynchronize(st)) != hipSuccess)
                rc = hip_error(e, "guide records readback");
        }
        nrec = std::min(nrec, g->rec_cap);
        if (rc == LMX_OK && nnz > 0 && hsum > 0.0f) {
            const float wnorm = (float)nnz / hsum;
            const uint32_t m = (uint32_t)std::min<uint64_t>(g->opts.batch, 4ull * nnz);
            for (uint32_t k = 0; k < g->opts.steps_per_pass && rc == LMX_OK; k++) {
                k_pick<<<(m + 255) / 256, 256, 0, st>>>(g->idx.p, m, g->nz.p, g->nz_count.p, g->rec_cap,
                                                         make_uint2((uint32_t)g->opts.seed, (uint32_t)(g->opts.seed >> 32)),
                                                         (uint32_t)g->step);
                rc = guide_train_step(g, g->recs.p, g->idx.p, m, wnorm, g->opts.wclamp, true, nullptr, st);
                float l = 0.0f;
                if (rc == LMX_OK && k + 1 == g->opts.steps_per_pass) {
                    if ((e = hipMemcpyAsync(&l, g->loss.p, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
                        (e = hipStreamSynchronize(st)) != hipSuccess)
                        rc = hip_error(e, "guide loss readback");
                    loss_sum = l;
                }
            }
        }
        nrec_nz = nnz;
    }
    if (rc == LMX_OK && (e = hipStreamSynchronize(st)) != hipSuccess) rc = hip_error(e, "lmx_render_guided: sync");
    const auto t2 = std::chrono::steady_clock::now();
    if (own) (void)hipStreamDestroy(own);
    if (stats) {
        stats->rays = rays;
        stats->iterations = iters;
        stats->queue_capacity = cap;
        stats->ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
        stats->samples = rc == LMX_OK ? p->num_samples : 0;
    }
    if (gs) {
        gs->records = nrec;
        gs->records_nonzero = nrec_nz;
        gs->steps = (uint32_t)g->step;
        gs->loss = loss_sum;
        gs->render_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        gs->train_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    }
    return rc;
}

}  // extern "C"
