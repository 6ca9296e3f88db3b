// Tile-cost predictor: a BF16 MLP (12 -> 64 -> 64 -> 1, ReLU) with FP32
// accumulation, run and trained on the device by hand-written MFMA kernels.
//
// Formulation.  Every layer is computed transposed, Z^T = W^T X^T, with the
// tiles (rows of X) as the N dimension of v_mfma_f32_16x16x16_bf16: the
// weights are the A operands (rows = output units), activations the B
// operands.  With the 16x16x16 operand layout
//   A[i][k]: lane i + 16 (k / 4), element k % 4
//   B[k][j]: lane j + 16 (k / 4), element k % 4
//   D[i][j]: lane j + 16 (i / 4), element i % 4
// the D tile of one layer is, element for element, the B operand of the
// next, so the forward pass needs no data movement between layers.  The
// weight-gradient products (sums over tiles) need the transposed
// activations and go through a small LDS tile per wavefront.
//
// Rounding points (mirrored by the CPU oracle of the network): every MFMA
// operand is rounded to bf16 (round to nearest even), every accumulation and
// every elementwise operation is fp32.  Gradients are those of the network
// as computed (bf16-rounded weights) with respect to the fp32 master
// weights.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "cpv_internal.h"

using namespace cpvrt;

namespace {

constexpr int kF = 12, kH = 64;
constexpr int oW1 = 0, ob1 = oW1 + kF * kH, oW2 = ob1 + kH, ob2 = oW2 + kH * kH, ow3 = ob2 + kH, ob3 = ow3 + kH;
constexpr int kParams = ob3 + 1;
static_assert(kParams == CPV_MLP_PARAMS, "parameter layout");

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return 0x7FC0u;   // NaN
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}

__device__ __forceinline__ float bf(float f) { return __uint_as_float(f2bf(f) << 16); }

__device__ __forceinline__ v4s pack4(float a, float b, float c, float d)
{
    v4s r;
    r.x = (short)f2bf(a);
    r.y = (short)f2bf(b);
    r.z = (short)f2bf(c);
    r.w = (short)f2bf(d);
    return r;
}

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

__device__ __forceinline__ v4f zero4() { v4f z = {0.0f, 0.0f, 0.0f, 0.0f}; return z; }

// The weights as one wavefront's MFMA operands.
struct WaveWeights {
    v4s w1[4];        // [rb]: W1[4 (l/16) + e][16 rb + l%16] (0 for features >= 12)
    v4s w2t[4][4];    // [rb][kb]: W2[16 kb + 4 (l/16) + e][16 rb + l%16]
    v4s w3[4];        // [kb]: lanes l%16 == 0: w3[16 kb + 4 (l/16) + e], else 0
    float b1[4][4];   // [rb][e]: b1[16 rb + 4 (l/16) + e]
    float b2[4][4];
    float b3;
};

__device__ void load_weights(const float* P, int lane, WaveWeights& W)
{
    const int r = lane & 15, q = lane >> 4;
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
        float v[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int k = 4 * q + e;
            v[e] = k < kF ? P[oW1 + k * kH + 16 * rb + r] : 0.0f;
        }
        W.w1[rb] = pack4(v[0], v[1], v[2], v[3]);
#pragma unroll
        for (int kb = 0; kb < 4; kb++) {
#pragma unroll
            for (int e = 0; e < 4; e++) v[e] = P[oW2 + (16 * kb + 4 * q + e) * kH + 16 * rb + r];
            W.w2t[rb][kb] = pack4(v[0], v[1], v[2], v[3]);
        }
#pragma unroll
        for (int e = 0; e < 4; e++) v[e] = r == 0 ? P[ow3 + 16 * rb + 4 * q + e] : 0.0f;
        W.w3[rb] = pack4(v[0], v[1], v[2], v[3]);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            W.b1[rb][e] = P[ob1 + 16 * rb + 4 * q + e];
            W.b2[rb][e] = P[ob2 + 16 * rb + 4 * q + e];
        }
    }
    W.b3 = P[ob3];
}

struct Activations {
    v4f z1[4], z2[4];   // pre-activations (bias added), D layout
    v4s h1[4], h2[4];   // bf16 post-ReLU, D layout = next B operand
};

// Forward pass of 16 rows; xb is the B operand of X^T (lane l: row l%16,
// features 4 (l/16) .. +3).  Returns y of row l%16 in lanes 0-15.
__device__ float forward(const WaveWeights& W, v4s xb, Activations& A)
{
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
        v4f z = mfma(W.w1[rb], xb, zero4());
        z.x += W.b1[rb][0]; z.y += W.b1[rb][1]; z.z += W.b1[rb][2]; z.w += W.b1[rb][3];
        A.z1[rb] = z;
        A.h1[rb] = pack4(fmaxf(z.x, 0.0f), fmaxf(z.y, 0.0f), fmaxf(z.z, 0.0f), fmaxf(z.w, 0.0f));
    }
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
        v4f z = zero4();
#pragma unroll
        for (int kb = 0; kb < 4; kb++) z = mfma(W.w2t[rb][kb], A.h1[kb], z);
        z.x += W.b2[rb][0]; z.y += W.b2[rb][1]; z.z += W.b2[rb][2]; z.w += W.b2[rb][3];
        A.z2[rb] = z;
        A.h2[rb] = pack4(fmaxf(z.x, 0.0f), fmaxf(z.y, 0.0f), fmaxf(z.z, 0.0f), fmaxf(z.w, 0.0f));
    }
    v4f y = zero4();
#pragma unroll
    for (int kb = 0; kb < 4; kb++) y = mfma(W.w3[kb], A.h2[kb], y);
    return y.x + W.b3;
}

__device__ __forceinline__ v4s load_x(const float* X, const uint32_t* idx, uint32_t rows, uint32_t g, int lane, uint32_t* row_out)
{
    const uint32_t slot = g * 16u + (uint32_t)(lane & 15);
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t row = 0xFFFFFFFFu;
    if (slot < rows) {
        row = idx ? idx[slot] : slot;
        v = *(const float4*)(X + (uint64_t)row * kFeatureStride + 4 * (lane >> 4));
    }
    *row_out = row;
    return pack4(v.x, v.y, v.z, v.w);
}

// Inference: wavefronts stride over groups of 16 rows.
__global__ __launch_bounds__(256) void mlp_predict_kernel(const float* P, const float* X, uint32_t rows, float* out)
{
    const int lane = threadIdx.x & 63;
    WaveWeights W;
    load_weights(P, lane, W);
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    const uint32_t groups = (rows + 15u) / 16u;
    for (uint32_t g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); g < groups; g += waves) {
        uint32_t row;
        const v4s xb = load_x(X, nullptr, rows, g, lane, &row);
        Activations A;
        const float y = forward(W, xb, A);
        if (lane < 16 && row != 0xFFFFFFFFu) out[row] = y;
    }
}

constexpr int kGradWaves = 4;

// Forward and backward pass over `rows` rows (gathered through idx when
// given); accumulates the gradient of sum((y - t)^2) * inv_n into grads and
// the summed squared error into loss[0].  Every wavefront of a workgroup
// runs the same number of iterations (the LDS exchange uses workgroup
// barriers); out-of-range rows contribute nothing.
__global__ __launch_bounds__(64 * kGradWaves) void mlp_grad_kernel(const float* P, const float* X, const float* T, const uint32_t* idx,
                                                                   uint32_t rows, float inv_n, float* grads, float* loss)
{
    __shared__ uint16_t s_dz2[kGradWaves][kH][16];
    __shared__ uint16_t s_h1[kGradWaves][kH][16];
    __shared__ uint16_t s_dz1[kGradWaves][kH][16];
    __shared__ uint16_t s_x[kGradWaves][16][16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    WaveWeights W;
    load_weights(P, lane, W);
    // W2 as the A operand of dH1^T = W2 dZ2^T: [rb][kb] lane l: W2[16 rb + l%16][16 kb + 4 (l/16) + e]
    v4s w2n[4][4];
    float w3d[4][4];   // bf16-rounded w3 in D layout: w3[16 rb + 4 (l/16) + e]
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
#pragma unroll
        for (int kb = 0; kb < 4; kb++) {
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; e++) v[e] = P[oW2 + (16 * rb + r) * kH + 16 * kb + 4 * q + e];
            w2n[rb][kb] = pack4(v[0], v[1], v[2], v[3]);
        }
#pragma unroll
        for (int e = 0; e < 4; e++) w3d[rb][e] = bf(P[ow3 + 16 * rb + 4 * q + e]);
    }
    v4f gw2t[4][4], gw1t[4];
    float gb1[4][4], gb2[4][4], gw3[4][4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
        gw1t[rb] = zero4();
#pragma unroll
        for (int cb = 0; cb < 4; cb++) gw2t[rb][cb] = zero4();
#pragma unroll
        for (int e = 0; e < 4; e++) gb1[rb][e] = gb2[rb][e] = gw3[rb][e] = 0.0f;
    }
    float gb3 = 0.0f, gl = 0.0f;

    const uint32_t groups = (rows + 15u) / 16u;
    const uint32_t wg_waves = gridDim.x * kGradWaves;
    const uint32_t iters = (groups + wg_waves - 1) / wg_waves;
    for (uint32_t it = 0; it < iters; it++) {
        const uint32_t g = it * wg_waves + blockIdx.x * kGradWaves + wv;
        uint32_t row;
        const v4s xb = load_x(X, idx, g < groups ? rows : 0u, g, lane, &row);
        Activations A;
        const float y = forward(W, xb, A);
        // dL/dy of the tile this lane's column holds
        const float yj = __shfl(y, r, 64);
        const bool valid = row != 0xFFFFFFFFu;
        const float d = valid ? yj - T[row] : 0.0f;
        const float dy = 2.0f * d * inv_n;
        if (lane < 16) {
            gb3 += dy;
            gl += d * d;
        }
        v4f dz2[4], dz1[4];
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
            const float h2[4] = {__uint_as_float((uint32_t)(uint16_t)A.h2[rb].x << 16), __uint_as_float((uint32_t)(uint16_t)A.h2[rb].y << 16),
                                 __uint_as_float((uint32_t)(uint16_t)A.h2[rb].z << 16), __uint_as_float((uint32_t)(uint16_t)A.h2[rb].w << 16)};
            const float z2[4] = {A.z2[rb].x, A.z2[rb].y, A.z2[rb].z, A.z2[rb].w};
            float dz[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                gw3[rb][e] += h2[e] * dy;
                dz[e] = z2[e] > 0.0f ? w3d[rb][e] * dy : 0.0f;
                gb2[rb][e] += dz[e];
            }
            dz2[rb].x = dz[0]; dz2[rb].y = dz[1]; dz2[rb].z = dz[2]; dz2[rb].w = dz[3];
        }
        v4s dz2b[4];
#pragma unroll
        for (int kb = 0; kb < 4; kb++) dz2b[kb] = pack4(dz2[kb].x, dz2[kb].y, dz2[kb].z, dz2[kb].w);
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
            v4f h = zero4();
#pragma unroll
            for (int kb = 0; kb < 4; kb++) h = mfma(w2n[rb][kb], dz2b[kb], h);
            const float z1[4] = {A.z1[rb].x, A.z1[rb].y, A.z1[rb].z, A.z1[rb].w};
            const float dh[4] = {h.x, h.y, h.z, h.w};
            float dz[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                dz[e] = z1[e] > 0.0f ? dh[e] : 0.0f;
                gb1[rb][e] += dz[e];
            }
            dz1[rb].x = dz[0]; dz1[rb].y = dz[1]; dz1[rb].z = dz[2]; dz1[rb].w = dz[3];
        }
        // transposes through LDS: D layout (unit 16 rb + 4 q + e, tile r)
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
            const v4s h1 = A.h1[rb];
            const short hv[4] = {h1.x, h1.y, h1.z, h1.w};
            const v4s z2b = dz2b[rb];
            const short zv[4] = {z2b.x, z2b.y, z2b.z, z2b.w};
            const float z1v[4] = {dz1[rb].x, dz1[rb].y, dz1[rb].z, dz1[rb].w};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int u = 16 * rb + 4 * q + e;
                s_h1[wv][u][r] = (uint16_t)hv[e];
                s_dz2[wv][u][r] = (uint16_t)zv[e];
                s_dz1[wv][u][r] = (uint16_t)f2bf(z1v[e]);
            }
        }
        {
            const short xv[4] = {xb.x, xb.y, xb.z, xb.w};
#pragma unroll
            for (int e = 0; e < 4; e++) s_x[wv][r][4 * q + e] = (uint16_t)xv[e];   // s_x[tile][feature]
        }
        __syncthreads();
        // dW2^T[u][k] += sum_t dZ2^T[u][t] H1^T[k][t]: A lane l = dZ2^T[16 rb + l%16][t = 4 (l/16) + e],
        // B lane l = H1^T[16 cb + l%16][t = 4 (l/16) + e]
        v4s a_dz2[4], b_h1[4], a_dz1[4];
#pragma unroll
        for (int blk = 0; blk < 4; blk++) {
            const int u = 16 * blk + r;
            a_dz2[blk] = *(const v4s*)&s_dz2[wv][u][4 * q];
            b_h1[blk] = *(const v4s*)&s_h1[wv][u][4 * q];
            a_dz1[blk] = *(const v4s*)&s_dz1[wv][u][4 * q];
        }
        // B operand of dW1^T: X[t = 4 (l/16) + e][f = l%16]
        v4s b_x;
        b_x.x = (short)s_x[wv][4 * q + 0][r];
        b_x.y = (short)s_x[wv][4 * q + 1][r];
        b_x.z = (short)s_x[wv][4 * q + 2][r];
        b_x.w = (short)s_x[wv][4 * q + 3][r];
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
#pragma unroll
            for (int cb = 0; cb < 4; cb++) gw2t[rb][cb] = mfma(a_dz2[rb], b_h1[cb], gw2t[rb][cb]);
            gw1t[rb] = mfma(a_dz1[rb], b_x, gw1t[rb]);
        }
        __syncthreads();
    }

    // ---- accumulate into the global gradient
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
        const float w1v[4] = {gw1t[rb].x, gw1t[rb].y, gw1t[rb].z, gw1t[rb].w};
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int u = 16 * rb + 4 * q + e;   // D row = unit, column = feature r
            if (r < kF) atomicAdd(&grads[oW1 + r * kH + u], w1v[e]);
        }
#pragma unroll
        for (int cb = 0; cb < 4; cb++) {
            const float v[4] = {gw2t[rb][cb].x, gw2t[rb][cb].y, gw2t[rb][cb].z, gw2t[rb][cb].w};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int u = 16 * rb + 4 * q + e;   // h2 unit
                const int k = 16 * cb + r;           // h1 unit
                atomicAdd(&grads[oW2 + k * kH + u], v[e]);
            }
        }
#pragma unroll
        for (int e = 0; e < 4; e++) {
            float a = gb1[rb][e], b = gb2[rb][e], c = gw3[rb][e];
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                a += __shfl_xor(a, off, 64);
                b += __shfl_xor(b, off, 64);
                c += __shfl_xor(c, off, 64);
            }
            if (r == 0) {
                const int u = 16 * rb + 4 * q + e;
                atomicAdd(&grads[ob1 + u], a);
                atomicAdd(&grads[ob2 + u], b);
                atomicAdd(&grads[ow3 + u], c);
            }
        }
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        gb3 += __shfl_xor(gb3, off, 64);
        gl += __shfl_xor(gl, off, 64);
    }
    if (lane == 0) {
        atomicAdd(&grads[ob3], gb3);
        atomicAdd(loss, gl);
    }
}

// Adam (Kingma & Ba) on the fp32 master parameters; the step count t
// starts at 1.
__global__ __launch_bounds__(256) void mlp_adam_kernel(float* P, const float* G, float* m, float* v, float lr, float c1, float c2)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kParams) return;
    const float g = G[i];
    const float mi = 0.9f * m[i] + 0.1f * g;
    const float vi = 0.999f * v[i] + 0.001f * g * g;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
}

// Minibatch row indices from a counter-based generator (splitmix64 of
// seed, step and slot).
__global__ __launch_bounds__(256) void mlp_sample_kernel(uint32_t* idx, uint32_t n, uint32_t rows, uint64_t seed, uint32_t step)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)step * 0x100000001ull + i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    idx[i] = (uint32_t)(z % rows);
}

}  // namespace

struct cpv_mlp {
    cpv_context* c = nullptr;
    float* params = nullptr;   // kParams fp32 master weights
    float* grads = nullptr;    // kParams + 1 (last: summed squared error)
    float* m = nullptr;
    float* v = nullptr;
    uint32_t steps_done = 0;
    std::mutex lock;
};

namespace {

// He-uniform initialisation from splitmix64 (host).
void init_params(uint64_t seed, std::vector<float>& p)
{
    p.assign(kParams, 0.0f);
    uint64_t s = seed;
    auto next = [&]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    const double a1 = std::sqrt(6.0 / kF), a2 = std::sqrt(6.0 / kH);
    for (int i = 0; i < kF * kH; i++) p[oW1 + i] = (float)((next() * 2.0 - 1.0) * a1);
    for (int i = 0; i < kH * kH; i++) p[oW2 + i] = (float)((next() * 2.0 - 1.0) * a2);
    for (int i = 0; i < kH; i++) p[ow3 + i] = (float)((next() * 2.0 - 1.0) * a2);
    for (int i = 0; i < kH; i++) p[ob1 + i] = 0.01f;
    for (int i = 0; i < kH; i++) p[ob2 + i] = 0.01f;
}

// Feature rows of the C-ABI (12 floats) padded to the device stride.
std::vector<float> pad_rows(const float* x, uint32_t rows)
{
    std::vector<float> out((size_t)rows * kFeatureStride, 0.0f);
    for (uint32_t i = 0; i < rows; i++) std::memcpy(&out[(size_t)i * kFeatureStride], x + (size_t)i * kF, kF * sizeof(float));
    return out;
}

struct DeviceRows {
    float* x = nullptr;
    float* t = nullptr;
    ~DeviceRows()
    {
        if (x) (void)hipFree(x);
        if (t) (void)hipFree(t);
    }
};

int upload_rows(const float* x, const float* t, uint32_t rows, DeviceRows* d, hipStream_t s)
{
    const std::vector<float> xp = pad_rows(x, rows);
    HIP_TRY(hipMalloc(&d->x, std::max<size_t>(xp.size(), 16) * 4));
    HIP_TRY(hipMemcpyAsync(d->x, xp.data(), xp.size() * 4, hipMemcpyHostToDevice, s));
    if (t) {
        HIP_TRY(hipMalloc(&d->t, std::max<size_t>(rows, 1) * 4));
        HIP_TRY(hipMemcpyAsync(d->t, t, (size_t)rows * 4, hipMemcpyHostToDevice, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    return CPV_OK;
}

int grad_blocks(uint32_t rows)
{
    const uint32_t groups = (rows + 15u) / 16u;
    // a few groups per wavefront amortise the weight loads and atomics
    return (int)std::max<uint32_t>(1, std::min<uint32_t>(256, (groups + 4 * kGradWaves - 1) / (4 * kGradWaves)));
}

void enqueue_grad(cpv_mlp* M, const float* x, const float* t, const uint32_t* idx, uint32_t rows, float inv_n, hipStream_t s)
{
    fill_u32(s, M->grads, 0u, kParams + 1);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(grad_blocks(rows)), dim3(64 * kGradWaves), 0, s, (const float*)M->params, x, t, idx,
                       rows, inv_n, M->grads, M->grads + kParams);
}

}  // namespace

int cpvrt::mlp_enqueue_predict(cpv_mlp* m, const float* features, uint32_t tiles, float* out, hipStream_t s)
{
    if (!m) return fail(CPV_ERROR_INVALID, "no predictor");
    const uint32_t groups = (tiles + 15u) / 16u;
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(2048, (groups + 3) / 4));
    hipLaunchKernelGGL(mlp_predict_kernel, dim3(blocks), dim3(256), 0, s, (const float*)m->params, features, tiles, out);
    HIP_TRY(hipGetLastError());
    return CPV_OK;
}

extern "C" int cpv_mlp_create(cpv_context* c, uint64_t seed, cpv_mlp** out)
{
    if (!c || !out) return fail(CPV_ERROR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    cpv_mlp* M = new cpv_mlp();
    M->c = c;
    std::vector<float> p;
    init_params(seed, p);
    if (hipMalloc(&M->params, kParams * 4) != hipSuccess || hipMalloc(&M->grads, (kParams + 1) * 4) != hipSuccess ||
        hipMalloc(&M->m, kParams * 4) != hipSuccess || hipMalloc(&M->v, kParams * 4) != hipSuccess) {
        cpv_mlp_destroy(M);
        return fail(CPV_ERROR_OUT_OF_MEMORY, "predictor buffers");
    }
    if (hipMemcpy(M->params, p.data(), kParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(M->m, 0, kParams * 4) != hipSuccess || hipMemset(M->v, 0, kParams * 4) != hipSuccess) {
        cpv_mlp_destroy(M);
        return fail(CPV_ERROR_DEVICE, "predictor initialisation");
    }
    *out = M;
    return CPV_OK;
}

extern "C" void cpv_mlp_destroy(cpv_mlp* M)
{
    if (!M) return;
    (void)hipSetDevice(M->c->device);
    (void)cpvrt::sync_context(M->c);
    if (M->params) (void)hipFree(M->params);
    if (M->grads) (void)hipFree(M->grads);
    if (M->m) (void)hipFree(M->m);
    if (M->v) (void)hipFree(M->v);
    delete M;
}

extern "C" int cpv_mlp_get_params(cpv_mlp* M, float* params, uint32_t count)
{
    if (!M || !params) return fail(CPV_ERROR_INVALID, "null argument");
    if (count != kParams) return fail(CPV_ERROR_INVALID, "expected %d parameters", kParams);
    HIP_TRY(hipSetDevice(M->c->device));
    HIP_TRY(cpvrt::sync_context(M->c));
    HIP_TRY(hipMemcpy(params, M->params, kParams * 4, hipMemcpyDeviceToHost));
    return CPV_OK;
}

extern "C" int cpv_mlp_set_params(cpv_mlp* M, const float* params, uint32_t count)
{
    if (!M || !params) return fail(CPV_ERROR_INVALID, "null argument");
    if (count != kParams) return fail(CPV_ERROR_INVALID, "expected %d parameters", kParams);
    HIP_TRY(hipSetDevice(M->c->device));
    HIP_TRY(cpvrt::sync_context(M->c));
    HIP_TRY(hipMemcpy(M->params, params, kParams * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(M->m, 0, kParams * 4));
    HIP_TRY(hipMemset(M->v, 0, kParams * 4));
    M->steps_done = 0;
    return CPV_OK;
}

extern "C" int cpv_mlp_predict(cpv_mlp* M, const float* x, uint32_t rows, float* y)
{
    if (!M || (rows && (!x || !y))) return fail(CPV_ERROR_INVALID, "null argument");
    if (rows == 0) return CPV_OK;
    std::lock_guard<std::mutex> guard(M->lock);
    HIP_TRY(hipSetDevice(M->c->device));
    HIP_TRY(cpvrt::sync_context(M->c));   // no in-flight batch reads the weights while they change
    hipStream_t s = M->c->stream;
    DeviceRows d;
    int rc = upload_rows(x, nullptr, rows, &d, s);
    if (rc) return rc;
    float* out = nullptr;
    HIP_TRY(hipMalloc(&out, (size_t)rows * 4));
    rc = mlp_enqueue_predict(M, d.x, rows, out, s);
    hipError_t e = hipMemcpyAsync(y, out, (size_t)rows * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(out);
    if (rc) return rc;
    if (e != hipSuccess) return fail(CPV_ERROR_DEVICE, "predict: %s", hipGetErrorString(e));
    return CPV_OK;
}

extern "C" int cpv_mlp_gradients(cpv_mlp* M, const float* x, const float* target, uint32_t rows, float* grads, float* loss)
{
    if (!M || !x || !target || !grads || rows == 0) return fail(CPV_ERROR_INVALID, "null argument or no rows");
    std::lock_guard<std::mutex> guard(M->lock);
    HIP_TRY(hipSetDevice(M->c->device));
    HIP_TRY(cpvrt::sync_context(M->c));   // no in-flight batch reads the weights while they change
    hipStream_t s = M->c->stream;
    DeviceRows d;
    int rc = upload_rows(x, target, rows, &d, s);
    if (rc) return rc;
    enqueue_grad(M, d.x, d.t, nullptr, rows, 1.0f / (float)rows, s);
    HIP_TRY(hipGetLastError());
    std::vector<float> g(kParams + 1);
    HIP_TRY(hipMemcpyAsync(g.data(), M->grads, g.size() * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::memcpy(grads, g.data(), kParams * 4);
    if (loss) *loss = g[kParams] / (float)rows;
    return CPV_OK;
}

extern "C" int cpv_mlp_train(cpv_mlp* M, const float* x, const float* target, uint32_t rows, uint32_t steps, uint32_t batch,
                             float learning_rate, uint64_t seed, float* loss)
{
    if (!M || !x || !target || rows == 0 || batch == 0) return fail(CPV_ERROR_INVALID, "null argument, no rows or empty batch");
    std::lock_guard<std::mutex> guard(M->lock);
    HIP_TRY(hipSetDevice(M->c->device));
    HIP_TRY(cpvrt::sync_context(M->c));   // no in-flight batch reads the weights while they change
    hipStream_t s = M->c->stream;
    DeviceRows d;
    int rc = upload_rows(x, target, rows, &d, s);
    if (rc) return rc;
    uint32_t* idx = nullptr;
    HIP_TRY(hipMalloc(&idx, (size_t)batch * 4));
    const float inv_b = 1.0f / (float)batch;
    for (uint32_t k = 0; k < steps; k++) {
        hipLaunchKernelGGL(mlp_sample_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, idx, batch, rows, seed, k);
        enqueue_grad(M, d.x, d.t, idx, batch, inv_b, s);
        const uint32_t t = ++M->steps_done;
        const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)t)));
        const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)t)));
        hipLaunchKernelGGL(mlp_adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, s, M->params, (const float*)M->grads, M->m,
                           M->v, learning_rate, c1, c2);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && loss) {
        // mean squared error over all rows with the trained weights
        enqueue_grad(M, d.x, d.t, nullptr, rows, 1.0f / (float)rows, s);
        float sq = 0.0f;
        e = hipMemcpyAsync(&sq, M->grads + kParams, 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        *loss = sq / (float)rows;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(idx);
    if (e != hipSuccess) return fail(CPV_ERROR_DEVICE, "train: %s", hipGetErrorString(e));
    return CPV_OK;
}
