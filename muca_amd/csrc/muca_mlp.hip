// muca_mlp.hip -- the weight surrogate: an MLP 4 -> 128 -> 128 -> 128 -> 1
// (SiLU) that predicts ln W(D)/N from (D/N, 1/L, beta, Delta) features, trained
// on converged weights and used to seed new weight runs (the init-weight
// boundary of multigpu_weight's `continuing` mode, muca_delta_multigpu.cu:223-235).
//
// Arithmetic: FP32 master parameters; every product uses BF16-rounded
// weights; the two 128 x 128 layers (forward, input-gradient and weight-
// gradient products) run on MFMA (v_mfma_f32_32x32x16_bf16) with BF16
// operands (activations and weights) and FP32 accumulation.  The 4-wide input
// layer and the 128 -> 1 output layer are VALU work in FP32 with BF16-rounded
// weights.  Adam on the FP32 masters.
//
// One workgroup (256 lanes = 4 wavefronts) owns a tile of 64 samples and runs
// the forward pass and, in training, the backward pass entirely in LDS; its
// parameter gradients go to its own slot of a partial-gradient buffer, which a
// second kernel sums in a fixed order (deterministic) before the Adam step.
#include "../../include/muca.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr int kH = 128;           // hidden width
constexpr int kIn = 4;            // input features
constexpr int kTile = 64;         // samples per workgroup
constexpr int kThreads = 256;

// parameter vector layout (FP32 masters, gradients, Adam moments)
constexpr int oW1 = 0;                       // [128][4]
constexpr int oB1 = oW1 + kH * kIn;          // [128]
constexpr int oW2 = oB1 + kH;                // [128][128], W2[o][k]
constexpr int oB2 = oW2 + kH * kH;
constexpr int oW3 = oB2 + kH;
constexpr int oB3 = oW3 + kH * kH;
constexpr int oW4 = oB3 + kH;                // [128]
constexpr int oB4 = oW4 + kH;                // [1]
constexpr int kParams = oB4 + 1;             // 33793

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// float -> bf16 bits, round to nearest even (inputs are finite)
__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t b)
{
    const uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ inline float silu(float z)
{
    return z / (1.0f + expf(-z));
}

__device__ inline float silu_grad(float z)
{
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// BF16 copies of the weights used by the products
struct MlpBf16 {
    uint16_t *w1;    // [128][4]
    uint16_t *w2;    // [128][128]  W2[o][k]
    uint16_t *w2t;   // [128][128]  W2[o][k] at [k][o]
    uint16_t *w3, *w3t;
    uint16_t *w4;    // [128]
};

struct TileArgs {
    const float *x;        // [n][4]
    const float *t;        // [n] targets (training)
    const float *w;        // [n] sample weights of the squared error (nullable: 1)
    float *y;              // [n] predictions (nullable)
    const float *p;        // FP32 parameters (biases are read from here)
    MlpBf16 wb;
    float *part;           // [tiles][kParams + 1] partial gradients and loss
    uint32_t n;            // samples
    float dy_scale;        // 2 / batch
};

// One 32 x 32 output tile of a wavefront: acc += A(r0 + i, k) * B(k, c0 + j)
// over k in [0, K), K % 16 == 0.  Operands are BF16 read through accessors.
// Lane l supplies A(r0 + l % 32, k0 + 8 (l / 32) + 0..7) and
// B(k0 + 8 (l / 32) + 0..7, c0 + l % 32); it receives C rows
// 8 (i / 4) + 4 (l / 32) + i % 4, column l % 32, for i = 0..15.
template <typename FA, typename FB>
__device__ inline f32x16 mma_tile(FA a_at, FB b_at, int r0, int c0, int K)
{
    const int lane = threadIdx.x & 63;
    const int rc = lane & 31, kg = (lane >> 5) * 8;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += 16) {
        u16x8 av, bv;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            av[j] = a_at(r0 + rc, k0 + kg + j);
            bv[j] = b_at(k0 + kg + j, c0 + rc);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av),
                                                      __builtin_bit_cast(bf16x8, bv), acc, 0, 0,
                                                      0);
    }
    return acc;
}

__device__ inline int tile_row(int i)
{
    return 8 * (i >> 2) + 4 * ((threadIdx.x & 63) >> 5) + (i & 3);
}

__device__ inline int tile_col()
{
    return threadIdx.x & 31;
}

template <bool TRAIN>
__global__ __launch_bounds__(kThreads) void k_mlp_tile(TileArgs g)
{
    __shared__ float xs[kTile][kIn];
    __shared__ float ts[kTile];
    __shared__ float dys[kTile];
    __shared__ uint16_t a1[kTile][kH];       // bf16 silu(z1)
    __shared__ uint16_t a2[kTile][kH];       // bf16 silu(z2)
    __shared__ float z2[kTile][kH];
    __shared__ float z3[kTile][kH];          // z3, later dz1 (fp32)
    __shared__ uint16_t dz3b[kTile][kH];     // bf16 dz3
    __shared__ uint16_t dz2b[kTile][kH];     // bf16 dz2
    __shared__ float red[kThreads];
    __shared__ float sq[kTile];              // squared errors

    const int tid = threadIdx.x, wave = tid >> 6;
    const uint32_t s0 = blockIdx.x * kTile;
    const float *P = g.p;

    for (int e = tid; e < kTile * kIn; e += kThreads) {
        const int s = e / kIn, j = e % kIn;
        xs[s][j] = (s0 + s < g.n) ? g.x[(uint64_t)(s0 + s) * kIn + j] : 0.0f;
    }
    if (TRAIN && tid < kTile) ts[tid] = (s0 + tid < g.n) ? g.t[s0 + tid] : 0.0f;
    __syncthreads();

    // layer 1 (VALU): z1 = b1 + sum_j W1[k][j] x[j], in the order j = 0..3
    for (int e = tid; e < kTile * kH; e += kThreads) {
        const int s = e / kH, k = e % kH;
        float z = P[oB1 + k];
#pragma unroll
        for (int j = 0; j < kIn; j++) z = z + bf2f(g.wb.w1[k * kIn + j]) * xs[s][j];
        a1[s][k] = f2bf(silu(z));
    }
    __syncthreads();

    // layers 2 and 3 (MFMA): wave w owns rows 32 (w & 1) .. +31, columns 64 (w >> 1) .. +63
    const int r0 = 32 * (wave & 1), cb = 64 * (wave >> 1);
    {
        auto A = [&](int r, int k) { return a1[r][k]; };
        auto B = [&](int k, int o) { return g.wb.w2[o * kH + k]; };
        for (int h = 0; h < 2; h++) {
            const int c0 = cb + 32 * h;
            const f32x16 acc = mma_tile(A, B, r0, c0, kH);
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int r = r0 + tile_row(i), c = c0 + tile_col();
                const float z = acc[i] + P[oB2 + c];
                z2[r][c] = z;
                a2[r][c] = f2bf(silu(z));
            }
        }
    }
    __syncthreads();
    {
        auto A = [&](int r, int k) { return a2[r][k]; };
        auto B = [&](int k, int o) { return g.wb.w3[o * kH + k]; };
        for (int h = 0; h < 2; h++) {
            const int c0 = cb + 32 * h;
            const f32x16 acc = mma_tile(A, B, r0, c0, kH);
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int r = r0 + tile_row(i), c = c0 + tile_col();
                z3[r][c] = acc[i] + P[oB3 + c];
            }
        }
    }
    __syncthreads();

    // output layer (VALU): y = b4 + sum_k w4[k] silu(z3[k]); four lanes per
    // sample each sum 32 consecutive k, then the four partials in order
    {
        const int s = tid >> 2, q = tid & 3;
        float acc = 0.0f;
        for (int k = 32 * q; k < 32 * q + 32; k++) acc = acc + bf2f(g.wb.w4[k]) * silu(z3[s][k]);
        red[tid] = acc;
    }
    __syncthreads();
    if (tid < kTile) {
        const float y = P[oB4] + (((red[4 * tid] + red[4 * tid + 1]) + red[4 * tid + 2]) +
                                  red[4 * tid + 3]);
        if (g.y && s0 + tid < g.n) g.y[s0 + tid] = y;
        if (TRAIN) {
            const bool valid = s0 + tid < g.n;
            const float wt = (valid && g.w) ? g.w[s0 + tid] : 1.0f;
            const float d = y - ts[tid];
            dys[tid] = valid ? g.dy_scale * (wt * d) : 0.0f;
            sq[tid] = valid ? wt * (d * d) : 0.0f;
        }
    }
    if (!TRAIN) return;
    __syncthreads();
    float *part = g.part + (uint64_t)blockIdx.x * (kParams + 1);
    if (tid == 0) {   // loss and db4, summed over the tile in order
        float l = 0.0f, db = 0.0f;
        for (int s = 0; s < kTile; s++) {
            l = l + sq[s];
            db = db + dys[s];
        }
        part[kParams] = l;
        part[oB4] = db;
    }
    // dw4[k] = sum_s dy[s] silu(z3[s][k]); dz3 = dy w4 silu'(z3); db3 = sum_s dz3
    if (tid < kH) {
        const int k = tid;
        const float w4 = bf2f(g.wb.w4[k]);
        float dw = 0.0f, db = 0.0f;
        for (int s = 0; s < kTile; s++) {
            const float z = z3[s][k];
            dw = dw + dys[s] * silu(z);
            const float dz = (dys[s] * w4) * silu_grad(z);
            db = db + dz;
            dz3b[s][k] = f2bf(dz);
        }
        part[oW4 + k] = dw;
        part[oB3 + k] = db;
    }
    __syncthreads();
    // dW3[o][k] = sum_s dz3[s][o] a2[s][k]: 128 x 128 = 16 tiles, four per wave
    {
        auto A = [&](int o, int s) { return dz3b[s][o]; };
        auto B = [&](int s, int k) { return a2[s][k]; };
        for (int tt = 0; tt < 4; tt++) {
            const int tix = 4 * wave + tt, o0 = 32 * (tix >> 2), k0 = 32 * (tix & 3);
            const f32x16 acc = mma_tile(A, B, o0, k0, kTile);
#pragma unroll
            for (int i = 0; i < 16; i++)
                part[oW3 + (o0 + tile_row(i)) * kH + k0 + tile_col()] = acc[i];
        }
    }
    // da2[s][k] = sum_o dz3[s][o] W3[o][k]; dz2 = da2 silu'(z2)
    {
        auto A = [&](int s, int o) { return dz3b[s][o]; };
        auto B = [&](int o, int k) { return g.wb.w3t[k * kH + o]; };
        for (int h = 0; h < 2; h++) {
            const int c0 = cb + 32 * h;
            const f32x16 acc = mma_tile(A, B, r0, c0, kH);
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int r = r0 + tile_row(i), c = c0 + tile_col();
                z2[r][c] = acc[i] * silu_grad(z2[r][c]);   // z2 now holds dz2 (fp32)
            }
        }
    }
    __syncthreads();
    if (tid < kH) {
        const int k = tid;
        float db = 0.0f;
        for (int s = 0; s < kTile; s++) {
            const float dz = z2[s][k];
            db = db + dz;
            dz2b[s][k] = f2bf(dz);
        }
        part[oB2 + k] = db;
    }
    __syncthreads();
    // dW2[o][k] = sum_s dz2[s][o] a1[s][k]
    {
        auto A = [&](int o, int s) { return dz2b[s][o]; };
        auto B = [&](int s, int k) { return a1[s][k]; };
        for (int tt = 0; tt < 4; tt++) {
            const int tix = 4 * wave + tt, o0 = 32 * (tix >> 2), k0 = 32 * (tix & 3);
            const f32x16 acc = mma_tile(A, B, o0, k0, kTile);
#pragma unroll
            for (int i = 0; i < 16; i++)
                part[oW2 + (o0 + tile_row(i)) * kH + k0 + tile_col()] = acc[i];
        }
    }
    // da1[s][k] = sum_o dz2[s][o] W2[o][k]; dz1 = da1 silu'(z1), z1 recomputed
    {
        auto A = [&](int s, int o) { return dz2b[s][o]; };
        auto B = [&](int o, int k) { return g.wb.w2t[k * kH + o]; };
        for (int h = 0; h < 2; h++) {
            const int c0 = cb + 32 * h;
            const f32x16 acc = mma_tile(A, B, r0, c0, kH);
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int r = r0 + tile_row(i), c = c0 + tile_col();
                float z = P[oB1 + c];
#pragma unroll
                for (int j = 0; j < kIn; j++) z = z + bf2f(g.wb.w1[c * kIn + j]) * xs[r][j];
                z3[r][c] = acc[i] * silu_grad(z);   // z3 now holds dz1 (fp32)
            }
        }
    }
    __syncthreads();
    // dW1[k][j] = sum_s dz1[s][k] x[s][j]; db1[k] = sum_s dz1[s][k]
    if (tid < kH) {
        const int k = tid;
        float db = 0.0f, dw[kIn] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int s = 0; s < kTile; s++) {
            const float dz = z3[s][k];
            db = db + dz;
#pragma unroll
            for (int j = 0; j < kIn; j++) dw[j] = dw[j] + dz * xs[s][j];
        }
        part[oB1 + k] = db;
#pragma unroll
        for (int j = 0; j < kIn; j++) part[oW1 + k * kIn + j] = dw[j];
    }
}

// Sums the tiles' partial gradients in tile order; optionally takes one Adam
// step and refreshes the BF16 copies.
__global__ __launch_bounds__(256) void k_mlp_reduce(const float *part, uint32_t tiles, float *grad,
                                                   float *loss, float *p, float *m, float *v,
                                                   MlpBf16 wb, int adam, float lr, float b1,
                                                   float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > kParams) return;
    float gsum = 0.0f;
    for (uint32_t t = 0; t < tiles; t++) gsum = gsum + part[(uint64_t)t * (kParams + 1) + i];
    if (i == kParams) {
        if (loss) *loss = gsum;
        return;
    }
    if (grad) grad[i] = gsum;
    if (!adam) return;
    const float mi = b1 * m[i] + (1.0f - b1) * gsum;
    const float vi = b2 * v[i] + (1.0f - b2) * gsum * gsum;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    const uint16_t h = f2bf(pi);
    if (i < oB1) {
        wb.w1[i - oW1] = h;
    } else if (i >= oW2 && i < oB2) {
        const int o = (i - oW2) / kH, k = (i - oW2) % kH;
        wb.w2[o * kH + k] = h;
        wb.w2t[k * kH + o] = h;
    } else if (i >= oW3 && i < oB3) {
        const int o = (i - oW3) / kH, k = (i - oW3) % kH;
        wb.w3[o * kH + k] = h;
        wb.w3t[k * kH + o] = h;
    } else if (i >= oW4 && i < oB4) {
        wb.w4[i - oW4] = h;
    }
}

__global__ void k_mlp_refresh_bf16(const float *p, MlpBf16 wb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    const uint16_t h = f2bf(p[i]);
    if (i < oB1) {
        wb.w1[i - oW1] = h;
    } else if (i >= oW2 && i < oB2) {
        const int o = (i - oW2) / kH, k = (i - oW2) % kH;
        wb.w2[o * kH + k] = h;
        wb.w2t[k * kH + o] = h;
    } else if (i >= oW3 && i < oB3) {
        const int o = (i - oW3) / kH, k = (i - oW3) % kH;
        wb.w3[o * kH + k] = h;
        wb.w3t[k * kH + o] = h;
    } else if (i >= oW4 && i < oB4) {
        wb.w4[i - oW4] = h;
    }
}

thread_local std::string g_merr;

int mfail(int code, const std::string &m)
{
    g_merr = m;
    return code;
}

#define MHIP(expr)                                                                      \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return mfail(MUCA_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

struct muca_mlp {
    int device = 0;
    hipStream_t st = nullptr;
    float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *loss = nullptr;
    float *part = nullptr;
    uint32_t part_tiles = 0;
    uint16_t *bf = nullptr;   // all BF16 copies in one allocation
    MlpBf16 wb{};
    float *x = nullptr, *t = nullptr, *y = nullptr;   // resident training data / scratch
    float *w = nullptr;                               // resident sample weights
    bool weighted = false;
    uint32_t n = 0, cap = 0;
    uint64_t step = 0;
};

namespace {

void mlp_free(muca_mlp *g)
{
    if (!g) return;
    (void)hipSetDevice(g->device);
    void *bufs[] = {g->p, g->m, g->v, g->grad, g->loss, g->part, g->bf, g->x, g->t, g->y, g->w};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (g->st) (void)hipStreamDestroy(g->st);
    delete g;
}

hipError_t ensure_part(muca_mlp *g, uint32_t tiles)
{
    if (tiles <= g->part_tiles) return hipSuccess;
    if (g->part) (void)hipFree(g->part);
    g->part = nullptr;
    g->part_tiles = 0;
    hipError_t e = hipMalloc((void **)&g->part, (size_t)tiles * (kParams + 1) * sizeof(float));
    if (e == hipSuccess) g->part_tiles = tiles;
    return e;
}

hipError_t ensure_data(muca_mlp *g, uint32_t n)
{
    if (n <= g->cap) return hipSuccess;
    for (float **b : {&g->x, &g->t, &g->y, &g->w})
        if (*b) {
            (void)hipFree(*b);
            *b = nullptr;
        }
    g->cap = 0;
    hipError_t e = hipMalloc((void **)&g->x, (size_t)n * kIn * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&g->t, (size_t)n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&g->y, (size_t)n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&g->w, (size_t)n * sizeof(float));
    if (e == hipSuccess) g->cap = n;
    return e;
}

TileArgs tile_args(muca_mlp *g, const float *x, const float *t, const float *w, float *y,
                   uint32_t n, float dy_scale)
{
    TileArgs a{};
    a.x = x;
    a.t = t;
    a.w = w;
    a.y = y;
    a.p = g->p;
    a.wb = g->wb;
    a.part = g->part;
    a.n = n;
    a.dy_scale = dy_scale;
    return a;
}

}  // namespace

extern "C" {

const char *muca_mlp_last_error(void)
{
    return g_merr.c_str();
}

uint32_t muca_mlp_param_count(void)
{
    return kParams;
}

int muca_mlp_create(int32_t device, muca_mlp **out)
{
    if (!out) return mfail(MUCA_E_ARG, "NULL output");
    *out = nullptr;
    int ndev = 0;
    MHIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return mfail(MUCA_E_ARG, "device not present");
    MHIP(hipSetDevice(device));
    muca_mlp *g = new muca_mlp();
    g->device = device;
    hipError_t e = hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&g->p, kParams * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&g->m, kParams * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&g->v, kParams * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&g->grad, kParams * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&g->loss, sizeof(float));
    const size_t nbf = kH * kIn + 4 * kH * kH + kH;
    if (e == hipSuccess) e = hipMalloc((void **)&g->bf, nbf * sizeof(uint16_t));
    if (e != hipSuccess) {
        mlp_free(g);
        return mfail(MUCA_E_NOMEM, std::string("mlp setup: ") + hipGetErrorString(e));
    }
    g->wb.w1 = g->bf;
    g->wb.w2 = g->wb.w1 + kH * kIn;
    g->wb.w2t = g->wb.w2 + kH * kH;
    g->wb.w3 = g->wb.w2t + kH * kH;
    g->wb.w3t = g->wb.w3 + kH * kH;
    g->wb.w4 = g->wb.w3t + kH * kH;
    *out = g;
    return MUCA_OK;
}

int muca_mlp_destroy(muca_mlp *g)
{
    mlp_free(g);
    return MUCA_OK;
}

int muca_mlp_set_params(muca_mlp *g, const float *params, uint32_t n)
{
    if (!g || !params || n != (uint32_t)kParams) return mfail(MUCA_E_ARG, "need kParams values");
    for (uint32_t i = 0; i < n; i++)
        if (!std::isfinite(params[i])) return mfail(MUCA_E_ARG, "parameters must be finite");
    MHIP(hipSetDevice(g->device));
    MHIP(hipMemcpy(g->p, params, kParams * sizeof(float), hipMemcpyHostToDevice));
    MHIP(hipMemset(g->m, 0, kParams * sizeof(float)));
    MHIP(hipMemset(g->v, 0, kParams * sizeof(float)));
    g->step = 0;
    hipLaunchKernelGGL(k_mlp_refresh_bf16, dim3((kParams + 255) / 256), dim3(256), 0, g->st, g->p,
                       g->wb);
    MHIP(hipGetLastError());
    MHIP(hipStreamSynchronize(g->st));
    return MUCA_OK;
}

int muca_mlp_get_params(muca_mlp *g, float *params, uint32_t n)
{
    if (!g || !params || n != (uint32_t)kParams) return mfail(MUCA_E_ARG, "need kParams values");
    MHIP(hipSetDevice(g->device));
    MHIP(hipStreamSynchronize(g->st));
    MHIP(hipMemcpy(params, g->p, kParams * sizeof(float), hipMemcpyDeviceToHost));
    return MUCA_OK;
}

int muca_mlp_predict(muca_mlp *g, const float *x, uint32_t n, float *y)
{
    if (!g || !x || !y || n == 0) return mfail(MUCA_E_ARG, "bad arguments");
    MHIP(hipSetDevice(g->device));
    MHIP(ensure_data(g, n));
    MHIP(hipMemcpy(g->x, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice));
    const uint32_t tiles = (n + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_mlp_tile<false>, dim3(tiles), dim3(kThreads), 0, g->st,
                       tile_args(g, g->x, nullptr, nullptr, g->y, n, 0.0f));
    MHIP(hipGetLastError());
    MHIP(hipStreamSynchronize(g->st));
    MHIP(hipMemcpy(y, g->y, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    return MUCA_OK;
}

int muca_mlp_gradient(muca_mlp *g, const float *x, const float *t, uint32_t n, float *grad,
                      float *loss)
{
    if (!g || !x || !t || n == 0) return mfail(MUCA_E_ARG, "bad arguments");
    MHIP(hipSetDevice(g->device));
    MHIP(ensure_data(g, n));
    const uint32_t tiles = (n + kTile - 1) / kTile;
    MHIP(ensure_part(g, tiles));
    MHIP(hipMemcpy(g->x, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice));
    MHIP(hipMemcpy(g->t, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_tile<true>, dim3(tiles), dim3(kThreads), 0, g->st,
                       tile_args(g, g->x, g->t, nullptr, nullptr, n, 2.0f / (float)n));
    MHIP(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kParams + 1 + 255) / 256), dim3(256), 0, g->st, g->part,
                       tiles, g->grad, g->loss, g->p, g->m, g->v, g->wb, 0, 0.0f, 0.0f, 0.0f, 0.0f,
                       1.0f, 1.0f);
    MHIP(hipGetLastError());
    MHIP(hipStreamSynchronize(g->st));
    if (grad) MHIP(hipMemcpy(grad, g->grad, kParams * sizeof(float), hipMemcpyDeviceToHost));
    if (loss) {
        float l = 0;
        MHIP(hipMemcpy(&l, g->loss, sizeof(float), hipMemcpyDeviceToHost));
        *loss = l / (float)n;
    }
    return MUCA_OK;
}

int muca_mlp_set_data(muca_mlp *g, const float *x, const float *t, const float *w, uint32_t n)
{
    if (!g || !x || !t || n == 0) return mfail(MUCA_E_ARG, "bad arguments");
    if (w)
        for (uint32_t i = 0; i < n; i++)
            if (!(w[i] >= 0.0f) || !std::isfinite(w[i]))
                return mfail(MUCA_E_ARG, "sample weights must be finite and >= 0");
    MHIP(hipSetDevice(g->device));
    MHIP(ensure_data(g, n));
    MHIP(hipMemcpy(g->x, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice));
    MHIP(hipMemcpy(g->t, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    if (w) MHIP(hipMemcpy(g->w, w, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    g->weighted = w != nullptr;
    g->n = n;
    return MUCA_OK;
}

int muca_mlp_train(muca_mlp *g, uint32_t steps, uint32_t batch, float lr, float lr_final,
                   float *loss_out, float *gpu_ms)
{
    if (!g || steps == 0 || batch == 0 || batch % kTile != 0)
        return mfail(MUCA_E_ARG, "batch must be a positive multiple of 64");
    if (g->n < batch) return mfail(MUCA_E_STATE, "set at least `batch` samples first");
    MHIP(hipSetDevice(g->device));
    const uint32_t tiles = batch / kTile;
    MHIP(ensure_part(g, tiles));
    struct Events {   // destroyed on every return path
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    MHIP(hipEventCreate(&ev.e[0]));
    MHIP(hipEventCreate(&ev.e[1]));
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
    MHIP(hipEventRecord(e0, g->st));
    const uint32_t nbatches = g->n / batch;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (uint32_t s = 0; s < steps; s++) {
        // minibatches walk the resident samples in order (shuffled by the caller)
        const uint64_t off = (uint64_t)(s % nbatches) * batch;
        // cosine schedule from lr to lr_final
        const float lr_s =
            lr_final + 0.5f * (lr - lr_final) * (1.0f + cosf(3.14159265f * (float)s / (float)steps));
        g->step++;
        const float c1 = 1.0f - powf(b1, (float)g->step), c2 = 1.0f - powf(b2, (float)g->step);
        hipLaunchKernelGGL(k_mlp_tile<true>, dim3(tiles), dim3(kThreads), 0, g->st,
                           tile_args(g, g->x + off * kIn, g->t + off,
                                     g->weighted ? g->w + off : nullptr, nullptr, batch,
                                     2.0f / (float)batch));
        MHIP(hipGetLastError());
        hipLaunchKernelGGL(k_mlp_reduce, dim3((kParams + 1 + 255) / 256), dim3(256), 0, g->st,
                           g->part, tiles, nullptr, g->loss, g->p, g->m, g->v, g->wb, 1, lr_s, b1,
                           b2, eps, c1, c2);
        MHIP(hipGetLastError());
    }
    MHIP(hipEventRecord(e1, g->st));
    MHIP(hipEventSynchronize(e1));
    float ms = 0;
    MHIP(hipEventElapsedTime(&ms, e0, e1));
    if (gpu_ms) *gpu_ms = ms;
    if (loss_out) {
        float l = 0;
        MHIP(hipMemcpy(&l, g->loss, sizeof(float), hipMemcpyDeviceToHost));
        *loss_out = l / (float)batch;
    }
    return MUCA_OK;
}

}  // extern "C"
