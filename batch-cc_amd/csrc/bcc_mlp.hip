// Learned step ordering: obstacle encoder, pair features and a 3-layer BF16
// MLP (96 -> kHid -> kHid -> 32 logits, kHid = BCC_MLP_HIDDEN, 64 by default) trained and evaluated with hand-written
// MFMA GEMMs and elementwise kernels.  The model only proposes an order in
// which the exact checker visits the sample bins of a pair; verdicts never
// depend on it.
//
// Bins: sample k = t*n + i of a pair (src/batch_cc.cu:114-128) lies in bin
// t = k / n, the t-th 1/32 of the edge.  Labels are 32-bit masks of bins that
// contain at least one colliding sample.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bcc.h"
#include "bcc_mlp.h"

namespace bcc {

// ------------------------------------------------------------ encoder
// 64 cells of the box [-1,1] x [-1,1] x [0,1.5]; every primitive contributes
// the fraction of each cell its axis-aligned bounding box covers; the
// environment's code is the max over primitives (a max-pooled set encoding).
__global__ void encode_env_kernel(const float* __restrict__ obs, const int* __restrict__ meta, int num_envs,
                                  float* __restrict__ enc)
{
    const int env = blockIdx.x;
    const int cell = threadIdx.x;   // 0..63
    if (env >= num_envs || cell >= kEnc)
        return;
    const int cx = cell & 3, cy = (cell >> 2) & 3, cz = cell >> 4;
    const float lo[3] = {-1.0f + 0.5f * cx, -1.0f + 0.5f * cy, 0.375f * cz};
    const float hi[3] = {lo[0] + 0.5f, lo[1] + 0.5f, lo[2] + 0.375f};
    const float ext[3] = {0.5f, 0.5f, 0.375f};
    const int* m = meta + env * 16;
    float best = 0.0f;
    auto cover = [&](const float* bl, const float* bh) {
        float f = 1.0f;
        for (int a = 0; a < 3; ++a) {
            const float o = fminf(bh[a], hi[a]) - fmaxf(bl[a], lo[a]);
            f *= fmaxf(o, 0.0f) / ext[a];
        }
        best = fmaxf(best, f);
    };
    // reference lists: spheres[4], capsules[8], z-capsules[8], cuboids[16], z-cuboids[16]
    for (int k = 0; k < m[1]; ++k) {
        const float* o = obs + m[0] + 4 * k;
        const float bl[3] = {o[0] - o[3], o[1] - o[3], o[2] - o[3]};
        const float bh[3] = {o[0] + o[3], o[1] + o[3], o[2] + o[3]};
        cover(bl, bh);
    }
    for (int list = 0; list < 2; ++list) {
        const int off = m[2 + 2 * list], n = m[3 + 2 * list];
        for (int k = 0; k < n; ++k) {
            const float* c = obs + off + 8 * k;
            float bl[3], bh[3];
            for (int a = 0; a < 3; ++a) {
                const float p = c[a], q = c[a] + c[3 + a];
                bl[a] = fminf(p, q) - c[6];
                bh[a] = fmaxf(p, q) + c[6];
            }
            cover(bl, bh);
        }
    }
    for (int list = 0; list < 2; ++list) {
        const int off = m[6 + 2 * list], n = m[7 + 2 * list];
        for (int k = 0; k < n; ++k) {
            const float* b = obs + off + 16 * k;
            float bl[3], bh[3];
            for (int a = 0; a < 3; ++a) {
                const float r = fabsf(b[3 + a]) * b[12] + fabsf(b[6 + a]) * b[13] + fabsf(b[9 + a]) * b[14];
                bl[a] = b[a] - r;
                bh[a] = b[a] + r;
            }
            cover(bl, bh);
        }
    }
    enc[(size_t)env * kEnc + cell] = best;
}

// ------------------------------------------------------------ features
// x[p] = [enc(env) | (s - lo) * scale - 1 | (e - lo) * scale - 1 | 0 ...], scale = 2 / range
// as BF16, rows for pairs p0 .. p0 + num_pairs - 1.  Pair p: env =
// pair_env ? pair_env[p] : p % E, edge = pair_edge ? pair_edge[p] : p / E
// (dense, edge-major).
__global__ void features_kernel(const float* __restrict__ enc, const float* __restrict__ edges, int dim,
                                const float* __restrict__ lo, const float* __restrict__ scale,
                                const int* __restrict__ pair_env, const int* __restrict__ pair_edge, int num_envs,
                                long long p0, long long num_pairs, uint16_t* __restrict__ x)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= num_pairs * kIn)
        return;
    const long long pl = idx / kIn;
    const int c = (int)(idx - pl * kIn);
    const long long p = p0 + pl;   // pair index in the batch
    const int env = pair_env ? pair_env[p] : (int)(p % num_envs);
    const long long edge = pair_edge ? (long long)pair_edge[p] : p / num_envs;
    float v = 0.0f;
    if (c < kEnc) {
        v = enc[(size_t)env * kEnc + c];
    } else if (c < kEnc + 2 * dim) {
        v = edge_feature(edges, dim, lo, scale, edge, c - kEnc);
    }
    x[idx] = f2bf(v);
}

// ------------------------------------------------------------ elementwise
// h = act(z + b) as BF16 (act = ReLU when relu != 0); z is [rows][cols] FP32.
__global__ void bias_act_kernel(const float* __restrict__ z, const float* __restrict__ b, int rows, int cols,
                                int relu, uint16_t* __restrict__ h, float* __restrict__ zb)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * cols)
        return;
    float v = z[i] + b[i % cols];
    if (zb)
        zb[i] = v;
    if (relu)
        v = fmaxf(v, 0.0f);
    if (h)
        h[i] = f2bf(v);
}

// Binary cross-entropy with logits over 32 bins: g = (sigmoid(z) - y) / rows
// (BF16), and the per-row loss summed into loss[0].
__global__ void bce_grad_kernel(const float* __restrict__ logits, const uint32_t* __restrict__ labels, int rows,
                                uint16_t* __restrict__ g, float* __restrict__ loss)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < (long long)rows * kOut) {
        const long long r = i / kOut;
        const int c = (int)(i - r * kOut);
        const float z = logits[i];
        const float y = (float)((labels[r] >> c) & 1u);
        const float s = 1.0f / (1.0f + expf(-z));
        g[i] = f2bf((s - y) / (float)rows);
        l = fmaxf(z, 0.0f) - z * y + log1pf(expf(-fabsf(z)));
    }
    // wave reduction then one atomic per wave
    for (int off = 32; off > 0; off >>= 1)
        l += __shfl_down(l, off);
    if ((threadIdx.x & 63) == 0)
        atomicAdd(loss, l / (float)rows);
}

// g = bf16(d * (zb > 0)): ReLU backward (d FP32 from a GEMM)
__global__ void relu_back_kernel(const float* __restrict__ d, const float* __restrict__ zb, long long n,
                                 uint16_t* __restrict__ g)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        g[i] = f2bf(zb[i] > 0.0f ? d[i] : 0.0f);
}

// column sums of a BF16 [rows][cols] matrix into FP32 db[cols]
__global__ void colsum_kernel(const uint16_t* __restrict__ g, int rows, int cols, float* __restrict__ db)
{
    const int c = blockIdx.x;
    float s = 0.0f;
    for (int r = threadIdx.x; r < rows; r += blockDim.x)
        s += bf2f(g[(size_t)r * cols + c]);
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        db[c] = red[0];
}

// Adam on FP32 master parameters; writes the BF16 copy used by the GEMMs.
__global__ void adam_kernel(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, long long n, float lr, float b1, float b2, float eps,
                            float c1, float c2, uint16_t* __restrict__ wb)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb)
        wb[i] = f2bf(wi);
}

__global__ void to_bf16_kernel(const float* __restrict__ a, long long n, uint16_t* __restrict__ b)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        b[i] = f2bf(a[i]);
}

// ------------------------------------------------------------ fused inference
// Inference over a dense batch in one kernel, computed transposed,
// Y^T = W^T X^T, so that an MFMA result tile (lane l: rows 4 (l/16) + v,
// column l % 16) of two stacked row tiles is already the next layer's B
// operand for one 32-wide k block (see kslot in bcc_mlp.h) and activations
// never leave registers.  Columns are pairs, 16 per MFMA tile, 48 per wave
// (64 would need more than the 256 registers of two waves per SIMD).
// Weights are read as prepacked A fragments: fragment (rt, kt) of a layer
// with W [K][N] holds for lane l the eight values
// W[32 kt + kslot(l/16, i)][16 rt + l % 16], i = 0..7.  Every product, its
// slot and k-block order and the bias/ReLU/rounding steps are those of the
// layer-by-layer path (gemm_bf16_kernel + bias_act_kernel), so logits and
// orders are identical to it.
constexpr int kFuseCols = 3;                     // 16-pair column tiles per wave
constexpr int kFusePairs = 16 * kFuseCols;       // pairs per wave
constexpr int kFuseBlockPairs = 4 * kFusePairs;  // pairs per workgroup (four waves)
static_assert(kIn % 32 == 0 && kHid % 32 == 0 && kOut % 16 == 0, "layer widths");

// Weight staging: the workgroup's four waves share each weight fragment,
// copied once per workgroup into LDS instead of being read from L2 by every
// wave.  Stage j holds layer 2's row tiles 2 j and 2 j + 1 (2 x 8 contiguous
// fragments) and layer 3's fragments at k block j, which consumes exactly
// those two row tiles (layers 2 and 3 hold three quarters of the weights).
constexpr int kStageFrags = 2 * kKT2 + kRT3;                  // 18 fragments
constexpr int kStageLoads = (kStageFrags * 64 + 255) / 256;   // per thread

__device__ __forceinline__ void stage_load(const bf16x8* __restrict__ frag, int j, bf16x8 (&pre)[kStageLoads])
{
#pragma unroll
    for (int i = 0; i < kStageLoads; ++i) {
        const int e = threadIdx.x + 256 * i;
        if (e < 2 * kKT2 * 64)
            pre[i] = frag[(kFragL2 + 2 * j * kKT2) * 64 + e];
        else if (e < kStageFrags * 64)
            pre[i] = frag[(kFragL3 + ((e >> 6) - 2 * kKT2) * kKT2 + j) * 64 + (e & 63)];
    }
}

__device__ __forceinline__ void stage_store(bf16x8* buf, const bf16x8 (&pre)[kStageLoads])
{
#pragma unroll
    for (int i = 0; i < kStageLoads; ++i) {
        const int e = threadIdx.x + 256 * i;
        if (e < kStageFrags * 64)
            buf[e] = pre[i];
    }
}

__global__ void pack_fragments_kernel(const uint16_t* __restrict__ W1, const uint16_t* __restrict__ W2,
                                      const uint16_t* __restrict__ W3, uint16_t* __restrict__ frag)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;   // (fragment, lane, i)
    if (idx >= kFrags * 64 * 8)
        return;
    const int i = idx & 7;
    const int lane = (idx >> 3) & 63;
    const int f = idx >> 9;
    const uint16_t* W;
    int N, KT, f0;
    if (f < kFragL2) {
        W = W1, N = kHid, KT = kKT1, f0 = kFragL1;
    } else if (f < kFragL3) {
        W = W2, N = kHid, KT = kKT2, f0 = kFragL2;
    } else {
        W = W3, N = kOut, KT = kKT2, f0 = kFragL3;
    }
    const int rt = (f - f0) / KT, kt = (f - f0) % KT;
    frag[idx] = W[(size_t)(32 * kt + kslot(lane >> 4, i)) * N + 16 * rt + (lane & 15)];
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void fused_order_kernel(
    const bf16x8* __restrict__ frag, const float* __restrict__ b1, const float* __restrict__ b2,
    const float* __restrict__ b3, const float* __restrict__ enc, const float* __restrict__ edges, int dim,
    const float* __restrict__ lo, const float* __restrict__ scale, int num_envs, const int* __restrict__ pair_env,
    const int* __restrict__ pair_edge, long long num_pairs, uint32_t* __restrict__ order, float* __restrict__ logits)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int col = lane & 15;
    const int g = lane >> 4;
    const int kq = 4 * g;
    const long long base = ((long long)blockIdx.x * 4 + wave) * kFusePairs;

    // layer-1 B operands: X^T blocks (k blocks x column tiles)
    bf16x8 xb[kKT1][kFuseCols];
    bool pair_ok[kFuseCols];
#pragma unroll
    for (int ct = 0; ct < kFuseCols; ++ct) {
        const long long p = base + 16 * ct + col;
        const bool ok = p < num_pairs;
        int env = 0;
        long long edge = 0;
        if (ok) {   // same pair mapping as features_kernel
            env = pair_env ? pair_env[p] : (int)(p % num_envs);
            edge = pair_edge ? (long long)pair_edge[p] : (pair_env ? p : p / num_envs);
        }
        pair_ok[ct] = ok;
        bf16x8 x[kKT1];
        mlp_pair_operands(enc + (size_t)env * kEnc, edges, dim, lo, scale, edge, ok, x);
#pragma unroll
        for (int kt = 0; kt < kKT1; ++kt)
            xb[kt][ct] = x[kt];
    }

    // layer 1: Y1^T (256 x 48) kept as BF16 B operands of layer 2: row tiles
    // 2 j and 2 j + 1 fill slots 0-3 and 4-7 of k block j (its weights are
    // read directly: staging them too, inside this fully unrolled loop, costs
    // the kernel its second resident workgroup)
    u32x4 y1[kKT2][kFuseCols];
#pragma unroll
    for (int rt = 0; rt < kRT12; ++rt) {
        f32x4 acc[kFuseCols];
#pragma unroll
        for (int ct = 0; ct < kFuseCols; ++ct)
            acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < kKT1; ++kt) {
            const bf16x8 a = frag[(kFragL1 + rt * kKT1 + kt) * 64 + lane];
#pragma unroll
            for (int ct = 0; ct < kFuseCols; ++ct)
                acc[ct] = mfma32(a, xb[kt][ct], acc[ct]);
        }
        const f32x4 bb = *(const f32x4*)(b1 + 16 * rt + kq);
#pragma unroll
        for (int ct = 0; ct < kFuseCols; ++ct) {
            float h[4];
#pragma unroll
            for (int v = 0; v < 4; ++v)
                h[v] = fmaxf(acc[ct][v] + bb[v], 0.0f);
            y1[rt >> 1][ct][2 * (rt & 1) + 0] = f2bf2(h[0], h[1]);
            y1[rt >> 1][ct][2 * (rt & 1) + 1] = f2bf2(h[2], h[3]);
        }
    }

    // layers 2 and 3: stage j computes layer 2's row tiles 2 j and 2 j + 1,
    // which form layer 3's k block j; the stage's weights come through LDS,
    // copied by the whole workgroup (the loads' latency is covered by the
    // other resident workgroup's MFMAs)
    __shared__ bf16x8 wbuf[kStageFrags * 64];
    f32x4 acc3[kRT3][kFuseCols];
#pragma unroll
    for (int r3 = 0; r3 < kRT3; ++r3)
#pragma unroll
        for (int ct = 0; ct < kFuseCols; ++ct)
            acc3[r3][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int j = 0; j < kKT2; ++j) {
        {
            bf16x8 pre[kStageLoads];
            stage_load(frag, j, pre);
            __syncthreads();   // every wave is done with the previous stage
            stage_store(wbuf, pre);
        }
        __syncthreads();
        u32x4 y2[kFuseCols];
#pragma unroll
        for (int h2 = 0; h2 < 2; ++h2) {
            f32x4 acc[kFuseCols];
#pragma unroll
            for (int ct = 0; ct < kFuseCols; ++ct)
                acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < kKT2; ++kt) {
                const bf16x8 a = wbuf[(h2 * kKT2 + kt) * 64 + lane];
#pragma unroll
                for (int ct = 0; ct < kFuseCols; ++ct)
                    acc[ct] = mfma32(a, __builtin_bit_cast(bf16x8, y1[kt][ct]), acc[ct]);
            }
            const f32x4 bb = *(const f32x4*)(b2 + 16 * (2 * j + h2) + kq);
#pragma unroll
            for (int ct = 0; ct < kFuseCols; ++ct) {
                float h[4];
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    h[v] = fmaxf(acc[ct][v] + bb[v], 0.0f);
                y2[ct][2 * h2 + 0] = f2bf2(h[0], h[1]);
                y2[ct][2 * h2 + 1] = f2bf2(h[2], h[3]);
            }
        }
#pragma unroll
        for (int r3 = 0; r3 < kRT3; ++r3) {
            const bf16x8 a = wbuf[(2 * kKT2 + r3) * 64 + lane];
#pragma unroll
            for (int ct = 0; ct < kFuseCols; ++ct)
                acc3[r3][ct] = mfma32(a, __builtin_bit_cast(bf16x8, y2[ct]), acc3[r3][ct]);
        }
    }

    // logits (bias added as in the layer-by-layer path), then the order word
#pragma unroll
    for (int ct = 0; ct < kFuseCols; ++ct) {
        float z[8];
#pragma unroll
        for (int r3 = 0; r3 < kRT3; ++r3) {
            const f32x4 bb = *(const f32x4*)(b3 + 16 * r3 + kq);
#pragma unroll
            for (int v = 0; v < 4; ++v)
                z[4 * r3 + v] = acc3[r3][ct][v] + bb[v];
        }
        const uint32_t word = spread_top4_lanes(z, g);
        const long long p = base + 16 * ct + col;
        if (pair_ok[ct]) {
            if (g == 0)
                order[p] = word;
            if (logits)
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    logits[p * kOut + 16 * (e >> 2) + kq + (e & 3)] = z[e];
        }
    }
}

}  // namespace bcc

// ============================================================ host side

#include <cstdio>
#include <fstream>

#include "bcc_internal.h"

using bcc_detail::fail;
using bcc_detail::hip_fail;

namespace {

using namespace bcc;

// parameter blob layout (FP32 master, same offsets for the BF16 copy,
// gradients and Adam moments)
constexpr long long P_W1 = 0;
constexpr long long P_B1 = P_W1 + (long long)kIn * kHid;
constexpr long long P_W2 = P_B1 + kHid;
constexpr long long P_B2 = P_W2 + (long long)kHid * kHid;
constexpr long long P_W3 = P_B2 + kHid;
constexpr long long P_B3 = P_W3 + (long long)kHid * kOut;
constexpr long long P_N = P_B3 + kOut;

unsigned blocks_for(long long n, int bs = 256)
{
    return (unsigned)((n + bs - 1) / bs);
}

// Split-K workspace: enough for 32 partial kHid x kHid tiles.
constexpr long long kSplitWs = 32LL * kHid * kHid;

// C = alpha op(A) op(B) + beta C.  Long, thin products (K >= 1024 and fewer
// than 64 output tiles: the weight gradients, K = batch rows) are split along
// K into the workspace `ws` (kSplitWs floats) and reduced in split order, so
// results do not depend on scheduling.
hipError_t gemm(bool ta, bool tb, const uint16_t* A, int lda, const uint16_t* B, int ldb, float* C, int ldc, int M,
                int N, int K, float alpha, float beta, hipStream_t s, float* ws = nullptr)
{
    if (M <= 0 || N <= 0)
        return hipSuccess;
    const int tiles = ((N + kGemmTile - 1) / kGemmTile) * ((M + kGemmTile - 1) / kGemmTile);
    int splits = 1;
    if (ws && K >= 1024 && tiles < 64) {
        splits = std::min<long long>(std::min(32, K / 256), kSplitWs / ((long long)M * N));
        splits = std::max(splits, 1);
    }
    const int kchunk = splits > 1 ? ((K + splits - 1) / splits + kGemmK - 1) / kGemmK * kGemmK : std::max(K, 1);
    splits = (K + kchunk - 1) / kchunk;
    float* out = splits > 1 ? ws : C;
    const int ldo = splits > 1 ? N : ldc;
    const float a = splits > 1 ? 1.0f : alpha, b = splits > 1 ? 0.0f : beta;
    const long long cstride = splits > 1 ? (long long)M * N : 0;
    const dim3 grid((N + kGemmTile - 1) / kGemmTile, (M + kGemmTile - 1) / kGemmTile, std::max(splits, 1));
    if (!ta && !tb)
        hipLaunchKernelGGL((gemm_bf16_kernel<false, false>), grid, dim3(256), 0, s, A, lda, B, ldb, out, ldo, M, N, K,
                           a, b, kchunk, cstride);
    else if (ta && !tb)
        hipLaunchKernelGGL((gemm_bf16_kernel<true, false>), grid, dim3(256), 0, s, A, lda, B, ldb, out, ldo, M, N, K,
                           a, b, kchunk, cstride);
    else if (!ta && tb)
        hipLaunchKernelGGL((gemm_bf16_kernel<false, true>), grid, dim3(256), 0, s, A, lda, B, ldb, out, ldo, M, N, K,
                           a, b, kchunk, cstride);
    else
        hipLaunchKernelGGL((gemm_bf16_kernel<true, true>), grid, dim3(256), 0, s, A, lda, B, ldb, out, ldo, M, N, K,
                           a, b, kchunk, cstride);
    if (splits > 1)
        hipLaunchKernelGGL(splitk_reduce_kernel, dim3(blocks_for((long long)M * N)), dim3(256), 0, s, ws, splits,
                           cstride, M, N, C, ldc, alpha, beta);
    return hipGetLastError();
}

struct Lcg {   // splitmix64
    uint64_t x;
    uint64_t next()
    {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    float uniform() { return (float)((next() >> 11) * (1.0 / 9007199254740992.0)); }
};

int dispatch_label(int robot, const LabelArgs& a, hipStream_t s)
{
    hipError_t e;
    switch (robot) {
    case BCC_PANDA: e = launch_label_panda(a, s); break;
    case BCC_FETCH: e = launch_label_fetch(a, s); break;
    case BCC_BAXTER: e = launch_label_baxter(a, s); break;
    case BCC_XARM7: e = launch_label_xarm7(a, s); break;
    default: return fail(BCC_ERR_INVALID, "unknown robot id");
    }
    if (e != hipSuccess)
        return hip_fail(e, "label kernel launch");
    return BCC_OK;
}

}  // namespace

struct bcc_model {
    int robot = 0;
    int dim = 0;
    int step = 0;
    uint64_t seed = 0;
    float* w = nullptr;       // FP32 master parameters [P_N]
    uint16_t* wb = nullptr;   // BF16 copy
    float* g = nullptr;       // gradients
    float* m = nullptr;       // Adam first moment
    float* v = nullptr;       // Adam second moment
    float* lo = nullptr;      // joint lower limits [dim]
    float* scale = nullptr;   // 2 / joint range [dim]
    uint16_t* frag = nullptr; // MFMA A fragments of the BF16 weights (fused inference)
    float* enc = nullptr;     // environment codes of the last scene ordered [enc_envs][64]
    int enc_envs = 0;
    float* ws = nullptr;      // split-K partial tiles (weight gradients)
    int cap = 0;              // scratch rows
    uint16_t *h1 = nullptr, *h2 = nullptr, *g1 = nullptr, *g2 = nullptr, *g3 = nullptr;
    float *z1 = nullptr, *z2 = nullptr, *z3 = nullptr, *d1 = nullptr, *d2 = nullptr, *loss = nullptr;
};

namespace {

static const float kLo[4][16] = {
    {-2.9671f, -1.8326f, -2.9671f, -3.1416f, -2.9671f, -0.0873f, -2.9671f},
    {0.0f, -1.6056f, -1.221f, -3.14159f, -2.251f, -3.14159f, -2.16f, -3.14159f},
    {-1.70167993878f, -2.147f, -3.05417993878f, -0.05f, -3.059f, -1.57079632679f, -3.059f, -1.70167993878f,
     -2.147f, -3.05417993878f, -0.05f, -3.059f, -1.57079632679f, -3.059f},
    {-2.8f, -1.7f, -2.8f, 0.0f, -2.8f, 0.0f, -2.8f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}};
static const float kRange[4][16] = {
    {5.9342f, 3.6652f, 5.9342f, 3.2289f, 5.9342f, 3.9095999999999997f, 5.9342f},
    {0.38615f, 3.2112f, 2.739f, 6.28318f, 4.502f, 6.28318f, 4.32f, 6.28318f},
    {3.40335987756f, 3.194f, 6.10835987756f, 2.6679999999999997f, 6.118f, 3.66479632679f, 6.118f,
     3.40335987756f, 3.194f, 6.10835987756f, 2.6679999999999997f, 6.118f, 3.66479632679f, 6.118f},
    {5.6f, 3.4f, 5.6f, 3.0f, 5.6f, 3.1f, 5.6f, 0.85f, 0.85f, 0.85f, 0.85f, 0.85f, 0.85f}};

void free_scratch(bcc_model* M)
{
    for (void* p : {(void*)M->h1, (void*)M->h2, (void*)M->g1, (void*)M->g2, (void*)M->g3, (void*)M->z1,
                    (void*)M->z2, (void*)M->z3, (void*)M->d1, (void*)M->d2})
        (void)hipFree(p);
    M->h1 = M->h2 = M->g1 = M->g2 = M->g3 = nullptr;
    M->z1 = M->z2 = M->z3 = M->d1 = M->d2 = nullptr;
    M->cap = 0;
}

int ensure_scratch(bcc_model* M, int rows)
{
    if (rows <= M->cap)
        return BCC_OK;
    free_scratch(M);
    const size_t r = (size_t)rows;
    hipError_t e = hipSuccess;
    auto al = [&](void** p, size_t bytes) {
        if (e == hipSuccess)
            e = hipMalloc(p, bytes);
    };
    al((void**)&M->h1, r * kHid * 2);
    al((void**)&M->h2, r * kHid * 2);
    al((void**)&M->g1, r * kHid * 2);
    al((void**)&M->g2, r * kHid * 2);
    al((void**)&M->g3, r * kOut * 2);
    al((void**)&M->z1, r * kHid * 4);
    al((void**)&M->z2, r * kHid * 4);
    al((void**)&M->z3, r * kOut * 4);
    al((void**)&M->d1, r * kHid * 4);
    al((void**)&M->d2, r * kHid * 4);
    if (e != hipSuccess) {
        free_scratch(M);
        return hip_fail(e, "model scratch");
    }
    M->cap = rows;
    return BCC_OK;
}

// logits of `rows` inputs x [rows][96] into M->z3 (also keeps h1, h2, z1, z2)
int forward(bcc_model* M, const uint16_t* x, int rows, hipStream_t s)
{
    const uint16_t* W1 = M->wb + P_W1;
    const uint16_t* W2 = M->wb + P_W2;
    const uint16_t* W3 = M->wb + P_W3;
    BCC_HIP(gemm(false, false, x, kIn, W1, kHid, M->z1, kHid, rows, kHid, kIn, 1.f, 0.f, s), "gemm1");
    hipLaunchKernelGGL(bias_act_kernel, dim3(blocks_for((long long)rows * kHid)), dim3(256), 0, s, M->z1,
                       M->w + P_B1, rows, kHid, 1, M->h1, M->z1);
    BCC_HIP(gemm(false, false, M->h1, kHid, W2, kHid, M->z2, kHid, rows, kHid, kHid, 1.f, 0.f, s), "gemm2");
    hipLaunchKernelGGL(bias_act_kernel, dim3(blocks_for((long long)rows * kHid)), dim3(256), 0, s, M->z2,
                       M->w + P_B2, rows, kHid, 1, M->h2, M->z2);
    BCC_HIP(gemm(false, false, M->h2, kHid, W3, kOut, M->z3, kOut, rows, kOut, kHid, 1.f, 0.f, s), "gemm3");
    hipLaunchKernelGGL(bias_act_kernel, dim3(blocks_for((long long)rows * kOut)), dim3(256), 0, s, M->z3,
                       M->w + P_B3, rows, kOut, 0, (uint16_t*)nullptr, M->z3);
    BCC_HIP(hipGetLastError(), "forward");
    return BCC_OK;
}

// gradients of the mean BCE over `rows` into grad (FP32 [P_N]); loss added to M->loss
int backward(bcc_model* M, const uint16_t* x, const uint32_t* labels, int rows, float* grad, hipStream_t s)
{
    const uint16_t* W2 = M->wb + P_W2;
    const uint16_t* W3 = M->wb + P_W3;
    hipLaunchKernelGGL(bce_grad_kernel, dim3(blocks_for((long long)rows * kOut)), dim3(256), 0, s, M->z3, labels,
                       rows, M->g3, M->loss);
    BCC_HIP(gemm(true, false, M->h2, kHid, M->g3, kOut, grad + P_W3, kOut, kHid, kOut, rows, 1.f, 0.f, s, M->ws), "dW3");
    hipLaunchKernelGGL(colsum_kernel, dim3(kOut), dim3(256), 0, s, M->g3, rows, kOut, grad + P_B3);
    BCC_HIP(gemm(false, true, M->g3, kOut, W3, kOut, M->d2, kHid, rows, kHid, kOut, 1.f, 0.f, s), "dH2");
    hipLaunchKernelGGL(relu_back_kernel, dim3(blocks_for((long long)rows * kHid)), dim3(256), 0, s, M->d2, M->z2,
                       (long long)rows * kHid, M->g2);
    BCC_HIP(gemm(true, false, M->h1, kHid, M->g2, kHid, grad + P_W2, kHid, kHid, kHid, rows, 1.f, 0.f, s, M->ws), "dW2");
    hipLaunchKernelGGL(colsum_kernel, dim3(kHid), dim3(256), 0, s, M->g2, rows, kHid, grad + P_B2);
    BCC_HIP(gemm(false, true, M->g2, kHid, W2, kHid, M->d1, kHid, rows, kHid, kHid, 1.f, 0.f, s), "dH1");
    hipLaunchKernelGGL(relu_back_kernel, dim3(blocks_for((long long)rows * kHid)), dim3(256), 0, s, M->d1, M->z1,
                       (long long)rows * kHid, M->g1);
    BCC_HIP(gemm(true, false, x, kIn, M->g1, kHid, grad + P_W1, kHid, kIn, kHid, rows, 1.f, 0.f, s, M->ws), "dW1");
    hipLaunchKernelGGL(colsum_kernel, dim3(kHid), dim3(256), 0, s, M->g1, rows, kHid, grad + P_B1);
    BCC_HIP(hipGetLastError(), "backward");
    return BCC_OK;
}

int upload_params(bcc_model* M, const float* host)
{
    BCC_HIP(hipMemcpy(M->w, host, sizeof(float) * P_N, hipMemcpyHostToDevice), "model params");
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(P_N)), dim3(256), 0, 0, M->w, (long long)P_N, M->wb);
    BCC_HIP(hipGetLastError(), "model params");
    BCC_HIP(hipDeviceSynchronize(), "model params");
    return BCC_OK;
}

void destroy_model(bcc_model* M)
{
    if (!M)
        return;
    free_scratch(M);
    for (void* p : {(void*)M->w, (void*)M->wb, (void*)M->g, (void*)M->m, (void*)M->v, (void*)M->lo,
                    (void*)M->scale, (void*)M->loss, (void*)M->frag, (void*)M->ws, (void*)M->enc})
        (void)hipFree(p);
    delete M;
}

int alloc_model(int robot, bcc_model** out)
{
    const int dim = bcc_robot_dimension(robot);
    if (dim < 0)
        return fail(BCC_ERR_INVALID, "unknown robot id");
    bcc_model* M = new bcc_model;
    M->robot = robot;
    M->dim = dim;
    hipError_t e = hipMalloc((void**)&M->w, sizeof(float) * P_N);
    if (e == hipSuccess) e = hipMalloc((void**)&M->wb, 2 * P_N);
    if (e == hipSuccess) e = hipMalloc((void**)&M->g, sizeof(float) * P_N);
    if (e == hipSuccess) e = hipMalloc((void**)&M->m, sizeof(float) * P_N);
    if (e == hipSuccess) e = hipMalloc((void**)&M->v, sizeof(float) * P_N);
    if (e == hipSuccess) e = hipMalloc((void**)&M->lo, sizeof(float) * 16);
    if (e == hipSuccess) e = hipMalloc((void**)&M->scale, sizeof(float) * 16);
    if (e == hipSuccess) e = hipMalloc((void**)&M->loss, sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&M->frag, sizeof(uint16_t) * kFrags * 64 * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&M->ws, sizeof(float) * kSplitWs);
    if (e == hipSuccess) e = hipMemset(M->m, 0, sizeof(float) * P_N);
    if (e == hipSuccess) e = hipMemset(M->v, 0, sizeof(float) * P_N);
    if (e == hipSuccess) e = hipMemcpy(M->lo, kLo[robot], sizeof(float) * 16, hipMemcpyHostToDevice);
    float scale[16] = {};
    for (int j = 0; j < dim; ++j)
        scale[j] = (float)(2.0 / (double)kRange[robot][j]);
    if (e == hipSuccess) e = hipMemcpy(M->scale, scale, sizeof(float) * 16, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        destroy_model(M);
        return hip_fail(e, "model allocation");
    }
    *out = M;
    return BCC_OK;
}

const char kMagic[8] = {'B', 'C', 'C', 'M', 'L', 'P', '0', '1'};

}  // namespace

extern "C" {

int bcc_model_num_params(void)
{
    return (int)P_N;
}

int bcc_model_create(int robot, uint64_t seed, bcc_model** out)
{
    if (!out)
        return fail(BCC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    bcc_model* M = nullptr;
    if (int rc = alloc_model(robot, &M))
        return rc;
    M->seed = seed;
    // He-uniform weights, zero biases (seeded, host-generated)
    std::vector<float> h(P_N, 0.0f);
    Lcg rng{seed ^ 0x6c62272e07bb0142ull};
    auto fill = [&](long long off, long long n, int fan_in) {
        const float a = std::sqrt(6.0f / (float)fan_in);
        for (long long i = 0; i < n; ++i)
            h[off + i] = (2.0f * rng.uniform() - 1.0f) * a;
    };
    fill(P_W1, (long long)kIn * kHid, kEnc + 2 * M->dim);
    fill(P_W2, (long long)kHid * kHid, kHid);
    fill(P_W3, (long long)kHid * kOut, kHid);
    if (int rc = upload_params(M, h.data())) {
        destroy_model(M);
        return rc;
    }
    *out = M;
    return BCC_OK;
}

int bcc_model_destroy(bcc_model* model)
{
    destroy_model(model);
    return BCC_OK;
}

int bcc_model_get_params(const bcc_model* M, float* host)
{
    if (!M || !host)
        return fail(BCC_ERR_INVALID, "bad arguments");
    BCC_HIP(hipMemcpy(host, M->w, sizeof(float) * P_N, hipMemcpyDeviceToHost), "model params");
    return BCC_OK;
}

int bcc_model_set_params(bcc_model* M, const float* host)
{
    if (!M || !host)
        return fail(BCC_ERR_INVALID, "bad arguments");
    return upload_params(M, host);
}

int bcc_model_save(const bcc_model* M, const char* path)
{
    if (!M || !path)
        return fail(BCC_ERR_INVALID, "bad arguments");
    std::vector<float> w(P_N), m(P_N), v(P_N);
    BCC_HIP(hipMemcpy(w.data(), M->w, sizeof(float) * P_N, hipMemcpyDeviceToHost), "model save");
    BCC_HIP(hipMemcpy(m.data(), M->m, sizeof(float) * P_N, hipMemcpyDeviceToHost), "model save");
    BCC_HIP(hipMemcpy(v.data(), M->v, sizeof(float) * P_N, hipMemcpyDeviceToHost), "model save");
    std::ofstream f(path, std::ios::binary);
    if (!f)
        return fail(BCC_ERR_INVALID, std::string("cannot open ") + path);
    const int32_t hdr[4] = {M->robot, (int32_t)P_N, M->step, 0};
    f.write(kMagic, 8);
    f.write((const char*)hdr, sizeof hdr);
    f.write((const char*)&M->seed, sizeof M->seed);
    f.write((const char*)w.data(), sizeof(float) * P_N);
    f.write((const char*)m.data(), sizeof(float) * P_N);
    f.write((const char*)v.data(), sizeof(float) * P_N);
    if (!f)
        return fail(BCC_ERR_INVALID, std::string("write failed: ") + path);
    return BCC_OK;
}

int bcc_model_load(const char* path, bcc_model** out)
{
    if (!path || !out)
        return fail(BCC_ERR_INVALID, "bad arguments");
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    char magic[8];
    int32_t hdr[4];
    uint64_t seed = 0;
    if (!f.read(magic, 8) || std::memcmp(magic, kMagic, 8) != 0 || !f.read((char*)hdr, sizeof hdr) ||
        !f.read((char*)&seed, sizeof seed) || hdr[1] != (int32_t)P_N)
        return fail(BCC_ERR_INVALID, std::string("not a model file: ") + path);
    std::vector<float> w(P_N), m(P_N), v(P_N);
    if (!f.read((char*)w.data(), sizeof(float) * P_N) || !f.read((char*)m.data(), sizeof(float) * P_N) ||
        !f.read((char*)v.data(), sizeof(float) * P_N))
        return fail(BCC_ERR_INVALID, std::string("truncated model file: ") + path);
    bcc_model* M = nullptr;
    if (int rc = alloc_model(hdr[0], &M))
        return rc;
    M->step = hdr[2];
    M->seed = seed;
    int rc = upload_params(M, w.data());
    if (rc == BCC_OK) {
        hipError_t e = hipMemcpy(M->m, m.data(), sizeof(float) * P_N, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(M->v, v.data(), sizeof(float) * P_N, hipMemcpyHostToDevice);
        if (e != hipSuccess)
            rc = hip_fail(e, "model load");
    }
    if (rc != BCC_OK) {
        destroy_model(M);
        return rc;
    }
    *out = M;
    return BCC_OK;
}

int bcc_encode_envs(const bcc_scene* scene, float* d_enc)
{
    if (!scene || (scene->num_envs > 0 && !d_enc))
        return fail(BCC_ERR_INVALID, "bad arguments");
    if (scene->num_envs == 0)
        return BCC_OK;
    hipLaunchKernelGGL(encode_env_kernel, dim3(scene->num_envs), dim3(64), 0, 0, scene->d_ref_obs,
                       scene->d_ref_meta, scene->num_envs, d_enc);
    BCC_HIP(hipGetLastError(), "encoder");
    BCC_HIP(hipDeviceSynchronize(), "encoder");
    return BCC_OK;
}

int bcc_model_features(const bcc_model* M, const bcc_scene* scene, const float* d_edges, int64_t num_pairs,
                       const int32_t* d_pair_env, const int32_t* d_pair_edge, uint16_t* d_x, void* stream)
{
    if (!M || !scene || num_pairs < 0 || (num_pairs > 0 && (!d_edges || !d_x)))
        return fail(BCC_ERR_INVALID, "bad arguments");
    if (num_pairs == 0)
        return BCC_OK;
    hipStream_t s = (hipStream_t)stream;
    float* enc = nullptr;
    BCC_HIP(hipMalloc((void**)&enc, sizeof(float) * kEnc * std::max(scene->num_envs, 1)), "encoder");
    hipLaunchKernelGGL(encode_env_kernel, dim3(std::max(scene->num_envs, 1)), dim3(64), 0, s, scene->d_ref_obs,
                       scene->d_ref_meta, scene->num_envs, enc);
    hipLaunchKernelGGL(features_kernel, dim3(blocks_for(num_pairs * kIn)), dim3(256), 0, s, enc, d_edges, M->dim,
                       M->lo, M->scale, d_pair_env, d_pair_edge, scene->num_envs, 0LL, (long long)num_pairs, d_x);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    (void)hipFree(enc);
    if (e != hipSuccess)
        return hip_fail(e, "features");
    return BCC_OK;
}

int bcc_label_pairs_steps(int robot, const bcc_scene* scene, const float* d_edges, int num_pairs,
                          const int32_t* d_pair_env, const int32_t* d_pair_edge, int resolution, int steps,
                          uint32_t* d_labels, void* stream)
{
    if (!scene || num_pairs < 0 || resolution < 1 || steps < 0 ||
        (num_pairs > 0 && (!d_edges || !d_pair_env || !d_pair_edge || !d_labels)))
        return fail(BCC_ERR_INVALID, "bad arguments");
    ModelDev md;
    if (int rc = bcc_detail::device_model(robot, &md))
        return rc;
    LabelArgs a{scene->d_ref_obs, scene->d_ref_meta, d_edges, num_pairs, resolution, steps, md, d_pair_env,
                d_pair_edge, d_labels};
    return dispatch_label(robot, a, (hipStream_t)stream);
}

int bcc_label_pairs(int robot, const bcc_scene* scene, const float* d_edges, int num_pairs,
                    const int32_t* d_pair_env, const int32_t* d_pair_edge, int resolution, uint32_t* d_labels,
                    void* stream)
{
    return bcc_label_pairs_steps(robot, scene, d_edges, num_pairs, d_pair_env, d_pair_edge, resolution, 0, d_labels,
                                 stream);
}

int bcc_model_grads(bcc_model* M, const uint16_t* d_x, const uint32_t* d_labels, int rows, float* d_grad,
                    float* loss)
{
    if (!M || rows <= 0 || !d_x || !d_labels || !d_grad)
        return fail(BCC_ERR_INVALID, "bad arguments");
    if (int rc = ensure_scratch(M, rows))
        return rc;
    BCC_HIP(hipMemset(M->loss, 0, sizeof(float)), "loss reset");
    if (int rc = forward(M, d_x, rows, nullptr))
        return rc;
    if (int rc = backward(M, d_x, d_labels, rows, d_grad, nullptr))
        return rc;
    BCC_HIP(hipDeviceSynchronize(), "grads");
    if (loss)
        BCC_HIP(hipMemcpy(loss, M->loss, sizeof(float), hipMemcpyDeviceToHost), "loss");
    return BCC_OK;
}

int bcc_model_train(bcc_model* M, const uint16_t* d_x, const uint32_t* d_labels, int num_rows, int epochs,
                    int batch, float lr, float* epoch_loss)
{
    if (!M || !d_x || !d_labels || num_rows <= 0 || epochs < 1 || batch < 1 || !(lr > 0.0f))
        return fail(BCC_ERR_INVALID, "bad arguments");
    batch = std::min(batch, num_rows);
    if (int rc = ensure_scratch(M, batch))
        return rc;
    const int steps = num_rows / batch;
    for (int ep = 0; ep < epochs; ++ep) {
        BCC_HIP(hipMemset(M->loss, 0, sizeof(float)), "loss reset");
        // rotate the batch boundaries each epoch
        const int shift = (int)(((long long)ep * 7919) % batch);
        for (int st = 0; st < steps; ++st) {
            const long long r0 = std::min<long long>((long long)st * batch + shift, (long long)num_rows - batch);
            const uint16_t* x = d_x + r0 * kIn;
            if (int rc = forward(M, x, batch, nullptr))
                return rc;
            if (int rc = backward(M, x, d_labels + r0, batch, M->g, nullptr))
                return rc;
            M->step += 1;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.0f - std::pow(b1, (float)M->step);
            const float c2 = 1.0f - std::pow(b2, (float)M->step);
            hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(P_N)), dim3(256), 0, 0, M->w, M->m, M->v, M->g,
                               (long long)P_N, lr, b1, b2, 1e-8f, c1, c2, M->wb);
            BCC_HIP(hipGetLastError(), "adam");
        }
        BCC_HIP(hipDeviceSynchronize(), "train");
        if (epoch_loss) {
            float l = 0.0f;
            BCC_HIP(hipMemcpy(&l, M->loss, sizeof(float), hipMemcpyDeviceToHost), "loss");
            epoch_loss[ep] = l / (float)std::max(steps, 1);
        }
    }
    return BCC_OK;
}

int bcc_model_logits(bcc_model* M, const uint16_t* d_x, int64_t rows, float* d_logits)
{
    if (!M || rows < 0 || (rows > 0 && (!d_x || !d_logits)))
        return fail(BCC_ERR_INVALID, "bad arguments");
    const int chunk = 65536;
    if (int rc = ensure_scratch(M, (int)std::min<int64_t>(std::max<int64_t>(rows, 1), chunk)))
        return rc;
    for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
        const int n = (int)std::min<int64_t>(chunk, rows - r0);
        if (int rc = forward(M, d_x + r0 * kIn, n, nullptr))
            return rc;
        BCC_HIP(hipMemcpyAsync(d_logits + r0 * kOut, M->z3, sizeof(float) * n * kOut, hipMemcpyDeviceToDevice, 0),
                "logits");
    }
    BCC_HIP(hipDeviceSynchronize(), "logits");
    return BCC_OK;
}

}  // extern "C"

namespace bcc_detail {

void joint_range(int robot, int j, float* lo, float* range)
{
    *lo = kLo[robot][j];
    *range = kRange[robot][j];
}

}  // namespace bcc_detail

namespace bcc_detail {

int model_robot(const bcc_model* M)
{
    return M->robot;
}

int model_mlp_args(bcc_model* M, const bcc_scene* scene, hipStream_t s, MlpArgs* out)
{
    if (M->enc_envs < scene->num_envs) {
        BCC_HIP(hipStreamSynchronize(s), "order scratch");
        (void)hipFree(M->enc);
        M->enc = nullptr;
        M->enc_envs = 0;
        BCC_HIP(hipMalloc((void**)&M->enc, sizeof(float) * kEnc * std::max(scene->num_envs, 1)), "order scratch");
        M->enc_envs = scene->num_envs;
    }
    hipLaunchKernelGGL(pack_fragments_kernel, dim3(blocks_for((long long)kFrags * 64 * 8)), dim3(256), 0, s,
                       M->wb + P_W1, M->wb + P_W2, M->wb + P_W3, M->frag);
    if (scene->num_envs > 0)
        hipLaunchKernelGGL(encode_env_kernel, dim3(scene->num_envs), dim3(64), 0, s, scene->d_ref_obs,
                           scene->d_ref_meta, scene->num_envs, M->enc);
    BCC_HIP(hipGetLastError(), "model operands");
    *out = MlpArgs{(const bf16x8*)M->frag, M->w + P_B1, M->w + P_B2, M->w + P_B3, M->enc, M->lo, M->scale,
                   0xffffffffu};
    return BCC_OK;
}

int model_mlp_snapshot(const bcc_model* M, const bcc_scene* scene, MlpArgs* out, void** owned)
{
    // [fragments | b1 | b2 | b3 | lo | scale | codes], each part 256-byte aligned
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_frag = 0;
    const size_t o_b1 = o_frag + al(sizeof(uint16_t) * kFrags * 64 * 8);
    const size_t o_b2 = o_b1 + al(sizeof(float) * kHid);
    const size_t o_b3 = o_b2 + al(sizeof(float) * kHid);
    const size_t o_lo = o_b3 + al(sizeof(float) * kOut);
    const size_t o_sc = o_lo + al(sizeof(float) * 16);
    const size_t o_enc = o_sc + al(sizeof(float) * 16);
    const size_t bytes = o_enc + al(sizeof(float) * kEnc * std::max(scene->num_envs, 1));
    char* d = nullptr;
    BCC_HIP(hipMalloc((void**)&d, bytes), "model snapshot");
    hipLaunchKernelGGL(pack_fragments_kernel, dim3(blocks_for((long long)kFrags * 64 * 8)), dim3(256), 0, 0,
                       M->wb + P_W1, M->wb + P_W2, M->wb + P_W3, (uint16_t*)(d + o_frag));
    hipError_t e = hipGetLastError();
    const std::pair<size_t, const void*> parts[] = {{o_b1, M->w + P_B1}, {o_b2, M->w + P_B2}};
    for (const auto& pr : parts)
        if (e == hipSuccess)
            e = hipMemcpy(d + pr.first, pr.second, sizeof(float) * kHid, hipMemcpyDeviceToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(d + o_b3, M->w + P_B3, sizeof(float) * kOut, hipMemcpyDeviceToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(d + o_lo, M->lo, sizeof(float) * 16, hipMemcpyDeviceToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(d + o_sc, M->scale, sizeof(float) * 16, hipMemcpyDeviceToDevice);
    if (e == hipSuccess && scene->num_envs > 0) {
        hipLaunchKernelGGL(encode_env_kernel, dim3(scene->num_envs), dim3(64), 0, 0, scene->d_ref_obs,
                           scene->d_ref_meta, scene->num_envs, (float*)(d + o_enc));
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        (void)hipFree(d);
        return hip_fail(e, "model snapshot");
    }
    *out = MlpArgs{(const bf16x8*)(d + o_frag), (const float*)(d + o_b1), (const float*)(d + o_b2),
                   (const float*)(d + o_b3), (const float*)(d + o_enc), (const float*)(d + o_lo),
                   (const float*)(d + o_sc), 0xffffffffu};
    *owned = d;
    return BCC_OK;
}

}  // namespace bcc_detail

namespace {

// Fused inference: the model's operands for the scene (weight fragments,
// environment codes), then one kernel for all layers and the top-4
// selection, all on stream s (no synchronization).  Pairs: every (edge, env)
// pair of a dense batch, edge-major (pair_env null), or the list
// (pair_env[p], pair_edge ? pair_edge[p] : p).  d_logits (optional) receives
// [pairs][32].
int order_fused(bcc_model* M, const bcc_scene* scene, const float* d_edges, long long pairs, const int* pair_env,
                const int* pair_edge, uint32_t* d_order, float* d_logits, hipStream_t s)
{
    if (pairs == 0 || scene->num_envs == 0)
        return BCC_OK;
    MlpArgs m;
    if (int rc = bcc_detail::model_mlp_args(M, scene, s, &m))
        return rc;
    hipLaunchKernelGGL(fused_order_kernel, dim3((unsigned)((pairs + kFuseBlockPairs - 1) / kFuseBlockPairs)), dim3(256), 0, s,
                       m.frag, m.b1, m.b2, m.b3, m.enc, d_edges, M->dim, m.lo, m.scale, scene->num_envs, pair_env,
                       pair_edge, pairs, d_order, d_logits);
    if (hipError_t e = hipGetLastError(); e != hipSuccess)
        return hip_fail(e, "fused order");
    return BCC_OK;
}

// Average over `iters` runs (after `warmup`) of what `step` launches on s,
// hipEvent-timed on s.
template <class F>
int time_on_stream(int warmup, int iters, hipStream_t s, double* avg_ms, F step)
{
    for (int i = 0; i < warmup; ++i)
        if (int rc = step())
            return rc;
    hipEvent_t t0, t1;
    BCC_HIP(hipEventCreate(&t0), "hipEventCreate");
    hipError_t e = hipEventCreate(&t1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(t0);
        return hip_fail(e, "hipEventCreate");
    }
    int rc = BCC_OK;
    e = hipEventRecord(t0, s);
    for (int i = 0; i < iters && e == hipSuccess && rc == BCC_OK; ++i)
        rc = step();
    if (e == hipSuccess)
        e = hipEventRecord(t1, s);
    if (e == hipSuccess)
        e = hipEventSynchronize(t1);
    float ms = 0.0f;
    if (e == hipSuccess)
        e = hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc != BCC_OK)
        return rc;
    if (e != hipSuccess)
        return hip_fail(e, "learned step timing");
    *avg_ms = (double)ms / iters;
    return BCC_OK;
}

}  // namespace

extern "C" {

int bcc_time_learned_dense(bcc_model* M, const bcc_scene* scene, const float* d_edges, int num_edges,
                           int resolution, uint32_t* d_order, uint8_t* d_results, int warmup, int iters,
                           void* stream, double* step_ms, double* order_ms)
{
    if (!M || !scene || num_edges < 0 || (num_edges > 0 && (!d_edges || !d_order || !d_results)) || iters < 1 ||
        warmup < 0 || !step_ms)
        return fail(BCC_ERR_INVALID, "bad arguments");
    hipStream_t s = (hipStream_t)stream;
    const long long pairs = (long long)num_edges * scene->num_envs;
    auto order = [&]() { return order_fused(M, scene, d_edges, pairs, nullptr, nullptr, d_order, nullptr, s); };
    auto step = [&]() {
        if (int rc = order())
            return rc;
        return bcc_check_dense_ex(M->robot, scene, d_edges, num_edges, resolution, d_order, nullptr, d_results,
                                  stream);
    };
    if (int rc = time_on_stream(warmup, iters, s, step_ms, step))
        return rc;
    if (order_ms)
        return time_on_stream(warmup, iters, s, order_ms, order);
    return BCC_OK;
}

int bcc_model_order_dense(bcc_model* M, const bcc_scene* scene, const float* d_edges, int num_edges,
                          uint32_t* d_order, void* stream)
{
    if (!M || !scene || num_edges < 0 || (num_edges > 0 && (!d_edges || !d_order)))
        return fail(BCC_ERR_INVALID, "bad arguments");
    return order_fused(M, scene, d_edges, (long long)num_edges * scene->num_envs, nullptr, nullptr, d_order, nullptr,
                       (hipStream_t)stream);
}

int bcc_model_logits_dense(bcc_model* M, const bcc_scene* scene, const float* d_edges, int num_edges,
                           uint32_t* d_order, float* d_logits, void* stream)
{
    if (!M || !scene || num_edges < 0 || (num_edges > 0 && (!d_edges || !d_order || !d_logits)))
        return fail(BCC_ERR_INVALID, "bad arguments");
    return order_fused(M, scene, d_edges, (long long)num_edges * scene->num_envs, nullptr, nullptr, d_order,
                       d_logits, (hipStream_t)stream);
}

int bcc_model_order_pairs(bcc_model* M, const bcc_scene* scene, const float* d_edges, int64_t num_pairs,
                          const int32_t* d_pair_env, const int32_t* d_pair_edge, uint32_t* d_order, float* d_logits,
                          void* stream)
{
    if (!M || !scene || num_pairs < 0 || (num_pairs > 0 && (!d_edges || !d_pair_env || !d_order)))
        return fail(BCC_ERR_INVALID, "bad arguments");
    if (num_pairs > 0 && scene->num_envs == 0)
        return fail(BCC_ERR_INVALID, "pairs given for a scene without environments");
    return order_fused(M, scene, d_edges, (long long)num_pairs, d_pair_env, d_pair_edge, d_order, d_logits,
                       (hipStream_t)stream);
}

int bcc_debug_gemm(int ta, int tb, const uint16_t* d_a, int lda, const uint16_t* d_b, int ldb, float* d_c, int ldc,
                   int m, int n, int k, float alpha, float beta)
{
    if (m < 0 || n < 0 || k < 0 || (m > 0 && n > 0 && (!d_a || !d_b || !d_c)))
        return fail(BCC_ERR_INVALID, "bad gemm arguments");
    float* ws = nullptr;   // split-K workspace, as the model's weight gradients use
    BCC_HIP(hipMalloc((void**)&ws, sizeof(float) * kSplitWs), "gemm workspace");
    hipError_t e = gemm(ta != 0, tb != 0, d_a, lda, d_b, ldb, d_c, ldc, m, n, k, alpha, beta, nullptr, ws);
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    (void)hipFree(ws);
    if (e != hipSuccess)
        return hip_fail(e, "gemm");
    return BCC_OK;
}

}  // extern "C"
