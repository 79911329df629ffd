// BF16 matrix-core (MFMA) building blocks for the learned step ordering.
//
// Matrices are row-major; BF16 values are stored as raw 16-bit patterns
// (uint16_t) and converted with round-to-nearest-even.  Products accumulate
// in FP32 on v_mfma_f32_16x16x32_bf16 (gfx950), whose operands for one
// 16 x 16 x 32 tile are
//   A (16x32): lane l holds A[l % 16][slots 8 (l/16) .. 8 (l/16) + 7]
//   B (32x16): lane l holds B[slots 8 (l/16) .. 8 (l/16) + 7][l % 16]
//   D (16x16): lane l holds D[4 (l/16) + v][l % 16], v = 0..3
// Slot 8 g + i (g = l/16, i = 0..7) of a 32-wide k block carries logical k
//   kslot(g, i) = 4 g + i (i < 4),  16 + 4 g + (i - 4) (i >= 4),
// the same in A and B (so the product is the plain sum over k).  With this
// assignment the D rows a lane holds of two stacked 16-row output tiles
// (rows 4 g + v of tile 2 j and of tile 2 j + 1) are exactly its B operand
// for the next layer's k block j: activations chain from layer to layer
// without lane exchanges.  Every MFMA in this file uses it, so the fused and
// the layer-by-layer paths add the same products in the same slots and agree
// bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bcc {

typedef short bf16x8 __attribute__((ext_vector_type(8)));    // eight BF16 bit patterns
typedef __bf16 bf16v8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));    // the same eight BF16, two per word

typedef __bf16 bf16v2 __attribute__((ext_vector_type(2)));
typedef float f32v2 __attribute__((ext_vector_type(2)));

// FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950).
__device__ __forceinline__ uint16_t f2bf(float f)
{
    return __builtin_bit_cast(uint16_t, (__bf16)f);
}

// Two values at once: low half = a, high half = b.
__device__ __forceinline__ uint32_t f2bf2(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32v2{a, b}), bf16v2));
}

// Eight FP32 values as one MFMA operand.
__device__ __forceinline__ bf16x8 pack8(const float (&v)[8])
{
    const u32x4 w{f2bf2(v[0], v[1]), f2bf2(v[2], v[3]), f2bf2(v[4], v[5]), f2bf2(v[6], v[7])};
    return __builtin_bit_cast(bf16x8, w);
}

__device__ __forceinline__ float bf2f(uint16_t h)
{
    return __uint_as_float((uint32_t)h << 16);
}

__device__ __forceinline__ constexpr int kslot(int g, int i)
{
    return i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4);
}

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16v8, a), __builtin_bit_cast(bf16v8, b), c, 0,
                                                   0, 0);
}

// Eight slots of one row of a 32-wide k block held contiguously (k = 0..31).
__device__ __forceinline__ bf16x8 load_slots(const uint16_t* row, int g)
{
    bf16x8 v;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        v[i] = (short)row[kslot(g, i)];
    return v;
}

// ------------------------------------------------------------ model shape
constexpr int kEnc = 64;      // occupancy cells (4 x 4 x 4)
constexpr int kIn = 96;       // padded input width: 64 + 2 * dim (dim <= 16)
#ifndef BCC_MLP_HIDDEN
#define BCC_MLP_HIDDEN 64
#endif
constexpr int kHid = BCC_MLP_HIDDEN;   // hidden width, a multiple of 32
static_assert(kHid % 32 == 0 && kHid >= 32, "hidden width: whole 32-wide k blocks");
constexpr int kOut = 32;

// Weight fragments of the fused inference kernels: fragment (rt, kt) of a
// layer with W [K][N] holds for lane l the eight values
// W[32 kt + kslot(l/16, i)][16 rt + l % 16], i = 0..7 (one MFMA A operand).
constexpr int kKT1 = kIn / 32;                   // k blocks of layer 1
constexpr int kKT2 = kHid / 32;                  // k blocks of layers 2 and 3
constexpr int kRT12 = kHid / 16;                 // row tiles of layers 1 and 2
constexpr int kRT3 = kOut / 16;                  // row tiles of layer 3
constexpr int kFragL1 = 0;
constexpr int kFragL2 = kFragL1 + kRT12 * kKT1;
constexpr int kFragL3 = kFragL2 + kRT12 * kKT2;
constexpr int kFrags = kFragL3 + kRT3 * kKT2;

// Joint coordinate idx of an edge ([2][dim]: start then end) normalised to
// [-1, 1] over the joint's limits: (q - lo) * scale - 1, scale = 2 / range.
__device__ __forceinline__ float edge_feature(const float* __restrict__ edges, int dim, const float* __restrict__ lo,
                                              const float* __restrict__ scale, long long edge, int idx)
{
    const int j = idx >= dim ? idx - dim : idx;
    const float q = edges[(size_t)edge * 2 * dim + idx];
    return (q - lo[j]) * scale[j] - 1.0f;
}

// The order word from a pair's 32 bin logits: four distinct bins, one byte
// each.  The first is the highest logit; each later one the highest logit
// among the bins at least kOrderSpread bins from every bin already chosen
// (any unchosen bin when none is that far), ties to the lower bin.  The
// probe checks the first bins' step-0 samples: neighbouring bins tend to
// collide together, so spreading the picks along the edge makes a miss of
// one less likely to repeat in the next.
#ifndef BCC_ORDER_SPREAD
#define BCC_ORDER_SPREAD 8
#endif
constexpr int kOrderSpread = BCC_ORDER_SPREAD;

// The selection from the layer-3 result registers.  Lane l holds the logits
// z[4 r3 + v] of bins 16 r3 + 4 g + v (r3 = 0, 1; v = 0..3; g = l / 16) of
// one pair; the four lanes l % 16 + 16 g' hold its 32 bins.  Each round every
// lane takes its best candidate, then the lanes 16 and 32 apart exchange
// theirs; all four lanes end with the same word.  Bins in taken0 are never
// picked (bins a probe has already visited) and the bins outside far0 count
// as near a chosen bin from the start.
__device__ __forceinline__ uint32_t spread_top4_lanes_from(const float (&z)[8], int g, uint32_t taken0,
                                                           uint32_t far0)
{
    uint32_t taken = taken0, far_from = far0, packed = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t cand = (far_from & ~taken) ? (far_from & ~taken) : ~taken;
        float bv = 0.0f;
        int bb = 64;   // none yet
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int bin = 16 * (e >> 2) + 4 * g + (e & 3);
            if (((cand >> bin) & 1u) && (bb == 64 || z[e] > bv || (z[e] == bv && bin < bb))) {
                bv = z[e];
                bb = bin;
            }
        }
#pragma unroll
        for (int off = 16; off <= 32; off <<= 1) {
            const float ov = __shfl_xor(bv, off);
            const int ob = __shfl_xor(bb, off);
            if (ob != 64 && (bb == 64 || ov > bv || (ov == bv && ob < bb))) {
                bv = ov;
                bb = ob;
            }
        }
        taken |= 1u << bb;
        packed |= (uint32_t)bb << (8 * k);
        if constexpr (kOrderSpread > 0) {
            // bins within kOrderSpread - 1 of the pick are no longer "far"
            const int lo = max(bb - (kOrderSpread - 1), 0);
            const int hi = min(bb + (kOrderSpread - 1), 31);
            const uint32_t width = (uint32_t)(hi - lo + 1);
            far_from &= ~((width >= 32 ? 0xffffffffu : ((1u << width) - 1u)) << lo);
        }
    }
    return packed;
}

__device__ __forceinline__ uint32_t spread_top4_lanes(const float (&z)[8], int g)
{
    return spread_top4_lanes_from(z, g, 0u, 0xffffffffu);
}

// Layer-1 B operands (X^T, three 32-wide k blocks) of the pair in column
// l % 16 of a 16-pair tile: slot i of k block kt is feature 32 kt + kslot(g,
// i); k blocks 0 and 1 are the environment code (slots 0-3 and 4-7 are four
// consecutive cells each, one 16-byte load), k block 2 the normalised edge
// coordinates and zero padding.  ok = false gives a zero column.
__device__ __forceinline__ void mlp_pair_operands(const float* __restrict__ code, const float* __restrict__ edges,
                                                  int dim, const float* __restrict__ lo,
                                                  const float* __restrict__ scale, long long edge, bool ok,
                                                  bf16x8 (&xb)[kKT1])
{
    static_assert(kEnc == 64 && kKT1 == 3, "feature layout");
    const int g = (threadIdx.x & 63) >> 4;
    const int kq = 4 * g;
    const float4* e4 = (const float4*)code;
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) {
        const float4 lo4 = ok ? e4[(32 * kt + kq) >> 2] : float4{0.f, 0.f, 0.f, 0.f};
        const float4 hi4 = ok ? e4[(32 * kt + 16 + kq) >> 2] : float4{0.f, 0.f, 0.f, 0.f};
        const float v[8] = {lo4.x, lo4.y, lo4.z, lo4.w, hi4.x, hi4.y, hi4.z, hi4.w};
        xb[kt] = pack8(v);
    }
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int idx = kslot(g, i);   // coordinate idx of the edge's [2][dim]
        v[i] = (ok && idx < 2 * dim) ? edge_feature(edges, dim, lo, scale, edge, idx) : 0.0f;
    }
    xb[2] = pack8(v);
}

// Raw buffer resource over `bytes` bytes at p.  Loads through it take a
// per-lane VGPR offset plus a wave-uniform SGPR or immediate offset, so a run
// of loads at constant distances from one base costs one VGPR of address
// (global loads keep a 64-bit VGPR address live per distinct offset).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ u32x4 load16(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff)
{
    return __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
}

// Logits of one 16-pair column tile, computed transposed (Y^T = W^T X^T) as
// in the fused inference kernel: the same MFMAs in the same k order, the
// same bias, ReLU and BF16 roundings, so the logits are bit-identical to it
// and to the layer-by-layer path.  Lane l ends with z[4 r3 + v] = logit of
// bin 16 r3 + 4 (l/16) + v of the pair in column l % 16.  Weight fragments
// are read straight from memory (L2-resident, 2 kFrags KiB) through a buffer
// resource; activations stay in registers: layer 1's kHid x 16 BF16 outputs
// are kHid / 8 VGPRs, layer 2 is produced two row tiles at a time and consumed at
// once by layer 3.  Every lane of the wave must be active (MFMA).
__device__ __forceinline__ void mlp_tile_logits(const bf16x8* __restrict__ frag, const float* __restrict__ b1,
                                                const float* __restrict__ b2, const float* __restrict__ b3,
                                                const bf16x8 (&xb)[kKT1], float (&z)[8])
{
    const int lane = threadIdx.x & 63;
    const __amdgpu_buffer_rsrc_t rf = raw_rsrc(frag, kFrags * 64 * 16);
    const __amdgpu_buffer_rsrc_t rb1 = raw_rsrc(b1, kHid * 4);
    const __amdgpu_buffer_rsrc_t rb2 = raw_rsrc(b2, kHid * 4);
    const __amdgpu_buffer_rsrc_t rb3 = raw_rsrc(b3, kOut * 4);
    const uint32_t vf = (uint32_t)lane * 16;          // lane's 16 bytes of a fragment
    const uint32_t vb = (uint32_t)(lane >> 4) * 16;   // bias elements 4 (l/16) .. +3 of a row tile
    auto A = [&](int f) { return __builtin_bit_cast(bf16x8, load16(rf, vf, f * 64 * 16)); };
    u32x4 y1[kKT2];
#pragma unroll
    for (int rt = 0; rt < kRT12; ++rt) {
        f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < kKT1; ++kt)
            acc = mfma32(A(kFragL1 + rt * kKT1 + kt), xb[kt], acc);
        const f32x4 bb = __builtin_bit_cast(f32x4, load16(rb1, vb, 64 * rt));
        float h[4];
#pragma unroll
        for (int v = 0; v < 4; ++v)
            h[v] = fmaxf(acc[v] + bb[v], 0.0f);
        y1[rt >> 1][2 * (rt & 1) + 0] = f2bf2(h[0], h[1]);
        y1[rt >> 1][2 * (rt & 1) + 1] = f2bf2(h[2], h[3]);
        // loads of later row tiles stay below this point: the kernels this
        // runs in are sized for their check program's registers
        __builtin_amdgcn_sched_barrier(0);
    }
    f32x4 acc3[kRT3];
#pragma unroll
    for (int r3 = 0; r3 < kRT3; ++r3)
        acc3[r3] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int j = 0; j < kKT2; ++j) {
        u32x4 y2;
#pragma unroll
        for (int h2 = 0; h2 < 2; ++h2) {
            f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < kKT2; ++kt) {
                acc = mfma32(A(kFragL2 + (2 * j + h2) * kKT2 + kt), __builtin_bit_cast(bf16x8, y1[kt]), acc);
                if (kt % 4 == 3)
                    __builtin_amdgcn_sched_barrier(0);   // at most four fragments in flight
            }
            const f32x4 bb = __builtin_bit_cast(f32x4, load16(rb2, vb, 64 * (2 * j + h2)));
            float h[4];
#pragma unroll
            for (int v = 0; v < 4; ++v)
                h[v] = fmaxf(acc[v] + bb[v], 0.0f);
            y2[2 * h2 + 0] = f2bf2(h[0], h[1]);
            y2[2 * h2 + 1] = f2bf2(h[2], h[3]);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int r3 = 0; r3 < kRT3; ++r3)
            acc3[r3] = mfma32(A(kFragL3 + r3 * kKT2 + j), __builtin_bit_cast(bf16x8, y2), acc3[r3]);
    }
#pragma unroll
    for (int r3 = 0; r3 < kRT3; ++r3) {
        const f32x4 bb = __builtin_bit_cast(f32x4, load16(rb3, vb, 64 * r3));
#pragma unroll
        for (int v = 0; v < 4; ++v)
            z[4 * r3 + v] = acc3[r3][v] + bb[v];
    }
}

// Generic GEMM: C[M][N] = alpha * op(A)[M][K] * op(B)[K][N] + beta * C
//   op(A)[m][k] = TA ? A[k*lda + m] : A[m*lda + k]
//   op(B)[k][n] = TB ? B[n*ldb + k] : B[k*ldb + n]
// 256 threads compute a 64x64 tile of C; each wave a 32x32 sub-tile
// (2x2 MFMA tiles).  K advances 32 per step through LDS; out-of-range
// elements are read as zero.  Each operand is staged with consecutive
// threads on its contiguous dimension.  Split K: block z covers
// k in [z*kchunk, (z+1)*kchunk) (kchunk a multiple of 32) and writes its tile
// to C + z*cstride (the caller reduces the partial tiles).
constexpr int kGemmTile = 64;
constexpr int kGemmK = 32;

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(const uint16_t* __restrict__ A, int lda,
                                                       const uint16_t* __restrict__ B, int ldb,
                                                       float* __restrict__ C, int ldc, int M, int N, int K,
                                                       float alpha, float beta, int kchunk, long long cstride)
{
    __shared__ uint16_t As[kGemmTile][kGemmK + 8];   // [m][k]
    __shared__ uint16_t Bs[kGemmTile][kGemmK + 8];   // [n][k]
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int m0 = blockIdx.y * kGemmTile;
    const int n0 = blockIdx.x * kGemmTile;
    const int wm = (wave >> 1) * 32;
    const int wn = (wave & 1) * 32;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(K, kbeg + kchunk);
    C += blockIdx.z * cstride;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
            acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = kbeg; k0 < kend; k0 += kGemmK) {
        // 64x32 elements of each operand, 8 per thread
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int idx = tid + 256 * r;   // 0..2047
            {
                const int mm = TA ? (idx & 63) : (idx >> 5);
                const int kk = TA ? (idx >> 6) : (idx & 31);
                const int gm = m0 + mm, gk = k0 + kk;
                uint16_t va = 0;
                if (gm < M && gk < kend)
                    va = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
                As[mm][kk] = va;
            }
            {
                const int nn = TB ? (idx >> 5) : (idx & 63);
                const int kk = TB ? (idx & 31) : (idx >> 6);
                const int gn = n0 + nn, gk = k0 + kk;
                uint16_t vb = 0;
                if (gn < N && gk < kend)
                    vb = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
                Bs[nn][kk] = vb;
            }
        }
        __syncthreads();
        const int g = lane >> 4;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const bf16x8 a = load_slots(&As[wm + 16 * i + (lane & 15)][0], g);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const bf16x8 b = load_slots(&Bs[wn + 16 * j + (lane & 15)][0], g);
                acc[i][j] = mfma32(a, b, acc[i][j]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gm = m0 + wm + 16 * i + 4 * (lane >> 4) + v;
                const int gn = n0 + wn + 16 * j + (lane & 15);
                if (gm < M && gn < N) {
                    float* c = C + (size_t)gm * ldc + gn;
                    *c = beta != 0.0f ? alpha * acc[i][j][v] + beta * *c : alpha * acc[i][j][v];
                }
            }
}

// Sum of split-K partial tiles in split order, then C = alpha * sum + beta * C.
static __global__ void splitk_reduce_kernel(const float* __restrict__ part, int splits, long long cstride, int M, int N,
                                     float* __restrict__ C, int ldc, float alpha, float beta)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)M * N)
        return;
    const int m = (int)(idx / N), n = (int)(idx - (long long)m * N);
    float s = 0.0f;
    for (int z = 0; z < splits; ++z)
        s += part[z * cstride + idx];
    float* c = C + (size_t)m * ldc + n;
    *c = beta != 0.0f ? alpha * s + beta * *c : alpha * s;
}

}  // namespace bcc
