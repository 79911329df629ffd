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
__global__ void splitk_reduce_kernel(const float* __restrict__ part, int splits, long long cstride, int M, int N,
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
