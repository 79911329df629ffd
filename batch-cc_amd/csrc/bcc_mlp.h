// BF16 matrix-core (MFMA) building blocks for the learned step ordering.
//
// Matrices are row-major; BF16 values are stored as raw 16-bit patterns
// (uint16_t) and converted with round-to-nearest-even.  Products accumulate
// in FP32 on v_mfma_f32_16x16x16_bf16, whose operand layout for a 16x16x16
// tile is:
//   A (16x16): lane l holds A[l % 16][4*(l/16) .. 4*(l/16)+3]
//   B (16x16): lane l holds B[4*(l/16) .. 4*(l/16)+3][l % 16]
//   D (16x16): lane l holds D[4*(l/16) + v][l % 16], v = 0..3
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bcc {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even (finite inputs)
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h)
{
    return __uint_as_float((uint32_t)h << 16);
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Generic GEMM: C[M][N] = alpha * op(A)[M][K] * op(B)[K][N] + beta * C
//   op(A)[m][k] = TA ? A[k*lda + m] : A[m*lda + k]
//   op(B)[k][n] = TB ? B[n*ldb + k] : B[k*ldb + n]
// 256 threads compute a 64x64 tile of C; each wave a 32x32 sub-tile
// (2x2 MFMA tiles).  K advances 16 per step through LDS; out-of-range
// elements are read as zero.  Each operand is staged with consecutive
// threads on its contiguous dimension.  Split K: block z covers
// k in [z*kchunk, (z+1)*kchunk) and writes its tile to C + z*cstride (the
// caller reduces the partial tiles).
constexpr int kGemmTile = 64;

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(const uint16_t* __restrict__ A, int lda,
                                                       const uint16_t* __restrict__ B, int ldb,
                                                       float* __restrict__ C, int ldc, int M, int N, int K,
                                                       float alpha, float beta, int kchunk, long long cstride)
{
    __shared__ uint16_t As[kGemmTile][16 + 4];   // [m][k]
    __shared__ uint16_t Bs[kGemmTile][16 + 4];   // [n][k]
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

    for (int k0 = kbeg; k0 < kend; k0 += 16) {
        // 64x16 elements of each operand, 4 per thread
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int idx = tid + 256 * r;   // 0..1023
            {
                const int mm = TA ? (idx & 63) : (idx >> 4);
                const int kk = TA ? (idx >> 6) : (idx & 15);
                const int gm = m0 + mm, gk = k0 + kk;
                uint16_t va = 0;
                if (gm < M && gk < kend)
                    va = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
                As[mm][kk] = va;
            }
            {
                const int nn = TB ? (idx >> 4) : (idx & 63);
                const int kk = TB ? (idx & 15) : (idx >> 6);
                const int gn = n0 + nn, gk = k0 + kk;
                uint16_t vb = 0;
                if (gn < N && gk < kend)
                    vb = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
                Bs[nn][kk] = vb;
            }
        }
        __syncthreads();
        const int kq = 4 * (lane >> 4);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const uint16_t* ar = &As[wm + 16 * i + (lane & 15)][kq];
            const bf16x4 a = bf16x4{(short)ar[0], (short)ar[1], (short)ar[2], (short)ar[3]};
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const uint16_t* br = &Bs[wn + 16 * j + (lane & 15)][kq];
                const bf16x4 b = bf16x4{(short)br[0], (short)br[1], (short)br[2], (short)br[3]};
                acc[i][j] = mfma16(a, b, acc[i][j]);
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
                    *c = alpha * acc[i][j][v] + (beta != 0.0f ? beta * *c : 0.0f);
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
    *c = alpha * s + (beta != 0.0f ? beta * *c : 0.0f);
}

}  // namespace bcc
