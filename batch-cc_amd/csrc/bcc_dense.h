// Dense plain-order kernel: every edge against every environment.
//
// Replaces batch_cc_kernel<Robot> (src/batch_cc.cu:74-133), which runs one
// 32-lane block per (env, edge) pair.  Here a 256-thread workgroup owns one
// environment and a chunk of edges; its 8 half-waves pull edges from the
// chunk through an LDS counter, so a half-wave whose pair ends early (first
// collision, or a short edge) immediately starts the next pair instead of
// idling.  Results are written edge-major, results[edge * E + env].
#pragma once

#include <algorithm>

#include "bcc_device.h"
#include "bcc_mlp.h"

namespace bcc {

constexpr int kBlock = 256;
static_assert(kBlock / 64 <= kMaxBlockWaves, "culling keeps one LDS slot row per wave");

// Step count and per-joint increment of one edge, then the configuration of
// lane t at step 0 (src/batch_cc.cu:103-116):
//   dist  = sqrt(sum (s-e)^2)        (fmaf accumulation)
//   n     = max(ceil(dist / 31 * resolution), 1), at most 2^22
//   delta = (e - s) / (31 n)
//   q     = fmaf(delta, t*n, s)
template <int D>
__device__ __forceinline__ int setup_pair(const float* __restrict__ se, int resolution, int t, float (&q)[D],
                                          float (&d)[D])
{
    float s[D], e[D];
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        s[j] = se[j];
        e[j] = se[D + j];
        const float diff = s[j] - e[j];
        acc = fmaf(diff, diff, acc);
    }
    const float dist = sqrtf(acc);
    // n is capped at 2^22 steps: never reached by a finite edge, it keeps a
    // non-finite input from spinning a half-wave for 2^31 steps
    const float nf = fminf(fmaxf(ceilf((dist / 31.0f) * (float)resolution), 1.0f), 4194304.0f);
    const int n = (int)nf;
    const float div = (float)(31 * n);
    const float tn = (float)(t * n);
#pragma unroll
    for (int j = 0; j < D; ++j) {
        d[j] = (e[j] - s[j]) / div;
        q[j] = fmaf(d[j], tn, s[j]);
    }
    return n;
}

// Kernel variants: kWaves is the occupancy target (waves per SIMD the
// register allocator must allow), kUnroll the obstacle-loop unroll factor.
// Each robot builds a high-occupancy variant for scenes with many obstacles
// per environment (the obstacle sweeps dominate and occupancy hides the
// scalar-load latency) and a low-occupancy one for few obstacles (forward
// kinematics and the self program dominate and want registers); the host
// picks one per launch from the scene's mean obstacle count.

template <int D>
__device__ __forceinline__ int setup_pair_sd(const float* __restrict__ se, int resolution, float (&s)[D],
                                             float (&d)[D])
{
    float e[D];
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        s[j] = se[j];
        e[j] = se[D + j];
        const float diff = s[j] - e[j];
        acc = fmaf(diff, diff, acc);
    }
    const float dist = sqrtf(acc);
    const float nf = fminf(fmaxf(ceilf((dist / 31.0f) * (float)resolution), 1.0f), 4194304.0f);
    const int n = (int)nf;
    const float div = (float)(31 * n);
#pragma unroll
    for (int j = 0; j < D; ++j)
        d[j] = (e[j] - s[j]) / div;
    return n;
}

// A learned order word names four distinct bins below 32, one byte each, the
// most likely to hold a colliding sample first; anything else means no
// prediction (0xffffffff).
__device__ __forceinline__ uint32_t valid_order(uint32_t top)
{
    const uint32_t taken = (1u << (top & 31)) | (1u << ((top >> 8) & 31)) | (1u << ((top >> 16) & 31)) |
                           (1u << ((top >> 24) & 31));
    return (__popc(taken) != 4 || (top & 0xe0e0e0e0u)) ? 0xffffffffu : top;
}

// Per-thread sample configuration lives in LDS, [joint][thread], not in
// registers: the check program reads q[j] once per joint (for its sin/cos)
// and the step loop adds delta once per step, so the configuration does not
// hold VGPRs across the whole program, which leaves the registers to the
// frames and sweep temporaries.
struct LdsRow {
    float* p;   // element j at p[j * kBlock]
    __device__ __forceinline__ float operator[](int j) const { return p[j * kBlock]; }
    __device__ __forceinline__ float& at(int j) const { return p[j * kBlock]; }
};

// Plain order, step 0 of lane t (src/batch_cc.cu:103-116) into the LDS row;
// returns the pair's step count.
template <int D>
__device__ __forceinline__ int plain_round0(const float* __restrict__ se, int resolution, int t, const LdsRow& Q)
{
    float qq[D], dd[D];
    const int n = setup_pair<D>(se, resolution, t, qq, dd);
#pragma unroll
    for (int j = 0; j < D; ++j)
        Q.at(j) = qq[j];
    return n;
}

// Plain order, the next step: config += delta (src/batch_cc.cu:126-128), with
// delta recomputed from the edge (the same arithmetic, so the same value) so
// that no increment has to stay resident between steps.  Linear in the step
// count: each step adds delta once to the carried configuration.
template <int D>
__device__ __forceinline__ void plain_next(const float* __restrict__ se, int resolution, const LdsRow& Q)
{
    float s[D], d[D];
    setup_pair_sd<D>(se, resolution, s, d);
#pragma unroll
    for (int j = 0; j < D; ++j)
        Q.at(j) = Q[j] + d[j];
}

// Block order for the step loops of the learned-order kernels: step 0 is the
// plain step 0 (lane t's first sample, the 32 evenly spaced samples that most
// often hold a collision); step k >= 1 covers the samples 32(k-1) .. 32k - 1
// of the remaining ones flattened as t (n-1) + (i-1), i >= 1 (lane t's i-th
// sample), which lie next to each other along the edge, so a half-wave's
// culling boxes (bcc_device.h) are tight and more obstacles are skipped.
// Sample (t, i) is fmaf(delta, t n, start) plus delta added i times: the
// plain order's own arithmetic, so every sample's bits, the sample set and
// the verdict are the plain order's; only the visiting order differs.  Pairs
// of more than kBlockOrderMaxSteps steps keep the plain order (the i
// additions per lane cost O(n) a step).
constexpr int kBlockOrderMaxSteps = 64;
template <int D>
__device__ __forceinline__ int block_step(const float* __restrict__ se, int resolution, int step, int t,
                                          const LdsRow& Q)
{
    float s[D], d[D];
    const int n = setup_pair_sd<D>(se, resolution, s, d);
    if (n > kBlockOrderMaxSteps) {
        if (step == 0) {
            const float tn = (float)(t * n);
#pragma unroll
            for (int j = 0; j < D; ++j)
                Q.at(j) = fmaf(d[j], tn, s[j]);
        } else {
#pragma unroll
            for (int j = 0; j < D; ++j)
                Q.at(j) = Q[j] + d[j];
        }
        return n;
    }
    int tt, i;
    if (step == 0) {
        tt = t;
        i = 0;
    } else {
        const int g = 32 * (step - 1) + t;
        tt = g / (n - 1);
        i = 1 + (g - tt * (n - 1));
    }
    const float tn = (float)(tt * n);
    float q[D];
#pragma unroll
    for (int j = 0; j < D; ++j)
        q[j] = fmaf(d[j], tn, s[j]);
    for (int k = 0; k < i; ++k) {
#pragma unroll
        for (int j = 0; j < D; ++j)
            q[j] = q[j] + d[j];
    }
#pragma unroll
    for (int j = 0; j < D; ++j)
        Q.at(j) = q[j];
    return n;
}

// Plain-order pair setup into LDS rows, configuration and increment (same
// arithmetic as setup_pair).
template <int D>
__device__ __forceinline__ int setup_pair_lds(const float* __restrict__ se, int resolution, int t, const LdsRow& q,
                                              const LdsRow& d)
{
    float qq[D], dd[D];
    const int n = setup_pair<D>(se, resolution, t, qq, dd);
#pragma unroll
    for (int j = 0; j < D; ++j) {
        q.at(j) = qq[j];
        d.at(j) = dd[j];
    }
    return n;
}

// rounds (optional): number of rounds (steps) each pair ran, [edge][env].
template <class Robot, int kWaves, int kUnroll>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kWaves, 8))) void dense_kernel(
    const float* __restrict__ obs, const int* __restrict__ meta, const float* __restrict__ edges, int num_edges,
    int num_envs, int resolution, int chunk, int chunks_per_env, uint8_t* __restrict__ rounds,
    uint8_t* __restrict__ out)
{
    constexpr int D = Robot::dim;
    const int env = blockIdx.x / chunks_per_env;
    const int ck = blockIdx.x - env * chunks_per_env;
    const int e_begin = ck * chunk;
    const int e_end = min(num_edges, e_begin + chunk);

    // steps taken from the edge (plain_next) or from an LDS increment row
    constexpr bool kFromEdge = Robot::kRoundsFromEdge;
    __shared__ int next_edge;
    __shared__ float qs[D][kBlock];
    __shared__ float ds[kFromEdge ? 1 : D][kBlock];
    if (threadIdx.x == 0)
        next_edge = e_begin;
    __syncthreads();

    const EnvView E = load_env(obs, meta, env);
    const int lane = threadIdx.x & 63;
    const int t = lane & 31;
    const int half0 = lane & 32;
    const LdsRow Q{&qs[0][threadIdx.x]};
    const LdsRow Dl{&ds[0][threadIdx.x]};

    int cur = -1, step = 0, n = 0;
    bool exhausted = false;

    for (;;) {
        const bool want = (cur < 0) && !exhausted;
        int g = 0;
        if (want && t == 0)
            g = atomicAdd(&next_edge, 1);
        g = __shfl(g, half0);
        if (want) {
            if (g < e_end) {
                cur = g;
                step = 0;
                if constexpr (kFromEdge)
                    n = plain_round0<D>(edges + (size_t)g * 2 * D, resolution, t, Q);
                else
                    n = setup_pair_lds<D>(edges + (size_t)g * 2 * D, resolution, t, Q, Dl);
            } else {
                exhausted = true;
            }
        }
        const uint64_t live = __ballot(cur >= 0);
        if (live == 0)
            break;
        const uint64_t hit = Robot::template check<kUnroll>(Q, E, live);
        if (cur >= 0) {
            const bool h = (hit >> lane) & 1ull;
            bool fin = h;
            if (!h) {
                ++step;
                fin = (step == n);
                if (!fin) {
                    if constexpr (kFromEdge) {
                        plain_next<D>(edges + (size_t)cur * 2 * D, resolution, Q);
                    } else {
#pragma unroll
                        for (int j = 0; j < D; ++j)
                            Q.at(j) = Q[j] + Dl[j];
                    }
                }
            }
            if (fin) {
                if (t == 0) {
                    out[(size_t)cur * num_envs + env] = h ? 1 : 0;
                    if (rounds)
                        rounds[(size_t)cur * num_envs + env] = (uint8_t)min(h ? step + 1 : step, 255);
                }
                cur = -1;
            }
        }
    }
}

// ------------------------------------------------------------ learned order
// A learned order changes which samples a pair visits first, never which
// samples it has: a probe pass checks K samples of each pair, one per lane,
// so 64 / K pairs share a wave (the check program decides lane groups of K
// instead of half-waves).  The probed samples are step 0 of the pair's first
// K predicted bins (bin b = lane b's samples {b n + i}; its step-0 sample is
// start + delta * (b n), exactly the plain one).  Pairs a probe leaves
// undecided run the plain step loop over all 32 n samples.  The verdict is an
// OR over the same sample set either way.  Without a prediction (an invalid
// order word) a probe takes K evenly spaced bins.
#ifndef BCC_PROBE1
#define BCC_PROBE1 2
#endif
constexpr int kProbe1 = BCC_PROBE1;   // predicted-probe samples per pair
static_assert(kProbe1 == 1 || kProbe1 == 2 || kProbe1 == 4, "the probe visits up to the four predicted bins");

constexpr int kOrderedChunkMax = 1024;   // edges per workgroup of the ordered dense kernel
constexpr int kOrderedTileMax = 256;     // pairs per tile / segment of the pipeline kernels

template <int K>
__device__ __forceinline__ int probe_bin(uint32_t top, int k)
{
    return top == 0xffffffffu ? (32 * k + 16) / K : (int)((top >> (8 * k)) & 31u);
}

// The step-0 sample of bin b of a pair, start + delta * (b n) (the plain
// step-0 sample of lane b, same arithmetic), into Q; returns the pair's step
// count n.
template <int D>
__device__ __forceinline__ int probe_at(const float* __restrict__ se, int resolution, int bin, const LdsRow& Q)
{
    float s[D], d[D];
    const int n = setup_pair_sd<D>(se, resolution, s, d);
    const float bn = (float)(bin * n);
#pragma unroll
    for (int j = 0; j < D; ++j)
        Q.at(j) = fmaf(d[j], bn, s[j]);
    return n;
}

// Lane k of a probe group: step 0 of the pair's k-th probed bin into Q.
template <int D, int K>
__device__ __forceinline__ void probe_sample(const float* __restrict__ se, int resolution, uint32_t top, int k,
                                             const LdsRow& Q)
{
    probe_at<D>(se, resolution, probe_bin<K>(top, k), Q);
}

// Step loop of one wave over a queue of pairs (the pairs a probe left
// undecided): each half-wave takes the next queued pair from the LDS counter
// *next and runs its steps in plain order, or in block order after step 0
// where the robot's kBlockSteps says so (block_step).  pair_edge / pair_env /
// pair_out map a queue entry to its edge, environment-independent verdict
// slot and rounds slot.
template <class Robot, int kUnroll, class Pair, class Queue>
__device__ __forceinline__ void plain_queue(const EnvView& E, const float* __restrict__ edges, int resolution,
                                            const Queue& queue, int nq, int* next, const LdsRow& Q, Pair pair)
{
    constexpr int D = Robot::dim;
    const int lane = threadIdx.x & 63;
    const int t = lane & 31;
    const int half0 = lane & 32;
    int cur = -1, step = 0, n = 0;
    const float* se = nullptr;
    bool exhausted = false;
    for (;;) {
        const bool want = (cur < 0) && !exhausted;
        int qi = 0;
        if (want && t == 0)
            qi = atomicAdd(next, 1);
        qi = __shfl(qi, half0);
        if (want) {
            if (qi < nq) {
                cur = queue[qi];
                step = 0;
                se = edges + pair.edge(cur) * 2 * D;
                if constexpr (Robot::kBlockSteps)
                    n = block_step<D>(se, resolution, 0, t, Q);
                else
                    n = plain_round0<D>(se, resolution, t, Q);
            } else {
                exhausted = true;
            }
        }
        const uint64_t live = __ballot(cur >= 0);
        if (live == 0)
            break;
        const uint64_t hit = Robot::template check<kUnroll>(Q, E, live);
        if (cur >= 0) {
            const bool h = (hit >> lane) & 1ull;
            bool fin = h;
            if (!h) {
                ++step;
                fin = (step == n);
                if (!fin) {
                    if constexpr (Robot::kBlockSteps)
                        block_step<D>(se, resolution, step, t, Q);
                    else
                        plain_next<D>(se, resolution, Q);
                }
            }
            if (fin) {
                if (t == 0)
                    pair.finish(cur, h, h ? step + 1 : step);
                cur = -1;
            }
        }
    }
}

struct DensePairs {
    const uint32_t* order;
    uint8_t* out;
    uint8_t* rounds;
    int num_envs, env;
    __device__ __forceinline__ size_t edge(int g) const { return (size_t)g; }
    __device__ __forceinline__ size_t slot(int g) const { return (size_t)g * num_envs + env; }
    __device__ __forceinline__ uint32_t top(int g) const { return valid_order(order[slot(g)]); }
    // rounds: 0 = decided by the probe, else 1 + plain steps run
    __device__ __forceinline__ void finish(int g, bool h, int steps) const
    {
        out[slot(g)] = h ? 1 : 0;
        if (rounds)
            rounds[slot(g)] = (uint8_t)min(steps + 1, 255);
    }
    __device__ __forceinline__ void probe_hit(int g) const
    {
        out[slot(g)] = 1;
        if (rounds)
            rounds[slot(g)] = 0;
    }
};

// Learned-order dense kernel: a workgroup owns one environment and a chunk of
// up to kOrderedChunkMax edges.  Probe pass: each wave takes 64 / K edges at
// a time and lane group e (K lanes) checks edge e's K probed samples; edges
// without a hit go to an LDS queue.  After a barrier the half-waves run the
// plain step loop over the queue.  order[edge][env] holds the learned top-4
// bins; rounds (optional) as DensePairs::finish.
template <class Robot, int K, int kWaves, int kUnroll>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kWaves, 8))) void ordered_dense_kernel(
    const float* __restrict__ obs, const int* __restrict__ meta, const float* __restrict__ edges, int num_edges,
    int num_envs, int resolution, int chunk, int chunks_per_env, const uint32_t* __restrict__ order,
    uint8_t* __restrict__ rounds, uint8_t* __restrict__ out)
{
    constexpr int D = Robot::dim;
    constexpr int kPairs = 64 / K;
    const int env = blockIdx.x / chunks_per_env;
    const int ck = blockIdx.x - env * chunks_per_env;
    const int e_begin = ck * chunk;
    const int e_end = min(num_edges, e_begin + chunk);   // chunk <= kOrderedChunkMax (host-checked)

    __shared__ int next_edge, queued, next_queued;
    __shared__ int queue[kOrderedChunkMax];
    __shared__ float qs[D][kBlock];
    if (threadIdx.x == 0) {
        next_edge = e_begin;
        queued = 0;
        next_queued = 0;
    }
    __syncthreads();

    const EnvView E = load_env(obs, meta, env);
    const int lane = threadIdx.x & 63;
    const int grp = lane / K;
    const int k = lane - grp * K;
    const LdsRow Q{&qs[0][threadIdx.x]};
    const DensePairs P{order, out, rounds, num_envs, env};

    for (;;) {
        int g0 = 0;
        if (lane == 0)
            g0 = atomicAdd(&next_edge, kPairs);
        g0 = __builtin_amdgcn_readfirstlane(__shfl(g0, 0));
        if (g0 >= e_end)
            break;
        const int g = g0 + grp;
        const bool valid = g < e_end;
        if (valid)
            probe_sample<D, K>(edges + (size_t)g * 2 * D, resolution, P.top(g), k, Q);
        const uint64_t live = __ballot(valid);
        const uint64_t hit = Robot::template check<kUnroll, K>(Q, E, live);
        if (valid && k == 0) {
            if ((hit >> lane) & 1ull)
                P.probe_hit(g);
            else
                queue[atomicAdd(&queued, 1)] = g;
        }
    }
    __syncthreads();
    plain_queue<Robot, kUnroll>(E, edges, resolution, queue, queued, &next_queued, Q, P);
}

// ------------------------------------------------------------ ragged batches
// Ragged batch: env e owns the contiguous pairs [env_off[e], env_off[e+1])
// (CSR); pair p checks edges[p] against environment e and writes out[p].
// The pairs are cut into tiles (env, first pair, end pair) on the host.  The
// kernel is persistent: every wave loops, taking the next tile from a global
// counter; its two half-waves pull pairs from the tile through a per-wave
// LDS counter, one step per iteration as in dense_kernel.  A wave grabs a new
// tile only when both halves are idle, so the two halves always share one
// environment (obstacle reads stay wave-uniform scalar loads).
struct RaggedTile {
    int env;
    int begin;   // first pair
    int end;     // one past the last pair
    int chunk;   // step chunk: with tile_steps > 0 the tile runs steps [chunk, chunk + 1) * tile_steps
};

// The step dimension: with tile_steps > 0 a tile covers steps [chunk *
// tile_steps, (chunk + 1) * tile_steps) of its pairs, so a pair of n steps
// is spread over ceil(n / tile_steps) tiles that waves take independently
// (batches with fewer pairs than half-waves, long edges).  Such a pair's
// samples are the same values: a chunk starts from the step-0 sample and
// adds delta once per step skipped, the reference's own sequence
// (src/batch_cc.cu:114-128).  Its verdict is an OR over the chunks: the
// plan zero-fills the verdicts, a chunk writes 1 on a collision, and a chunk
// reads the verdict before each step and stops once another chunk decided
// the pair.  A pair that fits one chunk is decided by it alone, as without
// the step dimension.
//
// Steps are taken from the edge (plain_next): on MI355X this is 4-6% faster
// for fetch and baxter batches than an LDS increment row and equal for panda.
template <class Robot, int kWaves, int kUnroll>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kWaves, 8))) void ragged_kernel(
    const float* __restrict__ obs, const int* __restrict__ meta, const float* __restrict__ edges,
    const RaggedTile* __restrict__ tiles, int num_tiles, int* __restrict__ tile_counter, int resolution,
    int tile_steps, uint8_t* __restrict__ out)
{
    constexpr int D = Robot::dim;
    constexpr int kBlockWaves = kBlock / 64;
    __shared__ int next_pair[kBlockWaves];
    __shared__ float qs[D][kBlock];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int t = lane & 31;
    const int half0 = lane & 32;
    const LdsRow Q{&qs[0][threadIdx.x]};

    for (;;) {
        // -- take a tile (wave-uniform)
        int tile = 0;
        if (lane == 0)
            tile = atomicAdd(tile_counter, 1);
        tile = __builtin_amdgcn_readfirstlane(__shfl(tile, 0));
        if (tile >= num_tiles)
            return;
        const RaggedTile T = tiles[tile];
        const int env = __builtin_amdgcn_readfirstlane(T.env);
        const int p_end = __builtin_amdgcn_readfirstlane(T.end);
        const int s_begin = tile_steps > 0 ? __builtin_amdgcn_readfirstlane(T.chunk) * tile_steps : 0;
        if (lane == 0)
            next_pair[wave] = T.begin;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        const EnvView E = load_env(obs, meta, env);
        int cur = -1, step = 0, s_end = 0;
        bool split = false;   // the current pair spans several chunks
        bool exhausted = false;
        for (;;) {
            bool want = (cur < 0) && !exhausted;
            while (want) {   // half-uniform: take pairs until one has steps in this chunk
                int g = 0;
                if (t == 0)
                    g = atomicAdd(&next_pair[wave], 1);
                g = __shfl(g, half0);
                if (g >= p_end) {
                    exhausted = true;
                    break;
                }
                const float* se = edges + (size_t)g * 2 * D;
                const int n = plain_round0<D>(se, resolution, t, Q);
                if (s_begin >= n)
                    continue;   // the pair ends before this chunk
                cur = g;
                step = s_begin;
                split = tile_steps > 0 && n > tile_steps;
                s_end = split ? min(n, s_begin + tile_steps) : n;
                if (s_begin > 0) {
                    // the chunk's first sample: step 0 plus s_begin increments
                    float s0[D], d[D];
                    setup_pair_sd<D>(se, resolution, s0, d);
#pragma unroll
                    for (int j = 0; j < D; ++j) {
                        float q = Q[j];
                        for (int i = 0; i < s_begin; ++i)
                            q = q + d[j];
                        Q.at(j) = q;
                    }
                }
                want = false;
            }
            const uint64_t live = __ballot(cur >= 0);
            if (live == 0)
                break;
            const uint64_t hit = Robot::template check<kUnroll>(Q, E, live);
            if (cur >= 0) {
                const bool h = (hit >> lane) & 1ull;
                bool fin = h;
                if (!h) {
                    ++step;
                    fin = (step == s_end);
                    if (!fin && split) {
                        // another chunk of this pair may have decided it
                        int decided = 0;
                        if (t == 0)
                            decided = __atomic_load_n(out + cur, __ATOMIC_RELAXED);
                        fin = __shfl(decided, half0) != 0;
                    }
                    if (!fin)
                        plain_next<D>(edges + (size_t)cur * 2 * D, resolution, Q);
                }
                if (fin) {
                    if (t == 0 && (h || !split))
                        out[cur] = h ? 1 : 0;
                    cur = -1;
                }
            }
        }
    }
}

// Resets the tile counter inside a captured graph (a kernel node rather than
// a memset node).
static __global__ void reset_counter_kernel(int* counter)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *counter = 0;
}

struct RaggedArgs {
    const float* obs;
    const int* meta;
    const float* edges;
    const RaggedTile* tiles;
    int num_tiles;
    int* tile_counter;
    int resolution;
    int grid;        // persistent workgroups
    uint8_t* out;
    int variant = 0;      // 1: high-occupancy kernel variant, 0: low
    int tile_steps = 0;   // > 0: tiles carry a step chunk of this many steps (verdicts zero-filled first)
};

// Kernel variant knobs of a robot: (waves per SIMD, obstacle-loop unroll) of
// its high- and low-occupancy variants.
struct VariantKnobs {
    int hi_waves, hi_unroll, lo_waves, lo_unroll;
};

// One persistent launch of one kernel variant (instantiated in its own
// translation unit, see bcc_robot_parts.h).
template <class Robot, int W, int U>
hipError_t launch_ragged_variant(const RaggedArgs& a, hipStream_t stream)
{
    hipLaunchKernelGGL((ragged_kernel<Robot, W, U>), dim3(a.grid), dim3(kBlock), 0, stream, a.obs, a.meta, a.edges,
                       a.tiles, a.num_tiles, a.tile_counter, a.resolution, a.tile_steps, a.out);
    return hipGetLastError();
}

template <class Robot, int HW, int HU, int LW, int LU>
hipError_t launch_ragged(const RaggedArgs& a, hipStream_t stream)
{
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, a.tile_counter);
    if (a.num_tiles == 0)
        return hipGetLastError();
    if (a.variant)
        return launch_ragged_variant<Robot, HW, HU>(a, stream);
    return launch_ragged_variant<Robot, LW, LU>(a, stream);
}

// ------------------------------------------------------------ learned-order plans
// A ragged plan with a learned order runs learn_finish_kernel over segments
// (one environment each) of a queue of pairs: each pair checks the step-0
// samples of its top kProbe1 predicted bins, and the pairs still undecided
// run the plain step loop over all 32 n samples.  Every verdict is an OR over
// the reference's sample set; the order only changes which samples are
// visited first.  For an ordered plan the segments are its tiles and the
// queue is the identity (set up on the host).
//
// Pairs are addressed by tiles (env, first, end) of an index range: index i
// is the row of the edge array and its verdict lives at out[i * stride + env
// * env_mult] (ragged: i).
struct PairSlots {
    int stride;
    int env_mult;
    __device__ __forceinline__ size_t slot(int i, int env) const
    {
        return (size_t)i * stride + (size_t)env * env_mult;
    }
};

// Queued pairs of one environment: queue entries [base, base + count).
struct QueueSeg {
    int env;
    int base;
    int count;   // at most kOrderedTileMax
    int pad;
};

// Device counters of one plan launch, zeroed (segment count: preset) at the
// start of every launch by reset_pipeline_kernel.
enum : int {
    kCtrSegs = 2,          // segments in the queue (preset)
    kCtrSegWork = 3,       // segment work counter
    kCtrLearnedHits = 4,   // pairs decided by the learned probe
    kCtrInts = 8
};

static __global__ void reset_pipeline_kernel(int* ctr, int preset_segs)
{
    if (blockIdx.x == 0 && threadIdx.x < kCtrInts)
        ctr[threadIdx.x] = threadIdx.x == kCtrSegs ? preset_segs : 0;
}

struct SlotPairs {
    PairSlots s;
    int env;
    uint8_t* out;
    __device__ __forceinline__ size_t edge(int i) const { return (size_t)i; }
    __device__ __forceinline__ void finish(int i, bool h, int) const { out[s.slot(i, env)] = h ? 1 : 0; }
};

// Persistent: each wave takes the next segment (the count of segments is
// read from ctr, preset by the plan's reset node), probes its pairs
// at the step-0 samples of their top K predicted bins (order[entry], as in
// valid_order; no order: no probe), then runs the plain step loop over the
// pairs still undecided with its two half-waves.
template <class Robot, int K, int kWaves, int kUnroll>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kWaves, 8))) void learn_finish_kernel(
    const float* __restrict__ obs, const int* __restrict__ meta, const float* __restrict__ edges,
    const QueueSeg* __restrict__ segs, int* __restrict__ ctr, int resolution, PairSlots slots,
    const uint32_t* __restrict__ order, const int* __restrict__ queue_idx, uint8_t* __restrict__ out)
{
    constexpr int D = Robot::dim;
    constexpr int kBlockWaves = kBlock / 64;
    constexpr int kPairs = 64 / K;
    __shared__ int queued[kBlockWaves], next_queued[kBlockWaves];
    __shared__ int queue[kBlockWaves][kOrderedTileMax];
    __shared__ float qs[D][kBlock];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int grp = lane / K;
    const int k = lane - grp * K;
    const LdsRow Q{&qs[0][threadIdx.x]};
    const int num_segs = __builtin_amdgcn_readfirstlane(__atomic_load_n(ctr + kCtrSegs, __ATOMIC_RELAXED));

    for (;;) {
        int sg = 0;
        if (lane == 0)
            sg = atomicAdd(ctr + kCtrSegWork, 1);
        sg = __builtin_amdgcn_readfirstlane(__shfl(sg, 0));
        if (sg >= num_segs)
            return;
        const QueueSeg S = segs[sg];
        const int env = __builtin_amdgcn_readfirstlane(S.env);
        const int base = __builtin_amdgcn_readfirstlane(S.base);
        const int count = __builtin_amdgcn_readfirstlane(S.count);
        if (lane == 0) {
            queued[wave] = 0;
            next_queued[wave] = 0;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        const EnvView E = load_env(obs, meta, env);
        if (order) {
            int hits = 0;
            for (int j0 = 0; j0 < count; j0 += kPairs) {
                const int j = j0 + grp;
                const bool valid = j < count;
                const int i = valid ? queue_idx[base + j] : 0;
                if (valid)
                    probe_sample<D, K>(edges + (size_t)i * 2 * D, resolution, valid_order(order[base + j]), k, Q);
                const uint64_t live = __ballot(valid);
                const uint64_t hit = Robot::template check<kUnroll, K>(Q, E, live);
                const bool h = (hit >> lane) & 1ull;
                if (valid && k == 0) {
                    if (h)
                        out[slots.slot(i, env)] = 1;
                    else
                        queue[wave][atomicAdd(&queued[wave], 1)] = i;
                }
                hits += __popcll(__ballot(valid && k == 0 && h));
            }
            if (lane == 0 && hits)
                atomicAdd(ctr + kCtrLearnedHits, hits);
        } else {
            for (int j = lane; j < count; j += 64)
                queue[wave][j] = queue_idx[base + j];
            if (lane == 0)
                queued[wave] = count;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        const int nq = __builtin_amdgcn_readfirstlane(queued[wave]);
        plain_queue<Robot, kUnroll>(E, edges, resolution, queue[wave], nq, &next_queued[wave], Q,
                                    SlotPairs{slots, env, out});
    }
}

struct LearnArgs {
    const float* obs;
    const int* meta;
    const float* edges;
    int* ctr;                  // [kCtrInts]
    int resolution;
    int grid;                  // persistent workgroups
    int variant;               // 1: high-occupancy kernel variant, 0: low
    PairSlots slots;
    uint8_t* out;
    int* queue_idx;
    QueueSeg* segs;
    const uint32_t* order;     // per queue entry, or null
};

template <class Robot, int W, int U>
hipError_t launch_learn_finish_variant(const LearnArgs& a, hipStream_t stream)
{
    hipLaunchKernelGGL((learn_finish_kernel<Robot, kProbe1, W, U>), dim3(a.grid), dim3(kBlock), 0, stream, a.obs,
                       a.meta, a.edges, a.segs, a.ctr, a.resolution, a.slots, a.order, a.queue_idx, a.out);
    return hipGetLastError();
}

// pipeline stage 3 (learned probe + plain finish) of one plan launch
template <class Robot, int HW, int HU, int LW, int LU>
hipError_t launch_learn(const LearnArgs& a, hipStream_t stream)
{
    return a.variant ? launch_learn_finish_variant<Robot, HW, HU>(a, stream)
                     : launch_learn_finish_variant<Robot, LW, LU>(a, stream);
}

// ------------------------------------------------------------ fused learned step
// The whole learned step of a batch in one kernel: inference runs inside the
// checking kernel, and only for the pairs that need it.  Each wave owns a
// tile (one environment, up to kLearnTile pairs; persistent waves take tiles
// from a counter) and runs four wave-local phases, with no workgroup
// barriers, so a wave that ends early takes its next tile at once:
//   1. probe: every pair's step-0 samples at K evenly spaced bins,
//      (32 k + 16) / K, 64 / K pairs per batch; a hit decides the pair, the
//      others are queued in LDS;
//   2. inference, for the queued pairs of at least min_steps steps only
//      (shorter pairs go straight to the step loop: their whole sample set
//      costs too little for a model to save more than it costs): the MLP on
//      the matrix cores, 16 pairs per MFMA column tile, giving each pair its
//      order word (the spread top-4 among the bins not yet probed) in LDS;
//   3. probes of the queued pairs at the word's first K bins (and, with
//      three passes, then at its next K); without a model: fixed bins between
//      the ones probed before; each pass compacts the queue in place;
//   4. the plain step loop, a half-wave per pair, over the pairs still
//      undecided.
// The probes only visit some of the reference's samples first; every verdict
// is still the OR over the whole sample set, so it equals plain order's for
// any model.  While one wave runs its MFMA phase the other waves on its SIMD
// run check code.
constexpr int kLearnTile = 256;    // pairs per tile
constexpr int kLearnPassesMax = 3;  // probe passes: the first, then up to 4 / K predicted ones
#ifndef BCC_LEARN_PASSES
#define BCC_LEARN_PASSES 2
#endif
constexpr int kLearnPassesDefault = BCC_LEARN_PASSES;

// Fixed bins of probe pass p (the first pass always; the later ones without a
// model): (32 k + offset) / K, k = 0..K-1 -- for K = 2: 8 and 24, 0 and 16,
// 4 and 20.
__host__ __device__ __forceinline__ constexpr int fixed_offset(int pass)
{
    return pass == 0 ? 16 : pass == 1 ? 0 : 8;
}

// Model operands of the fused learned kernel.
struct MlpArgs {
    const bf16x8* frag;   // packed weight fragments (kFrags x 64 lanes); null: no model
    const float* b1;
    const float* b2;
    const float* b3;
    const float* enc;     // [num_envs][kEnc] environment codes
    const float* lo;      // joint lower limits [dim]
    const float* scale;   // 2 / joint range [dim]
    uint32_t far0;        // bins that count as far from the first probe's bins
    int passes;           // probe passes, 1 .. kLearnPassesMax
    int min_steps;        // pairs of fewer steps skip inference and the later probes
};

// A wave's queue: `front` entries from the start of q, then the entries
// written from the end of the array (q[cap - 1], q[cap - 2], ...).
struct SplitQueue {
    const int* q;
    int front, cap;
    __device__ __forceinline__ int operator[](int i) const { return i < front ? q[i] : q[cap - 1 - (i - front)]; }
};

// Where tiles come from: an explicit list (ragged batches), or dense tiles,
// tile t = environment t / dense_chunks and edges [(t % dense_chunks) *
// dense_tile, + dense_tile): waves that take neighbouring tiles share an
// environment, so its obstacle records stay in the scalar cache.
struct LearnTiles {
    const RaggedTile* list;
    int num_tiles;
    int* counter;   // next-tile counter, zeroed before the launch
    int dense_chunks, dense_edges, dense_tile;
};

// Outputs of the fused learned kernel; pair index i is the edge row, its
// verdict lives at out[slots.slot(i, env)] (stage and words likewise).
struct LearnedOut {
    PairSlots slots;
    uint8_t* out;
    uint8_t* stage;    // optional: p = decided by probe pass p, else passes - 1 + plain steps run (<= 255)
    uint32_t* words;   // optional: the order word inferred for each pair that reached phase 2
    int* stats;        // optional [5]: pairs decided by probe passes 0, 1, 2, inferred, sent to the step loop
};

struct LearnedPairs {
    LearnedOut o;
    int env;
    int passes;
    __device__ __forceinline__ size_t edge(int i) const { return (size_t)i; }
    __device__ __forceinline__ void finish(int i, bool h, int steps) const
    {
        const size_t sl = o.slots.slot(i, env);
        o.out[sl] = h ? 1 : 0;
        if (o.stage)
            o.stage[sl] = (uint8_t)min(passes - 1 + steps, 255);
    }
};

// Order word of the pair in column l % 16 of a 16-pair tile (ok: the
// column holds a pair): operands, logits, and the spread top picks among the
// bins not in `probed`.  Not inlined: its registers (fragment offsets,
// activations) are allocated apart from the check program's.
template <int D>
__device__ __attribute__((noinline)) uint32_t infer_word(const MlpArgs& mlp, const float* __restrict__ code,
                                                         const float* __restrict__ edges, int i, bool ok,
                                                         uint32_t probed)
{
    bf16x8 xb[kKT1];
    mlp_pair_operands(code, edges, D, mlp.lo, mlp.scale, i, ok, xb);
    float z[8];
    mlp_tile_logits(mlp.frag, mlp.b1, mlp.b2, mlp.b3, xb, z);
    return spread_top4_lanes_from(z, (threadIdx.x & 63) >> 4, probed, mlp.far0);
}

template <class Robot, int K, int kWaves, int kUnroll>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kWaves, 8))) void learned_kernel(
    const float* __restrict__ obs, const int* __restrict__ meta, const float* __restrict__ edges, LearnTiles tl,
    int resolution, MlpArgs mlp, LearnedOut res)
{
    constexpr int D = Robot::dim;
    constexpr int kBlockWaves = kBlock / 64;
    constexpr int kPairs = 64 / K;
    // the predicted passes use the order word's four bins, K at a time
    static_assert(K <= 4, "a probe pass visits at most the order word's four bins");
    constexpr int kPassesK = 1 + 4 / K < kLearnPassesMax ? 1 + 4 / K : kLearnPassesMax;
    __shared__ int next_queued[kBlockWaves];
    __shared__ int queue[kBlockWaves][kLearnTile];
    __shared__ uint32_t word[kBlockWaves][kLearnTile];
    __shared__ float qs[D][kBlock];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int grp = lane / K;
    const int k = lane - grp * K;
    const uint64_t below = (1ull << lane) - 1ull;
    const LdsRow Q{&qs[0][threadIdx.x]};
    const int passes = min(max(mlp.passes, 1), kPassesK);
    uint32_t probed = 0;   // the first probe's bins
#pragma unroll
    for (int kk = 0; kk < K; ++kk)
        probed |= 1u << ((32 * kk + fixed_offset(0)) / K);

    for (;;) {
        // -- take a tile (wave-uniform)
        int t = 0;
        if (lane == 0)
            t = atomicAdd(tl.counter, 1);
        const int tile = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
        if (tile >= tl.num_tiles)
            return;
        int env, begin, end;
        if (tl.list) {
            const RaggedTile T = tl.list[tile];
            env = T.env;
            begin = T.begin;
            end = T.end;
        } else {
            env = tile / tl.dense_chunks;
            begin = (tile - env * tl.dense_chunks) * tl.dense_tile;
            end = min(tl.dense_edges, begin + tl.dense_tile);
        }
        env = __builtin_amdgcn_readfirstlane(env);
        begin = __builtin_amdgcn_readfirstlane(begin);
        end = __builtin_amdgcn_readfirstlane(end);
        const EnvView E = load_env(obs, meta, env);
        int nq = 0;   // queued pairs at the front of the queue (wave-uniform)
        int nshort = 0;   // pairs queued from the back, straight for the step loop
        int hits[kLearnPassesMax] = {0, 0, 0}, inferred = 0;
#pragma unroll 1
        for (int pass = 0; pass < passes; ++pass) {
            const int cnt = pass == 0 ? end - begin : nq;
            if (pass > 0 && cnt == 0)
                break;
            if (pass == 1 && mlp.frag) {
                // -- phase 2: inference over the queue, 16 pairs per column tile
                const float* code = mlp.enc + (size_t)env * kEnc;
                for (int c0 = 0; c0 < cnt; c0 += 16) {
                    const int j = c0 + (lane & 15);
                    const bool ok = j < cnt;
                    const int i = ok ? queue[wave][j] : 0;
                    const uint32_t w = infer_word<D>(mlp, code, edges, i, ok, probed);
                    if (ok && lane < 16) {
                        word[wave][j] = w;
                        if (res.words)
                            res.words[res.slots.slot(i, env)] = w;
                    }
                }
                inferred = cnt;
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            }
            // -- phases 1 and 3: probe K bins of each pair, queue the misses
            // (later passes compact in place: the entries a batch writes lie
            // below the end of the batch, which it has already read)
            nq = 0;
            for (int j0 = 0; j0 < cnt; j0 += kPairs) {
                const int j = j0 + grp;
                const bool valid = j < cnt;
                int i = 0, n = 0;
                uint32_t w = 0;   // the pair's order word (later passes, with a model)
                if (valid) {
                    int bin;
                    if (pass == 0) {
                        i = begin + j;
                        bin = (32 * k + fixed_offset(0)) / K;
                    } else {
                        i = queue[wave][j];
                        if (mlp.frag) {
                            w = word[wave][j];
                            bin = (int)((w >> (8 * (K * (pass - 1) + k))) & 31u);
                        } else {
                            bin = (32 * k + fixed_offset(pass)) / K;
                        }
                    }
                    n = probe_at<D>(edges + (size_t)i * 2 * D, resolution, bin, Q);
                }
                const uint64_t live = __ballot(valid);
                const uint64_t hit = Robot::template check<kUnroll, K>(Q, E, live);
                const bool h = (hit >> lane) & 1ull;
                const bool lead = valid && k == 0;
                const bool shortp = pass == 0 && n < mlp.min_steps;
                const uint64_t miss = __ballot(lead && !h && !shortp);
                const uint64_t miss_short = __ballot(lead && !h && shortp);
                if (lead) {
                    if (h) {
                        const size_t sl = res.slots.slot(i, env);
                        res.out[sl] = 1;
                        if (res.stage)
                            res.stage[sl] = (uint8_t)pass;
                    } else if (shortp) {
                        queue[wave][kLearnTile - 1 - (nshort + __popcll(miss_short & below))] = i;
                    } else {
                        // the word moves with its pair
                        const int pos = nq + __popcll(miss & below);
                        queue[wave][pos] = i;
                        word[wave][pos] = w;
                    }
                }
                nshort += __popcll(miss_short);
                const int decided = __popcll(__ballot(lead && h));
                if (pass == 0)
                    hits[0] += decided;
                else if (pass == 1)
                    hits[1] += decided;
                else
                    hits[2] += decided;
                nq += __popcll(miss);
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        }
        // -- phase 4: plain steps over the pairs still undecided
        if (lane == 0)
            next_queued[wave] = 0;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        plain_queue<Robot, kUnroll>(E, edges, resolution, SplitQueue{queue[wave], nq, kLearnTile}, nq + nshort,
                                    &next_queued[wave], Q, LearnedPairs{res, env, passes});
        if (res.stats && lane == 0) {
            atomicAdd(res.stats + 0, hits[0]);
            atomicAdd(res.stats + 1, hits[1]);
            atomicAdd(res.stats + 2, hits[2]);
            atomicAdd(res.stats + 3, inferred);
            atomicAdd(res.stats + 4, nq + nshort);
        }
    }
}

struct LearnedArgs {
    const float* obs;
    const int* meta;
    const float* edges;
    LearnTiles tiles;
    int resolution;
    int grid;      // persistent workgroups
    int variant;   // 1: high-occupancy kernel variant, 0: low
    MlpArgs mlp;
    LearnedOut out;
};

template <class Robot, int W, int U>
hipError_t launch_learned_variant(const LearnedArgs& a, hipStream_t stream)
{
    hipLaunchKernelGGL((learned_kernel<Robot, kProbe1, W, U>), dim3(a.grid), dim3(kBlock), 0, stream, a.obs, a.meta,
                       a.edges, a.tiles, a.resolution, a.mlp, a.out);
    return hipGetLastError();
}

template <class Robot, int HW, int HU, int LW, int LU>
hipError_t launch_learned(const LearnedArgs& a, hipStream_t stream)
{
    if (a.tiles.num_tiles == 0)
        return hipSuccess;
    return a.variant ? launch_learned_variant<Robot, HW, HU>(a, stream)
                     : launch_learned_variant<Robot, LW, LU>(a, stream);
}

// ------------------------------------------------------------ test hooks

template <class Robot>
__global__ void fk_kernel(const float* __restrict__ qs, int n, float* __restrict__ out)
{
    constexpr int D = Robot::dim;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float q[D];
#pragma unroll
    for (int j = 0; j < D; ++j)
        q[j] = qs[(size_t)i * D + j];
    Robot::spheres(q, out + (size_t)i * Robot::num_spheres * 4);
}

template <class Robot>
__global__ void samples_kernel(const float* __restrict__ edges, int m, int resolution, int* __restrict__ n_out,
                               float* __restrict__ cfg0, float* __restrict__ cfg_last)
{
    constexpr int D = Robot::dim;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int edge = gid >> 5;
    const int t = gid & 31;
    if (edge >= m)
        return;
    float q[D], d[D];
    const int n = setup_pair<D>(edges + (size_t)edge * 2 * D, resolution, t, q, d);
    if (t == 0)
        n_out[edge] = n;
#pragma unroll
    for (int j = 0; j < D; ++j)
        cfg0[((size_t)edge * 32 + t) * D + j] = q[j];
    for (int i = 1; i < n; ++i) {
#pragma unroll
        for (int j = 0; j < D; ++j)
            q[j] = q[j] + d[j];
    }
#pragma unroll
    for (int j = 0; j < D; ++j)
        cfg_last[((size_t)edge * 32 + t) * D + j] = q[j];
}

// Device copy of a robot's check-program tables (robots/bcc_model.h),
// uploaded once per device; used only by the work counter.
struct ModelDev {
    const float* sphere_r;
    const int* env_gate;
    const int* env_lone;
    const int* env_fine_off;
    const int* env_fine;
    const int* self_gate_a;
    const int* self_gate_b;
    const int* self_lone;
    const int* self_a_off;
    const int* self_a;
    const int* self_b_off;
    const int* self_b;
    const int* prog_kind;
    const int* prog_index;
    const int* prog_fk_ops;
    const int* prog_fk_trig;
    int num_program;
    int ref_fk_flops;
    int ref_fk_trig;
};

// One sample replayed in the reference's program order with its early
// returns (fkcc<Robot>, src/utils.cuh:342-5442 and the other robots): the
// generated FK fills every sphere centre, the program tables drive the tests.
template <class Robot>
__device__ __forceinline__ bool ref_sample_work(const float (&q)[Robot::dim], const RefEnvView& R, const ModelDev& m,
                                                uint64_t& w, uint64_t& tr)
{
    float P[Robot::num_spheres * 4];
    Robot::spheres(q, P);
    for (int it = 0; it < m.num_program; ++it) {
        const int g = m.prog_index[it];
        bool col = false;
        if (m.prog_kind[it] == 0) {
            const float* c = P + 4 * m.env_gate[g];
            if (bcc_ref_env_hit(R, c[0], c[1], c[2], (float)((double)c[3] + 0.005), w)) {
                if (m.env_lone[g]) {
                    col = true;
                } else {
                    for (int f = m.env_fine_off[g]; f < m.env_fine_off[g + 1] && !col; ++f) {
                        const float* cf = P + 4 * m.env_fine[f];
                        col = bcc_ref_env_hit(R, cf[0], cf[1], cf[2], (float)((double)cf[3] + 0.005), w);
                    }
                }
            }
        } else {
            const float* a = P + 4 * m.self_gate_a[g];
            const float* b = P + 4 * m.self_gate_b[g];
            w += 11;
            if (bcc_self_hit(a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3])) {
                if (m.self_lone[g]) {
                    col = true;
                } else {
                    for (int ia = m.self_a_off[g]; ia < m.self_a_off[g + 1] && !col; ++ia) {
                        const float* fa = P + 4 * m.self_a[ia];
                        for (int ib = m.self_b_off[g]; ib < m.self_b_off[g + 1] && !col; ++ib) {
                            const float* fb = P + 4 * m.self_b[ib];
                            w += 11;
                            col = bcc_self_hit(fa[0], fa[1], fa[2], fa[3], fb[0], fb[1], fb[2], fb[3]);
                        }
                    }
                }
            }
        }
        if (col) {
            w += (uint64_t)m.prog_fk_ops[it];
            tr += (uint64_t)m.prog_fk_trig[it];
            return true;
        }
    }
    w += (uint64_t)m.ref_fk_flops;
    tr += (uint64_t)m.ref_fk_trig;
    return false;
}

// Reference-order work counter: one half-wave per pair (grid-stride), every
// lane replays its sample, the step loop stops after the first step in which
// any lane collided (src/batch_cc.cu:118-129).  Adds each pair's operation
// and sin/cos counts into work[p] / trig[p] (zeroed by the caller) and writes
// its verdict.
template <class Robot>
__global__ __launch_bounds__(kBlock) void work_kernel(const float* __restrict__ ref_obs,
                                                      const int* __restrict__ ref_meta,
                                                      const float* __restrict__ edges, int num_edges, int num_envs,
                                                      int resolution, ModelDev model,
                                                      const int* __restrict__ pair_env,
                                                      unsigned long long* __restrict__ work,
                                                      unsigned long long* __restrict__ trig,
                                                      uint8_t* __restrict__ out)
{
    // dense: pair p = edge * num_envs + env; with pair_env (ragged): env =
    // pair_env[p], edge = p, and num_edges is the number of pairs
    constexpr int D = Robot::dim;
    const int lane = threadIdx.x & 63;
    const int t = lane & 31;
    const long long pairs = pair_env ? (long long)num_edges : (long long)num_edges * num_envs;
    const long long halves = ((long long)gridDim.x * kBlock) >> 5;
    for (long long p = ((long long)blockIdx.x * kBlock + threadIdx.x) >> 5; p < pairs; p += halves) {
        const int edge = pair_env ? (int)p : (int)(p / num_envs);
        const int env = pair_env ? pair_env[p] : (int)(p - (long long)edge * num_envs);
        const RefEnvView R = load_ref_env(ref_obs, ref_meta, env);
        float q[D], d[D];
        const int n = setup_pair<D>(edges + (size_t)edge * 2 * D, resolution, t, q, d);
        uint64_t w = 0, tr = 0;
        bool any = false;
        for (int i = 0; i < n && !any; ++i) {
            const bool c = ref_sample_work<Robot>(q, R, model, w, tr);
            any = ((__ballot(c) >> (lane & 32)) & 0xffffffffull) != 0;
#pragma unroll
            for (int j = 0; j < D; ++j)
                q[j] = q[j] + d[j];
        }
        atomicAdd(work + p, (unsigned long long)w);
        atomicAdd(trig + p, (unsigned long long)tr);
        if (t == 0)
            out[p] = any ? 1 : 0;
    }
}

// ------------------------------------------------------------ clearance audit
// Signed clearance (metres, negative = penetration) of one robot sphere of
// padded radius rp against one primitive: the distance the reference's
// squared-distance test compares, minus the radii.  FP32; the audit uses it
// only to find pairs near contact, which the FP64 oracle then decides.
__device__ __forceinline__ float clr_sphere(cfptr o, float x, float y, float z, float rp)
{
    const float xs = o[0] - x, ys = o[1] - y, zs = o[2] - z;
    return sqrtf(fmaf(zs, zs, fmaf(ys, ys, xs * xs))) - (o[3] + rp);
}

__device__ __forceinline__ float clr_capsule(cfptr c, float x, float y, float z, float rp, bool z_axis)
{
    const float dot = z_axis ? (z - c[2]) * c[5] : fmaf(z - c[2], c[5], fmaf(y - c[1], c[4], (x - c[0]) * c[3]));
    const float t = fmaxf(fminf(dot * c[7], 1.0f), 0.0f);
    const float xs = z_axis ? x - c[0] : x - fmaf(c[3], t, c[0]);
    const float ys = z_axis ? y - c[1] : y - fmaf(c[4], t, c[1]);
    const float zs = z - fmaf(c[5], t, c[2]);
    return sqrtf(fmaf(zs, zs, fmaf(ys, ys, xs * xs))) - (rp + c[6]);
}

__device__ __forceinline__ float clr_cuboid(cfptr b, float x, float y, float z, float rp, bool z_axis)
{
    const float xs = x - b[0], ys = y - b[1], zs = z - b[2];
    const float d1 = z_axis ? fmaf(b[4], ys, b[3] * xs) : fmaf(b[5], zs, fmaf(b[4], ys, b[3] * xs));
    const float d2 = z_axis ? fmaf(b[7], ys, b[6] * xs) : fmaf(b[8], zs, fmaf(b[7], ys, b[6] * xs));
    const float d3 = z_axis ? zs : fmaf(b[11], zs, fmaf(b[10], ys, b[9] * xs));
    const float a1 = fmaxf(0.0f, fabsf(d1) - b[12]);
    const float a2 = fmaxf(0.0f, fabsf(d2) - b[13]);
    const float a3 = fmaxf(0.0f, fabsf(d3) - b[14]);
    return sqrtf(fmaf(a3, a3, fmaf(a2, a2, a1 * a1))) - rp;
}

// Smallest clearance of one padded sphere over the environment, in the
// reference's lists (src/utils.cuh:213-245: no cylinders).
__device__ __forceinline__ float env_clearance(const RefEnvView& E, float x, float y, float z, float rp)
{
    float best = INFINITY;
    for (int k = 0; k < E.ns; ++k)
        best = fminf(best, clr_sphere(E.sph + 4 * k, x, y, z, rp));
    for (int k = 0; k < E.nc; ++k)
        best = fminf(best, clr_capsule(E.cap + 8 * k, x, y, z, rp, false));
    for (int k = 0; k < E.nzc; ++k)
        best = fminf(best, clr_capsule(E.zcap + 8 * k, x, y, z, rp, true));
    for (int k = 0; k < E.ncb; ++k)
        best = fminf(best, clr_cuboid(E.cub + 16 * k, x, y, z, rp, false));
    for (int k = 0; k < E.nzcb; ++k)
        best = fminf(best, clr_cuboid(E.zcub + 16 * k, x, y, z, rp, true));
    return best;
}

__device__ __forceinline__ float self_clearance(const float* a, const float* b)
{
    const float xs = a[0] - b[0], ys = a[1] - b[1], zs = a[2] - b[2];
    return sqrtf(fmaf(zs, zs, fmaf(ys, ys, xs * xs))) - (a[3] + b[3]);
}

// Gated clearance of one sample: the minimum over the check program's items
// of (lone item: its gate) or max(gate, smallest fine clearance) -- negative
// exactly when a gate and one of its fine spheres (pairs) both collide.  A
// gate no lower than the running minimum cannot lower it, so its fine spheres
// are skipped (the value is unchanged).
template <class Robot>
__device__ __forceinline__ float ref_sample_clearance(const float (&q)[Robot::dim], const RefEnvView& R,
                                                      const ModelDev& m)
{
    float P[Robot::num_spheres * 4];
    Robot::spheres(q, P);
    float best = INFINITY;
    for (int it = 0; it < m.num_program; ++it) {
        const int g = m.prog_index[it];
        if (m.prog_kind[it] == 0) {
            const float* c = P + 4 * m.env_gate[g];
            const float cg = env_clearance(R, c[0], c[1], c[2], (float)((double)c[3] + 0.005));
            if (m.env_lone[g] || cg >= best) {
                best = fminf(best, cg);
                continue;
            }
            float fmin = INFINITY;
            for (int f = m.env_fine_off[g]; f < m.env_fine_off[g + 1]; ++f) {
                const float* cf = P + 4 * m.env_fine[f];
                fmin = fminf(fmin, env_clearance(R, cf[0], cf[1], cf[2], (float)((double)cf[3] + 0.005)));
            }
            best = fminf(best, fmaxf(cg, fmin));
        } else {
            const float cg = self_clearance(P + 4 * m.self_gate_a[g], P + 4 * m.self_gate_b[g]);
            if (m.self_lone[g] || cg >= best) {
                best = fminf(best, cg);
                continue;
            }
            float fmin = INFINITY;
            for (int ia = m.self_a_off[g]; ia < m.self_a_off[g + 1]; ++ia)
                for (int ib = m.self_b_off[g]; ib < m.self_b_off[g + 1]; ++ib)
                    fmin = fminf(fmin, self_clearance(P + 4 * m.self_a[ia], P + 4 * m.self_b[ib]));
            best = fminf(best, fmaxf(cg, fmin));
        }
    }
    return best;
}

// Clearance audit: per pair, the minimum gated clearance over all 32 n
// samples (no early exit), one half-wave per pair as the work counter
// (grid-stride; dense: pair p = edge * num_envs + env, with pair_env:
// env = pair_env[p], edge = p).  A pair collides exactly when some sample's
// gated clearance is negative, so FP32 verdicts can differ from exact
// arithmetic only where |clearance| is within FP32 rounding of zero.
template <class Robot>
__global__ __launch_bounds__(kBlock) void audit_kernel(const float* __restrict__ ref_obs,
                                                       const int* __restrict__ ref_meta,
                                                       const float* __restrict__ edges, int num_edges, int num_envs,
                                                       int resolution, ModelDev model,
                                                       const int* __restrict__ pair_env,
                                                       float* __restrict__ clearance)
{
    constexpr int D = Robot::dim;
    const int lane = threadIdx.x & 63;
    const int t = lane & 31;
    const long long pairs = pair_env ? (long long)num_edges : (long long)num_edges * num_envs;
    const long long halves = ((long long)gridDim.x * kBlock) >> 5;
    for (long long p = ((long long)blockIdx.x * kBlock + threadIdx.x) >> 5; p < pairs; p += halves) {
        const int edge = pair_env ? (int)p : (int)(p / num_envs);
        const int env = pair_env ? pair_env[p] : (int)(p - (long long)edge * num_envs);
        const RefEnvView R = load_ref_env(ref_obs, ref_meta, env);
        float q[D], d[D];
        const int n = setup_pair<D>(edges + (size_t)edge * 2 * D, resolution, t, q, d);
        float best = INFINITY;
        for (int i = 0; i < n; ++i) {
            best = fminf(best, ref_sample_clearance<Robot>(q, R, model));
#pragma unroll
            for (int j = 0; j < D; ++j)
                q[j] = q[j] + d[j];
        }
#pragma unroll
        for (int o = 16; o > 0; o >>= 1)
            best = fminf(best, __shfl_xor(best, o));
        if (t == 0)
            clearance[p] = best;
    }
}

struct AuditArgs {
    const float* ref_obs;
    const int* ref_meta;
    const float* edges;
    int num_edges;
    int num_envs;
    int resolution;
    ModelDev model;
    const int* pair_env;
    float* clearance;
};

template <class Robot>
hipError_t launch_audit(const AuditArgs& a, hipStream_t stream)
{
    if (a.num_edges == 0 || a.num_envs == 0)
        return hipSuccess;
    const long long halves = a.pair_env ? (long long)a.num_edges : (long long)a.num_edges * a.num_envs;
    const long long blocks = std::min<long long>((halves + 7) / 8, 65536);
    hipLaunchKernelGGL(audit_kernel<Robot>, dim3((unsigned)blocks), dim3(kBlock), 0, stream, a.ref_obs, a.ref_meta,
                       a.edges, a.num_edges, a.num_envs, a.resolution, a.model, a.pair_env, a.clearance);
    return hipGetLastError();
}

// Training labels: for each pair, bit t of labels[p] is set when one of the
// first `steps` samples of bin t (lane t's samples t n + i, i < steps; steps
// = 0: all n) collides.  Every such sample is evaluated (no early exit) with
// the reference-order replay.  Pair p: env = pair_env[p], edge = pair_edge[p].
template <class Robot>
__global__ __launch_bounds__(kBlock) void label_kernel(const float* __restrict__ ref_obs,
                                                       const int* __restrict__ ref_meta,
                                                       const float* __restrict__ edges, int num_pairs,
                                                       int resolution, int steps, ModelDev model,
                                                       const int* __restrict__ pair_env,
                                                       const int* __restrict__ pair_edge,
                                                       uint32_t* __restrict__ labels)
{
    constexpr int D = Robot::dim;
    const int lane = threadIdx.x & 63;
    const int t = lane & 31;
    const long long halves = ((long long)gridDim.x * kBlock) >> 5;
    for (long long p = ((long long)blockIdx.x * kBlock + threadIdx.x) >> 5; p < num_pairs; p += halves) {
        const int env = pair_env[p];
        const int edge = pair_edge[p];
        const RefEnvView R = load_ref_env(ref_obs, ref_meta, env);
        float s[D], d[D], q[D];
        const int n = setup_pair_sd<D>(edges + (size_t)edge * 2 * D, resolution, s, d);
        const float tn = (float)(t * n);
#pragma unroll
        for (int j = 0; j < D; ++j)
            q[j] = fmaf(d[j], tn, s[j]);
        bool any = false;
        uint64_t w = 0, tr = 0;
        const int last = steps > 0 ? min(n, steps) : n;
        for (int i = 0; i < last; ++i) {
            any = ref_sample_work<Robot>(q, R, model, w, tr) || any;
#pragma unroll
            for (int j = 0; j < D; ++j)
                q[j] = q[j] + d[j];
        }
        const uint32_t bits = (uint32_t)(__ballot(any) >> (lane & 32));
        if (t == 0)
            labels[p] = bits;
    }
}

struct LabelArgs {
    const float* ref_obs;
    const int* ref_meta;
    const float* edges;
    int num_pairs;
    int resolution;
    int steps;   // samples per bin evaluated (0: all)
    ModelDev model;
    const int* pair_env;
    const int* pair_edge;
    uint32_t* labels;
};

template <class Robot>
hipError_t launch_label(const LabelArgs& a, hipStream_t stream)
{
    if (a.num_pairs == 0)
        return hipSuccess;
    const long long blocks = std::min<long long>(((long long)a.num_pairs + 7) / 8, 65536);
    hipLaunchKernelGGL(label_kernel<Robot>, dim3((unsigned)blocks), dim3(kBlock), 0, stream, a.ref_obs, a.ref_meta,
                       a.edges, a.num_pairs, a.resolution, a.steps, a.model, a.pair_env, a.pair_edge, a.labels);
    return hipGetLastError();
}

struct WorkArgs {
    const float* ref_obs;
    const int* ref_meta;
    const float* edges;
    int num_edges;
    int num_envs;
    int resolution;
    ModelDev model;
    const int* pair_env;
    unsigned long long* work;
    unsigned long long* trig;
    uint8_t* out;
};

template <class Robot>
hipError_t launch_work(const WorkArgs& a, hipStream_t stream)
{
    if (a.num_edges == 0 || a.num_envs == 0)
        return hipSuccess;
    const long long halves = a.pair_env ? (long long)a.num_edges : (long long)a.num_edges * a.num_envs;
    const long long blocks = std::min<long long>((halves + 7) / 8, 65536);
    hipLaunchKernelGGL(work_kernel<Robot>, dim3((unsigned)blocks), dim3(kBlock), 0, stream, a.ref_obs, a.ref_meta,
                       a.edges, a.num_edges, a.num_envs, a.resolution, a.model, a.pair_env, a.work, a.trig, a.out);
    return hipGetLastError();
}

struct DenseArgs {
    const float* obs;
    const int* meta;
    const float* edges;
    int num_edges;
    int num_envs;
    int resolution;
    uint8_t* out;
    const uint32_t* order = nullptr;   // learned top-4 bins per pair, or null (plain order)
    uint8_t* rounds = nullptr;         // rounds per pair, or null
    int variant = 0;                   // 1: high-occupancy kernel variant, 0: low
    int chunk = 64;                    // edges per workgroup
};

// One launch of one dense kernel variant (instantiated in its own
// translation unit, see bcc_robot_parts.h).
template <class Robot, bool kOrdered, int W, int U>
hipError_t launch_dense_variant(const DenseArgs& a, unsigned blocks, int chunk, int chunks_per_env,
                                hipStream_t stream)
{
    if constexpr (kOrdered)
        hipLaunchKernelGGL((ordered_dense_kernel<Robot, kProbe1, W, U>), dim3(blocks), dim3(kBlock), 0, stream, a.obs,
                           a.meta, a.edges, a.num_edges, a.num_envs, a.resolution, chunk, chunks_per_env, a.order,
                           a.rounds, a.out);
    else
        hipLaunchKernelGGL((dense_kernel<Robot, W, U>), dim3(blocks), dim3(kBlock), 0, stream, a.obs, a.meta, a.edges,
                           a.num_edges, a.num_envs, a.resolution, chunk, chunks_per_env, a.rounds, a.out);
    return hipGetLastError();
}

template <class Robot, int HW, int HU, int LW, int LU>
hipError_t launch_dense(const DenseArgs& a, hipStream_t stream)
{
    if (a.num_edges == 0 || a.num_envs == 0)
        return hipSuccess;
    const int chunk = a.chunk > 0 ? a.chunk : 64;
    if (a.order && chunk > kOrderedChunkMax)
        return hipErrorInvalidValue;
    const int chunks_per_env = (a.num_edges + chunk - 1) / chunk;
    const long long blocks = (long long)chunks_per_env * a.num_envs;
    if (blocks > 0x7fffffffLL)
        return hipErrorInvalidConfiguration;
    if (a.order && a.variant)
        return launch_dense_variant<Robot, true, HW, HU>(a, (unsigned)blocks, chunk, chunks_per_env, stream);
    if (a.order)
        return launch_dense_variant<Robot, true, LW, LU>(a, (unsigned)blocks, chunk, chunks_per_env, stream);
    if (a.variant)
        return launch_dense_variant<Robot, false, HW, HU>(a, (unsigned)blocks, chunk, chunks_per_env, stream);
    return launch_dense_variant<Robot, false, LW, LU>(a, (unsigned)blocks, chunk, chunks_per_env, stream);
}

template <class Robot>
hipError_t launch_fk(const float* qs, int n, float* out, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(fk_kernel<Robot>, dim3((n + 63) / 64), dim3(64), 0, stream, qs, n, out);
    return hipGetLastError();
}

template <class Robot>
hipError_t launch_samples(const float* edges, int m, int resolution, int* n_out, float* cfg0, float* cfg_last,
                          hipStream_t stream)
{
    if (m == 0)
        return hipSuccess;
    const long long threads = (long long)m * 32;
    hipLaunchKernelGGL(samples_kernel<Robot>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, stream, edges, m,
                       resolution, n_out, cfg0, cfg_last);
    return hipGetLastError();
}

// Per-robot entry points, one translation unit each (robot_<name>.hip).
#define BCC_DECLARE_ROBOT(name)                                                                                \
    hipError_t launch_dense_##name(const DenseArgs& a, hipStream_t stream);                                    \
    hipError_t launch_fk_##name(const float* qs, int n, float* out, hipStream_t stream);                       \
    hipError_t launch_samples_##name(const float* edges, int m, int resolution, int* n_out, float* cfg0,       \
                                     float* cfg_last, hipStream_t stream);                                     \
    hipError_t launch_work_##name(const WorkArgs& a, hipStream_t stream);                                     \
    hipError_t launch_ragged_##name(const RaggedArgs& a, hipStream_t stream);                                 \
    hipError_t launch_learn_##name(const LearnArgs& a, hipStream_t stream);                                   \
    hipError_t launch_learned_##name(const LearnedArgs& a, hipStream_t stream);                               \
    hipError_t launch_label_##name(const LabelArgs& a, hipStream_t stream);                                   \
    hipError_t launch_audit_##name(const AuditArgs& a, hipStream_t stream);                                   \
    extern const VariantKnobs knobs_##name;

BCC_DECLARE_ROBOT(panda)
BCC_DECLARE_ROBOT(fetch)
BCC_DECLARE_ROBOT(baxter)
BCC_DECLARE_ROBOT(xarm7)

}  // namespace bcc
