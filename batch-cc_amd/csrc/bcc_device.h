// Device-side building blocks of the gfx950 batch edge collision checker.
//
// Execution model: a 64-lane wavefront is two independent half-waves.  A
// half-wave owns one (environment, edge) pair at a time and evaluates one
// interpolation step of it per iteration: its 32 lanes are the reference's
// 32 samples {t*n + i} of step i (src/batch_cc.cu:114-128).  Both halves of a
// wave always belong to the same environment, so every obstacle parameter
// is wave-uniform and is read with scalar loads into SGPRs; the robot
// spheres live in VGPRs.  A pair is decided as soon as any of its samples
// collides (the verdict is the OR over the sample set), which is what the
// half-spread masks below track.
//
// Arithmetic follows the project's FP32 spec: the reference's primitive
// formulas (src/utils.cuh:60-210) with the products that feed an add or a
// subtract contracted into fmaf, the same order everywhere, compiled with
// -ffp-contract=off so nothing else fuses.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bcc {

// Obstacles are read through constant-address-space pointers: every lane of
// a wave reads the same obstacle, so the loads become scalar (s_load) and the
// parameters sit in SGPRs as VALU operands.
typedef const __attribute__((address_space(4))) float* cfptr;
typedef const __attribute__((address_space(4))) int* ciptr;

// Obstacles of one environment, device layout (floats).  The host sorts the
// reference's primitives into six lists; every list uses the reference's
// formula for its kind, and the two "exact special cases" are only used for
// primitives on which the special formula is bit-identical to the general one:
//   spheres  [ns][4]    x y z r                             (src/utils.cuh:60-73)
//   capsules [nc][8]    x1 y1 z1 xv yv zv r rdv             (src/utils.cuh:115-128)
//   zcaps    [nzc][8]   z-aligned capsules, and capsules with xv == yv == 0
//                       (src/utils.cuh:137-150)
//   cuboids  [ncb][16]  x y z a1(3) a2(3) a3(3) h1 h2 h3 pad  (src/utils.cuh:160-177)
//   zcubs    [nzcb][16] z-aligned cuboids                   (src/utils.cuh:186-203)
//   aabbs    [na][8]    x y z h1 h2 h3 pad pad: cuboids whose axes are exactly
//                       the identity and z-cuboids with axes (1,0),(0,1)
//   boxes    [nbox][8]  lo.x lo.y lo.z 0 hi.x hi.y hi.z 0: one bounding box
//                       per primitive above, in list order (nbox = the sum
//                       of the six counts); a primitive's test value is
//                       negative only for sphere centres within rp of its box
// Cylinders are never tested by the reference (src/utils.cuh:213-245) and are
// not uploaded.
// What a check program holds of its environment between sweeps: the
// obstacle buffer and the environment's meta record (4 SGPRs).  Each sweep
// reads the record again (bcc_env_lists), so the list pointers and counts do
// not occupy SGPRs through forward kinematics and the self tests.
struct EnvView {
    cfptr obs;
    ciptr meta;
};

// Per-environment record in the scene's meta array (int32): for list k in
// the order above, [2k] = offset in floats, [2k+1] = count; [14] = the
// culling threshold cull_min.
constexpr int kEnvMetaInts = 16;

__device__ __forceinline__ EnvView load_env(const float* obs_, const int* meta_, int env)
{
    return EnvView{(cfptr)obs_, (ciptr)meta_ + env * kEnvMetaInts};
}

struct EnvLists {
    cfptr sph;
    cfptr cap;
    cfptr zcap;
    cfptr cub;
    cfptr zcub;
    cfptr aabb;
    const float* box;   // read per lane (one box per lane), so a plain global pointer
    int ns, nc, nzc, ncb, nzcb, na, nbox;
    int cull_min;       // a sweep of N spheres culls when N * nbox >= cull_min
};

__device__ __forceinline__ EnvLists bcc_env_lists(const EnvView& E)
{
    ciptr m = E.meta;
    // an opaque copy of the pointer: the compiler cannot merge this sweep's
    // scalar loads with another sweep's and keep the values live in between
    asm volatile("" : "+s"(m));
    EnvLists v;
    v.sph = E.obs + m[0];
    v.ns = m[1];
    v.cap = E.obs + m[2];
    v.nc = m[3];
    v.zcap = E.obs + m[4];
    v.nzc = m[5];
    v.cub = E.obs + m[6];
    v.ncb = m[7];
    v.zcub = E.obs + m[8];
    v.nzcb = m[9];
    v.aabb = E.obs + m[10];
    v.na = m[11];
    v.box = (const float*)(uintptr_t)(E.obs + m[12]);
    v.nbox = m[13];
    v.cull_min = m[14];
    return v;
}

// ------------------------------------------------------------ sin/cos
// Cody-Waite reduction by pi/2 (three-float constant), Taylor polynomials of
// degree 9 / 10 on [-pi/4, pi/4].  Identical operation sequence on host and
// device, so FK results do not depend on the math library.
__device__ __forceinline__ void bcc_sincosf(float x, float& s, float& c)
{
    const float kf = __builtin_rintf(x * 0x1.45f306p-1f);
    float r = fmaf(-kf, 0x1.921fb6p+0f, x);
    r = fmaf(-kf, -0x1.777a5cp-25f, r);
    r = fmaf(-kf, -0x1.ee59dap-50f, r);
    const float r2 = r * r;
    float ps = fmaf(r2, (float)(1.0 / 362880.0), (float)(-1.0 / 5040.0));
    ps = fmaf(r2, ps, (float)(1.0 / 120.0));
    ps = fmaf(r2, ps, (float)(-1.0 / 6.0));
    const float r3 = r2 * r;
    const float sn = fmaf(r3, ps, r);
    float pc = fmaf(r2, (float)(-1.0 / 3628800.0), (float)(1.0 / 40320.0));
    pc = fmaf(r2, pc, (float)(-1.0 / 720.0));
    pc = fmaf(r2, pc, (float)(1.0 / 24.0));
    pc = fmaf(r2, pc, -0.5f);
    const float cs = fmaf(r2, pc, 1.0f);
    const int quad = ((int)kf) & 3;
    const float a = (quad & 1) ? cs : sn;
    const float b = (quad & 1) ? sn : cs;
    s = (quad & 2) ? -a : a;
    c = ((quad + 1) & 2) ? -b : b;
}

// ------------------------------------------------------------ masks
// Spread a lane mask to whole half-waves: a half is marked if any of its
// lanes is marked.
__device__ __forceinline__ uint64_t bcc_half_spread(uint64_t m)
{
    const uint32_t lo = (uint32_t)m;
    const uint32_t hi = (uint32_t)(m >> 32);
    return (lo ? 0xffffffffull : 0ull) | (hi ? 0xffffffff00000000ull : 0ull);
}

// Spread a lane mask to aligned groups of G lanes (G a power of two, at most
// 32): a group is marked if any of its lanes is marked.  G = 32 is the
// half-wave spread; G = 1 leaves the mask as it is.
template <int G>
__device__ __forceinline__ uint64_t bcc_group_spread(uint64_t m)
{
    static_assert(G >= 1 && G <= 32 && (G & (G - 1)) == 0, "group width must be a power of two up to 32");
    if constexpr (G == 32) {
        return bcc_half_spread(m);
    } else if constexpr (G == 1) {
        return m;
    } else {
        // OR each group's bits down into its lowest bit, keep those bits,
        // then copy each kept bit across its group (no carries: groups are
        // disjoint)
#pragma unroll
        for (int s = 1; s < G; s <<= 1)
            m |= m >> s;
        constexpr uint64_t kLow = ~0ull / ((1ull << G) - 1ull);   // bit 0 of every group
        return (m & kLow) * ((1ull << G) - 1ull);
    }
}

// ------------------------------------------------------------ primitives
// Each returns the reference's clearance value; < 0 means collision.

// sphere obstacle o = (x y z r) vs robot sphere padded radius rp  (src/utils.cuh:60-73)
__device__ __forceinline__ float bcc_sphere_val(float ox, float oy, float oz, float orad,
                                                float x, float y, float z, float rp)
{
    const float xs = ox - x, ys = oy - y, zs = oz - z;
    const float sum = fmaf(zs, zs, fmaf(ys, ys, xs * xs));
    const float rs = orad + rp;
    return fmaf(-rs, rs, sum);
}

// capsule (src/utils.cuh:115-128)
__device__ __forceinline__ float bcc_capsule_val(cfptr c, float x, float y, float z, float rp)
{
    const float x1 = c[0], y1 = c[1], z1 = c[2], xv = c[3], yv = c[4], zv = c[5], cr = c[6], rdv = c[7];
    const float dot = fmaf(z - z1, zv, fmaf(y - y1, yv, (x - x1) * xv));
    const float cdf = fmaxf(fminf(dot * rdv, 1.0f), 0.0f);
    const float xs = x - fmaf(xv, cdf, x1);
    const float ys = y - fmaf(yv, cdf, y1);
    const float zs = z - fmaf(zv, cdf, z1);
    const float sum = fmaf(zs, zs, fmaf(ys, ys, xs * xs));
    const float rs = rp + cr;
    return fmaf(-rs, rs, sum);
}

// z-aligned capsule (src/utils.cuh:137-150)
__device__ __forceinline__ float bcc_zcapsule_val(cfptr c, float x, float y, float z, float rp)
{
    const float x1 = c[0], y1 = c[1], z1 = c[2], zv = c[5], cr = c[6], rdv = c[7];
    const float dot = (z - z1) * zv;
    const float cdf = fmaxf(fminf(dot * rdv, 1.0f), 0.0f);
    const float xs = x - x1;
    const float ys = y - y1;
    const float zs = z - fmaf(zv, cdf, z1);
    const float sum = fmaf(zs, zs, fmaf(ys, ys, xs * xs));
    const float rs = rp + cr;
    return fmaf(-rs, rs, sum);
}

// oriented cuboid (src/utils.cuh:160-177); rsq = rp * rp
__device__ __forceinline__ float bcc_cuboid_val(cfptr b, float x, float y, float z, float rsq)
{
    const float xs = x - b[0], ys = y - b[1], zs = z - b[2];
    const float d1 = fmaf(b[5], zs, fmaf(b[4], ys, b[3] * xs));
    const float d2 = fmaf(b[8], zs, fmaf(b[7], ys, b[6] * xs));
    const float d3 = fmaf(b[11], zs, fmaf(b[10], ys, b[9] * xs));
    const float a1 = fmaxf(0.0f, fabsf(d1) - b[12]);
    const float a2 = fmaxf(0.0f, fabsf(d2) - b[13]);
    const float a3 = fmaxf(0.0f, fabsf(d3) - b[14]);
    const float sum = fmaf(a3, a3, fmaf(a2, a2, a1 * a1));
    return sum - rsq;
}

// z-aligned cuboid (src/utils.cuh:186-203)
__device__ __forceinline__ float bcc_zcuboid_val(cfptr b, float x, float y, float z, float rsq)
{
    const float xs = x - b[0], ys = y - b[1], zs = z - b[2];
    const float d1 = fmaf(b[4], ys, b[3] * xs);
    const float d2 = fmaf(b[7], ys, b[6] * xs);
    const float a1 = fmaxf(0.0f, fabsf(d1) - b[12]);
    const float a2 = fmaxf(0.0f, fabsf(d2) - b[13]);
    const float a3 = fmaxf(0.0f, fabsf(zs) - b[14]);
    const float sum = fmaf(a3, a3, fmaf(a2, a2, a1 * a1));
    return sum - rsq;
}

// axis-aligned cuboid: src/utils.cuh:160-177 with identity axes, where
// fmaf(0, ys, 1 * xs) == xs etc. make the projections exact copies of xs, ys, zs
__device__ __forceinline__ float bcc_aabb_val(cfptr b, float x, float y, float z, float rsq)
{
    const float xs = x - b[0], ys = y - b[1], zs = z - b[2];
    const float a1 = fmaxf(0.0f, fabsf(xs) - b[3]);
    const float a2 = fmaxf(0.0f, fabsf(ys) - b[4]);
    const float a3 = fmaxf(0.0f, fabsf(zs) - b[5]);
    const float sum = fmaf(a3, a3, fmaf(a2, a2, a1 * a1));
    return sum - rsq;
}

// robot self pair, unpadded (src/utils.cuh:85-88)
__device__ __forceinline__ bool bcc_self_hit(float ax, float ay, float az, float ar,
                                             float bx, float by, float bz, float br)
{
    const float xs = ax - bx, ys = ay - by, zs = az - bz;
    const float sum = fmaf(zs, zs, fmaf(ys, ys, xs * xs));
    const float rs = ar + br;
    return fmaf(-rs, rs, sum) < 0.0f;
}

// ------------------------------------------------------------ reference-order replay
// The work counter replays the reference's sequential semantics per lane, so
// it reads the environment in the reference's own lists and order
// (src/utils.cuh:213-245): spheres[4], capsules[8], z-capsules[8],
// cuboids[16], z-cuboids[16]; cylinders are not tested.
struct RefEnvView {
    cfptr sph;
    cfptr cap;
    cfptr zcap;
    cfptr cub;
    cfptr zcub;
    int ns, nc, nzc, ncb, nzcb;
};

__device__ __forceinline__ RefEnvView load_ref_env(const float* obs_, const int* meta_, int env)
{
    const cfptr obs = (cfptr)obs_;
    const ciptr m = (ciptr)meta_ + env * kEnvMetaInts;
    RefEnvView v;
    v.sph = obs + m[0];
    v.ns = m[1];
    v.cap = obs + m[2];
    v.nc = m[3];
    v.zcap = obs + m[4];
    v.nzc = m[5];
    v.cub = obs + m[6];
    v.ncb = m[7];
    v.zcub = obs + m[8];
    v.nzcb = m[9];
    return v;
}

// Reference operation weights: source-level FP32 operations of each call.
constexpr uint32_t kWCall = 2, kWSph = 11, kWCap = 28, kWZCap = 18, kWCub = 33, kWZCub = 24;

// sphere_environment_in_collision (src/utils.cuh:213-245) for one lane: the
// lists in order, each stopping at the first hit; adds its operation count.
__device__ __forceinline__ bool bcc_ref_env_hit(const RefEnvView& E, float x, float y, float z, float rp,
                                                uint64_t& w)
{
    const float rsq = rp * rp;
    bool h = false;
    w += kWCall;
    for (int k = 0; k < E.ns && !h; ++k) {
        const cfptr o = E.sph + 4 * k;
        w += kWSph;
        h = bcc_sphere_val(o[0], o[1], o[2], o[3], x, y, z, rp) < 0.0f;
    }
    for (int k = 0; k < E.nc && !h; ++k) {
        w += kWCap;
        h = bcc_capsule_val(E.cap + 8 * k, x, y, z, rp) < 0.0f;
    }
    for (int k = 0; k < E.nzc && !h; ++k) {
        w += kWZCap;
        h = bcc_zcapsule_val(E.zcap + 8 * k, x, y, z, rp) < 0.0f;
    }
    for (int k = 0; k < E.ncb && !h; ++k) {
        w += kWCub;
        h = bcc_cuboid_val(E.cub + 16 * k, x, y, z, rsq) < 0.0f;
    }
    for (int k = 0; k < E.nzcb && !h; ++k) {
        w += kWZCub;
        h = bcc_zcuboid_val(E.zcub + 16 * k, x, y, z, rsq) < 0.0f;
    }
    return h;
}

// ------------------------------------------------------------ exact culling
// A sweep tests N robot spheres of every lane in a lane mask against the
// environment.  When N times the environment's primitive count reaches the
// scene's threshold (below it the box reduction costs more than the tests it
// saves), the wave first computes for each half-wave the box
// of those spheres (centre -/+ padded radius over the masked lanes), widened
// by delta = 0.01 + 1e-5 * (largest coordinate) -- both halves' boxes are
// wave-uniform -- and then, 64 obstacles at a time, one obstacle per lane,
// which obstacles' boxes come within the widened box of either half
// (a ballot, kept in LDS).  Only those obstacles are tested.
//
// Why the verdicts cannot change: an obstacle is skipped only if, along some
// axis, its box is separated from a half's sphere box by more than delta.
// Then every masked sphere centre is farther than rp + delta from the box, so
// the exact Euclidean distance to the primitive exceeds rp + delta, and the
// test value (a squared distance minus (rp + R)^2, or for cuboids the
// distance formula with orthonormal axes) is at least delta^2 > 0 in exact
// arithmetic.  Its FP32 evaluation errs by a few ulp of the magnitudes
// involved (|centre - primitive| and the primitive's size), which the
// 1e-5-relative widenings on both sides exceed by orders of magnitude, and
// which the absolute 0.01 m exceeds for any scene smaller than 10 km.  NaN
// coordinates compare false and are never culled.  Lanes outside the mask
// do not shape the boxes; their flags are not used by the caller.
struct HalfBoxes {
    float lo0[3], hi0[3], lo1[3], hi1[3];   // wave-uniform
};

// Min / max over each 32-lane half-wave; every lane of a half ends with its
// half's value.  Inside each 16-lane row the partners come from DPP (quad
// permutations [1,0,3,2] and [2,3,0,1], then the half-row and row mirrors:
// after the two quad steps every lane holds its quad's value, the half-row
// mirror pairs the two quads of each 8 lanes, the row mirror the two halves
// of each row), which are plain VALU operands; only the last step, across
// the two rows of a half, goes through a lane shuffle.  All 64 lanes must be
// active (the sweeps run in wave-uniform control flow).
template <int kCtrl>
__device__ __forceinline__ float bcc_dpp(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xf, 0xf, false));
}

constexpr int kDppQuadSwap1 = 0xb1;      // quad_perm [1,0,3,2]
constexpr int kDppQuadSwap2 = 0x4e;      // quad_perm [2,3,0,1]
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppRowMirror = 0x140;

__device__ __forceinline__ float bcc_wave_min32(float v)
{
    v = fminf(v, bcc_dpp<kDppQuadSwap1>(v));
    v = fminf(v, bcc_dpp<kDppQuadSwap2>(v));
    v = fminf(v, bcc_dpp<kDppRowHalfMirror>(v));
    v = fminf(v, bcc_dpp<kDppRowMirror>(v));
    return fminf(v, __shfl_xor(v, 16, 64));
}

__device__ __forceinline__ float bcc_wave_max32(float v)
{
    v = fmaxf(v, bcc_dpp<kDppQuadSwap1>(v));
    v = fmaxf(v, bcc_dpp<kDppQuadSwap2>(v));
    v = fmaxf(v, bcc_dpp<kDppRowHalfMirror>(v));
    v = fmaxf(v, bcc_dpp<kDppRowMirror>(v));
    return fmaxf(v, __shfl_xor(v, 16, 64));
}

__device__ __forceinline__ float bcc_read_lane(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <int N>
__device__ __forceinline__ HalfBoxes bcc_half_boxes(const float (&x)[N], const float (&y)[N], const float (&z)[N],
                                                    const float (&rp)[N], uint64_t mask)
{
    const bool in = (mask >> (threadIdx.x & 63)) & 1ull;
    float lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = INFINITY;
        hi[a] = -INFINITY;
    }
    if (in) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            lo[0] = fminf(lo[0], x[i] - rp[i]);
            lo[1] = fminf(lo[1], y[i] - rp[i]);
            lo[2] = fminf(lo[2], z[i] - rp[i]);
            hi[0] = fmaxf(hi[0], x[i] + rp[i]);
            hi[1] = fmaxf(hi[1], y[i] + rp[i]);
            hi[2] = fmaxf(hi[2], z[i] + rp[i]);
        }
    }
    float big = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = bcc_wave_min32(lo[a]);
        hi[a] = bcc_wave_max32(hi[a]);
        big = fmaxf(big, fmaxf(fabsf(lo[a]), fabsf(hi[a])));
    }
    if (lo[0] <= hi[0]) {   // a half with masked lanes: widen (an empty half stays empty)
        const float delta = fmaf(big, 1e-5f, 0.01f);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] -= delta;
            hi[a] += delta;
        }
    }
    HalfBoxes W;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        W.lo0[a] = bcc_read_lane(lo[a], 0);
        W.hi0[a] = bcc_read_lane(hi[a], 0);
        W.lo1[a] = bcc_read_lane(lo[a], 32);
        W.hi1[a] = bcc_read_lane(hi[a], 32);
    }
    return W;
}

// Bit j: obstacle base + j (in box order) may touch a masked sphere.
__device__ __forceinline__ uint64_t bcc_near_mask(const EnvLists& E, int base, const HalfBoxes& W)
{
    const int j = base + (int)(threadIdx.x & 63);
    bool near = false;
    if (j < E.nbox) {
        const float4 l = *(const float4*)(E.box + 8 * (size_t)j);
        const float4 h = *(const float4*)(E.box + 8 * (size_t)j + 4);
        const bool far0 = h.x < W.lo0[0] || h.y < W.lo0[1] || h.z < W.lo0[2] || l.x > W.hi0[0] ||
                          l.y > W.hi0[1] || l.z > W.hi0[2];
        const bool far1 = h.x < W.lo1[0] || h.y < W.lo1[1] || h.z < W.lo1[2] || l.x > W.hi1[0] ||
                          l.y > W.hi1[1] || l.z > W.hi1[2];
        near = !(far0 && far1);
    }
    return __ballot(near);
}

// Near masks of a sweep's first kCullWindows 64-primitive windows, one LDS
// slot per window and wave (the kernels run at most kMaxBlockWaves waves per
// workgroup; sweeps of one wave never overlap).  Keeping the masks in LDS
// instead of the half-wave boxes in SGPRs frees those registers before the
// tests run.  Primitives beyond the covered windows are always tested.
constexpr int kCullWindows = 16;
constexpr int kMaxBlockWaves = 4;

__device__ __forceinline__ uint64_t* bcc_near_slots()
{
    __shared__ uint64_t slots[kMaxBlockWaves * kCullWindows];
    return slots + (threadIdx.x >> 6) * kCullWindows;
}

template <int N>
__device__ __forceinline__ void bcc_cull(const EnvLists& E, const float (&x)[N], const float (&y)[N],
                                         const float (&z)[N], const float (&rp)[N], uint64_t mask)
{
    const HalfBoxes W = bcc_half_boxes<N>(x, y, z, rp, mask);
    uint64_t* slots = bcc_near_slots();
    const int nwin = min((E.nbox + 63) >> 6, kCullWindows);
    for (int w = 0; w < nwin; ++w) {
        const uint64_t m = bcc_near_mask(E, 64 * w, W);
        if ((threadIdx.x & 63) == 0)
            slots[w] = m;
    }
}

__device__ __forceinline__ uint64_t bcc_window_near(int b0)
{
    if (b0 >= 64 * kCullWindows)
        return ~0ull;
    const uint64_t m = bcc_near_slots()[b0 >> 6];
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)m);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(m >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Bits of m for box indices [first, first + count) within the 64-obstacle
// window starting at base.
__device__ __forceinline__ uint64_t bcc_window_bits(uint64_t m, int base, int first, int count)
{
    const int lo = max(first - base, 0);
    const int hi = min(first + count - base, 64);
    if (lo >= hi)
        return 0;
    const uint64_t width = (hi - lo == 64) ? ~0ull : ((1ull << (hi - lo)) - 1ull);
    return m & (width << lo);
}

// ------------------------------------------------------------ env sweeps
// One pass over the environment's obstacles (all of them below the culling
// threshold, else the near ones) tests N robot spheres at once (N
// independent tests per obstacle keep the VALU pipeline full).  `rp` are the
// padded radii float(double(r) + 0.005) (src/utils.cuh:215).

// U: unroll factor of the obstacle loops, a tuning knob of each kernel
// variant.  The culled loop visits the set bits of the window's mask for one
// list.
#define BCC_SWEEP(first, count, stride, base, TEST)                              \
    {                                                                             \
        uint64_t mk = bcc_window_bits(near, b0, (first), (count));                \
        _Pragma("unroll U") for (; mk; mk &= mk - 1) {                            \
            const int k = b0 + (int)__builtin_ctzll(mk) - (first);                \
            const cfptr o = (base) + (stride) * k;                                \
            _Pragma("unroll") for (int i = 0; i < N; ++i) { TEST; }               \
            SETTLED_CHECK;                                                        \
        }                                                                         \
    }

// Every obstacle of one list in order (sweeps below the culling threshold).
#define BCC_SWEEP_ALL(count, stride, base, TEST)                                 \
    _Pragma("unroll U") for (int k = 0; k < (count); ++k) {                       \
        const cfptr o = (base) + (stride) * k;                                    \
        _Pragma("unroll") for (int i = 0; i < N; ++i) { TEST; }                   \
        SETTLED_CHECK;                                                            \
    }

#define BCC_SWEEP_LISTS(T_SPH, T_CAP, T_ZCAP, T_CUB, T_ZCUB, T_AABB)              \
    {                                                                             \
        BCC_SWEEP_ALL(E.ns, 4, E.sph, T_SPH)                                      \
        BCC_SWEEP_ALL(E.nc, 8, E.cap, T_CAP)                                      \
        BCC_SWEEP_ALL(E.nzc, 8, E.zcap, T_ZCAP)                                   \
        BCC_SWEEP_ALL(E.ncb, 16, E.cub, T_CUB)                                    \
        BCC_SWEEP_ALL(E.nzcb, 16, E.zcub, T_ZCUB)                                 \
        BCC_SWEEP_ALL(E.na, 8, E.aabb, T_AABB)                                    \
    }

// The six lists of one 64-obstacle window.
#define BCC_SWEEP_WINDOW(T_SPH, T_CAP, T_ZCAP, T_CUB, T_ZCUB, T_AABB)             \
    {                                                                             \
        const int f1 = E.ns, f2 = f1 + E.nc, f3 = f2 + E.nzc, f4 = f3 + E.ncb,    \
                  f5 = f4 + E.nzcb;                                               \
        BCC_SWEEP(0, E.ns, 4, E.sph, T_SPH)                                       \
        BCC_SWEEP(f1, E.nc, 8, E.cap, T_CAP)                                      \
        BCC_SWEEP(f2, E.nzc, 8, E.zcap, T_ZCAP)                                   \
        BCC_SWEEP(f3, E.ncb, 16, E.cub, T_CUB)                                    \
        BCC_SWEEP(f4, E.nzcb, 16, E.zcub, T_ZCUB)                                 \
        BCC_SWEEP(f5, E.na, 8, E.aabb, T_AABB)                                    \
    }

// Per-lane, per-sphere hit flags for the lanes in `live` (other lanes' flags
// are meaningless); no early exit (gate spheres and lone spheres: each lane
// needs its own answer for every sphere).
//
// kDual: sweeps below the culling threshold run the plain counted loops (a
// second copy of the code); otherwise the culled loop with every obstacle
// marked near.
template <int N, int U, bool kDual>
__device__ __forceinline__ void bcc_env_sweep(const EnvView& env, const float (&x)[N], const float (&y)[N],
                                              const float (&z)[N], const float (&rp)[N], uint64_t live, bool (&h)[N])
{
    const EnvLists E = bcc_env_lists(env);
    float rsq[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        h[i] = false;
        rsq[i] = rp[i] * rp[i];
    }
    if (E.nbox == 0)
        return;
#define SETTLED_CHECK
#define T_SPH h[i] = h[i] | (bcc_sphere_val(o[0], o[1], o[2], o[3], x[i], y[i], z[i], rp[i]) < 0.0f)
#define T_CAP h[i] = h[i] | (bcc_capsule_val(o, x[i], y[i], z[i], rp[i]) < 0.0f)
#define T_ZCAP h[i] = h[i] | (bcc_zcapsule_val(o, x[i], y[i], z[i], rp[i]) < 0.0f)
#define T_CUB h[i] = h[i] | (bcc_cuboid_val(o, x[i], y[i], z[i], rsq[i]) < 0.0f)
#define T_ZCUB h[i] = h[i] | (bcc_zcuboid_val(o, x[i], y[i], z[i], rsq[i]) < 0.0f)
#define T_AABB h[i] = h[i] | (bcc_aabb_val(o, x[i], y[i], z[i], rsq[i]) < 0.0f)
    const bool cull = N * E.nbox >= E.cull_min;
    if (kDual && !cull) {
        BCC_SWEEP_LISTS(T_SPH, T_CAP, T_ZCAP, T_CUB, T_ZCUB, T_AABB)
        return;
    }
    if (cull)
        bcc_cull<N>(E, x, y, z, rp, live);
    for (int b0 = 0; b0 < E.nbox; b0 += 64) {
        const uint64_t near = cull ? bcc_window_near(b0) : ~0ull;
        BCC_SWEEP_WINDOW(T_SPH, T_CAP, T_ZCAP, T_CUB, T_ZCUB, T_AABB)
    }
#undef T_SPH
#undef T_CAP
#undef T_ZCAP
#undef T_CUB
#undef T_ZCUB
#undef T_AABB
#undef SETTLED_CHECK
}

// Does any of the N spheres touch any obstacle?  Per-lane OR.  Only lanes in
// `need` matter; the sweep stops once every lane group of G lanes present in
// `need` has a hit (one colliding sample decides that group's pair).
template <int N, int U, bool kDual, int G = 32>
__device__ __forceinline__ bool bcc_env_any(const EnvView& env, const float (&x)[N], const float (&y)[N],
                                            const float (&z)[N], const float (&rp)[N], uint64_t need)
{
    const EnvLists E = bcc_env_lists(env);
    float rsq[N];
#pragma unroll
    for (int i = 0; i < N; ++i)
        rsq[i] = rp[i] * rp[i];
    bool h = false;
    if (E.nbox == 0)
        return h;
#define SETTLED_CHECK                                                         \
    if ((need & ~bcc_group_spread<G>(__ballot(h) & need)) == 0)               \
        return h;
#define T_SPH h = h | (bcc_sphere_val(o[0], o[1], o[2], o[3], x[i], y[i], z[i], rp[i]) < 0.0f)
#define T_CAP h = h | (bcc_capsule_val(o, x[i], y[i], z[i], rp[i]) < 0.0f)
#define T_ZCAP h = h | (bcc_zcapsule_val(o, x[i], y[i], z[i], rp[i]) < 0.0f)
#define T_CUB h = h | (bcc_cuboid_val(o, x[i], y[i], z[i], rsq[i]) < 0.0f)
#define T_ZCUB h = h | (bcc_zcuboid_val(o, x[i], y[i], z[i], rsq[i]) < 0.0f)
#define T_AABB h = h | (bcc_aabb_val(o, x[i], y[i], z[i], rsq[i]) < 0.0f)
    const bool cull = N * E.nbox >= E.cull_min;
    if (kDual && !cull) {
        BCC_SWEEP_LISTS(T_SPH, T_CAP, T_ZCAP, T_CUB, T_ZCUB, T_AABB)
        return h;
    }
    if (cull)
        bcc_cull<N>(E, x, y, z, rp, need);
    for (int b0 = 0; b0 < E.nbox; b0 += 64) {
        const uint64_t near = cull ? bcc_window_near(b0) : ~0ull;
        BCC_SWEEP_WINDOW(T_SPH, T_CAP, T_ZCAP, T_CUB, T_ZCUB, T_AABB)
    }
#undef T_SPH
#undef T_CAP
#undef T_ZCAP
#undef T_CUB
#undef T_ZCUB
#undef T_AABB
#undef SETTLED_CHECK
    return h;
}

#undef BCC_SWEEP_LISTS
#undef BCC_SWEEP_ALL
#undef BCC_SWEEP_WINDOW
#undef BCC_SWEEP

}  // namespace bcc
