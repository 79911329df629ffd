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
// Cylinders are never tested by the reference (src/utils.cuh:213-245) and are
// not uploaded.
struct EnvView {
    cfptr sph;
    cfptr cap;
    cfptr zcap;
    cfptr cub;
    cfptr zcub;
    cfptr aabb;
    int ns, nc, nzc, ncb, nzcb, na;
};

// Per-environment record in the scene's meta array (int32): for list k in
// the order above, [2k] = offset in floats, [2k+1] = count.
constexpr int kEnvMetaInts = 16;

__device__ __forceinline__ EnvView load_env(const float* obs_, const int* meta_, int env)
{
    const cfptr obs = (cfptr)obs_;
    const ciptr m = (ciptr)meta_ + env * kEnvMetaInts;
    EnvView v;
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
    v.aabb = obs + m[10];
    v.na = m[11];
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

// ------------------------------------------------------------ env sweeps
// One pass over the environment's obstacles tests N robot spheres at once
// (N independent tests per obstacle keep the VALU pipeline full).  `rp` are
// the padded radii float(double(r) + 0.005) (src/utils.cuh:215).

// U: obstacle loop unroll factor, a tuning knob of each kernel variant (more
// unrolling overlaps scalar loads with tests but keeps more temporaries live).
#define BCC_SWEEP(count, stride, base, TEST)                    \
    _Pragma("unroll U") for (int k = 0; k < (count); ++k) {     \
        const cfptr o = (base) + (stride) * k;                  \
        _Pragma("unroll") for (int i = 0; i < N; ++i) { TEST; } \
        SETTLED_CHECK;                                          \
    }

// Per-lane, per-sphere hit flags; no early exit (gate spheres and lone
// spheres: each lane needs its own answer for every sphere).
template <int N, int U>
__device__ __forceinline__ void bcc_env_sweep(const EnvView& E, const float (&x)[N], const float (&y)[N],
                                              const float (&z)[N], const float (&rp)[N], bool (&h)[N])
{
    float rsq[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        h[i] = false;
        rsq[i] = rp[i] * rp[i];
    }
#define SETTLED_CHECK
    BCC_SWEEP(E.ns, 4, E.sph, h[i] = h[i] | (bcc_sphere_val(o[0], o[1], o[2], o[3], x[i], y[i], z[i], rp[i]) < 0.0f))
    BCC_SWEEP(E.nc, 8, E.cap, h[i] = h[i] | (bcc_capsule_val(o, x[i], y[i], z[i], rp[i]) < 0.0f))
    BCC_SWEEP(E.nzc, 8, E.zcap, h[i] = h[i] | (bcc_zcapsule_val(o, x[i], y[i], z[i], rp[i]) < 0.0f))
    BCC_SWEEP(E.ncb, 16, E.cub, h[i] = h[i] | (bcc_cuboid_val(o, x[i], y[i], z[i], rsq[i]) < 0.0f))
    BCC_SWEEP(E.nzcb, 16, E.zcub, h[i] = h[i] | (bcc_zcuboid_val(o, x[i], y[i], z[i], rsq[i]) < 0.0f))
    BCC_SWEEP(E.na, 8, E.aabb, h[i] = h[i] | (bcc_aabb_val(o, x[i], y[i], z[i], rsq[i]) < 0.0f))
#undef SETTLED_CHECK
}

// Does any of the N spheres touch any obstacle?  Per-lane OR.  Only lanes in
// `need` matter; the sweep stops once every half-wave present in `need` has a
// hit (one colliding sample decides that half's pair).
template <int N, int U>
__device__ __forceinline__ bool bcc_env_any(const EnvView& E, const float (&x)[N], const float (&y)[N],
                                            const float (&z)[N], const float (&rp)[N], uint64_t need)
{
    float rsq[N];
#pragma unroll
    for (int i = 0; i < N; ++i)
        rsq[i] = rp[i] * rp[i];
    bool h = false;
#define SETTLED_CHECK                                                         \
    if ((need & ~bcc_half_spread(__ballot(h) & need)) == 0)                   \
        return h;
    BCC_SWEEP(E.ns, 4, E.sph, h = h | (bcc_sphere_val(o[0], o[1], o[2], o[3], x[i], y[i], z[i], rp[i]) < 0.0f))
    BCC_SWEEP(E.nc, 8, E.cap, h = h | (bcc_capsule_val(o, x[i], y[i], z[i], rp[i]) < 0.0f))
    BCC_SWEEP(E.nzc, 8, E.zcap, h = h | (bcc_zcapsule_val(o, x[i], y[i], z[i], rp[i]) < 0.0f))
    BCC_SWEEP(E.ncb, 16, E.cub, h = h | (bcc_cuboid_val(o, x[i], y[i], z[i], rsq[i]) < 0.0f))
    BCC_SWEEP(E.nzcb, 16, E.zcub, h = h | (bcc_zcuboid_val(o, x[i], y[i], z[i], rsq[i]) < 0.0f))
    BCC_SWEEP(E.na, 8, E.aabb, h = h | (bcc_aabb_val(o, x[i], y[i], z[i], rsq[i]) < 0.0f))
#undef SETTLED_CHECK
    return h;
}

#undef BCC_SWEEP

}  // namespace bcc
