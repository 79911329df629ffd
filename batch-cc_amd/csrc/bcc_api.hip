// Host side of the C ABI (include/bcc.h): scene upload, argument checks,
// launch dispatch per robot, timing, and the arithmetic test hooks.
//
// Replaces the host orchestration of src/batch_cc.cu:135-338.  Differences
// by design: environments are packed once into one device buffer per scene
// (one allocation instead of two per environment, src/batch_cc.cu:146-175),
// errors are returned as status codes with a thread-local message instead of
// exceptions/exit(), and nothing is printed.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <utility>

#include "../../include/bcc.h"
#include "bcc_dense.h"
#include "bcc_internal.h"
#include "bcc_tuning.h"
#include "robots/baxter_model.h"
#include "robots/xarm7_model.h"
#include "robots/fetch_model.h"
#include "robots/panda_model.h"

namespace bcc_detail {

thread_local std::string g_err = "no error";

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    return fail(BCC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace bcc_detail

using bcc_detail::fail;
using bcc_detail::g_err;
using bcc_detail::hip_fail;

namespace {

constexpr int kStride[6] = {4, 8, 8, 8, 15, 15};  // packed host layout

}  // namespace


namespace {

// Pack host environments into the device layout of bcc_device.h: per env
// six lists (spheres[4], capsules[8], z-capsules[8], cuboids[16],
// z-cuboids[16], aabbs[8]), each starting on a 16-float boundary.
// Primitives go to the list of their kind, except that capsules with
// xv == yv == 0 join the z-capsules and cuboids with exactly identity axes
// (z-cuboids with axes (1,0),(0,1)) become aabbs: for those inputs the
// special-case formulas are bit-identical to the general ones.  Cylinders
// are dropped: the reference never tests them (src/utils.cuh:213-245).
//
// A seventh list holds one bounding box per tested primitive, in list order
// ([lo.x lo.y lo.z 0 hi.x hi.y hi.z 0]), for the kernels' exact culling
// (bcc_device.h): every point where the primitive's test value can be
// negative for a robot sphere of padded radius rp lies within rp of the box.
// Boxes follow the formula each list uses, are computed in double, widened
// by 1e-5 (|centre| + size) and rounded outwards; primitives whose formula
// is not a Euclidean distance (cuboid axes not orthonormal to 4e-6) get an
// infinite box, i.e. are never culled.
namespace {

struct Box {
    double lo[3], hi[3];
};

void push_box(std::vector<float>& out, const Box& b, double scale)
{
    const double pad = 1e-5 * scale;
    float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        v[k] = std::nextafter((float)(b.lo[k] - pad), -INFINITY);
        v[4 + k] = std::nextafter((float)(b.hi[k] + pad), INFINITY);
    }
    out.insert(out.end(), v, v + 8);
}

void push_infinite_box(std::vector<float>& out)
{
    const float v[8] = {-INFINITY, -INFINITY, -INFINITY, 0.0f, INFINITY, INFINITY, INFINITY, 0.0f};
    out.insert(out.end(), v, v + 8);
}

// r[0..2], r[3..5] (, r[6..8]): rows of an axis matrix; |R R^T - I| <= 4e-6
bool orthonormal(const float* a, const float* b, const float* c, int n)
{
    const float* r[3] = {a, b, c};
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double d = 0.0;
            for (int k = 0; k < n; ++k)
                d += (double)r[i][k] * (double)r[j][k];
            if (!(std::fabs(d - (i == j ? 1.0 : 0.0)) <= 4e-6))
                return false;
        }
    return true;
}

// list k record -> box (see the list layouts in bcc_device.h)
void record_box(int k, const float* o, std::vector<float>& out)
{
    Box b;
    double scale = 0.0;
    switch (k) {
    case 0: {   // sphere x y z r
        const double r = std::fabs((double)o[3]);
        for (int i = 0; i < 3; ++i) {
            b.lo[i] = (double)o[i] - r;
            b.hi[i] = (double)o[i] + r;
            scale += std::fabs((double)o[i]);
        }
        scale += r;
        break;
    }
    case 1: {   // capsule: segment p1 .. p1 + v, radius r
        const double r = std::fabs((double)o[6]);
        for (int i = 0; i < 3; ++i) {
            const double p = o[i], q = (double)o[i] + (double)o[3 + i];
            b.lo[i] = std::min(p, q) - r;
            b.hi[i] = std::max(p, q) + r;
            scale += std::fabs(p) + std::fabs((double)o[3 + i]);
        }
        scale += r;
        break;
    }
    case 2: {   // z-capsule: segment (x1, y1, z1) .. (x1, y1, z1 + zv)
        const double r = std::fabs((double)o[6]);
        for (int i = 0; i < 2; ++i) {
            b.lo[i] = (double)o[i] - r;
            b.hi[i] = (double)o[i] + r;
        }
        const double p = o[2], q = (double)o[2] + (double)o[5];
        b.lo[2] = std::min(p, q) - r;
        b.hi[2] = std::max(p, q) + r;
        scale = std::fabs((double)o[0]) + std::fabs((double)o[1]) + std::fabs(p) + std::fabs((double)o[5]) + r;
        break;
    }
    case 3: {   // cuboid: centre, rows a1 a2 a3, half-extents
        if (!orthonormal(o + 3, o + 6, o + 9, 3)) {
            push_infinite_box(out);
            return;
        }
        for (int i = 0; i < 3; ++i) {
            double w = 0.0;
            for (int a = 0; a < 3; ++a)
                w += std::fabs((double)o[3 + 3 * a + i]) * std::fabs((double)o[12 + a]);
            b.lo[i] = (double)o[i] - w;
            b.hi[i] = (double)o[i] + w;
            scale += std::fabs((double)o[i]) + std::fabs((double)o[12 + i]);
        }
        break;
    }
    case 4: {   // z-cuboid: 2-D rows (a1x, a1y), (a2x, a2y), half-extents; z is axis-aligned
        if (!orthonormal(o + 3, o + 6, nullptr, 2)) {
            push_infinite_box(out);
            return;
        }
        for (int i = 0; i < 2; ++i) {
            const double w = std::fabs((double)o[3 + i]) * std::fabs((double)o[12]) +
                             std::fabs((double)o[6 + i]) * std::fabs((double)o[13]);
            b.lo[i] = (double)o[i] - w;
            b.hi[i] = (double)o[i] + w;
        }
        b.lo[2] = (double)o[2] - std::fabs((double)o[14]);
        b.hi[2] = (double)o[2] + std::fabs((double)o[14]);
        for (int i = 0; i < 3; ++i)
            scale += std::fabs((double)o[i]) + std::fabs((double)o[12 + i]);
        break;
    }
    default: {  // aabb: centre, half-extents
        for (int i = 0; i < 3; ++i) {
            const double h = std::fabs((double)o[3 + i]);
            b.lo[i] = (double)o[i] - h;
            b.hi[i] = (double)o[i] + h;
            scale += std::fabs((double)o[i]) + h;
        }
        break;
    }
    }
    push_box(out, b, scale);
}

}  // namespace

// Culling threshold stamped into each environment's record at scene
// creation: a sweep of N robot spheres culls when N * (primitives in the
// environment) reaches it (below, the box reduction costs more than the tests
// it saves).  -1: the tuning table's value.
std::atomic<int> g_cull_min{-1};

int build_scene(int num_envs, const std::vector<bcc_env_desc>& envs, bcc_scene** out)
{
    constexpr int kLists = 7;   // six primitive lists + their boxes
    const int stride[kLists] = {4, 8, 8, 16, 16, 8, 8};
    std::vector<std::vector<float>> lists((size_t)num_envs * kLists);
    for (int e = 0; e < num_envs; ++e) {
        const bcc_env_desc& d = envs[e];
        std::vector<float>* L = &lists[(size_t)e * kLists];
        for (uint32_t i = 0; i < d.num_spheres; ++i)
            L[0].insert(L[0].end(), d.spheres + 4 * (size_t)i, d.spheres + 4 * (size_t)i + 4);
        for (uint32_t i = 0; i < d.num_capsules; ++i) {
            const float* c = d.capsules + 8 * (size_t)i;
            std::vector<float>& dst = (c[3] == 0.0f && c[4] == 0.0f) ? L[2] : L[1];
            dst.insert(dst.end(), c, c + 8);
        }
        for (uint32_t i = 0; i < d.num_z_capsules; ++i)
            L[2].insert(L[2].end(), d.z_capsules + 8 * (size_t)i, d.z_capsules + 8 * (size_t)i + 8);
        for (uint32_t i = 0; i < d.num_cuboids; ++i) {
            const float* b = d.cuboids + 15 * (size_t)i;
            const bool ident = b[3] == 1.0f && b[4] == 0.0f && b[5] == 0.0f && b[6] == 0.0f && b[7] == 1.0f &&
                               b[8] == 0.0f && b[9] == 0.0f && b[10] == 0.0f && b[11] == 1.0f;
            if (ident) {
                const float a[8] = {b[0], b[1], b[2], b[12], b[13], b[14], 0.0f, 0.0f};
                L[5].insert(L[5].end(), a, a + 8);
            } else {
                L[3].insert(L[3].end(), b, b + 15);
                L[3].push_back(0.0f);
            }
        }
        for (uint32_t i = 0; i < d.num_z_cuboids; ++i) {
            const float* b = d.z_cuboids + 15 * (size_t)i;
            const bool ident = b[3] == 1.0f && b[4] == 0.0f && b[6] == 0.0f && b[7] == 1.0f;
            if (ident) {
                const float a[8] = {b[0], b[1], b[2], b[12], b[13], b[14], 0.0f, 0.0f};
                L[5].insert(L[5].end(), a, a + 8);
            } else {
                L[4].insert(L[4].end(), b, b + 15);
                L[4].push_back(0.0f);
            }
        }
        for (int k = 0; k < 6; ++k)
            for (size_t i = 0; i < L[k].size(); i += stride[k])
                record_box(k, &L[k][i], L[6]);
    }
    std::vector<int> meta((size_t)num_envs * bcc::kEnvMetaInts, 0);
    const int cull_set = g_cull_min.load();
    const int cull_min = cull_set >= 0 ? cull_set : kCullMinWork;
    size_t total = 0;
    auto align16 = [](size_t x) { return (x + 15) & ~(size_t)15; };
    for (int e = 0; e < num_envs; ++e) {
        for (int k = 0; k < kLists; ++k) {
            total = align16(total);
            if (total > 0x7fffffffULL)
                return fail(BCC_ERR_INVALID, "scene too large (more than 2^31 floats)");
            const std::vector<float>& l = lists[(size_t)e * kLists + k];
            meta[(size_t)e * bcc::kEnvMetaInts + 2 * k] = (int)total;
            meta[(size_t)e * bcc::kEnvMetaInts + 2 * k + 1] = (int)(l.size() / stride[k]);
            total += l.size();
        }
        meta[(size_t)e * bcc::kEnvMetaInts + 14] = cull_min;
    }
    total = align16(total) + 16;
    std::vector<float> host(total, 0.0f);
    for (int e = 0; e < num_envs; ++e)
        for (int k = 0; k < kLists; ++k) {
            const std::vector<float>& l = lists[(size_t)e * kLists + k];
            if (!l.empty())
                std::memcpy(&host[meta[(size_t)e * bcc::kEnvMetaInts + 2 * k]], l.data(), sizeof(float) * l.size());
        }
    // reference layout: the five tested lists in the reference's order
    std::vector<int> rmeta((size_t)num_envs * bcc::kEnvMetaInts, 0);
    std::vector<float> rhost;
    for (int e = 0; e < num_envs; ++e) {
        const bcc_env_desc& d = envs[e];
        const float* src[5] = {d.spheres, d.capsules, d.z_capsules, d.cuboids, d.z_cuboids};
        const uint32_t cnt[5] = {d.num_spheres, d.num_capsules, d.num_z_capsules, d.num_cuboids, d.num_z_cuboids};
        const int in_stride[5] = {4, 8, 8, 15, 15};
        const int out_stride[5] = {4, 8, 8, 16, 16};
        for (int k = 0; k < 5; ++k) {
            rhost.resize(align16(rhost.size()), 0.0f);
            if (rhost.size() > 0x7fffffffULL)
                return fail(BCC_ERR_INVALID, "scene too large (more than 2^31 floats)");
            rmeta[(size_t)e * bcc::kEnvMetaInts + 2 * k] = (int)rhost.size();
            rmeta[(size_t)e * bcc::kEnvMetaInts + 2 * k + 1] = (int)cnt[k];
            for (uint32_t i = 0; i < cnt[k]; ++i) {
                rhost.insert(rhost.end(), src[k] + (size_t)in_stride[k] * i, src[k] + (size_t)in_stride[k] * (i + 1));
                rhost.resize(rhost.size() + (out_stride[k] - in_stride[k]), 0.0f);
            }
        }
    }
    rhost.resize(align16(rhost.size()) + 16, 0.0f);

    bcc_scene* s = new bcc_scene;
    s->num_envs = num_envs;
    s->obs_floats = total;
    s->cull_min = cull_min;
    s->obstacles.resize(num_envs);
    for (int e = 0; e < num_envs; ++e) {
        const bcc_env_desc& d = envs[e];
        s->obstacles[e] = (int)(d.num_spheres + d.num_capsules + d.num_z_capsules + d.num_cuboids + d.num_z_cuboids);
    }
    const size_t meta_bytes = sizeof(int) * std::max<size_t>(meta.size(), 1);
    hipError_t err = hipMalloc((void**)&s->d_obs, sizeof(float) * total);
    if (err == hipSuccess)
        err = hipMalloc((void**)&s->d_meta, meta_bytes);
    if (err == hipSuccess)
        err = hipMalloc((void**)&s->d_ref_obs, sizeof(float) * rhost.size());
    if (err == hipSuccess)
        err = hipMalloc((void**)&s->d_ref_meta, meta_bytes);
    if (err == hipSuccess)
        err = hipMemcpy(s->d_obs, host.data(), sizeof(float) * total, hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemcpy(s->d_ref_obs, rhost.data(), sizeof(float) * rhost.size(), hipMemcpyHostToDevice);
    if (err == hipSuccess && !meta.empty())
        err = hipMemcpy(s->d_meta, meta.data(), sizeof(int) * meta.size(), hipMemcpyHostToDevice);
    if (err == hipSuccess && !rmeta.empty())
        err = hipMemcpy(s->d_ref_meta, rmeta.data(), sizeof(int) * rmeta.size(), hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        (void)hipFree(s->d_obs);
        (void)hipFree(s->d_meta);
        (void)hipFree(s->d_ref_obs);
        (void)hipFree(s->d_ref_meta);
        delete s;
        return hip_fail(err, "scene upload");
    }
    *out = s;
    return BCC_OK;
}

int check_env_desc(const bcc_env_desc& d, int e)
{
    const bool ok = (d.num_spheres == 0 || d.spheres) && (d.num_capsules == 0 || d.capsules) &&
                    (d.num_z_capsules == 0 || d.z_capsules) && (d.num_cylinders == 0 || d.cylinders) &&
                    (d.num_cuboids == 0 || d.cuboids) && (d.num_z_cuboids == 0 || d.z_cuboids);
    if (!ok)
        return fail(BCC_ERR_INVALID,
                    "Host environment has positive counts with null pointers (environment " + std::to_string(e) + ")");
    return BCC_OK;
}

// Forced kernel variant for tests (-1: chosen from the scene).
std::atomic<int> g_force_variant{-1};

// High-occupancy variant when the mean number of tested obstacles per
// environment (weighted by pairs) reaches the robot's threshold.
int choose_variant(int robot, double mean_obstacles)
{
    const int forced = g_force_variant.load();
    if (forced >= 0)
        return forced;
    return (mean_obstacles >= (double)kHiVariantMinObstacles[robot] &&
            mean_obstacles < (double)kHiVariantMaxObstacles[robot])
               ? 1
               : 0;
}

const bcc::VariantKnobs& knobs(int robot)
{
    switch (robot) {
    case BCC_FETCH: return bcc::knobs_fetch;
    case BCC_BAXTER: return bcc::knobs_baxter;
    case BCC_XARM7: return bcc::knobs_xarm7;
    default: return bcc::knobs_panda;
    }
}

double mean_obstacles(const bcc_scene* scene)
{
    if (scene->num_envs == 0)
        return 0.0;
    double sum = 0.0;
    for (int o : scene->obstacles)
        sum += o;
    return sum / scene->num_envs;
}

int device_cus();

// Edges per dense workgroup: 64, halved (down to 8) while the launch would
// have fewer than 8 workgroups per CU, so small batches still fill the GPU
// and end with a short tail.
std::atomic<int> g_dense_chunk{0};   // > 0: forced edges per workgroup (tuning hook)

int dense_chunk(int num_envs, int num_edges)
{
    if (const int forced = g_dense_chunk.load(); forced > 0)
        return forced;
    const long long target = 8LL * device_cus();
    int chunk = 64;
    while (chunk > 8 && (long long)num_envs * ((num_edges + chunk - 1) / chunk) < target)
        chunk /= 2;
    return chunk;
}

// Edges per workgroup of the learned-order kernel: its probe pass checks
// 64 / kProbe1 edges per wave at a time, so it wants larger chunks; 512,
// halved down to 32 while the launch has fewer than 8 workgroups per CU.
int ordered_chunk(int num_envs, int num_edges)
{
    if (const int forced = g_dense_chunk.load(); forced > 0)
        return std::min(forced, bcc::kOrderedChunkMax);
    const long long target = 8LL * device_cus();
    int chunk = 512;
    while (chunk > 32 && (long long)num_envs * ((num_edges + chunk - 1) / chunk) < target)
        chunk /= 2;
    return chunk;
}

int dispatch_dense(int robot, const bcc_scene* scene, bcc::DenseArgs a, hipStream_t s)
{
    a.variant = choose_variant(robot, mean_obstacles(scene));
    a.chunk = a.order ? ordered_chunk(a.num_envs, a.num_edges) : dense_chunk(a.num_envs, a.num_edges);
    hipError_t e;
    switch (robot) {
    case BCC_PANDA: e = bcc::launch_dense_panda(a, s); break;
    case BCC_FETCH: e = bcc::launch_dense_fetch(a, s); break;
    case BCC_BAXTER: e = bcc::launch_dense_baxter(a, s); break;
    case BCC_XARM7: e = bcc::launch_dense_xarm7(a, s); break;
    default: return fail(BCC_ERR_INVALID, "unknown robot id " + std::to_string(robot));
    }
    if (e != hipSuccess)
        return hip_fail(e, "dense kernel launch");
    return BCC_OK;
}

int check_dense_args(int robot, const bcc_scene* scene, const void* edges, int num_edges, int resolution,
                     const void* results)
{
    if (bcc_robot_dimension(robot) < 0)
        return fail(BCC_ERR_INVALID, "unknown robot id " + std::to_string(robot));
    if (!scene)
        return fail(BCC_ERR_INVALID, "scene is NULL");
    if (num_edges < 0)
        return fail(BCC_ERR_INVALID, "num_edges < 0");
    if (resolution < 1 || resolution > (1 << 20))
        return fail(BCC_ERR_INVALID, "resolution must be in [1, 2^20]");
    if (num_edges > 0 && scene->num_envs > 0 && (!edges || !results))
        return fail(BCC_ERR_INVALID, "edges/results buffer is NULL");
    return BCC_OK;
}

// ------------------------------------------------------------ test-hook kernels

__global__ void sincos_kernel(const float* __restrict__ x, int n, float* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float s, c;
    bcc::bcc_sincosf(x[i], s, c);
    out[2 * i] = s;
    out[2 * i + 1] = c;
}

// in[40]: x y z rp | sphere (4) | capsule (8) | cuboid (15 + pad) | aabb (8);
// out[6]: sphere, capsule, z-capsule, cuboid, z-cuboid and aabb test values.
// Every record is read from the global input buffer.
__global__ void prims_kernel(const float* __restrict__ in, int n, float* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float* p = in + (size_t)i * 40;
    const float x = p[0], y = p[1], z = p[2], rp = p[3];
    const bcc::cfptr cp = (bcc::cfptr)(p + 8);
    const bcc::cfptr bp = (bcc::cfptr)(p + 16);
    const bcc::cfptr ap = (bcc::cfptr)(p + 32);
    float* o = out + (size_t)i * 6;
    o[0] = bcc::bcc_sphere_val(p[4], p[5], p[6], p[7], x, y, z, rp);
    o[1] = bcc::bcc_capsule_val(cp, x, y, z, rp);
    o[2] = bcc::bcc_zcapsule_val(cp, x, y, z, rp);
    o[3] = bcc::bcc_cuboid_val(bp, x, y, z, rp * rp);
    o[4] = bcc::bcc_zcuboid_val(bp, x, y, z, rp * rp);
    o[5] = bcc::bcc_aabb_val(ap, x, y, z, rp * rp);
}

// in[64][4]: x y z rp per lane; out[12]: lo0 xyz, hi0 xyz, lo1 xyz, hi1 xyz,
// the widened half-wave boxes of the lanes in `mask` (one 64-thread wave).
__global__ void half_boxes_kernel(const float* __restrict__ in, uint64_t mask, float* __restrict__ out)
{
    const float* p = in + 4 * threadIdx.x;
    const float x[1] = {p[0]}, y[1] = {p[1]}, z[1] = {p[2]}, rp[1] = {p[3]};
    const bcc::HalfBoxes W = bcc::bcc_half_boxes<1>(x, y, z, rp, mask);
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; ++a) {
            out[a] = W.lo0[a];
            out[3 + a] = W.hi0[a];
            out[6 + a] = W.lo1[a];
            out[9 + a] = W.hi1[a];
        }
}

// Degenerate edges (e, e) of a batch: out[m] = (end of edge m, end of edge m).
__global__ void end_edges_kernel(const float* __restrict__ edges, int num_edges, int dim, float* __restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;   // (edge, half, joint)
    if (i >= (long long)num_edges * 2 * dim)
        return;
    const long long m = i / (2 * dim);
    const int j = (int)(i % dim);
    out[i] = edges[m * 2 * dim + dim + j];
}

__global__ void or_bytes_kernel(uint8_t* __restrict__ acc, const uint8_t* __restrict__ v, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        acc[i] = (uint8_t)(acc[i] | v[i]);
}

__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Known FP32 VALU work for calibrating the instruction and FLOP counters:
// lanes below `active` of every wave run kValuChains independent chains of
// `iters` fmaf each (then kValuChains - 1 adds and one store); the other
// lanes only store.  Lane values come from the thread index and the kernel
// arguments, so nothing folds.
constexpr int kValuChains = 8;

__global__ __launch_bounds__(256) void valu_kernel(float* __restrict__ out, int iters, int active, float m, float c)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    float r = 0.0f;
    if ((threadIdx.x & 63) < active) {
        float x[kValuChains];
#pragma unroll
        for (int k = 0; k < kValuChains; ++k)
            x[k] = (float)(tid + k) * 1e-7f;
        for (int i = 0; i < iters; ++i)
#pragma unroll
            for (int k = 0; k < kValuChains; ++k)
                x[k] = fmaf(x[k], m, c);
        r = x[0];
#pragma unroll
        for (int k = 1; k < kValuChains; ++k)
            r += x[k];
    }
    out[tid] = r;
}

}  // namespace

extern "C" {

const char* bcc_last_error(void)
{
    return g_err.c_str();
}

int bcc_robot_dimension(int robot)
{
    switch (robot) {
    case BCC_PANDA: return 7;
    case BCC_FETCH: return 8;
    case BCC_BAXTER: return 14;
    case BCC_XARM7: return 13;
    default: return -1;
    }
}

int bcc_robot_num_spheres(int robot)
{
    switch (robot) {
    case BCC_PANDA: return 70;
    case BCC_FETCH: return 125;
    case BCC_BAXTER: return 100;
    case BCC_XARM7: return 74;
    default: return -1;
    }
}

int bcc_device_count(int* count)
{
    if (!count)
        return fail(BCC_ERR_INVALID, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return hip_fail(e, "hipGetDeviceCount");
    }
    *count = n;
    return BCC_OK;
}

int bcc_scene_create(const bcc_env_desc* envs, int num_envs, bcc_scene** out)
{
    if (!out)
        return fail(BCC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (num_envs < 0 || (num_envs > 0 && !envs))
        return fail(BCC_ERR_INVALID, "bad environment array");
    std::vector<bcc_env_desc> v(envs, envs + num_envs);
    for (int e = 0; e < num_envs; ++e)
        if (int rc = check_env_desc(v[e], e))
            return rc;
    return build_scene(num_envs, v, out);
}

int bcc_scene_create_packed(const float* data, const int64_t* env_offsets, const int32_t* counts, int num_envs,
                            bcc_scene** out)
{
    if (!out)
        return fail(BCC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (num_envs < 0 || (num_envs > 0 && (!env_offsets || !counts)))
        return fail(BCC_ERR_INVALID, "bad packed scene arrays");
    std::vector<bcc_env_desc> v((size_t)num_envs);
    for (int e = 0; e < num_envs; ++e) {
        const float* base = data ? data + env_offsets[e] : nullptr;
        const float* ptr[6];
        for (int k = 0; k < 6; ++k) {
            if (counts[6 * e + k] < 0)
                return fail(BCC_ERR_INVALID, "negative primitive count in environment " + std::to_string(e));
            ptr[k] = (counts[6 * e + k] > 0) ? base : nullptr;
            if (counts[6 * e + k] > 0 && !base)
                return fail(BCC_ERR_INVALID, "Host environment has positive counts with null pointers");
            if (base)
                base += (size_t)counts[6 * e + k] * kStride[k];
        }
        v[e] = bcc_env_desc{ptr[0], (uint32_t)counts[6 * e + 0], ptr[1], (uint32_t)counts[6 * e + 1],
                            ptr[2], (uint32_t)counts[6 * e + 2], ptr[3], (uint32_t)counts[6 * e + 3],
                            ptr[4], (uint32_t)counts[6 * e + 4], ptr[5], (uint32_t)counts[6 * e + 5]};
    }
    return build_scene(num_envs, v, out);
}

int bcc_scene_destroy(bcc_scene* scene)
{
    if (!scene)
        return BCC_OK;
    hipError_t a = hipFree(scene->d_obs);
    hipError_t b = hipFree(scene->d_meta);
    hipError_t c = hipFree(scene->d_ref_obs);
    hipError_t d = hipFree(scene->d_ref_meta);
    delete scene;
    if (a == hipSuccess)
        a = c;
    if (b == hipSuccess)
        b = d;
    if (a != hipSuccess)
        return hip_fail(a, "hipFree(scene)");
    if (b != hipSuccess)
        return hip_fail(b, "hipFree(scene meta)");
    return BCC_OK;
}

int bcc_scene_num_envs(const bcc_scene* scene)
{
    return scene ? scene->num_envs : -1;
}

int bcc_check_dense_async(int robot, const bcc_scene* scene, const float* d_edges, int num_edges, int resolution,
                          uint8_t* d_results, void* stream)
{
    if (int rc = check_dense_args(robot, scene, d_edges, num_edges, resolution, d_results))
        return rc;
    bcc::DenseArgs a{scene->d_obs, scene->d_meta, d_edges, num_edges, scene->num_envs, resolution, d_results};
    return dispatch_dense(robot, scene, a, (hipStream_t)stream);
}

int bcc_check_dense_endpoints(int robot, const bcc_scene* scene, const float* d_edges, int num_edges,
                              int resolution, uint8_t* d_results, void* stream)
{
    if (int rc = check_dense_args(robot, scene, d_edges, num_edges, resolution, d_results))
        return rc;
    const size_t pairs = (size_t)num_edges * (size_t)scene->num_envs;
    if (pairs == 0)
        return BCC_OK;
    const hipStream_t s = (hipStream_t)stream;
    // the motion s -> e (its first sample, lane 0's fmaf(d, 0, s), is s
    // itself, so the reference's separate s -> s check adds nothing) ...
    bcc::DenseArgs a{scene->d_obs, scene->d_meta, d_edges, num_edges, scene->num_envs, resolution, d_results};
    if (int rc = dispatch_dense(robot, scene, a, s))
        return rc;
    // ... OR the motion e -> e: n = 1, every lane samples e exactly
    const int dim = bcc_robot_dimension(robot);
    const long long nf = (long long)num_edges * 2 * dim;
    float* d_end = nullptr;
    uint8_t* d_tmp = nullptr;
    hipError_t e = hipMallocAsync((void**)&d_end, sizeof(float) * (size_t)nf, s);
    if (e == hipSuccess)
        e = hipMallocAsync((void**)&d_tmp, pairs, s);
    int rc = BCC_OK;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(end_edges_kernel, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s, d_edges,
                           num_edges, dim, d_end);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        bcc::DenseArgs b{scene->d_obs, scene->d_meta, d_end, num_edges, scene->num_envs, resolution, d_tmp};
        rc = dispatch_dense(robot, scene, b, s);
        if (rc == BCC_OK) {
            hipLaunchKernelGGL(or_bytes_kernel, dim3(1024), dim3(256), 0, s, d_results, d_tmp, (long long)pairs);
            e = hipGetLastError();
        }
    }
    if (d_end)
        (void)hipFreeAsync(d_end, s);
    if (d_tmp)
        (void)hipFreeAsync(d_tmp, s);
    if (rc != BCC_OK)
        return rc;
    BCC_HIP(e, "endpoint check");
    return BCC_OK;
}

int bcc_check_dense(int robot, const bcc_scene* scene, const float* edges, int num_edges, int resolution,
                    uint8_t* results)
{
    if (int rc = check_dense_args(robot, scene, edges, num_edges, resolution, results))
        return rc;
    const size_t pairs = (size_t)num_edges * (size_t)scene->num_envs;
    if (pairs == 0)
        return BCC_OK;
    const int dim = bcc_robot_dimension(robot);
    for (size_t i = 0; i < (size_t)num_edges * 2 * dim; ++i)
        if (!std::isfinite(edges[i]))
            return fail(BCC_ERR_INVALID, "edge " + std::to_string(i / (2 * dim)) + " has a non-finite coordinate");
    float* d_edges = nullptr;
    uint8_t* d_out = nullptr;
    const size_t edge_bytes = sizeof(float) * 2 * dim * (size_t)num_edges;
    hipError_t e = hipMalloc((void**)&d_edges, edge_bytes);
    if (e == hipSuccess)
        e = hipMalloc((void**)&d_out, pairs);
    if (e == hipSuccess)
        e = hipMemcpy(d_edges, edges, edge_bytes, hipMemcpyHostToDevice);
    int rc = BCC_OK;
    if (e != hipSuccess) {
        rc = hip_fail(e, "dense check setup");
    } else {
        bcc::DenseArgs a{scene->d_obs, scene->d_meta, d_edges, num_edges, scene->num_envs, resolution, d_out};
        rc = dispatch_dense(robot, scene, a, nullptr);
        if (rc == BCC_OK) {
            e = hipMemcpy(results, d_out, pairs, hipMemcpyDeviceToHost);
            if (e != hipSuccess)
                rc = hip_fail(e, "dense check results");
        }
    }
    (void)hipFree(d_edges);
    (void)hipFree(d_out);
    return rc;
}

int bcc_check_dense_ex(int robot, const bcc_scene* scene, const float* d_edges, int num_edges, int resolution,
                       const uint32_t* d_order, uint8_t* d_rounds, uint8_t* d_results, void* stream)
{
    if (int rc = check_dense_args(robot, scene, d_edges, num_edges, resolution, d_results))
        return rc;
    bcc::DenseArgs a{scene->d_obs, scene->d_meta, d_edges, num_edges, scene->num_envs, resolution, d_results};
    a.order = d_order;
    a.rounds = d_rounds;
    return dispatch_dense(robot, scene, a, (hipStream_t)stream);
}

int bcc_time_dense_ex(int robot, const bcc_scene* scene, const float* d_edges, int num_edges, int resolution,
                      const uint32_t* d_order, uint8_t* d_results, int warmup, int iters, void* stream,
                      double* avg_ms);

int bcc_time_dense(int robot, const bcc_scene* scene, const float* d_edges, int num_edges, int resolution,
                   uint8_t* d_results, int warmup, int iters, void* stream, double* avg_ms)
{
    return bcc_time_dense_ex(robot, scene, d_edges, num_edges, resolution, nullptr, d_results, warmup, iters, stream,
                             avg_ms);
}

int bcc_time_dense_ex(int robot, const bcc_scene* scene, const float* d_edges, int num_edges, int resolution,
                      const uint32_t* d_order, uint8_t* d_results, int warmup, int iters, void* stream,
                      double* avg_ms)
{
    if (int rc = check_dense_args(robot, scene, d_edges, num_edges, resolution, d_results))
        return rc;
    if (!avg_ms || iters < 1 || warmup < 0)
        return fail(BCC_ERR_INVALID, "bad timing arguments");
    hipStream_t s = (hipStream_t)stream;
    bcc::DenseArgs a{scene->d_obs, scene->d_meta, d_edges, num_edges, scene->num_envs, resolution, d_results};
    a.order = d_order;
    for (int i = 0; i < warmup; ++i)
        if (int rc = dispatch_dense(robot, scene, a, s))
            return rc;
    hipEvent_t t0, t1;
    BCC_HIP(hipEventCreate(&t0), "hipEventCreate");
    BCC_HIP(hipEventCreate(&t1), "hipEventCreate");
    int rc = BCC_OK;
    hipError_t e = hipEventRecord(t0, s);
    for (int i = 0; i < iters && e == hipSuccess && rc == BCC_OK; ++i)
        rc = dispatch_dense(robot, scene, a, s);
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
        return hip_fail(e, "timing events");
    *avg_ms = (double)ms / iters;
    return BCC_OK;
}

}  // extern "C"

// ------------------------------------------------------------ fused learned step
namespace {

std::atomic<int> g_learned_tile{0};   // > 0: forced pairs per dense tile (tuning hook)
std::atomic<int> g_learned_far{0};    // first-probe exclusion radius of the second probe's picks
std::atomic<int> g_learned_passes{bcc::kLearnPassesDefault};   // probe passes (tuning hook)
// Per robot: the fewest steps a pair needs to be inferred and probed again
// (dense launches, learned plans) and the pairs per tile of learned plans.
// Inference pays where the pairs a first probe leaves undecided still
// collide often and have many steps left; on short edges it does not.
// Measured on MI355X: panda dense 3 (c3: 71.1 ms, 6: 73.3), fetch dense 6
// (c1: 4.53 ms, 3: 4.73); learned plans of c4's parts at 64-pair tiles and 6
// steps: panda 22.9 ms, fetch 23.9, baxter 43.5 (3 steps: 23.6, 24.5, 44.1;
// 32-pair tiles gain another 1% there but lose 6% on c2's Zipf batch, 16-pair
// tiles lose 8-20%).  Xarm7 takes panda's.
// The locality order of large dense launches (below) is on for panda only:
// c3 61.0 vs 66.7 ms; fetch's c1 ran 4.66 vs 4.50 ms with it (its sort and
// row copies cost more than its first probe gains); baxter and xarm7 have no
// dense benchmark batch to decide on.
struct LearnDefaults {
    int dense_min_steps;
    int plan_tile;
    int plan_min_steps;
    int locality;
};
constexpr LearnDefaults kLearnDefaults[4] = {{3, 64, 6, 1}, {6, 64, 6, 0}, {3, 64, 6, 0}, {3, 64, 6, 0}};
std::atomic<int> g_learned_min_steps{0};   // > 0: forced for every launch (tuning hook)
std::atomic<int> g_learned_locality{1};    // locality order: 0 never, 1 per-robot default, 2 always
// Dense learned launches of at most this many pairs (no per-pair
// diagnostics asked for) run the ordered dense kernel instead of the fused
// persistent kernel, with every pair's inferred word (with a model) or the
// invalid word (without one: the probe takes bins 8 and 24): at 1e5 pairs
// the persistent kernel has about two 8-pair tiles per resident wave and
// each wave's chain (probe, inference, probe, step loop on two half-waves)
// outlasts the whole plain launch, while inference of every pair costs
// ~0.04 ms (MI355X, 100 environments x 1000 panda edges: 0.36 ms with the
// model and 0.31 without, against 0.54 fused and 0.39 plain).  Measured on
// such batches of growing size, this route is the faster one up to 5.3e5
// pairs (1.01 ms against 1.18 fused) and the fused kernel from 2e6 (2.73
// against 3.44); at 1e6 the two tie (1.98, 2.02).
constexpr long long kLearnSmallPairs = 3LL << 18;
std::atomic<long long> g_learned_small{-1};   // >= 0: forced limit (tuning hook)

int learn_min_steps(int robot, bool plan)
{
    if (const int forced = g_learned_min_steps.load(); forced > 0)
        return forced;
    const LearnDefaults& d = kLearnDefaults[robot >= 0 && robot < 4 ? robot : 0];
    return plan ? d.plan_min_steps : d.dense_min_steps;
}

// Pairs per tile of a learned launch (a wave's tile): about 16 tiles per
// resident wave, a power of two in [8, 128] (MI355X: c3 runs best at 128,
// c1 at 64, and the 1e5-pair batches (c0, the README run) at 8 of 2 ... 256).
int learned_tile(long long pairs, long long resident_waves)
{
    if (const int forced = g_learned_tile.load(); forced > 0)
        return std::min(forced, bcc::kLearnTile);
    const double per = (double)pairs / (16.0 * (double)std::max<long long>(resident_waves, 1));
    int tile = 8;
    while (tile < 128 && 2.0 * tile <= per)
        tile *= 2;
    return tile;
}

// Bins at distance >= radius from every bin of the first probe: the second
// probe's picks prefer them.
uint32_t learned_far_mask(int radius)
{
    uint32_t far = 0xffffffffu;
    for (int k = 0; k < bcc::kProbe1; ++k) {
        const int b = (32 * k + bcc::fixed_offset(0)) / bcc::kProbe1;
        for (int d = -(radius - 1); d <= radius - 1; ++d)
            if (b + d >= 0 && b + d < 32)
                far &= ~(1u << (b + d));
    }
    return far;
}

int dispatch_learned(int robot, const bcc::LearnedArgs& a, hipStream_t s)
{
    hipError_t e;
    switch (robot) {
    case BCC_PANDA: e = bcc::launch_learned_panda(a, s); break;
    case BCC_FETCH: e = bcc::launch_learned_fetch(a, s); break;
    case BCC_BAXTER: e = bcc::launch_learned_baxter(a, s); break;
    case BCC_XARM7: e = bcc::launch_learned_xarm7(a, s); break;
    default: return fail(BCC_ERR_INVALID, "unknown robot id " + std::to_string(robot));
    }
    if (e != hipSuccess)
        return hip_fail(e, "learned kernel launch");
    return BCC_OK;
}

// Locality order of a dense learned launch: the first probe takes 32 pairs
// of one environment per batch, and its culling boxes (bcc_device.h) bound
// the robot spheres of all of them, so pairs whose arms point the same way
// should share a batch.  Key of an edge: the Morton interleave of the two
// joints that place the arm (panda, baxter, xarm7: 0 and 1; fetch: 1 and 2,
// after the torso) at both first-probe samples (8/31 and 24/31 of the way),
// 8 bits each over the joint's range (MI355X, c3 pre-sorted by candidate
// keys, learned step without a model: these four 58.6 ms, the two joints at
// bin 8 alone 60.0, three joints at bin 8 59.4, unsorted 64.4).  The launch
// checks the edges gathered in key
// order and scatters each result row back to its edge, so every output is
// where it would be without it.  Only for launches of at least
// kLocalityMinPairs pairs (sorting and the row copies cost ~0.1 ms) of the
// robots whose LearnDefaults ask for it.
constexpr int kOrderBits = 8;   // per coordinate; four coordinates
constexpr long long kLocalityMinPairs = 1LL << 22;
__global__ void locality_key_kernel(const float* __restrict__ edges, int dim, int m, int j0, int j1, float lo0,
                                    float inv0, float lo1, float inv1, uint32_t* __restrict__ keys,
                                    int* __restrict__ idx)
{
    const float top = (float)((1 << kOrderBits) - 1);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const float* se = edges + (size_t)i * 2 * dim;
        uint32_t u[4];
        for (int c = 0; c < 4; ++c) {
            const float f = (c < 2 ? 8.0f : 24.0f) / 31.0f;
            const int j = (c & 1) ? j1 : j0;
            const float lo = (c & 1) ? lo1 : lo0, inv = (c & 1) ? inv1 : inv0;
            const float q = se[j] + (se[dim + j] - se[j]) * f;
            // NaN compares false and lands on 0: any key is a valid order
            u[c] = (uint32_t)(fminf(fmaxf((q - lo) * inv, 0.0f), 1.0f) * top);
        }
        uint32_t key = 0;
        for (int b = 0; b < kOrderBits; ++b)
            for (int c = 0; c < 4; ++c)
                key |= ((u[c] >> b) & 1u) << (4 * b + c);
        keys[i] = key;
        idx[i] = i;
    }
}

// Row p of dst = row perm[p] of src (gather), or row perm[p] of dst = row p
// of src (scatter); rows of `width` elements.
template <class T>
__global__ void permute_rows_kernel(const T* __restrict__ src, T* __restrict__ dst, const int* __restrict__ perm,
                                   long long rows, int width, int scatter)
{
    const long long total = rows * width;
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < total;
         x += (long long)gridDim.x * blockDim.x) {
        const long long p = x / width;
        const long long c = x - p * width;
        const long long q = perm[p];
        if (scatter)
            dst[q * width + c] = src[x];
        else
            dst[x] = src[q * width + c];
    }
}

template <class T>
void permute_rows(const T* src, T* dst, const int* perm, long long rows, int width, bool scatter, hipStream_t s)
{
    const long long total = rows * width;
    const int blocks = (int)std::min<long long>((total + 255) / 256, 16384);
    hipLaunchKernelGGL((permute_rows_kernel<T>), dim3(blocks), dim3(256), 0, s, src, dst, perm, rows, width,
                       scatter ? 1 : 0);
}

// The edges' locality order on stream s: *order (sorted position -> edge
// row) points into *alloc, a stream-ordered allocation the caller frees with
// hipFreeAsync after the launches that read it.
int locality_order(int robot, const float* d_edges, int num_edges, hipStream_t s, void** alloc, const int** order)
{
    *alloc = nullptr;
    *order = nullptr;
    const int dim = bcc_robot_dimension(robot);
    const int j0 = robot == BCC_FETCH ? 1 : 0;
    const int j1 = j0 + 1;
    float lo0, r0, lo1, r1;
    bcc_detail::joint_range(robot, j0, &lo0, &r0);
    bcc_detail::joint_range(robot, j1, &lo1, &r1);
    size_t temp = 0;
    BCC_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (uint32_t*)nullptr, (uint32_t*)nullptr, (int*)nullptr,
                                               (int*)nullptr, num_edges, 0, 4 * kOrderBits, s),
            "locality sort size");
    // keys in, keys out, indices in, indices out, then the sort's workspace
    const size_t part = ((size_t)num_edges * 4 + 255) / 256 * 256;
    char* buf = nullptr;
    BCC_HIP(hipMallocAsync((void**)&buf, 4 * part + temp, s), "locality order buffers");
    uint32_t* keys = (uint32_t*)buf;
    uint32_t* keys_out = (uint32_t*)(buf + part);
    int* idx = (int*)(buf + 2 * part);
    int* idx_out = (int*)(buf + 3 * part);
    const int blocks = (int)std::min<long long>(((long long)num_edges + 255) / 256, 4096);
    hipLaunchKernelGGL(locality_key_kernel, dim3(blocks), dim3(256), 0, s, d_edges, dim, num_edges, j0, j1, lo0,
                       1.0f / r0, lo1, 1.0f / r1, keys, idx);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipcub::DeviceRadixSort::SortPairs(buf + 4 * part, temp, keys, keys_out, idx, idx_out, num_edges, 0,
                                               4 * kOrderBits, s);
    if (e != hipSuccess) {
        (void)hipFreeAsync(buf, s);
        return hip_fail(e, "locality order");
    }
    *alloc = buf;
    *order = idx_out;
    return BCC_OK;
}

__global__ void fill_words_kernel(uint32_t* __restrict__ p, long long n, uint32_t v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// One dense learned step on stream s: the model's operands (fragments and
// environment codes, two small kernels), then the fused learned kernel,
// persistent over the dense tiles (a wave that ends a tile takes the next
// one from a counter, so no workgroup waits for its slowest wave); the
// counter is a stream-ordered allocation of the call.
int dense_learned(int robot, bcc_model* model, const bcc_scene* scene, const float* d_edges, int num_edges,
                  int resolution, uint8_t* d_results, uint8_t* d_stage, uint32_t* d_words, int32_t* d_stats,
                  hipStream_t s)
{
    if (num_edges == 0 || scene->num_envs == 0)
        return BCC_OK;
    const long long small = g_learned_small.load();
    const long long pairs = (long long)scene->num_envs * num_edges;
    if (!d_stage && !d_words && !d_stats && pairs <= (small >= 0 ? small : kLearnSmallPairs)) {
        // small batch: every pair's order word, then the ordered dense
        // kernel (probe at the word's first bins, the workgroup's plain step
        // loop for the misses)
        uint32_t* d_order = nullptr;
        if (hipError_t e = hipMallocAsync((void**)&d_order, (size_t)pairs * sizeof(uint32_t), s); e != hipSuccess)
            return hip_fail(e, "small-batch order words");
        int rc = BCC_OK;
        if (model) {
            rc = bcc_model_order_dense(model, scene, d_edges, num_edges, d_order, s);
        } else {
            hipLaunchKernelGGL(fill_words_kernel, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0, s, d_order,
                               pairs, 0xffffffffu);
            if (hipError_t e = hipGetLastError(); e != hipSuccess)
                rc = hip_fail(e, "small-batch order words");
        }
        if (rc == BCC_OK)
            rc = bcc_check_dense_ex(robot, scene, d_edges, num_edges, resolution, d_order, nullptr, d_results, s);
        (void)hipFreeAsync(d_order, s);
        return rc;
    }
    bcc::MlpArgs m{};
    if (model) {
        if (int rc = bcc_detail::model_mlp_args(model, scene, s, &m))
            return rc;
    }
    m.far0 = learned_far_mask(g_learned_far.load());
    m.passes = g_learned_passes.load();
    m.min_steps = learn_min_steps(robot, false);
    const int variant = choose_variant(robot, mean_obstacles(scene));
    const bcc::VariantKnobs& k = knobs(robot);
    // one workgroup per resident slot: the variant's waves per SIMD
    const long long slots = (long long)device_cus() * (variant ? k.hi_waves : k.lo_waves);
    const int tile = learned_tile((long long)scene->num_envs * num_edges, slots * (bcc::kBlock / 64));
    const int chunks = (num_edges + tile - 1) / tile;
    const long long tiles = (long long)scene->num_envs * chunks;
    if (tiles > 0x7fffffffLL - 65536)
        return fail(BCC_ERR_INVALID, "too many tiles for one dense learned launch");
    // locality order: the launch reads gathered edges and writes its rows
    // to staging buffers that are scattered back afterwards
    const int E = scene->num_envs;
    const int width = 2 * bcc_robot_dimension(robot);
    void* order_alloc = nullptr;
    const int* order = nullptr;
    char* stage_buf = nullptr;
    const float* edges_in = d_edges;
    uint8_t* res_out = d_results;
    uint8_t* stage_out = d_stage;
    uint32_t* words_out = d_words;
    const int locality = g_learned_locality.load();
    if (num_edges >= 2 && (locality == 2 || (locality == 1 && kLearnDefaults[robot].locality &&
                                             (long long)E * num_edges >= kLocalityMinPairs))) {
        if (int rc = locality_order(robot, d_edges, num_edges, s, &order_alloc, &order))
            return rc;
        const size_t pairs = (size_t)E * num_edges;
        const size_t eb = ((size_t)num_edges * width * sizeof(float) + 255) / 256 * 256;
        const size_t rb = (pairs + 255) / 256 * 256;
        const size_t bytes = eb + rb + (d_stage ? rb : 0) + (d_words ? pairs * sizeof(uint32_t) : 0);
        if (hipError_t e = hipMallocAsync((void**)&stage_buf, bytes, s); e != hipSuccess) {
            (void)hipFreeAsync(order_alloc, s);
            return hip_fail(e, "locality staging buffers");
        }
        float* sorted = (float*)stage_buf;
        res_out = (uint8_t*)(stage_buf + eb);
        stage_out = d_stage ? (uint8_t*)(stage_buf + eb + rb) : nullptr;
        words_out = d_words ? (uint32_t*)(stage_buf + eb + rb + (d_stage ? rb : 0)) : nullptr;
        permute_rows<float>(d_edges, sorted, order, num_edges, width, false, s);
        // words the launch does not write keep the caller's values
        if (d_words)
            permute_rows<uint32_t>(d_words, words_out, order, num_edges, E, false, s);
        edges_in = sorted;
    }
    int* d_counter = nullptr;
    if (hipError_t e = hipMallocAsync((void**)&d_counter, sizeof(int), s); e != hipSuccess) {
        (void)hipFreeAsync(stage_buf, s);
        (void)hipFreeAsync(order_alloc, s);
        return hip_fail(e, "learned tile counter");
    }
    hipLaunchKernelGGL(bcc::reset_counter_kernel, dim3(1), dim3(64), 0, s, d_counter);
    bcc::LearnedArgs a{};
    a.obs = scene->d_obs;
    a.meta = scene->d_meta;
    a.edges = edges_in;
    a.tiles = bcc::LearnTiles{nullptr, (int)tiles, d_counter, chunks, num_edges, tile};
    a.resolution = resolution;
    a.variant = variant;
    a.grid = (int)std::min<long long>(slots, (tiles + bcc::kBlock / 64 - 1) / (bcc::kBlock / 64));
    a.mlp = m;
    a.out = bcc::LearnedOut{bcc::PairSlots{scene->num_envs, 1}, res_out, stage_out, words_out, d_stats};
    int rc = BCC_OK;
    if (hipError_t e = hipGetLastError(); e != hipSuccess)
        rc = hip_fail(e, "learned tile counter reset");
    if (rc == BCC_OK)
        rc = dispatch_learned(robot, a, s);
    if (rc == BCC_OK && order) {
        permute_rows<uint8_t>(res_out, d_results, order, num_edges, E, true, s);
        if (d_stage)
            permute_rows<uint8_t>(stage_out, d_stage, order, num_edges, E, true, s);
        if (d_words)
            permute_rows<uint32_t>(words_out, d_words, order, num_edges, E, true, s);
        if (hipError_t e = hipGetLastError(); e != hipSuccess)
            rc = hip_fail(e, "locality scatter");
    }
    (void)hipFreeAsync(d_counter, s);
    if (stage_buf)
        (void)hipFreeAsync(stage_buf, s);
    if (order_alloc)
        (void)hipFreeAsync(order_alloc, s);
    return rc;
}

int learned_args(int robot, const bcc_model* model, const bcc_scene* scene, const void* d_edges, int num_edges,
                 int resolution, const void* d_results)
{
    if (int rc = check_dense_args(robot, scene, d_edges, num_edges, resolution, d_results))
        return rc;
    if (model && bcc_detail::model_robot(model) != robot)
        return fail(BCC_ERR_INVALID, "the model was created for another robot");
    return BCC_OK;
}

}  // namespace

extern "C" {

int bcc_check_dense_learned(int robot, bcc_model* model, const bcc_scene* scene, const float* d_edges,
                            int num_edges, int resolution, uint8_t* d_results, uint8_t* d_stage, uint32_t* d_words,
                            int32_t* d_stats, void* stream)
{
    if (int rc = learned_args(robot, model, scene, d_edges, num_edges, resolution, d_results))
        return rc;
    return dense_learned(robot, model, scene, d_edges, num_edges, resolution, d_results, d_stage, d_words, d_stats,
                         (hipStream_t)stream);
}

int bcc_time_dense_learned(int robot, bcc_model* model, const bcc_scene* scene, const float* d_edges,
                           int num_edges, int resolution, uint8_t* d_results, int warmup, int iters, void* stream,
                           double* avg_ms)
{
    if (int rc = learned_args(robot, model, scene, d_edges, num_edges, resolution, d_results))
        return rc;
    if (!avg_ms || iters < 1 || warmup < 0)
        return fail(BCC_ERR_INVALID, "bad timing arguments");
    hipStream_t s = (hipStream_t)stream;
    auto step = [&]() {
        return dense_learned(robot, model, scene, d_edges, num_edges, resolution, d_results, nullptr, nullptr,
                             nullptr, s);
    };
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
        return hip_fail(e, "timing events");
    *avg_ms = (double)ms / iters;
    return BCC_OK;
}

int bcc_debug_learned_tile(int pairs)
{
    if (pairs < 0 || pairs > bcc::kLearnTile)
        return fail(BCC_ERR_INVALID, "learned tile must be in [0, 256]");
    g_learned_tile.store(pairs);
    return BCC_OK;
}

int bcc_debug_learned_passes(int passes)
{
    if (passes < 0 || passes > bcc::kLearnPassesMax)
        return fail(BCC_ERR_INVALID, "learned probe passes must be in [0, 3]");
    g_learned_passes.store(passes == 0 ? bcc::kLearnPassesDefault : passes);
    return BCC_OK;
}

int bcc_debug_learned_min_steps(int steps)
{
    if (steps < 0)
        return fail(BCC_ERR_INVALID, "learned minimum steps must be >= 0");
    g_learned_min_steps.store(steps);
    return BCC_OK;
}

int bcc_debug_learned_small(int64_t max_pairs)
{
    g_learned_small.store(max_pairs < 0 ? -1 : max_pairs);
    return BCC_OK;
}

int64_t bcc_learned_small_pairs(void)
{
    const long long small = g_learned_small.load();
    return small >= 0 ? small : kLearnSmallPairs;
}

int bcc_debug_learned_locality(int mode)
{
    if (mode < 0 || mode > 2)
        return fail(BCC_ERR_INVALID, "locality must be 0, 1 or 2");
    g_learned_locality.store(mode);
    return BCC_OK;
}

int bcc_debug_learned_far(int radius)
{
    if (radius < 0 || radius > 32)
        return fail(BCC_ERR_INVALID, "radius must be in [0, 32]");
    g_learned_far.store(radius);
    return BCC_OK;
}

}  // extern "C"

namespace bcc_detail {

const bcc_robot_model* host_model(int robot)
{
    switch (robot) {
    case BCC_PANDA: return &bcc_panda_model;
    case BCC_FETCH: return &bcc_fetch_model;
    case BCC_BAXTER: return &bcc_baxter_model;
    case BCC_XARM7: return &bcc_xarm7_model;
    default: return nullptr;
    }
}

// Program tables of a robot on the current device (uploaded on first use and
// kept for the life of the process).
int device_model(int robot, bcc::ModelDev* out)
{
    static std::mutex mu;
    static std::map<std::pair<int, int>, bcc::ModelDev> cache;
    int dev = 0;
    BCC_HIP(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({dev, robot});
    if (it != cache.end()) {
        *out = it->second;
        return BCC_OK;
    }
    const bcc_robot_model* m = host_model(robot);
    if (!m)
        return fail(BCC_ERR_INVALID, "unknown robot id");
    const int ne = m->num_env_groups, nsg = m->num_self_groups, np = m->num_program;
    std::vector<int> ints;
    std::vector<size_t> off;
    auto put = [&](const int* a, int n) {
        off.push_back(ints.size());
        ints.insert(ints.end(), a, a + n);
    };
    put(m->env_gate, ne);
    put(m->env_lone, ne);
    put(m->env_fine_off, ne + 1);
    put(m->env_fine, std::max(m->env_fine_off[ne], 1));
    put(m->self_gate_a, nsg);
    put(m->self_gate_b, nsg);
    put(m->self_lone, nsg);
    put(m->self_a_off, nsg + 1);
    put(m->self_a, std::max(m->self_a_off[nsg], 1));
    put(m->self_b_off, nsg + 1);
    put(m->self_b, std::max(m->self_b_off[nsg], 1));
    put(m->prog_kind, np);
    put(m->prog_index, np);
    put(m->prog_fk_ops, np);
    put(m->prog_fk_trig, np);
    int* d_ints = nullptr;
    float* d_r = nullptr;
    BCC_HIP(hipMalloc((void**)&d_ints, sizeof(int) * ints.size()), "model upload");
    BCC_HIP(hipMalloc((void**)&d_r, sizeof(float) * m->num_spheres), "model upload");
    BCC_HIP(hipMemcpy(d_ints, ints.data(), sizeof(int) * ints.size(), hipMemcpyHostToDevice), "model upload");
    BCC_HIP(hipMemcpy(d_r, m->sphere_r, sizeof(float) * m->num_spheres, hipMemcpyHostToDevice), "model upload");
    bcc::ModelDev d;
    d.sphere_r = d_r;
    const int* b = d_ints;
    d.env_gate = b + off[0];
    d.env_lone = b + off[1];
    d.env_fine_off = b + off[2];
    d.env_fine = b + off[3];
    d.self_gate_a = b + off[4];
    d.self_gate_b = b + off[5];
    d.self_lone = b + off[6];
    d.self_a_off = b + off[7];
    d.self_a = b + off[8];
    d.self_b_off = b + off[9];
    d.self_b = b + off[10];
    d.prog_kind = b + off[11];
    d.prog_index = b + off[12];
    d.prog_fk_ops = b + off[13];
    d.prog_fk_trig = b + off[14];
    d.num_program = np;
    d.ref_fk_flops = m->ref_fk_flops;
    d.ref_fk_trig = m->ref_fk_trig;
    cache[{dev, robot}] = d;
    *out = d;
    return BCC_OK;
}

}  // namespace bcc_detail

using bcc_detail::device_model;

extern "C" {

static int work_common(int robot, const bcc_scene* scene, const float* d_edges, int num_edges, const int* d_pair_env,
                       int resolution, uint64_t* d_work, uint64_t* d_trig, uint8_t* d_results, void* stream);

int bcc_work_dense(int robot, const bcc_scene* scene, const float* d_edges, int num_edges, int resolution,
                   uint64_t* d_work, uint64_t* d_trig, uint8_t* d_results, void* stream)
{
    return work_common(robot, scene, d_edges, num_edges, nullptr, resolution, d_work, d_trig, d_results, stream);
}

int bcc_work_pairs(int robot, const bcc_scene* scene, const float* d_edges, int num_pairs, const int32_t* d_pair_env,
                   int resolution, uint64_t* d_work, uint64_t* d_trig, uint8_t* d_results, void* stream)
{
    if (num_pairs > 0 && !d_pair_env)
        return fail(BCC_ERR_INVALID, "pair_env is NULL");
    return work_common(robot, scene, d_edges, num_pairs, d_pair_env, resolution, d_work, d_trig, d_results, stream);
}

static int work_common(int robot, const bcc_scene* scene, const float* d_edges, int num_edges, const int* d_pair_env,
                       int resolution, uint64_t* d_work, uint64_t* d_trig, uint8_t* d_results, void* stream)
{
    if (int rc = check_dense_args(robot, scene, d_edges, num_edges, resolution, d_results))
        return rc;
    const size_t pairs = d_pair_env ? (size_t)num_edges : (size_t)num_edges * (size_t)scene->num_envs;
    if (pairs == 0)
        return BCC_OK;
    if (!d_work || !d_trig)
        return fail(BCC_ERR_INVALID, "work/trig buffer is NULL");
    hipStream_t s = (hipStream_t)stream;
    BCC_HIP(hipMemsetAsync(d_work, 0, sizeof(uint64_t) * pairs, s), "work reset");
    BCC_HIP(hipMemsetAsync(d_trig, 0, sizeof(uint64_t) * pairs, s), "work reset");
    bcc::ModelDev md;
    if (int rc = device_model(robot, &md))
        return rc;
    bcc::WorkArgs a{scene->d_ref_obs, scene->d_ref_meta, d_edges, num_edges, scene->num_envs, resolution, md,
                    d_pair_env, (unsigned long long*)d_work, (unsigned long long*)d_trig, d_results};
    hipError_t e;
    switch (robot) {
    case BCC_PANDA: e = bcc::launch_work_panda(a, s); break;
    case BCC_FETCH: e = bcc::launch_work_fetch(a, s); break;
    case BCC_BAXTER: e = bcc::launch_work_baxter(a, s); break;
    case BCC_XARM7: e = bcc::launch_work_xarm7(a, s); break;
    default: return fail(BCC_ERR_INVALID, "unknown robot id");
    }
    if (e != hipSuccess)
        return hip_fail(e, "work kernel launch");
    return BCC_OK;
}

int bcc_audit_clearance(int robot, const bcc_scene* scene, const float* d_edges, int num_edges,
                        const int32_t* d_pair_env, int resolution, float* d_clearance, void* stream)
{
    if (int rc = check_dense_args(robot, scene, d_edges, num_edges, resolution, d_clearance))
        return rc;
    const size_t pairs = d_pair_env ? (size_t)num_edges : (size_t)num_edges * (size_t)scene->num_envs;
    if (pairs == 0)
        return BCC_OK;
    bcc::ModelDev md;
    if (int rc = device_model(robot, &md))
        return rc;
    bcc::AuditArgs a{scene->d_ref_obs, scene->d_ref_meta, d_edges, num_edges, scene->num_envs, resolution, md,
                     d_pair_env, d_clearance};
    const hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    switch (robot) {
    case BCC_PANDA: e = bcc::launch_audit_panda(a, s); break;
    case BCC_FETCH: e = bcc::launch_audit_fetch(a, s); break;
    case BCC_BAXTER: e = bcc::launch_audit_baxter(a, s); break;
    case BCC_XARM7: e = bcc::launch_audit_xarm7(a, s); break;
    default: return fail(BCC_ERR_INVALID, "unknown robot id");
    }
    if (e != hipSuccess)
        return hip_fail(e, "audit kernel launch");
    return BCC_OK;
}

}  // extern "C"

// A plan holds one batch's tile list, buffers and captured graph.  Two
// kinds:
//   plain    the ragged kernel over every tile;
//   ordered  a learned order given for every pair: learn_finish_kernel over
//            segments that are the tiles themselves (identity queue);
//   learned  the fused learned kernel, persistent over the tiles, with a
//            snapshot of the model taken when the plan is made.
enum PlanKind { kPlanPlain = 0, kPlanOrdered = 1, kPlanLearned = 2 };

struct bcc_ragged_plan {
    int robot = 0;
    int kind = kPlanPlain;
    const bcc_scene* scene = nullptr;
    bcc::RaggedTile* d_tiles = nullptr;
    int* d_counter = nullptr;    // plain: the tile counter; ordered: kCtrInts counters
    int num_tiles = 0;
    int64_t num_pairs = 0;
    int tile_pairs = 0;
    int groups_per_cu = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bcc::RaggedArgs args{};
    // ordered plans
    bcc::LearnArgs learn{};
    int preset_segs = 0;              // segments in the queue: every tile
    int* d_queue_idx = nullptr;       // [num_pairs] identity
    bcc::QueueSeg* d_segs = nullptr;  // [num_tiles]
    // learned plans
    bcc::LearnedArgs fused{};
    void* d_mlp = nullptr;            // the model snapshot (null: no model)
    uint8_t* d_results = nullptr;     // plain plans with a step dimension zero-fill these first
};

struct bcc_ragged_group {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int dispatch_ragged(int robot, const bcc::RaggedArgs& a, hipStream_t s)
{
    hipError_t e;
    switch (robot) {
    case BCC_PANDA: e = bcc::launch_ragged_panda(a, s); break;
    case BCC_FETCH: e = bcc::launch_ragged_fetch(a, s); break;
    case BCC_BAXTER: e = bcc::launch_ragged_baxter(a, s); break;
    case BCC_XARM7: e = bcc::launch_ragged_xarm7(a, s); break;
    default: return fail(BCC_ERR_INVALID, "unknown robot id " + std::to_string(robot));
    }
    if (e != hipSuccess)
        return hip_fail(e, "ragged kernel launch");
    return BCC_OK;
}

int dispatch_learn(int robot, const bcc::LearnArgs& a, hipStream_t s)
{
    hipError_t e;
    switch (robot) {
    case BCC_PANDA: e = bcc::launch_learn_panda(a, s); break;
    case BCC_FETCH: e = bcc::launch_learn_fetch(a, s); break;
    case BCC_BAXTER: e = bcc::launch_learn_baxter(a, s); break;
    case BCC_XARM7: e = bcc::launch_learn_xarm7(a, s); break;
    default: return fail(BCC_ERR_INVALID, "unknown robot id " + std::to_string(robot));
    }
    if (e != hipSuccess)
        return hip_fail(e, "learned-order kernel launch");
    return BCC_OK;
}

// Launch shape from the tuning table (bcc_tuning.h, written by
// batch-cc_amd/tune.py from timings on the GPU): per robot, workgroups per CU
// of the persistent grid and a target number of tiles per wave; the tile
// size follows from the batch, pairs per wave / tiles per wave, rounded down
// to a power of two in [1, 256], so short batches get small tiles (load
// balance at the end of the launch) and long ones large tiles (fewer
// tile fetches).
const bcc_tune_entry& tuned(int robot)
{
    for (const bcc_tune_entry& t : kTuneTable)
        if (t.robot == robot)
            return t;
    return kTuneTable[0];   // every robot has an entry (checked by the library tests)
}

int tile_for(int64_t total_pairs, int grid_waves, int tiles_per_wave)
{
    const double per = (double)total_pairs / ((double)std::max(grid_waves, 1) * std::max(tiles_per_wave, 1));
    int tile = 1;
    while (tile < 256 && 2.0 * tile <= per)
        tile *= 2;
    return tile;
}

int device_cus()
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return cus;
}

// Fills n bytes with v (a kernel node, not a memset node, when captured).
__global__ void fill_bytes_kernel(uint8_t* __restrict__ p, long long n, int v)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = (uint8_t)v;
}

// Every launch of a plan (captured once into its graph, or run eagerly).
int run_plan(const bcc_ragged_plan* p, hipStream_t s)
{
    if (p->kind == kPlanPlain) {
        if (p->args.tile_steps > 0 && p->num_pairs > 0) {
            // chunks of one pair only ever write a collision: start from 0
            hipLaunchKernelGGL(fill_bytes_kernel, dim3((unsigned)std::min<int64_t>((p->num_pairs + 255) / 256, 4096)),
                               dim3(256), 0, s, p->d_results, (long long)p->num_pairs, 0);
            if (hipError_t e = hipGetLastError(); e != hipSuccess)
                return hip_fail(e, "verdict zero-fill");
        }
        return dispatch_ragged(p->robot, p->args, s);
    }
    if (p->kind == kPlanLearned) {
        hipLaunchKernelGGL(bcc::reset_counter_kernel, dim3(1), dim3(64), 0, s, p->d_counter);
        if (hipError_t e = hipGetLastError(); e != hipSuccess)
            return hip_fail(e, "plan counter reset");
        return dispatch_learned(p->robot, p->fused, s);
    }
    hipLaunchKernelGGL(bcc::reset_pipeline_kernel, dim3(1), dim3(64), 0, s, p->d_counter, p->preset_segs);
    if (hipError_t e = hipGetLastError(); e != hipSuccess)
        return hip_fail(e, "plan counter reset");
    if (p->num_tiles == 0)
        return BCC_OK;
    return dispatch_learn(p->robot, p->learn, s);
}


}  // namespace

extern "C" {

void bcc_ragged_plan_destroy_impl(bcc_ragged_plan* p)
{
    if (!p)
        return;
    if (p->exec)
        (void)hipGraphExecDestroy(p->exec);
    if (p->graph)
        (void)hipGraphDestroy(p->graph);
    (void)hipFree(p->d_tiles);
    (void)hipFree(p->d_counter);
    (void)hipFree(p->d_queue_idx);
    (void)hipFree(p->d_segs);
    (void)hipFree(p->d_mlp);
    delete p;
}

int bcc_ragged_plan_create(int robot, const bcc_scene* scene, const int64_t* env_offsets, const float* d_edges,
                           int resolution, uint8_t* d_results, int tile_pairs, bcc_ragged_plan** out)
{
    return bcc_ragged_plan_create_tuned(robot, scene, env_offsets, d_edges, resolution, d_results, tile_pairs, 0,
                                        out);
}

}  // extern "C"

namespace {

// Batch description shared by the plan kinds: tiles (env, first, end) over
// an index range, the verdict slot of index i in environment env, and the
// pair-weighted mean obstacle count that picks the kernel variant.
struct PlanBatch {
    std::vector<bcc::RaggedTile> tiles;
    bcc::PairSlots slots{1, 0};
    int64_t total = 0;
    double mean_obstacles = 0.0;
    int tile_steps = 0;   // > 0: tiles carry step chunks of this many steps
};

// Captures the plan's launch sequence into its graph (the plan's buffers
// must be set); on failure destroys the plan.
int capture_plan(bcc_ragged_plan* p, bcc_ragged_plan** out)
{
    hipStream_t cap;
    hipError_t err = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
    if (err == hipSuccess) {
        err = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (err == hipSuccess) {
            const int rc = run_plan(p, cap);
            hipGraph_t g = nullptr;
            const hipError_t e2 = hipStreamEndCapture(cap, &g);
            p->graph = g;
            if (rc != BCC_OK) {
                (void)hipStreamDestroy(cap);
                bcc_ragged_plan_destroy_impl(p);
                return rc;
            }
            err = e2;
        }
        if (err == hipSuccess)
            err = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
        (void)hipStreamDestroy(cap);
    }
    if (err != hipSuccess) {
        bcc_ragged_plan_destroy_impl(p);
        return hip_fail(err, "plan graph capture");
    }
    *out = p;
    return BCC_OK;
}

// Plan of either kind over a batch.  d_order: ordered plans, one word per
// pair (index i).
// Learned plans: mlp holds the model's operands in the allocation mlp_owned,
// which the plan takes over (also on failure).
int create_plan_kind(int robot, const bcc_scene* scene, PlanBatch& b, const float* d_edges, int resolution,
                     int kind, const uint32_t* d_order, uint8_t* d_results, int tile_pairs, int groups_per_cu,
                     bcc_ragged_plan** out, const bcc::MlpArgs* mlp = nullptr, void* mlp_owned = nullptr)
{
    const int variant = choose_variant(robot, b.mean_obstacles);
    const bcc_tune_entry& tune = tuned(robot);
    if (groups_per_cu == 0)
        groups_per_cu = tune.groups_per_cu;
    bcc_ragged_plan* p = new bcc_ragged_plan;
    p->d_mlp = mlp_owned;
    p->robot = robot;
    p->kind = kind;
    p->scene = scene;
    p->num_tiles = (int)b.tiles.size();
    p->num_pairs = b.total;
    p->tile_pairs = tile_pairs;
    p->groups_per_cu = groups_per_cu;
    const size_t n = (size_t)std::max<int64_t>(b.total, 1);
    const size_t nt = std::max<size_t>(b.tiles.size(), 1);
    hipError_t err = hipMalloc((void**)&p->d_tiles, sizeof(bcc::RaggedTile) * nt);
    if (err == hipSuccess)
        err = hipMalloc((void**)&p->d_counter, sizeof(int) * bcc::kCtrInts);
    if (err == hipSuccess && !b.tiles.empty())
        err = hipMemcpy(p->d_tiles, b.tiles.data(), sizeof(bcc::RaggedTile) * b.tiles.size(), hipMemcpyHostToDevice);
    if (err == hipSuccess && kind == kPlanOrdered) {
        err = hipMalloc((void**)&p->d_queue_idx, sizeof(int) * n);
        if (err == hipSuccess)
            err = hipMalloc((void**)&p->d_segs, sizeof(bcc::QueueSeg) * nt);
    }
    if (err == hipSuccess && kind == kPlanOrdered) {
        // every tile is a segment of an identity queue
        std::vector<int> idx(n, 0);
        std::vector<bcc::QueueSeg> segs(b.tiles.size());
        size_t q = 0;
        for (size_t t = 0; t < b.tiles.size(); ++t) {
            segs[t] = bcc::QueueSeg{b.tiles[t].env, (int)q, b.tiles[t].end - b.tiles[t].begin, 0};
            for (int i = b.tiles[t].begin; i < b.tiles[t].end; ++i)
                idx[q++] = i;
        }
        p->preset_segs = (int)segs.size();
        err = hipMemcpy(p->d_queue_idx, idx.data(), sizeof(int) * n, hipMemcpyHostToDevice);
        if (err == hipSuccess && !segs.empty())
            err = hipMemcpy(p->d_segs, segs.data(), sizeof(bcc::QueueSeg) * segs.size(), hipMemcpyHostToDevice);
    }
    if (err != hipSuccess) {
        bcc_ragged_plan_destroy_impl(p);
        return hip_fail(err, "plan upload");
    }
    const int grid = device_cus() * groups_per_cu;
    if (kind == kPlanPlain) {
        p->args = bcc::RaggedArgs{scene->d_obs, scene->d_meta, d_edges, p->d_tiles, p->num_tiles, p->d_counter,
                                  resolution, grid, d_results};
        p->args.variant = variant;
        p->args.tile_steps = b.tile_steps;
        p->d_results = d_results;
    } else if (kind == kPlanLearned) {
        bcc::LearnedArgs& F = p->fused;
        F.obs = scene->d_obs;
        F.meta = scene->d_meta;
        F.edges = d_edges;
        F.tiles = bcc::LearnTiles{p->d_tiles, p->num_tiles, p->d_counter, 0, 0, 0};
        F.resolution = resolution;
        F.grid = grid;
        F.variant = variant;
        F.mlp = mlp ? *mlp : bcc::MlpArgs{};
        F.mlp.far0 = learned_far_mask(g_learned_far.load());
        F.mlp.passes = g_learned_passes.load();
        F.mlp.min_steps = learn_min_steps(robot, true);
        F.out = bcc::LearnedOut{b.slots, d_results, nullptr, nullptr, nullptr};
    } else {
        bcc::LearnArgs& L = p->learn;
        L.obs = scene->d_obs;
        L.meta = scene->d_meta;
        L.edges = d_edges;
        L.ctr = p->d_counter;
        L.resolution = resolution;
        L.grid = grid;
        L.variant = variant;
        L.slots = b.slots;
        L.out = d_results;
        L.queue_idx = p->d_queue_idx;
        L.segs = p->d_segs;
        L.order = d_order;
    }
    return capture_plan(p, out);
}

// Tiles of a CSR batch: env e owns pairs [env_offsets[e], env_offsets[e+1]).
int ragged_batch(const bcc_scene* scene, const int64_t* env_offsets, int robot, int tile_pairs, int groups_per_cu,
                 int max_tile, PlanBatch& b, int* tile_out)
{
    if (!env_offsets)
        return fail(BCC_ERR_INVALID, "env_offsets is NULL");
    const int E = scene->num_envs;
    if (env_offsets[0] != 0)
        return fail(BCC_ERR_INVALID, "env_offsets[0] must be 0");
    for (int e = 0; e < E; ++e)
        if (env_offsets[e + 1] < env_offsets[e])
            return fail(BCC_ERR_INVALID, "env_offsets must be non-decreasing");
    const int64_t total = env_offsets[E];
    if (total > 0x7fffffffLL)
        return fail(BCC_ERR_INVALID, "more than 2^31 pairs in one ragged batch");
    if (tile_pairs < -1024 || groups_per_cu < 0 || groups_per_cu > 32)
        return fail(BCC_ERR_INVALID, "tile_pairs must be >= -1024 and groups_per_cu in 0..32");
    double weighted = 0.0;   // mean obstacles per environment, weighted by pairs
    for (int e = 0; e < E; ++e)
        weighted += (double)scene->obstacles[e] * (double)(env_offsets[e + 1] - env_offsets[e]);
    b.total = total;
    b.mean_obstacles = total > 0 ? weighted / (double)total : 0.0;
    b.slots = bcc::PairSlots{1, 0};
    const int variant = choose_variant(robot, b.mean_obstacles);
    const bcc_tune_entry& tune = tuned(robot);
    if (groups_per_cu == 0)
        groups_per_cu = tune.groups_per_cu;
    if (tile_pairs <= 0) {
        // waves that can be resident at once: the grid, capped by the variant's occupancy
        const bcc::VariantKnobs& k = knobs(robot);
        const int resident = std::min(groups_per_cu, variant ? k.hi_waves : k.lo_waves);
        const int per_wave = tile_pairs < 0 ? -tile_pairs : tune.tiles_per_wave;
        tile_pairs = tile_for(total, device_cus() * resident * (bcc::kBlock / 64), per_wave);
    }
    tile_pairs = std::min(tile_pairs, max_tile);
    for (int e = 0; e < E; ++e)
        for (int64_t i = env_offsets[e]; i < env_offsets[e + 1]; i += tile_pairs)
            b.tiles.push_back(bcc::RaggedTile{e, (int)i, (int)std::min<int64_t>(i + tile_pairs, env_offsets[e + 1]), 0});
    // heavy environments first: more obstacles, then larger tiles
    std::stable_sort(b.tiles.begin(), b.tiles.end(), [&](const bcc::RaggedTile& x, const bcc::RaggedTile& y) {
        const int64_t wx = (int64_t)scene->obstacles[x.env] * (x.end - x.begin);
        const int64_t wy = (int64_t)scene->obstacles[y.env] * (y.end - y.begin);
        return wx > wy;
    });
    *tile_out = tile_pairs;
    return BCC_OK;
}

// Plain or ordered plan over one CSR layout.
// Step counts n of every pair of a ragged batch (the arithmetic of
// setup_pair_sd, so the same n the kernels compute), into n_out.
__global__ void pair_steps_kernel(const float* __restrict__ edges, int dim, long long pairs, int resolution,
                                  int* __restrict__ n_out)
{
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < pairs;
         p += (long long)gridDim.x * blockDim.x) {
        const float* se = edges + p * 2 * dim;
        float acc = 0.0f;
        for (int j = 0; j < dim; ++j) {
            const float diff = se[j] - se[dim + j];
            acc = fmaf(diff, diff, acc);
        }
        const float dist = sqrtf(acc);
        n_out[p] = (int)fminf(fmaxf(ceilf((dist / 31.0f) * (float)resolution), 1.0f), 4194304.0f);
    }
}

// The step dimension of a plain plan's tiles: tile_steps > 0 splits every
// tile whose pairs run more steps than that into chunks of tile_steps steps;
// 0 asks the tuning table's steps per tile, applied only when the batch has
// fewer pairs than twice the grid's half-waves (with more, the pairs alone
// fill the GPU); < 0 never splits.  Splitting reads every pair's step count
// from the device once, here.
int split_steps(int robot, const float* d_edges, int resolution, int tile_steps, int groups_per_cu, PlanBatch& b)
{
    if (tile_steps == 0) {
        const bcc_tune_entry& tune = tuned(robot);
        const bcc::VariantKnobs& k = knobs(robot);
        const int groups = groups_per_cu > 0 ? groups_per_cu : tune.groups_per_cu;
        const int variant = choose_variant(robot, b.mean_obstacles);
        const long long halves = (long long)device_cus() * std::min(groups, variant ? k.hi_waves : k.lo_waves) *
                                 (bcc::kBlock / 32);
        tile_steps = (tune.steps_per_tile > 0 && b.total < 2 * halves) ? tune.steps_per_tile : -1;
    }
    if (tile_steps < 0 || b.total == 0 || b.tiles.empty())
        return BCC_OK;
    const int dim = bcc_robot_dimension(robot);
    int* d_n = nullptr;
    BCC_HIP(hipMalloc((void**)&d_n, sizeof(int) * (size_t)b.total), "step counts");
    hipLaunchKernelGGL(pair_steps_kernel, dim3((unsigned)std::min<int64_t>((b.total + 255) / 256, 4096)), dim3(256),
                       0, 0, d_edges, dim, (long long)b.total, resolution, d_n);
    std::vector<int> n((size_t)b.total);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpy(n.data(), d_n, sizeof(int) * (size_t)b.total, hipMemcpyDeviceToHost);
    (void)hipFree(d_n);
    if (e != hipSuccess)
        return hip_fail(e, "step counts");
    std::vector<bcc::RaggedTile> tiles;
    for (const bcc::RaggedTile& t : b.tiles) {
        int nmax = 1;
        for (int i = t.begin; i < t.end; ++i)
            nmax = std::max(nmax, n[(size_t)i]);
        const int chunks = (nmax + tile_steps - 1) / tile_steps;
        for (int c = 0; c < chunks; ++c)
            tiles.push_back(bcc::RaggedTile{t.env, t.begin, t.end, c});
    }
    b.tiles.swap(tiles);
    b.tile_steps = tile_steps;
    return BCC_OK;
}

int create_plan(int robot, const bcc_scene* scene, const int64_t* env_offsets, const float* d_edges, int resolution,
                bool ordered, const uint32_t* d_order, uint8_t* d_results, int tile_pairs, int groups_per_cu,
                bcc_ragged_plan** out, int tile_steps = -1)
{
    if (!out)
        return fail(BCC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (int rc = check_dense_args(robot, scene, d_edges, 0, resolution, d_results))
        return rc;
    PlanBatch b;
    // the pipeline kernels queue a tile's undecided pairs in LDS
    const int max_tile = ordered ? bcc::kOrderedTileMax : 1 << 30;
    if (int rc = ragged_batch(scene, env_offsets, robot, tile_pairs, groups_per_cu, max_tile, b, &tile_pairs))
        return rc;
    if (b.total > 0 && (!d_edges || !d_results))
        return fail(BCC_ERR_INVALID, "edges/results buffer is NULL");
    if (ordered && b.total > 0 && !d_order)
        return fail(BCC_ERR_INVALID, "d_order is NULL");
    if (!ordered)
        if (int rc = split_steps(robot, d_edges, resolution, tile_steps, groups_per_cu, b))
            return rc;
    return create_plan_kind(robot, scene, b, d_edges, resolution, ordered ? kPlanOrdered : kPlanPlain, d_order,
                            d_results, tile_pairs, groups_per_cu, out);
}

}  // namespace

extern "C" {

int bcc_ragged_plan_create_tuned(int robot, const bcc_scene* scene, const int64_t* env_offsets,
                                 const float* d_edges, int resolution, uint8_t* d_results, int tile_pairs,
                                 int groups_per_cu, bcc_ragged_plan** out)
{
    return create_plan(robot, scene, env_offsets, d_edges, resolution, false, nullptr, d_results, tile_pairs,
                       groups_per_cu, out, 0);
}

int bcc_ragged_plan_create_shaped(int robot, const bcc_scene* scene, const int64_t* env_offsets,
                                  const float* d_edges, int resolution, uint8_t* d_results, int tile_pairs,
                                  int tile_steps, int groups_per_cu, bcc_ragged_plan** out)
{
    if (tile_steps < -1 || tile_steps > (1 << 22))
        return fail(BCC_ERR_INVALID, "tile_steps must be -1, 0 or in 1..2^22");
    return create_plan(robot, scene, env_offsets, d_edges, resolution, false, nullptr, d_results, tile_pairs,
                       groups_per_cu, out, tile_steps);
}

int bcc_ragged_plan_tile_steps(const bcc_ragged_plan* plan)
{
    return plan && plan->kind == kPlanPlain ? plan->args.tile_steps : 0;
}

int bcc_ragged_plan_create_ordered(int robot, const bcc_scene* scene, const int64_t* env_offsets,
                                   const float* d_edges, int resolution, const uint32_t* d_order,
                                   uint8_t* d_results, int tile_pairs, int groups_per_cu, bcc_ragged_plan** out)
{
    return create_plan(robot, scene, env_offsets, d_edges, resolution, true, d_order, d_results, tile_pairs,
                       groups_per_cu, out);
}

int bcc_ragged_plan_create_learned(int robot, const bcc_model* model, const bcc_scene* scene,
                                   const int64_t* env_offsets, const float* d_edges, int resolution,
                                   uint8_t* d_results, int tile_pairs, int groups_per_cu, bcc_ragged_plan** out)
{
    if (!out)
        return fail(BCC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (int rc = learned_args(robot, model, scene, d_edges, 0, resolution, d_results))
        return rc;
    PlanBatch b;
    if (tile_pairs == 0)
        tile_pairs = std::min(kLearnDefaults[robot].plan_tile, bcc::kLearnTile);
    // the fused learned kernel queues a tile's undecided pairs in LDS
    if (int rc = ragged_batch(scene, env_offsets, robot, tile_pairs, groups_per_cu, bcc::kLearnTile, b, &tile_pairs))
        return rc;
    if (b.total > 0 && (!d_edges || !d_results))
        return fail(BCC_ERR_INVALID, "edges/results buffer is NULL");
    bcc::MlpArgs m{};
    void* owned = nullptr;
    if (model) {
        if (int rc = bcc_detail::model_mlp_snapshot(model, scene, &m, &owned))
            return rc;
    }
    return create_plan_kind(robot, scene, b, d_edges, resolution, kPlanLearned, nullptr, d_results, tile_pairs,
                            groups_per_cu, out, &m, owned);
}

int bcc_ragged_plan_launch(bcc_ragged_plan* plan, void* stream)
{
    if (!plan)
        return fail(BCC_ERR_INVALID, "plan is NULL");
    BCC_HIP(hipGraphLaunch(plan->exec, (hipStream_t)stream), "hipGraphLaunch");
    return BCC_OK;
}

int bcc_ragged_plan_launch_eager(bcc_ragged_plan* plan, void* stream)
{
    if (!plan)
        return fail(BCC_ERR_INVALID, "plan is NULL");
    return run_plan(plan, (hipStream_t)stream);
}

int bcc_ragged_plan_num_tiles(const bcc_ragged_plan* plan)
{
    return plan ? plan->num_tiles : -1;
}

int bcc_ragged_plan_config(const bcc_ragged_plan* plan, int* tile_pairs, int* groups_per_cu)
{
    if (!plan)
        return fail(BCC_ERR_INVALID, "plan is NULL");
    if (tile_pairs)
        *tile_pairs = plan->tile_pairs;
    if (groups_per_cu)
        *groups_per_cu = plan->groups_per_cu;
    return BCC_OK;
}

}  // extern "C"

namespace {

// Average time of `iters` replays of an instantiated graph, after `warmup`.
int time_graph(hipGraphExec_t exec, int warmup, int iters, hipStream_t s, double* avg_ms)
{
    if (!avg_ms || iters < 1 || warmup < 0)
        return fail(BCC_ERR_INVALID, "bad timing arguments");
    for (int i = 0; i < warmup; ++i)
        BCC_HIP(hipGraphLaunch(exec, s), "hipGraphLaunch");
    hipEvent_t t0, t1;
    BCC_HIP(hipEventCreate(&t0), "hipEventCreate");
    hipError_t e = hipEventCreate(&t1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(t0);
        return hip_fail(e, "hipEventCreate");
    }
    e = hipEventRecord(t0, s);
    for (int i = 0; i < iters && e == hipSuccess; ++i)
        e = hipGraphLaunch(exec, s);
    if (e == hipSuccess)
        e = hipEventRecord(t1, s);
    if (e == hipSuccess)
        e = hipEventSynchronize(t1);
    float ms = 0.0f;
    if (e == hipSuccess)
        e = hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (e != hipSuccess)
        return hip_fail(e, "graph timing");
    *avg_ms = (double)ms / iters;
    return BCC_OK;
}

}  // namespace

extern "C" {

int bcc_ragged_plan_time(bcc_ragged_plan* plan, int warmup, int iters, void* stream, double* avg_ms)
{
    if (!plan)
        return fail(BCC_ERR_INVALID, "plan is NULL");
    return time_graph(plan->exec, warmup, iters, (hipStream_t)stream, avg_ms);
}

int bcc_ragged_plan_destroy(bcc_ragged_plan* plan)
{
    bcc_ragged_plan_destroy_impl(plan);
    return BCC_OK;
}

int bcc_ragged_group_destroy(bcc_ragged_group* group)
{
    if (!group)
        return BCC_OK;
    if (group->exec)
        (void)hipGraphExecDestroy(group->exec);
    if (group->graph)
        (void)hipGraphDestroy(group->graph);
    delete group;
    return BCC_OK;
}

int bcc_ragged_group_create(bcc_ragged_plan* const* plans, int num_plans, bcc_ragged_group** out)
{
    if (!out)
        return fail(BCC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!plans || num_plans < 1)
        return fail(BCC_ERR_INVALID, "a group needs at least one plan");
    for (int i = 0; i < num_plans; ++i) {
        if (!plans[i] || !plans[i]->graph)
            return fail(BCC_ERR_INVALID, "plan " + std::to_string(i) + " is NULL");
        for (int j = 0; j < i; ++j)
            if (plans[j] == plans[i])
                return fail(BCC_ERR_INVALID, "a plan appears twice in one group (its launches would overlap)");
    }
    bcc_ragged_group* g = new bcc_ragged_group;
    hipError_t e = hipGraphCreate(&g->graph, 0);
    for (int i = 0; i < num_plans && e == hipSuccess; ++i) {
        hipGraphNode_t node;
        e = hipGraphAddChildGraphNode(&node, g->graph, nullptr, 0, plans[i]->graph);
    }
    if (e == hipSuccess)
        e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        bcc_ragged_group_destroy(g);
        return hip_fail(e, "ragged group graph");
    }
    *out = g;
    return BCC_OK;
}

int bcc_ragged_group_launch(bcc_ragged_group* group, void* stream)
{
    if (!group)
        return fail(BCC_ERR_INVALID, "group is NULL");
    BCC_HIP(hipGraphLaunch(group->exec, (hipStream_t)stream), "hipGraphLaunch");
    return BCC_OK;
}

int bcc_ragged_group_time(bcc_ragged_group* group, int warmup, int iters, void* stream, double* avg_ms)
{
    if (!group)
        return fail(BCC_ERR_INVALID, "group is NULL");
    return time_graph(group->exec, warmup, iters, (hipStream_t)stream, avg_ms);
}

int bcc_check_ragged(int robot, const bcc_scene* scene, const int64_t* env_offsets, const float* edges,
                     int resolution, uint8_t* results)
{
    if (int rc = check_dense_args(robot, scene, edges, 0, resolution, results))
        return rc;
    if (!env_offsets)
        return fail(BCC_ERR_INVALID, "env_offsets is NULL");
    const int64_t total = env_offsets[scene->num_envs];
    if (total <= 0)
        return BCC_OK;
    if (!edges || !results)
        return fail(BCC_ERR_INVALID, "edges/results buffer is NULL");
    const int dim = bcc_robot_dimension(robot);
    for (int64_t i = 0; i < total * 2 * dim; ++i)
        if (!std::isfinite(edges[i]))
            return fail(BCC_ERR_INVALID, "edge " + std::to_string(i / (2 * dim)) + " has a non-finite coordinate");
    float* d_edges = nullptr;
    uint8_t* d_out = nullptr;
    const size_t edge_bytes = sizeof(float) * 2 * dim * (size_t)total;
    hipError_t e = hipMalloc((void**)&d_edges, edge_bytes);
    if (e == hipSuccess)
        e = hipMalloc((void**)&d_out, (size_t)total);
    if (e == hipSuccess)
        e = hipMemcpy(d_edges, edges, edge_bytes, hipMemcpyHostToDevice);
    int rc = BCC_OK;
    bcc_ragged_plan* plan = nullptr;
    if (e != hipSuccess)
        rc = hip_fail(e, "ragged check setup");
    else
        rc = bcc_ragged_plan_create(robot, scene, env_offsets, d_edges, resolution, d_out, 0, &plan);
    if (rc == BCC_OK) {
        rc = bcc_ragged_plan_launch_eager(plan, nullptr);
        if (rc == BCC_OK) {
            e = hipMemcpy(results, d_out, (size_t)total, hipMemcpyDeviceToHost);
            if (e != hipSuccess)
                rc = hip_fail(e, "ragged check results");
        }
    }
    bcc_ragged_plan_destroy_impl(plan);
    (void)hipFree(d_edges);
    (void)hipFree(d_out);
    return rc;
}

}  // extern "C"

// ------------------------------------------------------------ allocation-free ragged check
// bcc_check_ragged_noalloc runs the ragged plan's persistent kernel over a CSR
// batch whose offsets are already on the device.  Its only memory beyond the
// caller's inputs and outputs is `scratch`: a header of kNoAllocHeaderInts
// ints (the persistent kernel's tile counter at [0]) followed by at most
// kNoAllocMaxTiles tiles, which a one-workgroup kernel cuts from the offsets
// on the device.  Each device holds its own copy for the life of the
// process; every call on a device reuses it, so calls that overlap on one
// device race (issue them on one stream).

constexpr int kNoAllocHeaderInts = 64;   // 256 bytes
constexpr int kNoAllocMaxTiles = 1 << 18;
constexpr int kNoAllocTileInts = (int)(sizeof(bcc::RaggedTile) / sizeof(int));
static_assert(sizeof(bcc::RaggedTile) % sizeof(int) == 0 && alignof(bcc::RaggedTile) <= 16, "tile layout");

static __device__ __align__(256) int scratch[kNoAllocHeaderInts + kNoAllocMaxTiles * kNoAllocTileInts];

namespace {

constexpr int kNoAllocBlock = 1024;
constexpr int kNoAllocMaxDevices = 64;

// Device address of `scratch` per device, taken once with that device
// current (each device's copy lives at its own address).
std::mutex g_scratch_mu;
int* g_scratch_ptr[kNoAllocMaxDevices] = {};

int scratch_on_current_device(int** out)
{
    int dev = 0;
    BCC_HIP(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        return fail(BCC_ERR_INVALID, "device ordinal " + std::to_string(dev) + " is beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        BCC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<int*>(p);
    }
    *out = g_scratch_ptr[dev];
    return BCC_OK;
}

// Cuts a CSR batch into tiles of at most `tile` pairs: one workgroup walks
// the environments kNoAllocBlock at a time; each thread counts its
// environment's tiles, a workgroup scan places them after those of the
// environments before it, and the thread writes them.  Offsets are clamped
// to [0, total] and writes to the first `cap` slots, so any offsets array
// keeps every access in bounds; slots [count, cap) get empty tiles, which the
// persistent kernel takes and skips.  Also zeroes the tile counter.
__global__ __launch_bounds__(kNoAllocBlock) void noalloc_tiles_kernel(const int64_t* __restrict__ off, int num_envs,
                                                                      int64_t total, int tile, int cap,
                                                                      bcc::RaggedTile* __restrict__ tiles,
                                                                      int* __restrict__ counter)
{
    constexpr int kWaves = kNoAllocBlock / 64;
    __shared__ int wave_total[kWaves];
    __shared__ int placed;   // tiles of the environments walked so far (capped at cap)
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    if (t == 0) {
        placed = 0;
        *counter = 0;
    }
    __syncthreads();
    for (int e0 = 0; e0 < num_envs; e0 += kNoAllocBlock) {
        const int e = e0 + t;
        int64_t begin = 0, end = 0;
        if (e < num_envs) {
            begin = min(max(off[e], (int64_t)0), total);
            end = min(max(off[e + 1], begin), total);
        }
        const int count = (int)min((end - begin + tile - 1) / tile, (int64_t)cap);
        int incl = count;   // inclusive scan over the wave
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(incl, d);
            if (lane >= d)
                incl += y;
        }
        if (lane == 63)
            wave_total[wave] = incl;
        __syncthreads();
        int before = placed;
        for (int w = 0; w < wave; ++w)
            before += wave_total[w];
        before += incl - count;
        for (int j = 0; j < count; ++j) {
            const int slot = before + j;
            if (slot >= cap)
                break;
            const int64_t b = begin + (int64_t)j * tile;
            tiles[slot] = bcc::RaggedTile{e, (int)b, (int)min(b + tile, end), 0};
        }
        __syncthreads();
        if (t == 0) {
            int sum = placed;
            for (int w = 0; w < kWaves; ++w)
                sum = min(sum + wave_total[w], cap);
            placed = sum;
        }
        __syncthreads();
    }
    for (int slot = placed + t; slot < cap; slot += kNoAllocBlock)
        tiles[slot] = bcc::RaggedTile{0, 0, 0, 0};
}

}  // namespace

extern "C" {

int bcc_check_ragged_noalloc(int robot, const bcc_scene* scene, const int64_t* d_env_offsets, const float* d_edges,
                             int64_t num_pairs, int resolution, uint8_t* d_results, void* stream)
{
    if (int rc = check_dense_args(robot, scene, d_edges, 0, resolution, d_results))
        return rc;
    if (num_pairs < 0 || num_pairs > 0x7fffffffLL)
        return fail(BCC_ERR_INVALID, "num_pairs must be in [0, 2^31 - 1]");
    if (num_pairs == 0)
        return BCC_OK;
    const int E = scene->num_envs;
    if (E < 1)
        return fail(BCC_ERR_INVALID, "pairs given for a scene without environments");
    if (!d_env_offsets || !d_edges || !d_results)
        return fail(BCC_ERR_INVALID, "d_env_offsets/d_edges/d_results is NULL");
    // tile and grid as a ragged plan picks them; the pair weights of the
    // obstacle mean are on the device, so the variant uses the scene's mean
    const int variant = choose_variant(robot, mean_obstacles(scene));
    const bcc_tune_entry& tune = tuned(robot);
    const bcc::VariantKnobs& k = knobs(robot);
    const int resident = std::min(tune.groups_per_cu, variant ? k.hi_waves : k.lo_waves);
    const int cus = device_cus();
    int64_t tile = tile_for(num_pairs, cus * resident * (bcc::kBlock / 64), tune.tiles_per_wave);
    // at most E - 1 partial tiles beyond ceil(pairs / tile): grow the tile
    // until the list fits the scratch array
    auto bound = [&](int64_t t) { return (num_pairs + t - 1) / t + E; };
    while (bound(tile) > kNoAllocMaxTiles && tile < num_pairs)
        tile *= 2;
    tile = std::min(tile, num_pairs);   // one tile per environment is the coarsest list
    if (bound(tile) > kNoAllocMaxTiles)
        return fail(BCC_ERR_INVALID, "an allocation-free ragged batch holds at most " +
                                         std::to_string(kNoAllocMaxTiles - 1) + " environments, got " +
                                         std::to_string(E));
    const int cap = (int)bound(tile);
    int* base = nullptr;
    if (int rc = scratch_on_current_device(&base))
        return rc;
    int* counter = base;
    bcc::RaggedTile* tiles = reinterpret_cast<bcc::RaggedTile*>(base + kNoAllocHeaderInts);
    const hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(noalloc_tiles_kernel, dim3(1), dim3(kNoAllocBlock), 0, s, d_env_offsets, E, num_pairs,
                       (int)tile, cap, tiles, counter);
    BCC_HIP(hipGetLastError(), "tile kernel launch");
    bcc::RaggedArgs a{scene->d_obs, scene->d_meta, d_edges, tiles, cap, counter, resolution,
                      cus * tune.groups_per_cu, d_results};
    a.variant = variant;
    return dispatch_ragged(robot, a, s);
}

int bcc_debug_dense_chunk(int edges)
{
    if (edges < 0 || edges > 4096)
        return fail(BCC_ERR_INVALID, "edges per workgroup must be in 0..4096");
    g_dense_chunk.store(edges);
    return BCC_OK;
}

int bcc_debug_cull_min(int work)
{
    g_cull_min.store(work < 0 ? -1 : work);
    return BCC_OK;
}

int bcc_debug_scene_cull_min(const bcc_scene* scene, int* out)
{
    if (!scene || !out)
        return fail(BCC_ERR_INVALID, "null argument");
    *out = scene->cull_min;
    return BCC_OK;
}

int bcc_debug_force_variant(int variant)
{
    if (variant < -1 || variant > 1)
        return fail(BCC_ERR_INVALID, "variant must be -1 (automatic), 0 or 1");
    g_force_variant.store(variant);
    return BCC_OK;
}

int bcc_malloc(void** d_ptr, size_t bytes)
{
    if (!d_ptr)
        return fail(BCC_ERR_INVALID, "d_ptr is NULL");
    BCC_HIP(hipMalloc(d_ptr, bytes ? bytes : 1), "hipMalloc");
    return BCC_OK;
}

int bcc_free(void* d_ptr)
{
    BCC_HIP(hipFree(d_ptr), "hipFree");
    return BCC_OK;
}

int bcc_memcpy_htod(void* d_dst, const void* src, size_t bytes)
{
    BCC_HIP(hipMemcpy(d_dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
    return BCC_OK;
}

int bcc_memcpy_dtoh(void* dst, const void* d_src, size_t bytes)
{
    BCC_HIP(hipMemcpy(dst, d_src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return BCC_OK;
}

int bcc_memset(void* d_dst, int value, size_t bytes)
{
    BCC_HIP(hipMemset(d_dst, value, bytes), "hipMemset");
    return BCC_OK;
}

int bcc_synchronize(void)
{
    BCC_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return BCC_OK;
}

int bcc_debug_fk(int robot, const float* d_q, int n, float* d_out)
{
    if (n < 0 || (n > 0 && (!d_q || !d_out)))
        return fail(BCC_ERR_INVALID, "bad fk arguments");
    hipError_t e;
    switch (robot) {
    case BCC_PANDA: e = bcc::launch_fk_panda(d_q, n, d_out, nullptr); break;
    case BCC_FETCH: e = bcc::launch_fk_fetch(d_q, n, d_out, nullptr); break;
    case BCC_BAXTER: e = bcc::launch_fk_baxter(d_q, n, d_out, nullptr); break;
    case BCC_XARM7: e = bcc::launch_fk_xarm7(d_q, n, d_out, nullptr); break;
    default: return fail(BCC_ERR_INVALID, "unknown robot id");
    }
    if (e != hipSuccess)
        return hip_fail(e, "fk kernel");
    BCC_HIP(hipDeviceSynchronize(), "fk kernel");
    return BCC_OK;
}

int bcc_debug_sincos(const float* d_x, int n, float* d_out)
{
    if (n < 0 || (n > 0 && (!d_x || !d_out)))
        return fail(BCC_ERR_INVALID, "bad sincos arguments");
    if (n == 0)
        return BCC_OK;
    hipLaunchKernelGGL(sincos_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, d_x, n, d_out);
    BCC_HIP(hipGetLastError(), "sincos kernel");
    BCC_HIP(hipDeviceSynchronize(), "sincos kernel");
    return BCC_OK;
}

int bcc_debug_samples(int robot, const float* d_edges, int m, int resolution, int32_t* d_n, float* d_cfg,
                      float* d_last)
{
    if (m < 0 || resolution < 1 || (m > 0 && (!d_edges || !d_n || !d_cfg || !d_last)))
        return fail(BCC_ERR_INVALID, "bad samples arguments");
    hipError_t e;
    switch (robot) {
    case BCC_PANDA: e = bcc::launch_samples_panda(d_edges, m, resolution, d_n, d_cfg, d_last, nullptr); break;
    case BCC_FETCH: e = bcc::launch_samples_fetch(d_edges, m, resolution, d_n, d_cfg, d_last, nullptr); break;
    case BCC_BAXTER: e = bcc::launch_samples_baxter(d_edges, m, resolution, d_n, d_cfg, d_last, nullptr); break;
    case BCC_XARM7: e = bcc::launch_samples_xarm7(d_edges, m, resolution, d_n, d_cfg, d_last, nullptr); break;
    default: return fail(BCC_ERR_INVALID, "unknown robot id");
    }
    if (e != hipSuccess)
        return hip_fail(e, "samples kernel");
    BCC_HIP(hipDeviceSynchronize(), "samples kernel");
    return BCC_OK;
}

int bcc_debug_copy(const void* d_src, void* d_dst, size_t bytes)
{
    if ((bytes % 16) != 0 || (bytes && (!d_src || !d_dst)))
        return fail(BCC_ERR_INVALID, "copy size must be a multiple of 16 bytes");
    const size_t n = bytes / 16;
    if (n == 0)
        return BCC_OK;
    hipLaunchKernelGGL(copy_kernel, dim3(2048), dim3(256), 0, 0, (const uint4*)d_src, (uint4*)d_dst, n);
    BCC_HIP(hipGetLastError(), "copy kernel");
    BCC_HIP(hipDeviceSynchronize(), "copy kernel");
    return BCC_OK;
}

int bcc_debug_valu(int blocks, int iters, int active_lanes, float* d_out, double* ms)
{
    if (blocks < 1 || iters < 0 || active_lanes < 0 || active_lanes > 64 || !d_out)
        return fail(BCC_ERR_INVALID, "bad VALU calibration arguments");
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (ms) {
        BCC_HIP(hipEventCreate(&t0), "hipEventCreate");
        BCC_HIP(hipEventCreate(&t1), "hipEventCreate");
        BCC_HIP(hipEventRecord(t0, 0), "hipEventRecord");
    }
    hipLaunchKernelGGL(valu_kernel, dim3(blocks), dim3(256), 0, 0, d_out, iters, active_lanes, 0.999f, 1e-3f);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && ms)
        e = hipEventRecord(t1, 0);
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    float t = 0.0f;
    if (e == hipSuccess && ms)
        e = hipEventElapsedTime(&t, t0, t1);
    if (t0)
        (void)hipEventDestroy(t0);
    if (t1)
        (void)hipEventDestroy(t1);
    BCC_HIP(e, "VALU calibration kernel");
    if (ms)
        *ms = t;
    return BCC_OK;
}

int bcc_debug_half_boxes(const float* d_in, uint64_t mask, float* d_out)
{
    if (!d_in || !d_out)
        return fail(BCC_ERR_INVALID, "bad half-box arguments");
    hipLaunchKernelGGL(half_boxes_kernel, dim3(1), dim3(64), 0, 0, d_in, mask, d_out);
    BCC_HIP(hipGetLastError(), "half-box kernel");
    BCC_HIP(hipDeviceSynchronize(), "half-box kernel");
    return BCC_OK;
}

int bcc_debug_prims(const float* d_in, int n, float* d_out)
{
    if (n < 0 || (n > 0 && (!d_in || !d_out)))
        return fail(BCC_ERR_INVALID, "bad prims arguments");
    if (n == 0)
        return BCC_OK;
    hipLaunchKernelGGL(prims_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, d_in, n, d_out);
    BCC_HIP(hipGetLastError(), "prims kernel");
    BCC_HIP(hipDeviceSynchronize(), "prims kernel");
    return BCC_OK;
}

}  // extern "C"
