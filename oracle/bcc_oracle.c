/*
 * CPU oracle for the batch edge collision checker.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or the CPU baseline -- never as the product path.
 *
 * It restates the reference's plain method with plain loops:
 *   - sample set and step loop: src/batch_cc.cu:96-132
 *     (n = max(ceil(dist/31 * resolution), 1), delta = (end-start)/(31 n),
 *      lane t starts at start + delta*(t*n) and adds delta once per step;
 *      32*n samples, lane 31 runs past `end` for steps >= 1)
 *   - environment test with +0.005 padding and the loop order spheres ->
 *     capsules -> z-capsules -> cuboids -> z-cuboids, cylinders never
 *     tested: src/utils.cuh:213-245
 *   - primitives: src/utils.cuh:60-88 (sphere, self), :115-156 (capsules),
 *     :160-210 (cuboids)
 *   - bounding-sphere gated env and self programs of fkcc<Robot>:
 *     src/utils.cuh:814-1173 and :1197-5434 (Panda; Fetch :10382-24992,
 *     Baxter :24994-38241, Xarm7 :38243-44112), as tables in
 *     csrc/robots/<robot>_model.h
 *
 * Three evaluation modes:
 *   FP32   the project's fixed FP32 arithmetic (explicit fmaf, written to
 *          match the nvcc-style contraction of the reference formulas);
 *          the HIP kernels must reproduce its verdicts bit for bit.
 *   FP64   exact-math sample positions, libm sin/cos, float64 everywhere;
 *          returns per pair the signed contact margin (metres): the minimum
 *          over samples of the gated clearance, negative iff colliding.
 *   WORK   FP32 with the reference's per-lane early exits, counting the
 *          reference's source-level FP32 operations in plain order.
 *
 * FK spec (shared with the generated device code, written independently
 * here): for joint j with parent frame (Rp, tp) and fixed transform (G, g)
 *   A[i][k] = sum_l Rp[i][l] G[l][k] over G[l][k] != 0 (first term a
 *             product, then fmaf), t[i] = tp[i] + fmaf-chain of Rp[i][l] g[l]
 *   revolute:  R[:,0] = fmaf(s, A[:,1], c*A[:,0]); R[:,1] = fmaf(-s, A[:,0], c*A[:,1]); R[:,2] = A[:,2]
 *   prismatic: R = A, t += q*A[:,2] (fmaf)
 *   sphere centre P[i] = t[i] + fmaf-chain of R[i][l] p[l] over p[l] != 0
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../batch-cc_amd/csrc/robots/bcc_model.h"
#include "../batch-cc_amd/csrc/robots/panda_model.h"
#include "../batch-cc_amd/csrc/robots/fetch_model.h"
#include "../batch-cc_amd/csrc/robots/baxter_model.h"
#include "../batch-cc_amd/csrc/robots/xarm7_model.h"

#define ORC_MAX_DIM 16
#define ORC_MAX_SPHERES 256

/* ------------------------------------------------------------ scene */

/* Flat host scene: per env e, counts[6e + k] for k = spheres, capsules,
 * z-capsules, cylinders, cuboids, z-cuboids; the env's floats start at
 * data + env_off[e] in that order with 4, 8, 8, 8, 15, 15 floats each.
 * Field order follows src/collision/shapes.hh:30-245. */
enum { K_SPH = 0, K_CAP, K_ZCAP, K_CYL, K_CUB, K_ZCUB, K_NUM };
static const int k_stride[K_NUM] = {4, 8, 8, 8, 15, 15};

typedef struct {
    const float *p[K_NUM];
    int n[K_NUM];
} orc_env;

static orc_env env_view(const float *data, const int64_t *env_off, const int32_t *counts, int e)
{
    orc_env v;
    const float *base = data + env_off[e];
    for (int k = 0; k < K_NUM; ++k) {
        v.n[k] = counts[6 * e + k];
        v.p[k] = base;
        base += (size_t)v.n[k] * k_stride[k];
    }
    return v;
}

const bcc_robot_model *orc_model(int robot)
{
    switch (robot) {
    case 0: return &bcc_panda_model;
    case 1: return &bcc_fetch_model;
    case 2: return &bcc_baxter_model;
    case 3: return &bcc_xarm7_model;
    default: return NULL;
    }
}

int orc_robot_dim(int robot)
{
    const bcc_robot_model *m = orc_model(robot);
    return m ? m->dim : -1;
}

int orc_robot_num_spheres(int robot)
{
    const bcc_robot_model *m = orc_model(robot);
    return m ? m->num_spheres : -1;
}

/* ------------------------------------------------------------ sin/cos */

/* Cody-Waite reduction by pi/2 with a three-float constant, then Taylor
 * polynomials of degree 9 (sin) and 10 (cos) on [-pi/4, pi/4]; truncation
 * error < 3e-9 relative there.  Constants are pi/2 split into float parts
 * and the Taylor coefficients 1/k! rounded to float. */
static const float ORC_2_PI = 0x1.45f306p-1f;     /* 2/pi */
static const float ORC_P1 = 0x1.921fb6p+0f;       /* float(pi/2) */
static const float ORC_P2 = -0x1.777a5cp-25f;     /* float(pi/2 - P1) */
static const float ORC_P3 = -0x1.ee59dap-50f;     /* float(pi/2 - P1 - P2) */

void orc_sincosf(float x, float *s, float *c)
{
    const float kf = rintf(x * ORC_2_PI);
    float r = fmaf(-kf, ORC_P1, x);
    r = fmaf(-kf, ORC_P2, r);
    r = fmaf(-kf, ORC_P3, r);
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
    switch (((int)kf) & 3) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
    }
}

/* ------------------------------------------------------------ sample set */

/* src/batch_cc.cu:103-111; the sum of squares is accumulated with fmaf as
 * nvcc contracts `ans += diff * diff` (src/utils.cuh:102-112). */
int orc_steps(const float *s, const float *e, int dim, int resolution, float *delta)
{
    float acc = 0.0f;
    for (int j = 0; j < dim; ++j) {
        const float diff = s[j] - e[j];
        acc = fmaf(diff, diff, acc);
    }
    const float dist = sqrtf(acc);
    /* capped at 2^22 steps (the GPU kernels apply the same guard against
     * non-finite input; no finite edge within joint limits reaches it) */
    const float nf = fminf(fmaxf(ceilf((dist / 31.0f) * (float)resolution), 1.0f), 4194304.0f);
    const int n = (int)nf;
    if (delta) {
        const float div = (float)(31 * n);
        for (int j = 0; j < dim; ++j)
            delta[j] = (e[j] - s[j]) / div;
    }
    return n;
}

/* Sample (t, i): start + delta*(t*n) (fmaf), then i sequential additions
 * (src/batch_cc.cu:114-128). */
void orc_sample(const float *s, const float *delta, int n, int t, int i, int dim, float *out)
{
    for (int j = 0; j < dim; ++j) {
        float c = fmaf(delta[j], (float)(t * n), s[j]);
        for (int k = 0; k < i; ++k)
            c = c + delta[j];
        out[j] = c;
    }
}

/* ------------------------------------------------------------ FK (FP32) */

typedef struct {
    float R[ORC_MAX_DIM + 1][9];
    float t[ORC_MAX_DIM + 1][3];
} orc_frames;  /* index f+1, f = -1 is the base */

static void fk32(const bcc_robot_model *m, const float *q, orc_frames *F)
{
    for (int a = 0; a < 9; ++a)
        F->R[0][a] = (a % 4 == 0) ? 1.0f : 0.0f;
    F->t[0][0] = F->t[0][1] = F->t[0][2] = 0.0f;
    for (int oi = 0; oi < m->dim; ++oi) {
        const int j = m->joint_order[oi];
        const float *Rp = F->R[m->joint_parent[j] + 1];
        const float *tp = F->t[m->joint_parent[j] + 1];
        const float *G = m->joint_R[j];
        const float *g = m->joint_t[j];
        float A[9], t[3];
        for (int i = 0; i < 3; ++i) {
            for (int k = 0; k < 3; ++k) {
                int first = 1;
                float acc = 0.0f;
                for (int l = 0; l < 3; ++l) {
                    const float gv = G[3 * l + k];
                    if (gv == 0.0f)
                        continue;
                    acc = first ? Rp[3 * i + l] * gv : fmaf(Rp[3 * i + l], gv, acc);
                    first = 0;
                }
                A[3 * i + k] = acc;
            }
            float acc = tp[i];
            for (int l = 0; l < 3; ++l) {
                if (g[l] == 0.0f)
                    continue;
                acc = fmaf(Rp[3 * i + l], g[l], acc);
            }
            t[i] = acc;
        }
        float *R = F->R[j + 1];
        float *T = F->t[j + 1];
        if (m->joint_type[j] == 0) {
            float sn, cs;
            orc_sincosf(q[j], &sn, &cs);
            for (int i = 0; i < 3; ++i) {
                R[3 * i + 0] = fmaf(sn, A[3 * i + 1], cs * A[3 * i + 0]);
                R[3 * i + 1] = fmaf(-sn, A[3 * i + 0], cs * A[3 * i + 1]);
                R[3 * i + 2] = A[3 * i + 2];
                T[i] = t[i];
            }
        } else {
            for (int i = 0; i < 3; ++i) {
                R[3 * i + 0] = A[3 * i + 0];
                R[3 * i + 1] = A[3 * i + 1];
                R[3 * i + 2] = A[3 * i + 2];
                T[i] = fmaf(q[j], A[3 * i + 2], t[i]);
            }
        }
    }
}

static void centre32(const bcc_robot_model *m, const orc_frames *F, int sidx, float *P)
{
    const int f = m->sphere_frame[sidx] + 1;
    const float *p = m->sphere_p[sidx];
    for (int i = 0; i < 3; ++i) {
        float acc = F->t[f][i];
        for (int l = 0; l < 3; ++l) {
            if (p[l] == 0.0f)
                continue;
            acc = fmaf(F->R[f][3 * i + l], p[l], acc);
        }
        P[i] = acc;
    }
}

/* All sphere centres + radii, [num_spheres][4] (FP32 spec). */
int orc_fk_spheres(int robot, const float *q, float *out)
{
    const bcc_robot_model *m = orc_model(robot);
    if (!m)
        return -1;
    orc_frames F;
    fk32(m, q, &F);
    for (int s = 0; s < m->num_spheres; ++s) {
        centre32(m, &F, s, out + 4 * s);
        out[4 * s + 3] = m->sphere_r[s];
    }
    return 0;
}

/* ------------------------------------------------------------ primitives (FP32) */

/* env radius padding: float(double(r) + 0.005), src/utils.cuh:215 */
float orc_pad(float r)
{
    return (float)((double)r + 0.005);
}

/* src/utils.cuh:60-73 with the obstacle as `a` */
float orc_sphere_val(const float *o, float x, float y, float z, float rp)
{
    const float xs = o[0] - x, ys = o[1] - y, zs = o[2] - z;
    const float sum = fmaf(zs, zs, fmaf(ys, ys, xs * xs));
    const float rs = o[3] + rp;
    return fmaf(-rs, rs, sum);
}

/* src/utils.cuh:115-128 */
float orc_capsule_val(const float *c, float x, float y, float z, float rp)
{
    const float dot = fmaf(z - c[2], c[5], fmaf(y - c[1], c[4], (x - c[0]) * c[3]));
    const float cdf = fmaxf(fminf(dot * c[7], 1.0f), 0.0f);
    const float xs = x - fmaf(c[3], cdf, c[0]);
    const float ys = y - fmaf(c[4], cdf, c[1]);
    const float zs = z - fmaf(c[5], cdf, c[2]);
    const float sum = fmaf(zs, zs, fmaf(ys, ys, xs * xs));
    const float rs = rp + c[6];
    return fmaf(-rs, rs, sum);
}

/* src/utils.cuh:137-150 */
float orc_zcapsule_val(const float *c, float x, float y, float z, float rp)
{
    const float dot = (z - c[2]) * c[5];
    const float cdf = fmaxf(fminf(dot * c[7], 1.0f), 0.0f);
    const float xs = x - c[0];
    const float ys = y - c[1];
    const float zs = z - fmaf(c[5], cdf, c[2]);
    const float sum = fmaf(zs, zs, fmaf(ys, ys, xs * xs));
    const float rs = rp + c[6];
    return fmaf(-rs, rs, sum);
}

/* src/utils.cuh:160-177; fields cx cy cz a1x a1y a1z a2x a2y a2z a3x a3y a3z h1 h2 h3 */
float orc_cuboid_val(const float *b, float x, float y, float z, float rsq)
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

/* src/utils.cuh:186-203 */
float orc_zcuboid_val(const float *b, float x, float y, float z, float rsq)
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

/* src/utils.cuh:85-88 (unpadded) */
float orc_self_val(const float *a, const float *b)
{
    const float xs = a[0] - b[0], ys = a[1] - b[1], zs = a[2] - b[2];
    const float sum = fmaf(zs, zs, fmaf(ys, ys, xs * xs));
    const float rs = a[3] + b[3];
    return fmaf(-rs, rs, sum);
}

/* Reference operation weights (source-level FP32 ops of each test). */
enum { W_CALL = 2, W_SPH = 11, W_CAP = 28, W_ZCAP = 18, W_CUB = 33, W_ZCUB = 24, W_SELF = 11 };

/* src/utils.cuh:213-245: sequential loops, first hit ends the test.
 * `work`, when non-NULL, accumulates the reference operation count. */
static int env_hit32(const orc_env *E, float x, float y, float z, float r, uint64_t *work)
{
    const float rp = orc_pad(r);
    const float rsq = rp * rp;
    uint64_t w = W_CALL;
    int hit = 0;
    for (int i = 0; i < E->n[K_SPH] && !hit; ++i, w += W_SPH)
        hit = orc_sphere_val(E->p[K_SPH] + 4 * i, x, y, z, rp) < 0.0f;
    for (int i = 0; i < E->n[K_CAP] && !hit; ++i, w += W_CAP)
        hit = orc_capsule_val(E->p[K_CAP] + 8 * i, x, y, z, rp) < 0.0f;
    for (int i = 0; i < E->n[K_ZCAP] && !hit; ++i, w += W_ZCAP)
        hit = orc_zcapsule_val(E->p[K_ZCAP] + 8 * i, x, y, z, rp) < 0.0f;
    for (int i = 0; i < E->n[K_CUB] && !hit; ++i, w += W_CUB)
        hit = orc_cuboid_val(E->p[K_CUB] + 15 * i, x, y, z, rsq) < 0.0f;
    for (int i = 0; i < E->n[K_ZCUB] && !hit; ++i, w += W_ZCUB)
        hit = orc_zcuboid_val(E->p[K_ZCUB] + 15 * i, x, y, z, rsq) < 0.0f;
    if (work)
        *work += w;
    return hit;
}

/* One sample, reference program order with early return (fkcc<Robot>).
 * Returns 1 if the sample collides.  work/trig accumulate reference ops. */
static int sample_collides32(const bcc_robot_model *m, const orc_env *E, const float *q,
                             uint64_t *work, uint64_t *trig)
{
    orc_frames F;
    fk32(m, q, &F);
    float P[ORC_MAX_SPHERES][4];
    unsigned char have[ORC_MAX_SPHERES];
    memset(have, 0, (size_t)m->num_spheres);
#define SPH(s) (have[s] ? P[s] : (centre32(m, &F, (s), P[s]), P[s][3] = m->sphere_r[s], have[s] = 1, P[s]))
    for (int it = 0; it < m->num_program; ++it) {
        const int g = m->prog_index[it];
        int collided = 0;
        if (m->prog_kind[it] == 0) {
            const float *c = SPH(m->env_gate[g]);
            if (env_hit32(E, c[0], c[1], c[2], c[3], work)) {
                if (m->env_lone[g]) {
                    collided = 1;
                } else {
                    for (int f = m->env_fine_off[g]; f < m->env_fine_off[g + 1] && !collided; ++f) {
                        const float *cf = SPH(m->env_fine[f]);
                        collided = env_hit32(E, cf[0], cf[1], cf[2], cf[3], work);
                    }
                }
            }
        } else {
            const float *a = SPH(m->self_gate_a[g]);
            const float *b = SPH(m->self_gate_b[g]);
            if (work)
                *work += W_SELF;
            if (orc_self_val(a, b) < 0.0f) {
                if (m->self_lone[g]) {
                    collided = 1;
                } else {
                    for (int ia = m->self_a_off[g]; ia < m->self_a_off[g + 1] && !collided; ++ia) {
                        for (int ib = m->self_b_off[g]; ib < m->self_b_off[g + 1] && !collided; ++ib) {
                            const float *fa = SPH(m->self_a[ia]);
                            const float *fb = SPH(m->self_b[ib]);
                            if (work)
                                *work += W_SELF;
                            collided = orc_self_val(fa, fb) < 0.0f;
                        }
                    }
                }
            }
        }
        if (collided) {
            if (work)
                *work += (uint64_t)m->prog_fk_ops[it];
            if (trig)
                *trig += (uint64_t)m->prog_fk_trig[it];
            return 1;
        }
    }
#undef SPH
    if (work)
        *work += (uint64_t)m->ref_fk_flops;
    if (trig)
        *trig += (uint64_t)m->ref_fk_trig;
    return 0;
}

/* Verdict of one sample configuration in one env (FP32 spec). */
int orc_config_collides(int robot, const float *data, const int64_t *env_off, const int32_t *counts,
                        int env, const float *q)
{
    const bcc_robot_model *m = orc_model(robot);
    if (!m)
        return -1;
    orc_env E = env_view(data, env_off, counts, env);
    return sample_collides32(m, &E, q, NULL, NULL);
}

/* One (env, edge) pair in plain order: steps 0..n-1, 32 lanes each, stop
 * after the first step with any colliding lane (src/batch_cc.cu:118-129). */
static int pair32(const bcc_robot_model *m, const orc_env *E, const float *s, const float *e, int resolution,
                  uint64_t *work, uint64_t *trig)
{
    float delta[ORC_MAX_DIM], cfg[32][ORC_MAX_DIM];
    const int dim = m->dim;
    const int n = orc_steps(s, e, dim, resolution, delta);
    for (int t = 0; t < 32; ++t)
        for (int j = 0; j < dim; ++j)
            cfg[t][j] = fmaf(delta[j], (float)(t * n), s[j]);
    for (int i = 0; i < n; ++i) {
        int any = 0;
        for (int t = 0; t < 32; ++t) {
            if (work) {
                any |= sample_collides32(m, E, cfg[t], work, trig);
            } else if (sample_collides32(m, E, cfg[t], NULL, NULL)) {
                return 1;
            }
        }
        if (any)
            return 1;
        for (int t = 0; t < 32; ++t)
            for (int j = 0; j < dim; ++j)
                cfg[t][j] = cfg[t][j] + delta[j];
    }
    return 0;
}

/* ------------------------------------------------------------ threads */

typedef struct {
    const bcc_robot_model *m;
    const float *data;
    const int64_t *env_off;
    const int32_t *counts;
    int E, M, resolution, mode;
    const float *edges;
    const int32_t *pair_env;   /* optional explicit pair list */
    const int32_t *pair_edge;
    int64_t num_pairs;
    uint8_t *out;
    double *margin;
    uint64_t *work;            /* per pair (WORK mode) */
    uint64_t *trig;
    int64_t next;
    pthread_mutex_t lock;
} orc_job;

static int pair64(const bcc_robot_model *m, const orc_env *E, const float *s, const float *e, int resolution,
                  double *margin);

static void run_pair(orc_job *J, int64_t p)
{
    int env, edge;
    if (J->pair_env) {
        env = J->pair_env[p];
        edge = J->pair_edge[p];
    } else {
        edge = (int)(p / J->E);
        env = (int)(p % J->E);
    }
    const int dim = J->m->dim;
    const float *s = J->edges + (size_t)edge * 2 * dim;
    const float *e = s + dim;
    orc_env E = env_view(J->data, J->env_off, J->counts, env);
    if (J->mode == 0) {
        J->out[p] = (uint8_t)pair32(J->m, &E, s, e, J->resolution, NULL, NULL);
    } else if (J->mode == 1) {
        double mg = 0.0;
        J->out[p] = (uint8_t)pair64(J->m, &E, s, e, J->resolution, &mg);
        if (J->margin)
            J->margin[p] = mg;
    } else {
        uint64_t w = 0, tr = 0;
        J->out[p] = (uint8_t)pair32(J->m, &E, s, e, J->resolution, &w, &tr);
        J->work[p] = w;
        J->trig[p] = tr;
    }
}

static void *worker(void *arg)
{
    orc_job *J = (orc_job *)arg;
    const int64_t chunk = 64;
    for (;;) {
        pthread_mutex_lock(&J->lock);
        const int64_t b = J->next;
        J->next += chunk;
        pthread_mutex_unlock(&J->lock);
        if (b >= J->num_pairs)
            break;
        const int64_t end = b + chunk < J->num_pairs ? b + chunk : J->num_pairs;
        for (int64_t p = b; p < end; ++p)
            run_pair(J, p);
    }
    return NULL;
}

static int run_job(orc_job *J, int nthreads)
{
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > 256)
        nthreads = 256;
    J->next = 0;
    pthread_mutex_init(&J->lock, NULL);
    pthread_t th[256];
    int started = 0;
    for (int i = 1; i < nthreads; ++i) {
        if (pthread_create(&th[i], NULL, worker, J) != 0)
            break;
        ++started;
    }
    worker(J);
    for (int i = 1; i <= started; ++i)
        pthread_join(th[i], NULL);
    pthread_mutex_destroy(&J->lock);
    return 0;
}

static int job_init(orc_job *J, int robot, const float *data, const int64_t *env_off, const int32_t *counts,
                    int E, const float *edges, int M, int resolution)
{
    memset(J, 0, sizeof(*J));
    J->m = orc_model(robot);
    if (!J->m || E < 0 || M < 0 || resolution < 1)
        return -1;
    J->data = data;
    J->env_off = env_off;
    J->counts = counts;
    J->E = E;
    J->M = M;
    J->resolution = resolution;
    J->edges = edges;
    J->num_pairs = (int64_t)E * M;
    return 0;
}

/* Dense verdicts, out[edge*E + env] (src/batch_cc.cu:131), FP32 spec. */
int orc_check_dense(int robot, const float *data, const int64_t *env_off, const int32_t *counts, int E,
                    const float *edges, int M, int resolution, uint8_t *out, int nthreads)
{
    orc_job J;
    if (job_init(&J, robot, data, env_off, counts, E, edges, M, resolution))
        return -1;
    J.mode = 0;
    J.out = out;
    return run_job(&J, nthreads);
}

/* Explicit pair list (ragged batches): out[p] for (pair_env[p], pair_edge[p]). */
int orc_check_pairs(int robot, const float *data, const int64_t *env_off, const int32_t *counts, int E,
                    const float *edges, int M, int resolution, const int32_t *pair_env, const int32_t *pair_edge,
                    int64_t num_pairs, int mode, uint8_t *out, double *margin, uint64_t *work, uint64_t *trig,
                    int nthreads)
{
    orc_job J;
    if (job_init(&J, robot, data, env_off, counts, E, edges, M, resolution))
        return -1;
    J.pair_env = pair_env;
    J.pair_edge = pair_edge;
    J.num_pairs = num_pairs;
    J.mode = mode;
    J.out = out;
    J.margin = margin;
    J.work = work;
    J.trig = trig;
    if (mode == 2 && (!work || !trig))
        return -1;
    return run_job(&J, nthreads);
}

/* ------------------------------------------------------------ FP64 mode */

typedef struct {
    double R[ORC_MAX_DIM + 1][9];
    double t[ORC_MAX_DIM + 1][3];
} orc_frames64;

static void fk64(const bcc_robot_model *m, const double *q, orc_frames64 *F)
{
    for (int a = 0; a < 9; ++a)
        F->R[0][a] = (a % 4 == 0) ? 1.0 : 0.0;
    F->t[0][0] = F->t[0][1] = F->t[0][2] = 0.0;
    for (int oi = 0; oi < m->dim; ++oi) {
        const int j = m->joint_order[oi];
        const double *Rp = F->R[m->joint_parent[j] + 1];
        const double *tp = F->t[m->joint_parent[j] + 1];
        const double *G = m->joint_R64[j];
        const double *g = m->joint_t64[j];
        double A[9], t[3];
        for (int i = 0; i < 3; ++i) {
            for (int k = 0; k < 3; ++k)
                A[3 * i + k] = Rp[3 * i] * G[k] + Rp[3 * i + 1] * G[3 + k] + Rp[3 * i + 2] * G[6 + k];
            t[i] = tp[i] + Rp[3 * i] * g[0] + Rp[3 * i + 1] * g[1] + Rp[3 * i + 2] * g[2];
        }
        double *R = F->R[j + 1];
        double *T = F->t[j + 1];
        if (m->joint_type[j] == 0) {
            const double sn = sin(q[j]), cs = cos(q[j]);
            for (int i = 0; i < 3; ++i) {
                R[3 * i + 0] = cs * A[3 * i + 0] + sn * A[3 * i + 1];
                R[3 * i + 1] = cs * A[3 * i + 1] - sn * A[3 * i + 0];
                R[3 * i + 2] = A[3 * i + 2];
                T[i] = t[i];
            }
        } else {
            for (int i = 0; i < 3; ++i) {
                for (int k = 0; k < 3; ++k)
                    R[3 * i + k] = A[3 * i + k];
                T[i] = t[i] + q[j] * A[3 * i + 2];
            }
        }
    }
}

static void centre64(const bcc_robot_model *m, const orc_frames64 *F, int sidx, double *P)
{
    const int f = m->sphere_frame[sidx] + 1;
    const double *p = m->sphere_p64[sidx];
    for (int i = 0; i < 3; ++i)
        P[i] = F->t[f][i] + F->R[f][3 * i] * p[0] + F->R[f][3 * i + 1] * p[1] + F->R[f][3 * i + 2] * p[2];
}

int orc_fk_spheres64(int robot, const double *q, double *out)
{
    const bcc_robot_model *m = orc_model(robot);
    if (!m)
        return -1;
    orc_frames64 F;
    fk64(m, q, &F);
    for (int s = 0; s < m->num_spheres; ++s) {
        centre64(m, &F, s, out + 4 * s);
        out[4 * s + 3] = (double)m->sphere_r[s];
    }
    return 0;
}

static double clampd(double v, double lo, double hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

/* Signed distances (metres), negative = penetration. */
double orc_sphere_dist64(const float *o, double x, double y, double z, double rp)
{
    const double dx = o[0] - x, dy = o[1] - y, dz = o[2] - z;
    return sqrt(dx * dx + dy * dy + dz * dz) - (rp + (double)o[3]);
}

double orc_capsule_dist64(const float *c, double x, double y, double z, double rp, int z_aligned)
{
    double dot;
    if (z_aligned)
        dot = (z - c[2]) * (double)c[5];
    else
        dot = (x - c[0]) * (double)c[3] + (y - c[1]) * (double)c[4] + (z - c[2]) * (double)c[5];
    const double t = clampd(dot * (double)c[7], 0.0, 1.0);
    const double px = z_aligned ? c[0] : c[0] + (double)c[3] * t;
    const double py = z_aligned ? c[1] : c[1] + (double)c[4] * t;
    const double pz = c[2] + (double)c[5] * t;
    const double dx = x - px, dy = y - py, dz = z - pz;
    return sqrt(dx * dx + dy * dy + dz * dz) - (rp + (double)c[6]);
}

double orc_cuboid_dist64(const float *b, double x, double y, double z, double rp, int z_aligned)
{
    const double xs = x - b[0], ys = y - b[1], zs = z - b[2];
    double d1, d2, d3;
    if (z_aligned) {
        d1 = b[3] * xs + b[4] * ys;
        d2 = b[6] * xs + b[7] * ys;
        d3 = zs;
    } else {
        d1 = b[3] * xs + b[4] * ys + b[5] * zs;
        d2 = b[6] * xs + b[7] * ys + b[8] * zs;
        d3 = b[9] * xs + b[10] * ys + b[11] * zs;
    }
    const double a1 = fmax(0.0, fabs(d1) - b[12]);
    const double a2 = fmax(0.0, fabs(d2) - b[13]);
    const double a3 = fmax(0.0, fabs(d3) - b[14]);
    return sqrt(a1 * a1 + a2 * a2 + a3 * a3) - rp;
}

static double env_clear64(const orc_env *E, const double *c, float r)
{
    const double rp = (double)r + 0.005;
    double best = INFINITY, d;
    for (int i = 0; i < E->n[K_SPH]; ++i)
        if ((d = orc_sphere_dist64(E->p[K_SPH] + 4 * i, c[0], c[1], c[2], rp)) < best) best = d;
    for (int i = 0; i < E->n[K_CAP]; ++i)
        if ((d = orc_capsule_dist64(E->p[K_CAP] + 8 * i, c[0], c[1], c[2], rp, 0)) < best) best = d;
    for (int i = 0; i < E->n[K_ZCAP]; ++i)
        if ((d = orc_capsule_dist64(E->p[K_ZCAP] + 8 * i, c[0], c[1], c[2], rp, 1)) < best) best = d;
    for (int i = 0; i < E->n[K_CUB]; ++i)
        if ((d = orc_cuboid_dist64(E->p[K_CUB] + 15 * i, c[0], c[1], c[2], rp, 0)) < best) best = d;
    for (int i = 0; i < E->n[K_ZCUB]; ++i)
        if ((d = orc_cuboid_dist64(E->p[K_ZCUB] + 15 * i, c[0], c[1], c[2], rp, 1)) < best) best = d;
    return best;
}

static double self_clear64(const double *a, float ra, const double *b, float rb)
{
    const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return sqrt(dx * dx + dy * dy + dz * dz) - ((double)ra + (double)rb);
}

/* Gated clearance of one sample: min over program items of
 * (lone: gate) or max(gate, min fine); collides iff < 0. */
static double sample_clear64(const bcc_robot_model *m, const orc_env *E, const double *q)
{
    orc_frames64 F;
    fk64(m, q, &F);
    double P[ORC_MAX_SPHERES][3];
    for (int s = 0; s < m->num_spheres; ++s)
        centre64(m, &F, s, P[s]);
    double best = INFINITY;
    for (int g = 0; g < m->num_env_groups; ++g) {
        const double cg = env_clear64(E, P[m->env_gate[g]], m->sphere_r[m->env_gate[g]]);
        if (m->env_lone[g]) {
            if (cg < best) best = cg;
            continue;
        }
        if (cg >= best)
            continue;
        double fmin = INFINITY;
        for (int f = m->env_fine_off[g]; f < m->env_fine_off[g + 1]; ++f) {
            const double cf = env_clear64(E, P[m->env_fine[f]], m->sphere_r[m->env_fine[f]]);
            if (cf < fmin) fmin = cf;
        }
        const double v = cg > fmin ? cg : fmin;
        if (v < best) best = v;
    }
    for (int g = 0; g < m->num_self_groups; ++g) {
        const int a = m->self_gate_a[g], b = m->self_gate_b[g];
        const double cg = self_clear64(P[a], m->sphere_r[a], P[b], m->sphere_r[b]);
        if (m->self_lone[g]) {
            if (cg < best) best = cg;
            continue;
        }
        if (cg >= best)
            continue;
        double fmin = INFINITY;
        for (int ia = m->self_a_off[g]; ia < m->self_a_off[g + 1]; ++ia)
            for (int ib = m->self_b_off[g]; ib < m->self_b_off[g + 1]; ++ib) {
                const int x = m->self_a[ia], y = m->self_b[ib];
                const double c = self_clear64(P[x], m->sphere_r[x], P[y], m->sphere_r[y]);
                if (c < fmin) fmin = c;
            }
        const double v = cg > fmin ? cg : fmin;
        if (v < best) best = v;
    }
    return best;
}

/* FP64: same n (from the FP32 rule), exact sample positions
 * start + (end-start) * (t*n + i) / (31 n). */
static int pair64(const bcc_robot_model *m, const orc_env *E, const float *s, const float *e, int resolution,
                  double *margin)
{
    const int dim = m->dim;
    const int n = orc_steps(s, e, dim, resolution, NULL);
    double best = INFINITY;
    double q[ORC_MAX_DIM];
    for (int k = 0; k < 32 * n; ++k) {
        const double u = (double)k / (31.0 * n);
        for (int j = 0; j < dim; ++j)
            q[j] = (double)s[j] + ((double)e[j] - (double)s[j]) * u;
        const double c = sample_clear64(m, E, q);
        if (c < best) best = c;
    }
    *margin = best;
    return best < 0.0;
}

/* Dense FP64 verdicts and margins, out[edge*E + env]. */
int orc_check_dense64(int robot, const float *data, const int64_t *env_off, const int32_t *counts, int E,
                      const float *edges, int M, int resolution, uint8_t *out, double *margin, int nthreads)
{
    orc_job J;
    if (job_init(&J, robot, data, env_off, counts, E, edges, M, resolution))
        return -1;
    J.mode = 1;
    J.out = out;
    J.margin = margin;
    return run_job(&J, nthreads);
}

/* Dense reference-order work counts, per pair. */
int orc_work_dense(int robot, const float *data, const int64_t *env_off, const int32_t *counts, int E,
                   const float *edges, int M, int resolution, uint8_t *out, uint64_t *work, uint64_t *trig,
                   int nthreads)
{
    orc_job J;
    if (job_init(&J, robot, data, env_off, counts, E, edges, M, resolution))
        return -1;
    J.mode = 2;
    J.out = out;
    J.work = work;
    J.trig = trig;
    return run_job(&J, nthreads);
}

/* Joint frames (float64): R [dim][9] row-major, t [dim][3], world frame. */
int orc_fk_frames64(int robot, const double *q, double *R, double *t)
{
    const bcc_robot_model *m = orc_model(robot);
    if (!m)
        return -1;
    orc_frames64 F;
    fk64(m, q, &F);
    for (int j = 0; j < m->dim; ++j) {
        memcpy(R + 9 * j, F.R[j + 1], sizeof(double) * 9);
        memcpy(t + 3 * j, F.t[j + 1], sizeof(double) * 3);
    }
    return 0;
}
