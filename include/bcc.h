/*
 * bcc.h -- C ABI of the MI355X (gfx950) batch edge collision checker.
 *
 * This is the drop-in boundary for the reference's entry point
 *   template <typename Robot>
 *   void batch_cc::batch_cc(std::vector<Environment<float>>& h_envs,
 *                           std::vector<std::array<Robot::Configuration, 2>>& edges,
 *                           int resolution, std::vector<bool>& results);
 * declared at src/batch_cc.hh:4-7 and defined at src/batch_cc.cu:244-338.
 * The header-only C++ shim include/batch_cc.hh implements exactly that
 * signature on top of these functions; other languages bind these directly.
 *
 * Conventions
 *   - Every function returns an int status: BCC_OK (0) or a BCC_ERR_* code;
 *     bcc_last_error() returns a thread-local message for the last failure.
 *     Nothing here prints or exits (the reference prints timings and calls
 *     exit() on CUDA errors, src/utils.cuh:252-258).
 *   - Robots: BCC_PANDA (7 DoF), BCC_FETCH (8), BCC_BAXTER (14), BCC_XARM7 (13:
 *     seven arm joints and the six gripper joints its generated model reads).
 *   - Edges: float [num_edges][2][dim] = (start, end) joint configurations
 *     (src/batch_cc.cu:292-295).
 *   - Results: uint8 [num_edges][num_envs], 1 = the edge collides in that
 *     environment, edge-major as the reference writes them
 *     (src/batch_cc.cu:131, :319-321).
 *   - Every call uses the current HIP device.  Async entry points take a
 *     hipStream_t passed as void* (NULL = default stream).
 */
#ifndef BCC_H
#define BCC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    BCC_OK = 0,
    BCC_ERR_INVALID = 1,      /* bad argument (reference: std::runtime_error, src/batch_cc.cu:45-56) */
    BCC_ERR_HIP = 2,          /* HIP runtime error */
    BCC_ERR_NOMEM = 3,
    BCC_ERR_UNSUPPORTED = 4
};

enum { BCC_PANDA = 0, BCC_FETCH = 1, BCC_BAXTER = 2, BCC_XARM7 = 3 };

typedef struct bcc_scene bcc_scene;

/*
 * One environment as flat float arrays: the numeric fields of
 * ppln::collision::Environment<float> (src/collision/environment.hh:11-35)
 * in the field order of src/collision/shapes.hh:
 *   spheres    [n][4]   x y z r                              (shapes.hh:225-245)
 *   capsules   [n][8]   x1 y1 z1 xv yv zv r rdv              (shapes.hh:131-222)
 *   z_capsules [n][8]   same layout, axis along z
 *   cylinders  [n][8]   same layout; accepted and stored, never tested, as in
 *                       the reference (src/utils.cuh:213-245)
 *   cuboids    [n][15]  x y z a1x a1y a1z a2x a2y a2z a3x a3y a3z r1 r2 r3  (shapes.hh:30-126)
 *   z_cuboids  [n][15]  same layout, axis 3 = z
 * A positive count with a NULL pointer is rejected (src/batch_cc.cu:45-56).
 */
typedef struct bcc_env_desc {
    const float *spheres;
    uint32_t num_spheres;
    const float *capsules;
    uint32_t num_capsules;
    const float *z_capsules;
    uint32_t num_z_capsules;
    const float *cylinders;
    uint32_t num_cylinders;
    const float *cuboids;
    uint32_t num_cuboids;
    const float *z_cuboids;
    uint32_t num_z_cuboids;
} bcc_env_desc;

/* Thread-local text of the last error (never NULL). */
const char *bcc_last_error(void);

/* Degrees of freedom of a robot, or -1 (src/Robots.hh:13,69,125). */
int bcc_robot_dimension(int robot);

/* Number of visible HIP devices. */
int bcc_device_count(int *count);

/*
 * Upload environments once; the handle keeps a device-resident copy
 * (replaces the per-call setup_environment_on_device, src/batch_cc.cu:135-235).
 */
int bcc_scene_create(const bcc_env_desc *envs, int num_envs, bcc_scene **out);

/*
 * Same from one packed host buffer: env e holds counts[6e+k] primitives of
 * kind k = spheres, capsules, z-capsules, cylinders, cuboids, z-cuboids,
 * stored from data + env_offsets[e] in that order with 4, 8, 8, 8, 15, 15
 * floats each.
 */
int bcc_scene_create_packed(const float *data, const int64_t *env_offsets, const int32_t *counts, int num_envs,
                            bcc_scene **out);

int bcc_scene_destroy(bcc_scene *scene);
int bcc_scene_num_envs(const bcc_scene *scene);

/*
 * Plain dense check, synchronous, host buffers: every edge against every
 * environment (src/batch_cc.cu:244-338 semantics; results[edge*E + env]).
 */
int bcc_check_dense(int robot, const bcc_scene *scene, const float *edges, int num_edges, int resolution,
                    uint8_t *results);

/* Plain dense check on device buffers, enqueued on `stream`. */
int bcc_check_dense_async(int robot, const bcc_scene *scene, const float *d_edges, int num_edges, int resolution,
                          uint8_t *d_results, void *stream);

/*
 * Endpoint-inclusive dense check on device buffers (the reference
 * validator's vamp semantics, scripts/batch_cc_validate.cpp:627-646: an edge
 * is valid when the motions s -> e, s -> s and e -> e all validate):
 * d_results = plain verdict of (s, e) OR plain verdict of (e, e).  (s, s)
 * needs no pass of its own: the first sample of (s, e) is s exactly.
 * Asynchronous on `stream`; allocates its temporaries stream-ordered.
 */
int bcc_check_dense_endpoints(int robot, const bcc_scene *scene, const float *d_edges, int num_edges,
                              int resolution, uint8_t *d_results, void *stream);

/*
 * Average device time (ms) of `iters` launches of the dense kernel after
 * `warmup` untimed ones, measured with hipEvents on the launch stream.
 */
int bcc_time_dense(int robot, const bcc_scene *scene, const float *d_edges, int num_edges, int resolution,
                   uint8_t *d_results, int warmup, int iters, void *stream, double *avg_ms);

/*
 * Reference-order work accounting (no reference counterpart; feeds the
 * roofline): replays every pair with the reference's per-lane early exits in
 * plain step order (src/batch_cc.cu:118-129, src/utils.cuh:213-245) and
 * writes, per pair, the reference's source-level FP32 operation count
 * (d_work), its sin/cos count (d_trig) and the verdict (d_results).  All
 * three are [num_edges][num_envs] device arrays; the verdicts equal those of
 * bcc_check_dense.
 */
int bcc_work_dense(int robot, const bcc_scene *scene, const float *d_edges, int num_edges, int resolution,
                   uint64_t *d_work, uint64_t *d_trig, uint8_t *d_results, void *stream);

/* Same accounting for an explicit pair list: pair p checks edge p against
 * environment d_pair_env[p]; arrays are [num_pairs]. */
int bcc_work_pairs(int robot, const bcc_scene *scene, const float *d_edges, int num_pairs, const int32_t *d_pair_env,
                   int resolution, uint64_t *d_work, uint64_t *d_trig, uint8_t *d_results, void *stream);

/*
 * Clearance audit (no reference counterpart; the reference's validator only
 * diffs verdicts, scripts/batch_cc_validate.cpp:1074-1103): per pair, the
 * smallest gated clearance in metres over all 32 n samples, evaluated without
 * early exit in FP32 -- per sample, the minimum over the check program's
 * items of (lone item: its gate) or max(gate, smallest fine clearance), where
 * an environment clearance is the distance to the nearest primitive minus the
 * padded radius (r + 0.005) and a self clearance the distance minus both
 * radii.  Negative exactly when the pair collides; pairs within FP32 rounding
 * of zero are the ones whose verdict can depend on the arithmetic.  Dense when
 * d_pair_env is NULL (d_clearance[edge * num_envs + env]); else pair p checks
 * edge p against environment d_pair_env[p] (num_edges = number of pairs).
 */
int bcc_audit_clearance(int robot, const bcc_scene *scene, const float *d_edges, int num_edges,
                        const int32_t *d_pair_env, int resolution, float *d_clearance, void *stream);

/*
 * Ragged batches (no reference counterpart: the reference only takes the
 * dense cross product, src/batch_cc.cu:244-261).  Environment e owns pairs
 * [env_offsets[e], env_offsets[e+1]) (CSR, host array of num_envs + 1);
 * pair p checks edges[p] ([2][dim] floats) against environment e and writes
 * results[p] (1 = collision).  Same verdict semantics as bcc_check_dense.
 */
int bcc_check_ragged(int robot, const bcc_scene *scene, const int64_t *env_offsets, const float *edges,
                     int resolution, uint8_t *results);

/*
 * Allocation-free ragged check.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops).  Same verdicts as a ragged plan over the same CSR
 * batch; everything is device-resident and asynchronous on `stream`:
 * d_env_offsets int64 [num_envs + 1] (offsets[0] = 0, non-decreasing,
 * offsets[num_envs] = num_pairs), d_edges [num_pairs][2][dim], d_results
 * [num_pairs].  The tile list lives in a static per-device scratch array of
 * 4 MiB + 256 bytes, kept until the process ends, so a batch holds at most
 * 2^18 - 1 environments and 2^31 - 1 pairs (larger: BCC_ERR_INVALID).  Calls
 * that overlap on one device race on that array: issue them on one stream.
 * Offsets that break the rules above give unspecified verdicts for the
 * affected pairs, never an out-of-bounds access.
 */
int bcc_check_ragged_noalloc(int robot, const bcc_scene *scene, const int64_t *d_env_offsets, const float *d_edges,
                             int64_t num_pairs, int resolution, uint8_t *d_results, void *stream);

/*
 * A ragged plan owns the tile list (env, pair range) of one CSR layout and a
 * captured hipGraph (tile-counter reset + one persistent kernel launch) over
 * caller-owned device buffers d_edges / d_results.  tile_pairs = 0 picks the
 * tuned size.  The plan keeps pointers to the scene and buffers: they
 * must outlive it.  Launches of one plan must not overlap (they share the
 * tile counter).
 */
typedef struct bcc_ragged_plan bcc_ragged_plan;
int bcc_ragged_plan_create(int robot, const bcc_scene *scene, const int64_t *env_offsets, const float *d_edges,
                           int resolution, uint8_t *d_results, int tile_pairs, bcc_ragged_plan **out);
/* Same, with the persistent grid given as workgroups per CU (one wave per
 * SIMD each).  tile_pairs > 0 is the tile size; 0 sizes tiles from the tuning
 * table's target tiles per wave for this robot; -T targets T tiles per wave
 * (tile = pairs per resident wave / T, a power of two in [1, 256]).
 * groups_per_cu = 0 takes the table's grid. */
int bcc_ragged_plan_create_tuned(int robot, const bcc_scene *scene, const int64_t *env_offsets,
                                 const float *d_edges, int resolution, uint8_t *d_results, int tile_pairs,
                                 int groups_per_cu, bcc_ragged_plan **out);
/* Same, with a learned visiting order per pair: d_order [total pairs] on the
 * device, uint32 as in bcc_check_dense_ex (d_order[p] for pair p).  Each pair
 * first checks the step-0 samples of its first predicted bins, then runs the
 * plain step loop; an order that does not name four distinct bins below 32
 * probes evenly spaced bins instead.  Tiles hold at most 256 pairs.  The plan
 * reads d_order at every launch.  Verdicts equal the plain plan's.  No
 * reference counterpart (learned ordering). */
int bcc_ragged_plan_create_ordered(int robot, const bcc_scene *scene, const int64_t *env_offsets,
                                   const float *d_edges, int resolution, const uint32_t *d_order,
                                   uint8_t *d_results, int tile_pairs, int groups_per_cu, bcc_ragged_plan **out);
/* The same with the tile shape's step dimension (no reference counterpart;
 * the reference runs a pair's steps serially in one block,
 * src/batch_cc.cu:118-129): tile_steps > 0 splits each tile whose pairs run
 * more steps into chunks of tile_steps steps that waves take independently
 * (the samples are the same values; a pair's verdict is the OR of its
 * chunks, and a chunk stops once another decided the pair); 0 applies the
 * tuning table's chunk to batches with fewer pairs than twice the grid's
 * half-waves; -1 never splits.  bcc_ragged_plan_tile_steps gives the chunk a
 * plan uses (0: none).  Verdicts equal the plain plan's. */
int bcc_ragged_plan_create_shaped(int robot, const bcc_scene *scene, const int64_t *env_offsets,
                                  const float *d_edges, int resolution, uint8_t *d_results, int tile_pairs,
                                  int tile_steps, int groups_per_cu, bcc_ragged_plan **out);
int bcc_ragged_plan_tile_steps(const bcc_ragged_plan *plan);
int bcc_ragged_plan_launch(bcc_ragged_plan *plan, void *stream);        /* graph replay */
int bcc_ragged_plan_launch_eager(bcc_ragged_plan *plan, void *stream);  /* same kernels, no graph */
int bcc_ragged_plan_num_tiles(const bcc_ragged_plan *plan);
/* tile size and grid the plan was built with */
int bcc_ragged_plan_config(const bcc_ragged_plan *plan, int *tile_pairs, int *groups_per_cu);
int bcc_ragged_plan_time(bcc_ragged_plan *plan, int warmup, int iters, void *stream, double *avg_ms);
int bcc_ragged_plan_destroy(bcc_ragged_plan *plan);

/*
 * Mixed-robot batches (no reference counterpart: the reference checks one
 * robot type per call, src/batch_cc.cu:244-261,340-342).  A group replays
 * several ragged plans - typically one per robot, each over its own scene and
 * buffers - as one hipGraph whose plans are independent branches.  The plans
 * must outlive the group and must not be launched while a group launch that
 * contains them runs.
 */
typedef struct bcc_ragged_group bcc_ragged_group;
int bcc_ragged_group_create(bcc_ragged_plan *const *plans, int num_plans, bcc_ragged_group **out);
int bcc_ragged_group_launch(bcc_ragged_group *group, void *stream);
int bcc_ragged_group_time(bcc_ragged_group *group, int warmup, int iters, void *stream, double *avg_ms);
int bcc_ragged_group_destroy(bcc_ragged_group *group);

/*
 * Dense check with options (device buffers):
 *   d_order  (optional) learned visiting order per pair, [num_edges][num_envs]
 *            uint32: four distinct sample bins (0..31), one byte per rank,
 *            most likely to collide first (bin b = lane b's samples b n + i).
 *            The pair first checks the step-0 samples of its first predicted
 *            bins (a probe, several pairs per wave), then, if none collided,
 *            runs the plain step loop over all 32 n samples.  Any other value
 *            probes evenly spaced bins.  Verdicts do not depend on the order.
 *   d_rounds (optional) uint8 per pair: without d_order the rounds (32-sample
 *            steps) the pair ran; with d_order 0 if the probe decided the
 *            pair, else 1 + the plain steps it ran.
 */
int bcc_check_dense_ex(int robot, const bcc_scene *scene, const float *d_edges, int num_edges, int resolution,
                       const uint32_t *d_order, uint8_t *d_rounds, uint8_t *d_results, void *stream);
int bcc_time_dense_ex(int robot, const bcc_scene *scene, const float *d_edges, int num_edges, int resolution,
                      const uint32_t *d_order, uint8_t *d_results, int warmup, int iters, void *stream,
                      double *avg_ms);

/*
 * Learned step ordering (no reference counterpart).  A bcc_model is a
 * 3-layer MLP (96 -> H -> H -> 32; H = 64, the build's BCC_MLP_HIDDEN, read
 * back as the root of bcc_model_num_params() = H^2 + 130 H + 32) over a 64-cell max-pooled obstacle
 * occupancy code of the environment and the normalised edge endpoints,
 * predicting which of the 32 sample bins of a pair hold a collision.  BF16
 * weights and activations, FP32 accumulation on MFMA, FP32 master weights
 * and Adam state; trained on the GPU from labels the exact checker writes.
 * Parameter order: W1 [96][H], b1, W2 [H][H], b2, W3 [H][32], b3.
 */
typedef struct bcc_model bcc_model;
int bcc_model_num_params(void);
int bcc_model_create(int robot, uint64_t seed, bcc_model **out);
int bcc_model_destroy(bcc_model *model);
int bcc_model_get_params(const bcc_model *model, float *host_params);
int bcc_model_set_params(bcc_model *model, const float *host_params);
int bcc_model_save(const bcc_model *model, const char *path);      /* params + Adam state */
int bcc_model_load(const char *path, bcc_model **out);
/* 64-cell occupancy code per environment: d_enc [num_envs][64] */
int bcc_encode_envs(const bcc_scene *scene, float *d_enc);
/* BF16 input rows [num_pairs][96] for pairs (d_pair_env[p], d_pair_edge[p]) */
int bcc_model_features(const bcc_model *model, const bcc_scene *scene, const float *d_edges, int64_t num_pairs,
                       const int32_t *d_pair_env, const int32_t *d_pair_edge, uint16_t *d_x, void *stream);
/* training labels: bit t of d_labels[p] = bin t of the pair holds a colliding sample */
int bcc_label_pairs(int robot, const bcc_scene *scene, const float *d_edges, int num_pairs,
                    const int32_t *d_pair_env, const int32_t *d_pair_edge, int resolution, uint32_t *d_labels,
                    void *stream);
/* the same over the first `steps` samples of each bin only (steps = 1: the
 * step-0 samples, the ones the ordered kernels' probe checks; 0: all) */
int bcc_label_pairs_steps(int robot, const bcc_scene *scene, const float *d_edges, int num_pairs,
                          const int32_t *d_pair_env, const int32_t *d_pair_edge, int resolution, int steps,
                          uint32_t *d_labels, void *stream);
/* one forward + backward pass (mean binary cross-entropy), FP32 gradients into d_grad, no update */
int bcc_model_grads(bcc_model *model, const uint16_t *d_x, const uint32_t *d_labels, int rows, float *d_grad,
                    float *loss);
/* Adam training over rows of d_x / d_labels; epoch_loss (host, optional) receives the mean loss per epoch */
int bcc_model_train(bcc_model *model, const uint16_t *d_x, const uint32_t *d_labels, int num_rows, int epochs,
                    int batch, float lr, float *epoch_loss);
int bcc_model_logits(bcc_model *model, const uint16_t *d_x, int64_t rows, float *d_logits);
/* top-4 bins of every (edge, env) pair of a dense batch into d_order [num_edges][num_envs]
 * (one fused kernel: features, three MFMA layers and the selection; same
 * values as bcc_model_logits on bcc_model_features rows).  The bcc_model_order_*
 * and _logits_ calls run asynchronously on `stream` (results are ready when
 * the stream is); they allocate only when a scene has more environments than
 * any the model ordered before. */
int bcc_model_order_dense(bcc_model *model, const bcc_scene *scene, const float *d_edges, int num_edges,
                          uint32_t *d_order, void *stream);
/* the same, also writing every pair's 32 logits into d_logits [num_edges][num_envs][32] */
int bcc_model_logits_dense(bcc_model *model, const bcc_scene *scene, const float *d_edges, int num_edges,
                           uint32_t *d_order, float *d_logits, void *stream);
/* the same for a list of pairs (d_pair_env[p], d_pair_edge[p]), e.g. a ragged
 * batch (d_pair_edge = NULL: pair p uses edge p); environment and edge indices
 * must be in range.  d_order [num_pairs]; d_logits [num_pairs][32] or NULL. */
int bcc_model_order_pairs(bcc_model *model, const bcc_scene *scene, const float *d_edges, int64_t num_pairs,
                          const int32_t *d_pair_env, const int32_t *d_pair_edge, uint32_t *d_order,
                          float *d_logits, void *stream);
/* One learned-order step over a dense batch, timed: bcc_model_order_dense
 * into d_order, then bcc_check_dense_ex with that order, `iters` times after
 * `warmup`, hipEvents on `stream`; step_ms = average step, order_ms
 * (optional) = average of the inference alone, timed the same way. */
int bcc_time_learned_dense(bcc_model *model, const bcc_scene *scene, const float *d_edges, int num_edges,
                           int resolution, uint32_t *d_order, uint8_t *d_results, int warmup, int iters,
                           void *stream, double *step_ms, double *order_ms);

/*
 * Fused learned step (no reference counterpart; replaces the plain step loop
 * of src/batch_cc.cu:118-129 as the order in which one pair's samples are
 * visited).  One kernel does the whole step: each wave owns a tile of pairs
 * of one environment and
 *   1. probes every pair's step-0 samples at 2 evenly spaced bins (8, 24)
 *      (panda launches of 2^22 pairs or more: the edges in a locality order
 *      sorted on the device, the results scattered back to their rows);
 *   2. runs the model (BF16 MFMA) for the pairs the probe left undecided
 *      only, and of those only for pairs of at least 3 steps (fetch: 6;
 *      shorter ones go to step 4 directly): its order word, the spread top-4
 *      bins among the other 30;
 *   3. probes the word's first 2 bins (with three probe passes, see
 *      bcc_debug_learned_passes, then its next 2; model NULL: the fixed bins 0
 *      and 16, then 4 and 20);
 *   4. runs the plain step loop over all 32 n samples of the pairs still
 *      undecided.
 * Verdicts equal bcc_check_dense's for any model.  The model must have been
 * created for `robot`.  Optional outputs ([num_edges][num_envs] like the
 * results): d_stage p = decided by probe p + 1, else passes - 1 + the plain
 * steps run (capped at 255; passes = 2 by default); d_words the order word (as
 * bcc_check_dense_ex) inferred for each pair that reached step 2 (others
 * untouched); d_stats int32[5] += pairs decided by probes 1, 2 and 3,
 * inferred, sent to the step loop.  Asynchronous on `stream`; allocates the
 * model's environment codes as bcc_model_order_dense does and a tile
 * counter stream-ordered.  Small batches (bcc_learned_small_pairs() pairs or
 * fewer) that ask for none of the three optional outputs run instead as
 * inference of every pair (model NULL: the invalid word) followed by
 * bcc_check_dense_ex with those words.
 */
int bcc_check_dense_learned(int robot, bcc_model *model, const bcc_scene *scene, const float *d_edges,
                            int num_edges, int resolution, uint8_t *d_results, uint8_t *d_stage, uint32_t *d_words,
                            int32_t *d_stats, void *stream);
/* Average device time (ms) of `iters` learned steps (operands + fused
 * kernel) after `warmup`, hipEvents on `stream`. */
int bcc_time_dense_learned(int robot, bcc_model *model, const bcc_scene *scene, const float *d_edges,
                           int num_edges, int resolution, uint8_t *d_results, int warmup, int iters, void *stream,
                           double *avg_ms);
/* A ragged plan running the fused learned kernel persistently over its
 * tiles (one environment and at most 256 pairs each, a wave per tile;
 * tile_pairs = 0 takes 64; pairs of fewer than 6 steps skip inference),
 * captured in a hipGraph like the plain plans.  It snapshots the model when
 * it is made (later training does not
 * change it; model NULL: fixed second-probe bins).  Verdicts equal the plain
 * plan's.  No reference counterpart. */
int bcc_ragged_plan_create_learned(int robot, const bcc_model *model, const bcc_scene *scene,
                                   const int64_t *env_offsets, const float *d_edges, int resolution,
                                   uint8_t *d_results, int tile_pairs, int groups_per_cu, bcc_ragged_plan **out);
/* Tuning hooks (no reference counterpart; verdicts never depend on them):
 * pairs per tile of dense learned launches (0 = chosen per launch), and the
 * radius around the first probe's bins inside which the second probe's picks
 * are taken only when no farther bin is left (0 = no preference), and the
 * number of probe passes (1 .. 3; 0 = the default, 2), and the fewest steps
 * n a pair needs to be inferred and probed again (0 = the per-robot
 * defaults above).  All
 * are read when a launch or plan is made. */
int bcc_debug_learned_tile(int pairs);
int bcc_debug_learned_far(int radius);
int bcc_debug_learned_passes(int passes);
int bcc_debug_learned_min_steps(int steps);
/* Tuning hook (no reference counterpart): panda's dense learned launches
 * of at least 2^22 pairs check the edges in a locality order (sorted on the
 * device by where the first probe places the arm, so each probe batch's
 * culling boxes are small) and scatter the results back: mode 1 (the
 * default); 0 = never, 2 = every launch of every robot.  Verdicts and
 * per-pair outputs never depend on it. */
int bcc_debug_learned_locality(int mode);
/* Tuning hook (no reference counterpart): dense learned launches of at most
 * bcc_learned_small_pairs() pairs (3 x 2^18 by default; max_pairs >= 0 sets it,
 * negative restores the default) that ask for no stage, words or statistics
 * run as bcc_model_order_dense for every pair (without a model: the invalid
 * word, so the probe takes bins 8 and 24) followed by bcc_check_dense_ex
 * with those words (one stream-ordered words buffer); larger ones, and
 * every launch that asks for those outputs, run the fused kernel.  Verdicts
 * never depend on it. */
int bcc_debug_learned_small(int64_t max_pairs);
int64_t bcc_learned_small_pairs(void);

/* test hook: C = alpha * op(A) op(B) + beta * C on the MFMA GEMM (BF16 in, FP32 out) */
int bcc_debug_gemm(int ta, int tb, const uint16_t *d_a, int lda, const uint16_t *d_b, int ldb, float *d_c, int ldc,
                   int m, int n, int k, float alpha, float beta);

/* Test hook (no reference counterpart): force the dense/ragged kernel
 * variant, 0 = low occupancy, 1 = high occupancy, -1 = chosen per launch from
 * the scene's mean obstacle count (the default).  Verdicts never depend on it. */
int bcc_debug_force_variant(int variant);

/* Tuning hook (no reference counterpart): scenes created afterwards cull
 * distant obstacles in a sweep of N robot spheres only when N times the
 * environment's primitive count is at least `work` (0 = always, a huge value
 * = never, -1 = the built-in tuned value).  Verdicts never depend on it. */
int bcc_debug_cull_min(int work);

/* Test hook (no reference counterpart): the culling threshold stamped into
 * `scene` when it was created (see bcc_debug_cull_min). */
int bcc_debug_scene_cull_min(const bcc_scene *scene, int *out);

/* Tuning hook (no reference counterpart): edges per workgroup of dense
 * launches (0 = chosen per launch: 64, halved down to 8 while the launch has
 * fewer than 8 workgroups per CU).  Verdicts never depend on it. */
int bcc_debug_dense_chunk(int edges);

/* Test hook (no reference counterpart): the culling step's half-wave boxes
 * for one wave.  d_in [64][4] = x y z padded-radius per lane; d_out [12] =
 * lo xyz, hi xyz of lanes 0-31 then of lanes 32-63 among the lanes in `mask`,
 * each half widened by 0.01 + 1e-5 * its largest |coordinate| (an empty half
 * gives lo = +inf, hi = -inf). */
int bcc_debug_half_boxes(const float *d_in, uint64_t mask, float *d_out);

/* Device memory helpers for bindings that do not bring their own allocator. */
int bcc_malloc(void **d_ptr, size_t bytes);
int bcc_free(void *d_ptr);
int bcc_memcpy_htod(void *d_dst, const void *src, size_t bytes);
int bcc_memcpy_dtoh(void *dst, const void *d_src, size_t bytes);
int bcc_memset(void *d_dst, int value, size_t bytes);
int bcc_synchronize(void);

/*
 * Test hooks (no reference counterpart): expose the device arithmetic so
 * tests can compare it bit for bit with a host restatement.
 *   bcc_debug_fk       : d_q [n][dim] -> d_out [n][num_spheres][4] sphere centres + radius
 *   bcc_debug_sincos   : d_x [n] -> d_out [n][2] (sin, cos)
 *   bcc_debug_samples  : d_edges [m][2][dim] -> d_n [m], d_cfg [m][32][dim] step-0 lane configs,
 *                        d_last [m][32][dim] configurations of the last step (i = n-1)
 *   bcc_debug_prims    : d_in [n][40] = x y z rp, sphere (4), capsule (8), cuboid (15 + 1 pad),
 *                        aabb (x y z h1 h2 h3 + 2 pad) -> d_out [n][6] values of the sphere,
 *                        capsule, z-capsule (same capsule record), cuboid, z-cuboid (same
 *                        cuboid record) and aabb tests
 */
int bcc_robot_num_spheres(int robot);
int bcc_debug_fk(int robot, const float *d_q, int n, float *d_out);
int bcc_debug_sincos(const float *d_x, int n, float *d_out);
int bcc_debug_samples(int robot, const float *d_edges, int m, int resolution, int32_t *d_n, float *d_cfg,
                      float *d_last);
int bcc_debug_prims(const float *d_in, int n, float *d_out);
/* Plain device copy of `bytes` (a multiple of 16): a known-traffic kernel used
 * to calibrate hardware byte counters. */
int bcc_debug_copy(const void *d_src, void *d_dst, size_t bytes);
/* Known FP32 VALU work: `blocks` workgroups of 256 threads; in every wave the
 * lanes below active_lanes (0..64) run 8 independent chains of `iters` FMAs
 * (then 7 adds); every thread writes one float to d_out[blocks * 256].  A
 * known-work kernel used to calibrate the hardware VALU/FLOP counters and to
 * measure the FP32 FMA rate; ms (optional) receives its hipEvent time. */
int bcc_debug_valu(int blocks, int iters, int active_lanes, float *d_out, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* BCC_H */
