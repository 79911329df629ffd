// Launch table: the culling threshold, per robot the kernel-variant
// threshold and the ragged launch shape (persistent workgroups per CU and
// the target number of tiles per wave, from which a plan sizes its tiles).
// Written by batch-cc_amd/tune.py.
#pragma once

struct bcc_tune_entry {
    int robot;            // BCC_PANDA / BCC_FETCH / BCC_BAXTER / BCC_XARM7
    int tiles_per_wave;   // target tiles per persistent wave
    int groups_per_cu;    // persistent workgroups per CU
    int steps_per_tile;   // step chunk of small batches' tiles (0: no step dimension)
};

// Exact culling of distant obstacles: a sweep of N robot spheres culls when
// N * (primitives in the environment) is at least this.
static const int kCullMinWork = 32;

// Kernel variant per launch: the high-occupancy variant when the mean number
// of tested obstacles per environment (weighted by pairs) is at least the
// first value and below the second, per robot (panda, fetch, baxter, xarm7).
static const int kHiVariantMinObstacles[4] = {20, 0, 1073741824, 1073741824};
static const int kHiVariantMaxObstacles[4] = {163, 1073741824, 1073741824, 1073741824};

// source: tune.py on gfx950, 2026-12-04, seed 0 (step dimension; launch shapes kept)
static const bcc_tune_entry kTuneTable[] = {
    {0, 8, 6, 2},   // panda: score 1.015
    {1, 4, 8, 2},   // fetch: score 1.003
    {2, 16, 4, 0},   // baxter: score 1.000
    {3, 16, 4, 1},   // xarm7: baxter's setting (not tuned)
};
