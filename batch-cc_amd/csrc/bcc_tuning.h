// Ragged-batch launch table: per robot, the persistent grid (workgroups per
// CU) and the target number of tiles per wave, from which a plan sizes its
// tiles; and the kernel-variant thresholds.  Written by batch-cc_amd/tune.py.
#pragma once

struct bcc_tune_entry {
    int robot;            // BCC_PANDA / BCC_FETCH / BCC_BAXTER
    int tiles_per_wave;   // target tiles per persistent wave
    int groups_per_cu;    // persistent workgroups per CU
};

// Kernel variant per launch: the high-occupancy variant when the mean number
// of tested obstacles per environment (weighted by pairs) is at least this,
// per robot (panda, fetch, baxter).
static const int kHiVariantMinObstacles[3] = {32, 16, 32};

// source: tune.py on gfx950, 2026-12-04, seed 0
static const bcc_tune_entry kTuneTable[] = {
    {0, 16, 6},   // panda: score 1.003
    {1, 8, 8},   // fetch: score 1.016
    {2, 16, 6},   // baxter: score 1.001
};
