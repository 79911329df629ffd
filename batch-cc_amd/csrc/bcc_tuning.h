// Ragged-batch launch table: per robot, by mean pairs per environment, the
// tile size (pairs per tile) and the persistent grid (workgroups per CU).
// Written by batch-cc_amd/tune.py; an entry applies when the batch's mean
// pairs per environment is at most max_mean_pairs (the last entry of each
// robot has no bound).
#pragma once

struct bcc_tune_entry {
    int robot;            // BCC_PANDA / BCC_FETCH / BCC_BAXTER
    long long max_mean_pairs;
    int tile_pairs;
    int groups_per_cu;
};

// Kernel variant per launch: the high-occupancy variant when the mean number
// of tested obstacles per environment (weighted by pairs) is at least this,
// per robot (panda, fetch, baxter).
static const int kHiVariantMinObstacles[3] = {32, 32, 64};

// source: tune.py on gfx950, 2026-12-03, seed 0
static const bcc_tune_entry kTuneTable[] = {
    {0, 8LL, 2, 4},   // panda: 2.459 ms
    {0, 64LL, 4, 8},   // panda: 7.052 ms
    {0, 512LL, 4, 6},   // panda: 6.983 ms
    {0, 0x7fffffffffffffffLL, 2, 6},   // panda: 6.513 ms
    {1, 8LL, 4, 6},   // fetch: 1.822 ms
    {1, 64LL, 2, 8},   // fetch: 8.045 ms
    {1, 512LL, 4, 8},   // fetch: 7.915 ms
    {1, 0x7fffffffffffffffLL, 4, 6},   // fetch: 7.565 ms
    {2, 8LL, 4, 8},   // baxter: 3.921 ms
    {2, 64LL, 2, 6},   // baxter: 9.504 ms
    {2, 512LL, 2, 8},   // baxter: 9.196 ms
    {2, 0x7fffffffffffffffLL, 2, 6},   // baxter: 9.126 ms
};
