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

// source: tune.py on gfx950, 2026-12-03, seed 0
static const bcc_tune_entry kTuneTable[] = {
    {0, 8LL, 4, 6},   // panda: 2.648 ms
    {0, 64LL, 4, 4},   // panda: 7.817 ms
    {0, 512LL, 4, 6},   // panda: 7.448 ms
    {0, 0x7fffffffffffffffLL, 4, 4},   // panda: 6.839 ms
    {1, 8LL, 4, 6},   // fetch: 1.967 ms
    {1, 64LL, 8, 8},   // fetch: 8.943 ms
    {1, 512LL, 4, 6},   // fetch: 8.667 ms
    {1, 0x7fffffffffffffffLL, 4, 6},   // fetch: 8.294 ms
    {2, 8LL, 16, 6},   // baxter: 5.154 ms
    {2, 64LL, 4, 8},   // baxter: 13.452 ms
    {2, 512LL, 4, 6},   // baxter: 12.489 ms
    {2, 0x7fffffffffffffffLL, 4, 6},   // baxter: 11.833 ms
};
