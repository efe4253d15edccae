// Decision table of the batched SPD solvers: {t16_max, wg_per_cu} for FP64
// (D) and FP32 (S).  Untuned defaults; the empirical search in
// math-atlas_amd/tune/batched_search.py overwrites this file with its winners.
#pragma once
#define ATLG_BATCH_TUNE_D {96, 2}
#define ATLG_BATCH_TUNE_S {128, 2}
