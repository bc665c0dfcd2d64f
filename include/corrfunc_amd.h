/*
 * corrfunc_amd.h -- extensions of the MI355X engine beyond the reference C API.
 *
 * None of these has a counterpart in the reference (Corrfunc v2.5.3); they exist
 * for callers that benchmark or steer the GPU engine. The reference entry points
 * in corrfunc_theory.h work without them.
 *
 * Input residency: every entry point of corrfunc_theory.h accepts position and
 * weight pointers in host memory (as the reference does; they are copied to the
 * GPU) or in GPU memory (detected with hipPointerGetAttributes and used in place).
 */
#ifndef CORRFUNC_AMD_H
#define CORRFUNC_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "corrfunc_theory.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Pair-kernel variants. TUNED (default): FP32 filter on cell-local coordinates,
 * FP64 replay of the surviving pairs, persistent waves. PLAIN: every pair of
 * every cell pair decided in FP64, one workgroup per cell pair. Both give the
 * same counts. The environment variable CORRFUNC_AMD_ENGINE=plain selects PLAIN
 * at load time. */
#define CF_AMD_ENGINE_TUNED 0
#define CF_AMD_ENGINE_PLAIN 1

typedef struct {
    int nmesh[3];            /* grid used by the last call */
    int refine[3];           /* bin refine factors used by the last call */
    int64_t num_cell_pairs;  /* neighbour-cell pairs in the work list */
    int64_t num_work_items;  /* work items of the tuned kernel (0 for PLAIN) */
    double tested_pairs;     /* sum over cell pairs of N1*N2 (N(N-1)/2 for a cell with itself) */
    double kernel_ms;        /* pair-kernel time, HIP events on the engine's stream */
    double grid_ms;          /* host wall time from entry to the end of gridding */
    double total_ms;         /* host wall time of the engine part of the call */
    int engine;              /* CF_AMD_ENGINE_* used */
    int pad;
} cf_amd_stats;

/* Selects the pair kernel for subsequent calls; returns EXIT_SUCCESS/EXIT_FAILURE. */
int cf_amd_set_engine(int kind);
/* Returns the pair kernel currently selected. */
int cf_amd_get_engine(void);
/* Copies the statistics of the last successful call. */
int cf_amd_last_stats(cf_amd_stats *out);
/* Library identification string. */
const char *cf_amd_version(void);

/* DD(r) auto-correlations of `ncat` independent catalogs (mocks, jackknife or
 * bootstrap resamples) in one call: every catalog is gridded on the GPU, their
 * cell pairs form one work list and one persistent kernel launch counts them all.
 * results[k] equals what countpairs(ND[k], X[k], Y[k], Z[k], 0, ..., autocorr=1,
 * binfile, &results[k], &copy_of_options, extra) returns; release each with
 * free_results. W[k] are the weights of catalog k (used when
 * extra->weight_method = PAIR_PRODUCT; W may be NULL otherwise). Pointers may be
 * host or device memory. All catalogs must share the periodic wraps. */
int cf_amd_countpairs_batch(int ncat, const int64_t *ND, void **X, void **Y, void **Z, void **W,
                            const int numthreads, const char *binfile, results_countpairs *results,
                            struct config_options *options, struct extra_options *extra);

/* Device memory on the engine's GPU, from the engine's own HIP runtime, for
 * callers that keep catalogs resident in HBM across calls. */
void *cf_amd_device_alloc(size_t bytes);            /* NULL on failure */
int cf_amd_device_free(void *p);
#define CF_AMD_H2D 0
#define CF_AMD_D2H 1
#define CF_AMD_D2D 2
int cf_amd_memcpy(void *dst, const void *src, size_t bytes, int kind);  /* synchronous */

#ifdef __cplusplus
}
#endif

#endif /* CORRFUNC_AMD_H */
