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

/* Pair-kernel variants. TUNED (default): integer filter on quantized cell-local
 * positions, FP64 replay of the surviving pairs, persistent waves. PLAIN: every pair of
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
    int sched;               /* CF_AMD_SCHED_* that ordered the work list */
    double list_ms;          /* cost prediction, ordering and work-list construction (device events) */
    int hist_copies;         /* LDS histogram replication of the TUNED kernel */
    int autotuned;           /* 0: autotuner off, 1: cached choice used, 2: tuned during this call */
    int queue_fetch;         /* work items per queue atomic of the TUNED kernel (0 for PLAIN) */
    int global_hist;         /* 1: the histogram was too large for LDS and was kept in global memory */
    int drain_pairs;         /* TUNED kernel: survivors each lane decides per drain (1 or 2; 0 for PLAIN) */
    int workgroup_threads;   /* TUNED kernel: threads per workgroup (256, 512 or 1024; 0 for PLAIN) */
} cf_amd_stats;

/* Autotuner. When on, a call whose options carry the default (non-custom) bin
 * refine factors uses the refine factors and LDS histogram replication that ran
 * fastest for its workload signature (mode, size octaves of both sets, rmax and
 * pimax relative to the box, periodic axes, weights, averages, precision). On
 * the first call of a signature every candidate (refine (1,1,1), (2,2,1),
 * (2,2,2), (3,3,1), (3,3,2), (4,4,2); then replication 8 and 2 on the best) is
 * timed on the call's own inputs, and the winner is cached in-process and, if
 * the environment variable CORRFUNC_AMD_TUNE_CACHE names a file, appended to
 * it (read back at the first tuned call of a later process). The batched entry
 * point tunes per catalog size octave. Counts follow the reference's rules for
 * the chosen refine factors (the reference's results do not depend on them).
 * CORRFUNC_AMD_AUTOTUNE=1 turns it on at load time. */
int cf_amd_set_autotune(int on);
int cf_amd_get_autotune(void);
int cf_amd_autotune_clear(void);  /* forgets the in-process choices (the file is read again) */

/* Work-list schedulers of the TUNED kernel. Its persistent waves pull work items
 * (<= 64 primaries of a cell pair, or a slice of their secondaries) in list order.
 *   NONE       cell-pair list order, one item per 64 primaries;
 *   HEURISTIC  cell pairs sorted by descending N1*N2, items heavier than
 *              1/(4 * resident waves) of the total split into slices;
 *   LEARNED    the same with the cost predicted by the cost model's network
 *              (falls back to HEURISTIC while no model is trained or loaded).
 * Counts are identical under every scheduler. Default LEARNED; the environment
 * variable CORRFUNC_AMD_SCHED=none|heuristic selects another at load time. */
#define CF_AMD_SCHED_NONE 0
#define CF_AMD_SCHED_HEURISTIC 1
#define CF_AMD_SCHED_LEARNED 2
int cf_amd_set_scheduler(int kind);
int cf_amd_get_scheduler(void);

/* The cost model: an MLP (CF_AMD_MODEL_FEATURES inputs, two hidden layers of 32
 * ReLU units, one output) that predicts log2 of the device time of a cell pair
 * from features the engine computes per cell pair. Training samples are recorded
 * by the engine itself: while recording is on, every TUNED call times each cell
 * pair (wall clock of the waves on its items) and appends (features, log2 ticks)
 * to a device store of CF_AMD_MODEL_STORE samples. Training (Adam on the MSE of
 * standardised targets) and inference run as HIP kernels on the GPU.
 * Parameter layout: W1[32][F] | b1[32] | W2[32][32] | b2[32] | W3[32] | b3;
 * normalisation: mean[F] | inv_sd[F] | y_mean | y_sd. */
#define CF_AMD_MODEL_FEATURES 10
#define CF_AMD_MODEL_PARAMS 1441
#define CF_AMD_MODEL_NORM 22
#define CF_AMD_MODEL_STORE (1 << 21)
int cf_amd_record_costs(int on);
/* Copies up to `cap` recorded samples (X: [cap][F] row-major, y: log2 ticks);
 * returns the number stored (X, y may be NULL to query it), or -1 on error. */
int64_t cf_amd_cost_samples(float *X, float *y, int64_t cap);
int cf_amd_clear_cost_samples(void);
/* Trains from fresh weights (seeded) on n samples (X == NULL: the recorded
 * ones); writes the final loss over all samples to *loss when non-NULL. */
int cf_amd_model_train(const float *X, const float *y, int64_t n, int epochs, int batch, float lr, unsigned seed,
                       float *loss);
int cf_amd_model_get(float *params, float *norm);
int cf_amd_model_set(const float *params, const float *norm);  /* installs a trained model */
int cf_amd_model_reset(void);                                 /* forgets it */
int cf_amd_model_trained(void);
int cf_amd_model_predict(const float *X, int64_t n, float *out);  /* log2 cost of n feature rows */
/* Loss L = 1/(2n) sum (net(x) - (y - y_mean)/y_sd)^2 of n <= 2^20 samples and its
 * gradient with respect to `params` (the GPU's forward and backward kernels). */
int cf_amd_model_grad(const float *X, const float *y, int64_t n, const float *params, const float *norm,
                      float *grad, float *loss);
/* One Adam step of the training kernel (beta1 0.9, beta2 0.999, eps 1e-8,
 * bias-corrected, step >= 1) on CF_AMD_MODEL_PARAMS host values, in place. */
int cf_amd_model_adam_step(float *params, const float *grad, float *m, float *v, int step, float lr);

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

/* The batched DD(r) of cf_amd_countpairs_batch captured once into a HIP graph
 * and replayed on new positions of the same catalog sizes (jackknife and
 * bootstrap resamples written into the same device arrays). No counterpart in
 * the reference. Every decision the batch takes from the positions -- each
 * catalog's extent and grid, its cell pairs and the work list -- is taken on the
 * device, so a launch counts whatever the arrays hold when it runs; results[k]
 * equals countpairs() on catalog k then (release each with free_results).
 * X[k], Y[k], Z[k] (and W[k] with PAIR_PRODUCT weights) must be device memory
 * that stays valid while the graph exists. The box must be periodic with its
 * size given on every axis. Capacities are fixed at create: cells from the box,
 * cell pairs and work items from the capture-time positions plus 25 %; a launch
 * whose positions need more fails with a message (create the graph again). The
 * work list is ordered and split by the scheduler selected at create (the cost
 * model's parameters are read at each launch). */
typedef struct cf_amd_batch_graph cf_amd_batch_graph;
int cf_amd_batch_graph_create(int ncat, const int64_t *ND, void **X, void **Y, void **Z, void **W,
                              const int numthreads, const char *binfile, struct config_options *options,
                              struct extra_options *extra, cf_amd_batch_graph **graph);
int cf_amd_batch_graph_launch(cf_amd_batch_graph *graph, results_countpairs *results);
int cf_amd_batch_graph_destroy(cf_amd_batch_graph *graph);

/* countpairs() for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). It has
 * no counterpart in the reference. Same arguments and results as countpairs()
 * (DD(r), theory/DD/countpairs.h:28-35) with three differences:
 *  - positions and weights must be in GPU memory;
 *  - results is the caller's: results->npairs, rupp, rpavg and weightavg point
 *    to arrays of results->nbin entries on entry (at least the number of bin
 *    edges); on return results->nbin holds the number of edges. Nothing is
 *    malloc'ed: do not call free_results on it;
 *  - every buffer the kernels use lives in one fixed device array of
 *    CF_AMD_NOALLOC_SCRATCH_BYTES per GPU, held for the life of the process:
 *    about 70 B per particle of each set and 64 B per neighbour-cell pair (about
 *    38 per cell in an autocorrelation, 75 in a cross-correlation, at the
 *    default refine factors); a call that needs more fails with a message.
 *    The call allocates no device or pinned memory.
 * The scratch array is shared by every call on a GPU: calls that overlap on one
 * GPU race (the library serialises calls from threads of one process).
 * The autotuner and cost recording are not used by this entry point. */
#define CF_AMD_NOALLOC_SCRATCH_BYTES (64u << 20)
int cf_amd_countpairs_noalloc(const int64_t ND1, void *X1, void *Y1, void *Z1, const int64_t ND2, void *X2,
                              void *Y2, void *Z2, const int numthreads, const int autocorr, const char *binfile,
                              results_countpairs *results, struct config_options *options,
                              struct extra_options *extra);

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
