/*
 * corrfunc_theory.h -- C ABI of the MI355X theory pair-counting engine.
 *
 * Drop-in replacement for the public C interface of Corrfunc's theory pair
 * counters (reference v2.5.3). Every type below has the byte layout of its
 * reference counterpart and every entry point has the reference signature and
 * error behaviour (EXIT_SUCCESS / EXIT_FAILURE, message on stderr, results
 * arrays malloc'ed by the library and released with the matching
 * free_results* call, which uses libc free()).
 *
 *   reference                                 this header
 *   utils/defs.h:37-45   api_cell_timings     struct api_cell_timings
 *   utils/defs.h:53-156  config_options       struct config_options (1024 B)
 *   utils/defs.h:236-346 get_config_options   get_config_options()
 *   utils/defs.h:353-419 weights/extra        weight_struct, weight_method_t,
 *                                             struct extra_options (1024 B)
 *   utils/cpu_features.h:20-35 isa enum       enum isa
 *   theory/DD/countpairs.h:20-35              results_countpairs, countpairs()
 *   theory/DDrppi/countpairs_rp_pi.h:19-37    results_countpairs_rp_pi, countpairs_rp_pi()
 *   theory/DDsmu/countpairs_s_mu.h:19-39      results_countpairs_s_mu, countpairs_s_mu()
 *   theory/xi/countpairs_xi.h:20-36           results_countpairs_xi, countpairs_xi()
 *   theory/wp/countpairs_wp.h:20-37           results_countpairs_wp, countpairs_wp()
 *
 * The arithmetic runs on the GPU; `numthreads` is accepted and bounds the host
 * threads used for host-side steps. `instruction_set` selects which FP64
 * rounding order of the reference kernels the engine reproduces exactly
 * (FALLBACK/SSE42/AVX/AVX2: the non-FMA orders; AVX512F: the FMA orders;
 * -1 resolves to the highest x86 ISA of the calling host, as the reference's
 * 'fastest' does).
 */
#ifndef CORRFUNC_THEORY_H
#define CORRFUNC_THEORY_H

#include <stdint.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The reference stringifies a quoted macro (defs.h:24 API_VERSION STR("2.5.3")),
 * so options.version holds the 7 characters "2.5.3" including the quotes. */
#define CORRFUNC_API_VERSION "\"2.5.3\""
#define NLATMAX 100
#define BOXSIZE_NOTGIVEN (-2.)
#define MAX_FAST_DIVIDE_NR_STEPS 3
#define OPTIONS_HEADER_SIZE 1024
#define EXTRA_OPTIONS_HEADER_SIZE 1024
#define MAX_NUM_WEIGHTS 10

#define BINNING_REF_MASK 0x0000000F
#define BINNING_ORD_MASK 0x000000F0
#define BINNING_DFL 0x0
#define BINNING_CUST 0x1

/* x86 instruction-set identifiers (reference utils/cpu_features.h:20-35). */
typedef enum {
    FALLBACK = 0,
    SSE = 1, SSE2 = 2, SSE3 = 3, SSSE3 = 4, SSE4 = 5,
    SSE42 = 6,
    AVX = 7,
    AVX2 = 8,
    AVX512F = 9,
    ARM64 = 10,
    NUM_ISA
} isa;

struct api_cell_timings {
    int64_t N1;
    int64_t N2;
    int64_t time_in_ns;
    int first_cellindex;
    int second_cellindex;
    int tid;
};

struct config_options {
    union { double boxsize; double boxsize_x; };
    double boxsize_y;
    double boxsize_z;
    /* cosmology block, used only by the mocks engines */
    struct { double OMEGA_M, OMEGA_B, OMEGA_L, HUBBLE, LITTLE_H, SIGMA_8, NS; };
    double c_api_time;
    struct api_cell_timings *cell_timings;
    int64_t totncells_timings;
    size_t float_type;
    int32_t instruction_set;
    char version[32];
    uint8_t verbose;
    uint8_t c_api_timer;
    uint8_t c_cell_timer;
    uint8_t need_avg_sep;
    uint8_t autocorr;
    uint8_t periodic;
    uint8_t sort_on_z;
    uint8_t is_comoving_dist;
    uint8_t link_in_dec;
    uint8_t link_in_ra;
    uint8_t fast_divide_and_NR_steps;
    uint8_t fast_acos;
    uint8_t enable_min_sep_opt;
    int8_t bin_refine_factors[3];
    uint16_t max_cells_per_dim;
    uint8_t copy_particles;
    uint8_t use_heap_sort;
    union { uint32_t binning_flags; uint8_t bin_masks[4]; };
    uint8_t reserved[OPTIONS_HEADER_SIZE - 172];
};

typedef struct {
    void *weights[MAX_NUM_WEIGHTS];   /* weights[w] points at N values of the position dtype */
    int64_t num_weights;
} weight_struct;

typedef enum {
    NONE = -42,
    PAIR_PRODUCT = 0,
    NUM_WEIGHT_TYPE
} weight_method_t;

struct extra_options {
    weight_struct weights0;
    weight_struct weights1;
    weight_method_t weight_method;
    uint8_t reserved[EXTRA_OPTIONS_HEADER_SIZE - 2 * sizeof(weight_struct) - sizeof(weight_method_t)];
};

/* Layout checks: offsets measured on the reference header (x86-64, LP64). */
#define CF_STATIC_ASSERT(c, m) typedef char cf_static_assert_##m[(c) ? 1 : -1]
CF_STATIC_ASSERT(sizeof(struct config_options) == 1024, config_options_size);
CF_STATIC_ASSERT(sizeof(struct extra_options) == 1024, extra_options_size);
CF_STATIC_ASSERT(sizeof(struct api_cell_timings) == 40, cell_timings_size);
CF_STATIC_ASSERT(offsetof(struct config_options, c_api_time) == 80, off_c_api_time);
CF_STATIC_ASSERT(offsetof(struct config_options, float_type) == 104, off_float_type);
CF_STATIC_ASSERT(offsetof(struct config_options, instruction_set) == 112, off_isa);
CF_STATIC_ASSERT(offsetof(struct config_options, version) == 116, off_version);
CF_STATIC_ASSERT(offsetof(struct config_options, verbose) == 148, off_verbose);
CF_STATIC_ASSERT(offsetof(struct config_options, enable_min_sep_opt) == 160, off_minsep);
CF_STATIC_ASSERT(offsetof(struct config_options, bin_refine_factors) == 161, off_refine);
CF_STATIC_ASSERT(offsetof(struct config_options, max_cells_per_dim) == 164, off_maxcells);
CF_STATIC_ASSERT(offsetof(struct config_options, binning_flags) == 168, off_flags);
CF_STATIC_ASSERT(offsetof(struct extra_options, weights1) == 88, off_weights1);
CF_STATIC_ASSERT(offsetof(struct extra_options, weight_method) == 176, off_method);

/* ---- option helpers (same semantics as reference utils/defs.h:158-230) ---- */
static inline void set_bin_refine_scheme(struct config_options *o, const int8_t flag)
{
    o->binning_flags = (o->binning_flags & ~(uint32_t)BINNING_REF_MASK) | ((uint32_t)flag & BINNING_REF_MASK);
}
static inline void reset_bin_refine_scheme(struct config_options *o) { set_bin_refine_scheme(o, BINNING_DFL); }
static inline int8_t get_bin_refine_scheme(struct config_options *o) { return (int8_t)(o->binning_flags & BINNING_REF_MASK); }
static inline void set_bin_refine_factors(struct config_options *o, const int f[3])
{
    for (int i = 0; i < 3; i++) {
        int8_t v = (int8_t)f[i];
        if (f[i] > 127) {
            fprintf(stderr, "Warning: bin refine factor[%d] can be at most 127. Found %d instead\n", i, f[i]);
            v = 1;
        }
        o->bin_refine_factors[i] = v;
    }
    reset_bin_refine_scheme(o);
}
static inline void set_custom_bin_refine_factors(struct config_options *o, const int f[3])
{
    set_bin_refine_factors(o, f);
    set_bin_refine_scheme(o, BINNING_CUST);
}
static inline void reset_bin_refine_factors(struct config_options *o)
{
    o->bin_refine_factors[0] = 2;
    o->bin_refine_factors[1] = 2;
    o->bin_refine_factors[2] = 1;
    reset_bin_refine_scheme(o);
}
static inline void set_max_cells(struct config_options *o, const int max)
{
    if (max <= 0) {
        fprintf(stderr, "Warning: Max. cells per dimension was requested to be set to a negative number = %d...returning\n", max);
        return;
    }
    if (max > 32767)
        fprintf(stderr, "Warning: Max cells per dimension is a 2-byte integer and can not hold supplied value of %d\n", max);
    o->max_cells_per_dim = (uint16_t)max;
}
static inline void reset_max_cells(struct config_options *o) { o->max_cells_per_dim = NLATMAX; }

/* Defaults of the reference built with theory.options (PERIODIC,
 * ENABLE_MIN_SEP_OPT, COPY_PARTICLES) and the library's float type = float. */
static inline struct config_options get_config_options(void)
{
    struct config_options o;
    memset(&o, 0, sizeof o);
    snprintf(o.version, sizeof(o.version) - 1, "%s", CORRFUNC_API_VERSION);
    o.boxsize_x = BOXSIZE_NOTGIVEN;
    o.boxsize_y = BOXSIZE_NOTGIVEN;
    o.boxsize_z = BOXSIZE_NOTGIVEN;
#ifdef DOUBLE_PREC
    o.float_type = sizeof(double);
#else
    o.float_type = sizeof(float);
#endif
#ifndef SILENT
    o.verbose = 1;
#endif
#ifdef OUTPUT_RPAVG
    o.need_avg_sep = 1;
#endif
    o.periodic = 1;
#if defined(__AVX512F__)
    o.instruction_set = AVX512F;
#elif defined(__AVX2__)
    o.instruction_set = AVX2;
#elif defined(__AVX__)
    o.instruction_set = AVX;
#elif defined(__SSE4_2__)
    o.instruction_set = SSE42;
#else
    o.instruction_set = FALLBACK;
#endif
    o.enable_min_sep_opt = 1;
    o.copy_particles = 1;
    o.totncells_timings = 0;
    o.cell_timings = NULL;
    reset_max_cells(&o);
    reset_bin_refine_factors(&o);
    return o;
}

static inline int get_num_weights_by_method(const weight_method_t m) { return m == PAIR_PRODUCT ? 1 : 0; }

static inline int get_weight_method_by_name(const char *name, weight_method_t *m)
{
    if (name == NULL || strcmp(name, "") == 0) { *m = NONE; return EXIT_SUCCESS; }
    if (strcmp(name, "pair_product") == 0 || strcmp(name, "p") == 0) { *m = PAIR_PRODUCT; return EXIT_SUCCESS; }
    return EXIT_FAILURE;
}

static inline struct extra_options get_extra_options(const weight_method_t m)
{
    struct extra_options e;
    memset(&e, 0, sizeof e);
    e.weight_method = m;
    e.weights0.num_weights = get_num_weights_by_method(m);
    e.weights1.num_weights = e.weights0.num_weights;
    return e;
}

static inline void free_cell_timings(struct config_options *o)
{
    if (o->totncells_timings > 0 && o->cell_timings != NULL) free(o->cell_timings);
    o->cell_timings = NULL;
    o->totncells_timings = 0;
}

/* ---- results structs (reference theory/<mode>/countpairs*.h) ---- */
typedef struct {
    uint64_t *npairs;
    double *rupp;
    double *rpavg;
    double *weightavg;
    int nbin;
} results_countpairs;

typedef struct {
    uint64_t *npairs;
    double *rupp;
    double *rpavg;
    double *weightavg;
    double pimax;
    int nbin;
    int npibin;
} results_countpairs_rp_pi;

typedef struct {
    uint64_t *npairs;
    double *supp;
    double *savg;
    double mu_max;
    double mu_min;
    double *weightavg;
    int nsbin;
    int nmu_bins;
} results_countpairs_s_mu;

typedef struct {
    uint64_t *npairs;
    double *xi;
    double *rupp;
    double *ravg;
    double *weightavg;
    int nbin;
} results_countpairs_xi;

typedef struct {
    uint64_t *npairs;
    double *wp;
    double *rupp;
    double *rpavg;
    double *weightavg;
    double pimax;
    int nbin;
} results_countpairs_wp;

/* ---- entry points ---- */

/* 3-D pair counts DD(r). Replaces theory/DD/countpairs.h:28-35 (countpairs.c:34-77). */
int countpairs(const int64_t ND1, void *X1, void *Y1, void *Z1,
               const int64_t ND2, void *X2, void *Y2, void *Z2,
               const int numthreads, const int autocorr, const char *binfile,
               results_countpairs *results, struct config_options *options,
               struct extra_options *extra);
/* Replaces theory/DD/countpairs.c:19-31. */
void free_results(results_countpairs *results);

/* 2-D counts DD(rp, pi). Replaces theory/DDrppi/countpairs_rp_pi.h:29-37. */
int countpairs_rp_pi(const int64_t ND1, void *X1, void *Y1, void *Z1,
                     const int64_t ND2, void *X2, void *Y2, void *Z2,
                     const int numthreads, const int autocorr, const char *binfile,
                     const double pimax, results_countpairs_rp_pi *results,
                     struct config_options *options, struct extra_options *extra);
/* Replaces theory/DDrppi/countpairs_rp_pi.c (free_results_rp_pi). */
void free_results_rp_pi(results_countpairs_rp_pi *results);

/* 2-D counts DD(s, mu). Replaces theory/DDsmu/countpairs_s_mu.h:30-39. */
int countpairs_s_mu(const int64_t ND1, void *X1, void *Y1, void *Z1,
                    const int64_t ND2, void *X2, void *Y2, void *Z2,
                    const int numthreads, const int autocorr, const char *sbinfile,
                    const double mu_max, const int nmu_bins,
                    results_countpairs_s_mu *results, struct config_options *options,
                    struct extra_options *extra);
/* Replaces theory/DDsmu/countpairs_s_mu.c (free_results_s_mu). */
void free_results_s_mu(results_countpairs_s_mu *results);

/* Periodic auto-correlation xi(r) with analytic RR. Replaces theory/xi/countpairs_xi.h:30-36. */
int countpairs_xi(const int64_t ND1, void *X1, void *Y1, void *Z1,
                  const double boxsize, const int numthreads, const char *binfile,
                  results_countpairs_xi *results, struct config_options *options,
                  struct extra_options *extra);
/* Replaces theory/xi/countpairs_xi.h:29 (free_results_xi). */
void free_results_xi(results_countpairs_xi *results);

/* Projected correlation w_p(rp) with analytic RR. Replaces theory/wp/countpairs_wp.h:30-37. */
int countpairs_wp(const int64_t ND1, void *X1, void *Y1, void *Z1,
                  const double boxsize, const int numthreads, const char *binfile,
                  const double pimax, results_countpairs_wp *result,
                  struct config_options *options, struct extra_options *extra);
/* Replaces theory/wp/countpairs_wp.c (free_results_wp). */
void free_results_wp(results_countpairs_wp *results);

#ifdef __cplusplus
}
#endif

#endif /* CORRFUNC_THEORY_H */
