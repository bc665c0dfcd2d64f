// Command-line programs DD, DDrppi, DDsmu, wp, xi, DDrppi_mocks and DDsmu_mocks
// on top of the engine library: same arguments, defaults, checks and output
// columns as the reference's theory/DD/DD.c, DDrppi/DDrppi.c, DDsmu/DDsmu.c,
// wp/wp.c, xi/xi.c (built with theory.options: periodic, positions read as float
// unless DOUBLE_PREC, averages printed only with OUTPUT_RPAVG) and
// mocks/DDrppi_mocks/DDrppi_mocks.c, mocks/DDsmu_mocks/DDsmu_mocks.c (built with
// mocks.options: DOUBLE_PREC and OUTPUT_RPAVG). One source, built once per
// program with CF_CLI set to the program's number.
#include <cassert>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sys/time.h>

#include "../../include/corrfunc_io.h"
#include "../../include/corrfunc_mocks.h"
#include "../../include/corrfunc_theory.h"

#define CLI_DD 1
#define CLI_DDRPPI 2
#define CLI_DDSMU 3
#define CLI_WP 4
#define CLI_XI 5
#define CLI_DDRPPI_MOCKS 6
#define CLI_DDSMU_MOCKS 7
#ifndef CF_CLI
#error "CF_CLI selects the program (1 DD, 2 DDrppi, 3 DDsmu, 4 wp, 5 xi, 6 DDrppi_mocks, 7 DDsmu_mocks)"
#endif

#ifdef DOUBLE_PREC
typedef double DOUBLE;
#else
typedef float DOUBLE;
#endif

namespace {

double seconds_since(const timeval& t0)
{
    timeval t1;
    gettimeofday(&t1, nullptr);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-6 * (double)(t1.tv_usec - t0.tv_usec);
}

#if CF_CLI == CLI_DD || CF_CLI == CLI_DDRPPI || CF_CLI == CLI_DDSMU || CF_CLI == CLI_DDRPPI_MOCKS || \
    CF_CLI == CLI_DDSMU_MOCKS
constexpr bool kTwoSets = true;
#else
constexpr bool kTwoSets = false;
#endif

struct Program {
    const char* name;
    const char* usage;
    const char* args[9];
    int nargs;
};

#if CF_CLI == CLI_DD
const Program kProg = {"DD",
                       "DD file1 format1 file2 format2 binfile boxsize numthreads [weight_method weights_file1 "
                       "weights_format1 [weights_file2 weights_format2]] > DDfile",
                       {"file1", "format1", "file2", "format2", "binfile", "boxsize", "Nthreads"},
                       7};
#elif CF_CLI == CLI_DDRPPI
const Program kProg = {"DDrppi",
                       "DDrppi file1 format1 file2 format2 binfile pimax boxsize numthreads [weight_method "
                       "weights_file1 weights_format1 [weights_file2 weights_format2]] > DDfile",
                       {"file1", "format1", "file2", "format2", "binfile", "pimax", "boxsize", "Nthreads"},
                       8};
#elif CF_CLI == CLI_DDSMU
const Program kProg = {"DDsmu",
                       "DDsmu file1 format1 file2 format2 sbinfile mu_max nmu_bins boxsize numthreads "
                       "[weight_method weights_file1 weights_format1 [weights_file2 weights_format2]] > DDfile",
                       {"file1", "format1", "file2", "format2", "sbinfile", "mu_max", "nmu_bins", "boxsize",
                        "Nthreads"},
                       9};
#elif CF_CLI == CLI_DDRPPI_MOCKS
const Program kProg = {"DDrppi_mocks",
                       "DDrppi_mocks file1 format1 file2 format2 binfile pimax cosmology numthreads [weight_method "
                       "weights_file1 weights_format1 [weights_file2 weights_format2]] > DDfile",
                       {"file1", "format1", "file2", "format2", "binfile", "pimax", "cosmology flag", "numthreads"},
                       8};
#elif CF_CLI == CLI_DDSMU_MOCKS
const Program kProg = {"DDsmu_mocks",
                       "DDsmu_mocks file1 format1 file2 format2 sbinfile mu_max nmu_bins cosmology numthreads "
                       "[weight_method weights_file1 weights_format1 [weights_file2 weights_format2]] > DDfile",
                       {"file1", "format1", "file2", "format2", "sbinfile", "mu_max", "nmu_bins", "cosmology flag",
                        "numthreads"},
                       9};
#elif CF_CLI == CLI_WP
const Program kProg = {"wp",
                       "wp boxsize file format binfile pimax Nthreads [weight_method weights_file weights_format] "
                       "> wpfile",
                       {"boxsize", "file", "format", "binfile", "pimax", "Nthreads"},
                       6};
#else
const Program kProg = {"xi",
                       "xi boxsize file format binfile Nthreads [weight_method weights_file weights_format] > xifile",
                       {"boxsize", "file", "format", "binfile", "Nthreads"},
                       5};
#endif

const char* const kOptArgs[] = {"weight_method", "weights_file1", "weights_format1", "weights_file2",
                                "weights_format2"};

void print_help()
{
    fprintf(stderr, "=========================================================================\n");
    fprintf(stderr, "   --- %s\n", kProg.usage);
    fprintf(stderr, "     formats: a=ascii, c=csv, f=fast-food; weight_method: 'pair_product' or 'none'\n");
#ifdef OUTPUT_RPAVG
    fprintf(stderr, "\tOutput RPAVG = True\n");
#else
    fprintf(stderr, "\tOutput RPAVG = False\n");
#endif
#ifdef DOUBLE_PREC
    fprintf(stderr, "\tPrecision = double\n");
#else
    fprintf(stderr, "\tPrecision = float\n");
#endif
    fprintf(stderr, "=========================================================================\n");
}

// Reads num_weights weight columns of a catalog; false (message printed) when the
// row count differs from the positions' (DD.c:171-181).
bool read_weights(const char* file, const char* format, int num_weights, int64_t nd, const char* posfile,
                  DOUBLE** w, double& read_time)
{
    timeval t0;
    gettimeofday(&t0, nullptr);
    const int64_t n = read_columns_into_array(file, format, sizeof(DOUBLE), num_weights, (void**)w);
    read_time += seconds_since(t0);
    if (n != nd) {
        fprintf(stderr, "Error: read %" PRId64 " lines from %s, but read %" PRId64 " from %s\n", n, file, nd,
                posfile);
        return false;
    }
    return true;
}

}  // namespace

int main(int argc, char* argv[])
{
    timeval t_start, t0;
    gettimeofday(&t_start, nullptr);
    const int nargs = kProg.nargs;
    if (argc < nargs + 1) {
        print_help();
        fprintf(stderr, "\nFound: %d parameters\n ", argc - 1);
        for (int i = 1; i < argc; i++) {
            if (i <= nargs)
                fprintf(stderr, "\t\t %s = `%s' \n", kProg.args[i - 1], argv[i]);
            else
                fprintf(stderr, "\t\t <> = `%s' \n", argv[i]);
        }
        fprintf(stderr, "\nMissing required parameters \n");
        for (int i = argc; i <= nargs; i++) fprintf(stderr, "\t\t %s = `?'\n", kProg.args[i - 1]);
        return EXIT_FAILURE;
    }
    const int nopt = argc - (nargs + 1);
    if (nopt != 0 && nopt != 3 && !(kTwoSets && nopt == 5)) {
        print_help();
        fprintf(stderr, "\nFound: %d optional arguments; must be 0 (no weights), 3 (for one set of weights)%s\n ",
                nopt, kTwoSets ? " or 5 (for two sets)" : "");
        return EXIT_FAILURE;
    }

    // positional arguments
    const char *file1, *format1, *file2, *format2, *binfile;
    double boxsize;
    int nthreads;
    DOUBLE pimax = 40.0;
    double mu_max = 1.0;
    int nmu_bins = 1;
    int cosmology = 1;
    (void)pimax;
    (void)mu_max;
    (void)nmu_bins;
    (void)cosmology;
#if CF_CLI == CLI_WP || CF_CLI == CLI_XI
    boxsize = atof(argv[1]);
    file1 = file2 = argv[2];
    format1 = format2 = argv[3];
    binfile = argv[4];
#if CF_CLI == CLI_WP
    sscanf(argv[5], sizeof(DOUBLE) == 8 ? "%lf" : "%f", &pimax);
#endif
    nthreads = atoi(argv[nargs]);
#else
    file1 = argv[1];
    format1 = argv[2];
    file2 = argv[3];
    format2 = argv[4];
    binfile = argv[5];
#if CF_CLI == CLI_DD
    boxsize = atof(argv[6]);
#elif CF_CLI == CLI_DDRPPI_MOCKS
    sscanf(argv[6], sizeof(DOUBLE) == 8 ? "%lf" : "%f", &pimax);
    cosmology = atoi(argv[7]);
    boxsize = 0.0;
#elif CF_CLI == CLI_DDSMU_MOCKS
    sscanf(argv[6], "%lf", &mu_max);
    sscanf(argv[7], "%d", &nmu_bins);
    cosmology = atoi(argv[8]);
    boxsize = 0.0;
#elif CF_CLI == CLI_DDRPPI
    sscanf(argv[6], sizeof(DOUBLE) == 8 ? "%lf" : "%f", &pimax);
    boxsize = atof(argv[7]);
#else
    sscanf(argv[6], "%lf", &mu_max);
    nmu_bins = atoi(argv[7]);
    boxsize = atof(argv[8]);
#endif
    nthreads = atoi(argv[nargs]);
#endif
    if (nthreads < 1) {
        fprintf(stderr, "Error: Nthreads must be at least 1...returning\n");
        return EXIT_FAILURE;
    }

    weight_method_t weight_method = NONE;
    int num_weights = 0;
    const char *wfile1 = nullptr, *wformat1 = nullptr, *wfile2 = nullptr, *wformat2 = nullptr;
    if (nopt >= 3) {
        if (get_weight_method_by_name(argv[nargs + 1], &weight_method) != EXIT_SUCCESS) {
            fprintf(stderr, "Error: Unknown weight method \"%s\"\n", argv[nargs + 1]);
            return EXIT_FAILURE;
        }
        num_weights = get_num_weights_by_method(weight_method);
        wfile1 = argv[nargs + 2];
        wformat1 = argv[nargs + 3];
    }
    if (nopt >= 5) {
        wfile2 = argv[nargs + 4];
        wformat2 = argv[nargs + 5];
    }

    fprintf(stderr, "Running `%s' with the parameters \n", argv[0]);
    fprintf(stderr, "\n\t\t -------------------------------------\n");
    for (int i = 1; i < argc; i++)
        fprintf(stderr, "\t\t %-10s = %s \n", i <= nargs ? kProg.args[i - 1] : kOptArgs[i - nargs - 1], argv[i]);
    fprintf(stderr, "\t\t -------------------------------------\n");

    // catalogs
    double read_time = 0.0;
    DOUBLE *x1 = nullptr, *y1 = nullptr, *z1 = nullptr, *x2 = nullptr, *y2 = nullptr, *z2 = nullptr;
    DOUBLE *w1[MAX_NUM_WEIGHTS] = {nullptr}, *w2[MAX_NUM_WEIGHTS] = {nullptr};
    gettimeofday(&t0, nullptr);
    const int64_t ND1 = read_positions(file1, format1, sizeof(DOUBLE), 3, &x1, &y1, &z1);
    read_time += seconds_since(t0);
    if (ND1 < 0) return EXIT_FAILURE;
    const int autocorr = kTwoSets ? (strcmp(file1, file2) == 0) : 1;
#if CF_CLI == CLI_XI
    for (int64_t i = 0; i < ND1; i++) {
        assert(x1[i] >= 0.0 && x1[i] <= boxsize && "xpos is within limits [0, boxsize]");
        assert(y1[i] >= 0.0 && y1[i] <= boxsize && "ypos is within limits [0, boxsize]");
        assert(z1[i] >= 0.0 && z1[i] <= boxsize && "zpos is within limits [0, boxsize]");
    }
#endif
    if (wfile1 && !read_weights(wfile1, wformat1, num_weights, ND1, file1, w1, read_time)) return EXIT_FAILURE;
    int64_t ND2 = ND1;
    if (!autocorr) {
        gettimeofday(&t0, nullptr);
        ND2 = read_positions(file2, format2, sizeof(DOUBLE), 3, &x2, &y2, &z2);
        read_time += seconds_since(t0);
        if (ND2 < 0) return EXIT_FAILURE;
        if (wfile2 && !read_weights(wfile2, wformat2, num_weights, ND2, file2, w2, read_time)) return EXIT_FAILURE;
    } else {
        x2 = x1;
        y2 = y1;
        z2 = z1;
        for (int w = 0; w < MAX_NUM_WEIGHTS; w++) w2[w] = w1[w];
    }

    // pair counts
    gettimeofday(&t0, nullptr);
    struct config_options options = get_config_options();
    options.boxsize = boxsize;
#if CF_CLI == CLI_WP
    options.verbose = 1;
#endif
    struct extra_options extra = get_extra_options(weight_method);
    for (int w = 0; w < num_weights; w++) {
        extra.weights0.weights[w] = (void*)w1[w];
        extra.weights1.weights[w] = (void*)w2[w];
    }
    int status;
#if CF_CLI == CLI_DD
    results_countpairs res;
    status = countpairs(ND1, x1, y1, z1, ND2, x2, y2, z2, nthreads, autocorr, binfile, &res, &options, &extra);
#elif CF_CLI == CLI_DDRPPI
    results_countpairs_rp_pi res;
    status = countpairs_rp_pi(ND1, x1, y1, z1, ND2, x2, y2, z2, nthreads, autocorr, binfile, pimax, &res, &options,
                              &extra);
#elif CF_CLI == CLI_DDSMU
    results_countpairs_s_mu res;
    status = countpairs_s_mu(ND1, x1, y1, z1, ND2, x2, y2, z2, nthreads, autocorr, binfile, mu_max, nmu_bins, &res,
                             &options, &extra);
#elif CF_CLI == CLI_DDRPPI_MOCKS
    results_countpairs_mocks res;
    status = countpairs_mocks(ND1, x1, y1, z1, ND2, x2, y2, z2, nthreads, autocorr, binfile, pimax, cosmology, &res,
                              &options, &extra);
#elif CF_CLI == CLI_DDSMU_MOCKS
    results_countpairs_mocks_s_mu res;
    status = countpairs_mocks_s_mu(ND1, x1, y1, z1, ND2, x2, y2, z2, nthreads, autocorr, binfile, mu_max, nmu_bins,
                                   cosmology, &res, &options, &extra);
#elif CF_CLI == CLI_WP
    results_countpairs_wp res;
    status = countpairs_wp(ND1, x1, y1, z1, boxsize, nthreads, binfile, pimax, &res, &options, &extra);
#else
    results_countpairs_xi res;
    status = countpairs_xi(ND1, x1, y1, z1, boxsize, nthreads, binfile, &res, &options, &extra);
#endif
    free(x1);
    free(y1);
    free(z1);
    for (int w = 0; w < num_weights; w++) free(w1[w]);
    if (!autocorr) {
        free(x2);
        free(y2);
        free(z2);
        for (int w = 0; w < num_weights; w++) free(w2[w]);
    }
    if (status != EXIT_SUCCESS) return status;
    const double pair_time = seconds_since(t0);

    // results, one line per reported bin (columns as the reference programs print them)
#if CF_CLI == CLI_DD
    double rlow = res.rupp[0];
    for (int i = 1; i < res.nbin; i++) {
        fprintf(stdout, "%e\t%e\t%e\t%12" PRIu64 "\t%e\n", rlow, res.rupp[i], res.rpavg[i], res.npairs[i],
                res.weightavg[i]);
        rlow = res.rupp[i];
    }
    free_results(&res);
#elif CF_CLI == CLI_DDRPPI
    const double dpi = pimax / (double)res.npibin;
    for (int i = 1; i < res.nbin; i++) {
        const double logrp = log10(res.rupp[i]);
        for (int j = 0; j < res.npibin; j++) {
            const int k = i * (res.npibin + 1) + j;
            fprintf(stdout, "%e\t%e\t%e\t%12" PRIu64 "\t%e\n", logrp, (j + 1) * dpi, res.rpavg[k], res.npairs[k],
                    res.weightavg[k]);
        }
    }
    free_results_rp_pi(&res);
#elif CF_CLI == CLI_DDSMU
    double smin = res.supp[0];
    const double dmu = mu_max / (double)nmu_bins;
    for (int i = 1; i < res.nsbin; i++) {
        const double smax = res.supp[i];
        for (int j = 0; j < nmu_bins; j++) {
            const int k = i * (nmu_bins + 1) + j;
            fprintf(stdout, "%e\t%e\t%e\t%12" PRIu64 "\t%e\n", smin, smax, (j + 1) * dmu, res.npairs[k],
                    res.weightavg[k]);
        }
        smin = smax;
    }
    free_results_s_mu(&res);
#elif CF_CLI == CLI_DDRPPI_MOCKS
    const double dpi = pimax / (double)res.npibin;
    for (int i = 1; i < res.nbin; i++) {
        const double logrp = log10(res.rupp[i]);
        for (int j = 0; j < res.npibin; j++) {
            const int k = i * (res.npibin + 1) + j;
            fprintf(stdout, "%10" PRIu64 " %20.8lf %20.8lf  %20.8lf %20.8lf \n", res.npairs[k], res.rpavg[k], logrp,
                    (j + 1) * dpi, res.weightavg[k]);
        }
    }
    free_results_mocks(&res);
#elif CF_CLI == CLI_DDSMU_MOCKS
    const double dmu = mu_max / (double)res.nmu_bins;
    for (int i = 1; i < res.nsbin; i++) {
        const double log_supp = log10(res.supp[i]);
        for (int j = 0; j < res.nmu_bins; j++) {
            const int k = i * (res.nmu_bins + 1) + j;
            fprintf(stdout, "%10" PRIu64 " %20.8lf %20.8lf  %20.8lf %20.8lf \n", res.npairs[k], res.savg[k], log_supp,
                    (j + 1) * dmu, res.weightavg[k]);
        }
    }
    free_results_mocks_s_mu(&res);
#elif CF_CLI == CLI_WP
    double rlow = res.rupp[0];
    for (int i = 1; i < res.nbin; ++i) {
        fprintf(stdout, "%e\t%e\t%e\t%12" PRIu64 "\t%e\t%e\n", rlow, res.rupp[i], res.rpavg[i], res.npairs[i],
                res.wp[i], res.weightavg[i]);
        rlow = res.rupp[i];
    }
    free_results_wp(&res);
#else
    double rlow = res.rupp[0];
    for (int i = 1; i < res.nbin; ++i) {
        fprintf(stdout, "%e\t%e\t%e\t%12" PRIu64 "\t%e\t%e\n", rlow, res.rupp[i], res.ravg[i], res.npairs[i],
                res.xi[i], res.weightavg[i]);
        rlow = res.rupp[i];
    }
    free_results_xi(&res);
#endif
    fprintf(stderr,
            "%s> Done -  ND1=%" PRId64 " ND2=%" PRId64 ". Time taken = %6.2lf seconds. read-in time = %6.2lf seconds "
            "pair-counting time = %6.2lf sec\n",
            kProg.name, ND1, ND2, seconds_since(t_start), read_time, pair_time);
    return EXIT_SUCCESS;
}
