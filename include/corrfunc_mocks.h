/* On-sky (mock catalog) pair counters of the engine library (libcorrfunc_amd.so):
   DD(rp, pi) and DD(s, mu) of catalogs given as (RA, DEC, cz) or (RA, DEC,
   comoving distance), with the line of sight of each pair along the mean of
   the two position vectors.

   Same entry points, structs, argument meaning and error behaviour as the
   reference's mocks/DDrppi_mocks/countpairs_rp_pi_mocks.h:19-41 and
   mocks/DDsmu_mocks/countpairs_s_mu_mocks.h:19-45. Options and weights are the
   struct config_options / struct extra_options of corrfunc_theory.h
   (options->is_comoving_dist: CZ arrays hold comoving distances in Mpc/h;
   otherwise cz in km/s, converted with the cosmology `cosmology`: 1 LasDamas
   (Omega_m 0.25), 2 Planck (Omega_m 0.302), flat).

   NOTE: as in the reference, the input RA / DEC / CZ arrays are modified in
   place when they need fixing (RA in [-180, 180] is shifted by 180, DEC in
   [0, 180] by -90, and redshifts z (max < 10) are multiplied by c). */
#ifndef CORRFUNC_MOCKS_H
#define CORRFUNC_MOCKS_H

#include "corrfunc_theory.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t *npairs;   /* [(nbin+1)*(npibin+1)], index k*(npibin+1)+j */
    double *rupp;       /* [nbin] bin edges */
    double *rpavg;      /* as npairs; 0 unless options->need_avg_sep */
    double *weightavg;  /* as npairs */
    double pimax;
    int nbin;
    int npibin;
} results_countpairs_mocks;

typedef struct {
    uint64_t *npairs;   /* [(nsbin+1)*(nmu_bins+1)], index k*(nmu_bins+1)+j */
    double *supp;       /* [nsbin] bin edges */
    double *savg;
    double mu_max;
    double mu_min;      /* always 0 */
    double *weightavg;
    int nsbin;
    int nmu_bins;
} results_countpairs_mocks_s_mu;

/* Replaces countpairs_mocks (mocks/DDrppi_mocks/countpairs_rp_pi_mocks.h:29-37). */
int countpairs_mocks(const int64_t ND1, void *theta1, void *phi1, void *czD1,
                     const int64_t ND2, void *theta2, void *phi2, void *czD2,
                     const int numthreads, const int autocorr, const char *binfile, const double pimax,
                     const int cosmology, results_countpairs_mocks *results, struct config_options *options,
                     struct extra_options *extra);
/* Replaces free_results_mocks (countpairs_rp_pi_mocks.h:39). */
void free_results_mocks(results_countpairs_mocks *results);

/* Replaces countpairs_mocks_s_mu (mocks/DDsmu_mocks/countpairs_s_mu_mocks.h:30-40). */
int countpairs_mocks_s_mu(const int64_t ND1, void *theta1, void *phi1, void *czD1,
                          const int64_t ND2, void *theta2, void *phi2, void *czD2,
                          const int numthreads, const int autocorr, const char *sbinfile, const double mu_max,
                          const int nmu_bins, const int cosmology, results_countpairs_mocks_s_mu *results,
                          struct config_options *options, struct extra_options *extra);
/* Replaces free_results_mocks_s_mu (countpairs_s_mu_mocks.h:42). */
void free_results_mocks_s_mu(results_countpairs_mocks_s_mu *results);

#ifdef __cplusplus
}
#endif
#endif
