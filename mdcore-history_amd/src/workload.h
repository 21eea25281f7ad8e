/* Synthetic systems for drivers, tests and the benchmark (workload.c). */
#ifndef MD_WORKLOAD_H
#define MD_WORKLOAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

double md_uniform(uint64_t *state);
int md_gen_fcc(int nx, int ny, int nz, double a, double offset, double *x);
int md_gen_maxwell(int N, const double *mass, double kT, uint64_t seed, double *v);
int md_gen_water(int nmol, const double *box, int nlx, int nly, int nlz, double roh, double theta, double jitter,
                 uint64_t seed, double *x);
int md_gen_rsa(int N, const double *box, double rmin, uint64_t seed, long long max_trials, double *x);
int md_gen_types(int N, int ntypes, int first, uint64_t seed, int *type);
int md_gen_uniform(int n, double lo, double hi, uint64_t seed, double *out);
long long md_sample_pairs(const double *x, const int *type, int N, const double *box, double cutoff,
                          long long nmax, uint64_t seed, double *r, int *ti, int *tj);

#ifdef __cplusplus
}
#endif

#endif
