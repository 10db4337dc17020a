/*
 * scft_oracle.h — CPU FP64 restatement of SCFT.c, TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load this code, and only as the reference results are compared with
 * or as the timed CPU baseline.
 * The product (libscft) never links or calls it.
 */
#ifndef SCFT_ORACLE_H
#define SCFT_ORACLE_H

#include <complex.h>
#include "../include/scft.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_state;

/* SCFT.c:84-396, 718-734 (returns SCFT_EVOLUME when the volume check fails). */
int  orc_derive(const scft_params *p, scft_derived *d);

/* FFTW semantics: unnormalised c2c DFT, sign -1 forward, +1 backward,
 * row-major (n0, n1, n2) with the last index fastest. */
void orc_fft3d(int n0, int n1, int n2, double complex *data, int sign);
/* O(n^2) long-double DFT of one line, for checking orc_fft3d. */
void orc_dft_brute(int n, const double complex *in, double complex *out, int sign);

orc_state *orc_create(const scft_params *p);
void orc_destroy(orc_state *st);
const scft_derived *orc_get_derived(const orc_state *st);
void orc_set_fields(orc_state *st, const double *w13, const double *pfield);
void orc_get_fields(const orc_state *st, double *w13, double *pfield);
void orc_set_mask(orc_state *st, const double *mask3);
void orc_get_densities(const orc_state *st, double *phi13, double *phi_tot);
/* q (which = 0) or q-dagger (which = 1) of species sp at contour index s. */
int  orc_get_propagator(const orc_state *st, int sp, int which, int s, double *out);

/* POSIX drand48 state (seed48 order: low, mid, high 16-bit words). */
void orc_seed48(orc_state *st, const unsigned short seed[3]);
void orc_get_seed(const orc_state *st, unsigned short seed[3]);
double orc_drand48(orc_state *st);
double orc_gauss(orc_state *st);     /* funGaussDev, SCFT.c:1769-1782 */
void orc_gauss_n(orc_state *st, long n, double *out);   /* n successive deviates (out may be NULL) */

/* One iteration of SCFT.c:1090-1580; iter is the value after iIter++. */
int  orc_iterate(orc_state *st, unsigned long iter, scft_iter_out *out);

/* Single propagator block (funPDESolver, SCFT.c:2310-2412) on a caller
 * array q (ns_eff*V doubles, layout [s][idx]), for unit checks.
 * seed_mode: 0 continue from q[start], 1 = funPDESolver's seedBool. */
void orc_pde_block(orc_state *st, int start, int stop, int ns_eff, double *q,
                   const double *w, int seed_mode, unsigned long iter);

/* The whole reference main() with its file I/O in directory dir:
 * iterations run from counter.dat's value + 1 up to iter_max + 1
 * (SCFT.c:1088-1090).  Returns 0 or a negative error. */
int  orc_main(const scft_params *p, unsigned long iter_max, int d_wrt,
              int d_wrt_energies, const char *dir);

#ifdef __cplusplus
}
#endif
#endif
