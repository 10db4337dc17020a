/*
 * scft_oracle.c — CPU FP64 restatement of the reference SCFT.c
 * (gahme-glitch/stimuli-responsive-nanodroplets).
 *
 * TEST INFRASTRUCTURE ONLY: the reference that libscft is compared with, and
 * the CPU baseline that bench.py times.  The product never links it.
 *
 * It follows the reference's algorithm literally (complex work arrays that
 * carry across contour steps, the 2(NZ+1) odd extension, the same loop
 * orders for every reduction), with two differences that do not change
 * results beyond round-off: the run parameters are a runtime struct instead
 * of macros, and the DFT is this file's own mixed-radix FFT instead of FFTW
 * (FFTW3 is not installed here).  Propagators are stored [s][idx] instead
 * of the reference's [idx][s]; the layout is not part of the algorithm.
 *
 * Each function cites the reference lines it restates.
 */
#define _GNU_SOURCE
#include "scft_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------ */
/* Setup: SCFT.c:84-396 and the coating rescale of SCFT.c:718-734.     */
/* ------------------------------------------------------------------ */

static int grid_points(double L, double dr)
{
    /* SCFT.c:103-105 */
    if (L == 1.0) return 1;
    return (int)((L / dr) + 0.5);
}

int orc_derive(const scft_params *p, scft_derived *d)
{
    memset(d, 0, sizeof(*d));
    const double eps = 0.000001;                          /* SCFT.c:89 */
    double vsum = 0.0;                                    /* SCFT.c:85-87 */
    for (int i = 0; i < 4; i++) vsum += p->vol_dbk[i];
    for (int i = 0; i < 10; i++) vsum += p->vol_homo[i];
    if (fabs(1.0 - vsum) >= eps) return SCFT_EVOLUME;     /* SCFT.c:91-101 */

    d->nx = grid_points(p->lx, p->dr_target);
    d->ny = grid_points(p->ly, p->dr_target);
    d->nz = grid_points(p->lz, p->dr_target);
    d->nzext = p->mask ? 2 * (d->nz + 1) : d->nz;         /* SCFT.c:680 */

    const int NS = p->ns;
    const double DS = 1.0 / (NS - 1);                     /* parameters.h:20 */

    for (int b = 0; b < 4; b++) {                         /* SCFT.c:120-162 */
        if (p->vol_dbk[b] > 0.0) {
            int n = (int)floor(p->len_dbk[b] * NS);
            if (n % 2 == 0) n = n - 1;
            int nx = (int)floor(p->f_dbk[b] * n) + 1;
            if (nx % 2 == 0) nx = nx - 1;
            d->ns_sp[b] = n;
            d->ns_x[b] = nx;
            d->ns_y[b] = n - nx;
            d->alpha[b] = (double)((n - 1) * DS);
        }
    }
    for (int h = 0; h < 10; h++) {                        /* SCFT.c:165-253 */
        if (p->vol_homo[h] > 0.0) {
            int n = (int)floor(p->len_homo[h] * NS);
            if (n % 2 == 0 || n > NS) n = n - 1;
            d->ns_sp[4 + h] = n;
            d->alpha[4 + h] = (double)((n - 1) * DS);
        }
    }
    if (p->mask) {                                        /* SCFT.c:255-288 */
        for (int s = 0; s < 3; s++) {
            if (p->vol_s[s] > 0.0) {
                int n = (int)floor(p->len_s[s] * NS);
                if (n % 2 == 0) n = n - 1;
                d->ns_sp[14 + s] = n;
                d->alpha[14 + s] = (double)((n - 1) * DS);
            }
        }
    }

    /* phi averages, SCFT.c:292-301 */
    double *pa = d->phi_avg;
    pa[0] = p->vol_homo[0] + p->f_dbk[0] * p->vol_dbk[0];
    pa[1] = p->vol_homo[1] + (1.0 - p->f_dbk[0]) * p->vol_dbk[0];
    pa[2] = p->vol_homo[2];
    pa[3] = p->vol_homo[3] + p->f_dbk[1] * p->vol_dbk[1];
    pa[4] = p->vol_homo[4] + (1.0 - p->f_dbk[1]) * p->vol_dbk[1];
    pa[5] = p->vol_homo[5] + p->f_dbk[2] * p->vol_dbk[2];
    pa[6] = p->vol_homo[6] + (1.0 - p->f_dbk[2]) * p->vol_dbk[2];
    pa[7] = p->vol_homo[7] + p->f_dbk[3] * p->vol_dbk[3];
    pa[8] = p->vol_homo[8] + (1.0 - p->f_dbk[3]) * p->vol_dbk[3];
    pa[9] = p->vol_homo[9];

    /* chi_eff: copy, then zero every chi of a chemistry that is absent
     * (SCFT.c:107-119, 303-396).  A..J are zeroed when phi_avg <= eps;
     * S1..S3 when MASK is 0 or their volume is <= eps. */
    for (int f = 0; f < 13; f++)
        for (int g = 0; g < 13; g++)
            d->chi_eff[f][g] = (f == g) ? 0.0 : p->chi[f][g];
    int zero_field[13] = {0};
    for (int f = 0; f < 10; f++)
        if (fabs(pa[f]) <= eps) zero_field[f] = 1;
    if (p->mask) {
        for (int s = 0; s < 3; s++) {
            if (fabs(p->vol_s[s]) <= eps) zero_field[10 + s] = 1;
            else pa[10 + s] = p->vol_s[s];
        }
    } else {
        for (int s = 0; s < 3; s++) { zero_field[10 + s] = 1; pa[10 + s] = 0.0; }
    }
    for (int f = 0; f < 13; f++)
        if (zero_field[f])
            for (int g = 0; g < 13; g++) { d->chi_eff[f][g] = 0.0; d->chi_eff[g][f] = 0.0; }

    /* coating rescale, SCFT.c:718-734 */
    d->coating_vol = 0.0;
    if (p->mask) {
        d->coating_vol = p->vol_s[0] + p->vol_s[1] + p->vol_s[2];
        for (int f = 0; f < 10; f++) pa[f] = pa[f] * (1.0 - d->coating_vol);
    }

    /* active set of funUpdate, SCFT.c:2001-2042 (brush volumes count even
     * when MASK is 0) */
    double fvol[13];
    for (int f = 0; f < 10; f++) fvol[f] = p->vol_homo[f];
    for (int s = 0; s < 3; s++) fvol[10 + s] = p->vol_s[s];
    int act[13] = {0};
    if (p->vol_dbk[0] > 0.0) { act[0] = 1; act[1] = 1; }
    if (p->vol_dbk[1] > 0.0) { act[3] = 1; act[4] = 1; }
    if (p->vol_dbk[2] > 0.0) { act[5] = 1; act[6] = 1; }
    if (p->vol_dbk[3] > 0.0) { act[7] = 1; act[8] = 1; }
    d->n_active = 0;
    for (int f = 0; f < 13; f++) {
        if (fvol[f] > 0.0 || act[f]) { d->active[f] = 1; d->n_active++; }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* FFT with FFTW's conventions (stands in for fftw_execute_dft,         */
/* SCFT.c:2493, 2510, 2595, 2612).  Mixed-radix decimation in time,     */
/* generic O(p^2) butterflies for every prime factor p.                 */
/* ------------------------------------------------------------------ */

typedef struct {
    int n;
    cplx *w;      /* w[e] = exp(sign*2*pi*i*e/n) */
    int sign;
} fft_plan1;

static void plan1_init(fft_plan1 *pl, int n, int sign)
{
    pl->n = n;
    pl->sign = sign;
    pl->w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int e = 0; e < n; e++) {
        double a = 2.0 * M_PI * (double)e / (double)n;
        pl->w[e] = cos(a) + sign * sin(a) * I;
    }
}

static int smallest_factor(int n)
{
    if (n % 4 == 0) return 4;
    for (int p = 2; p * p <= n; p++)
        if (n % p == 0) return p;
    return n;
}

/* y[0..n) = DFT_n of x[0], x[xs], x[2xs], ...; step = N/n into the table */
static void fft_rec(const fft_plan1 *pl, int n, const cplx *x, int xs, cplx *y, int step)
{
    if (n == 1) { y[0] = x[0]; return; }
    int p = smallest_factor(n);
    int m = n / p;
    for (int r = 0; r < p; r++)
        fft_rec(pl, m, x + (size_t)r * xs, xs * p, y + (size_t)r * m, step * p);
    const int N = pl->n;
    if (p == 2) {
        for (int k = 0; k < m; k++) {
            cplx a = y[k], b = y[m + k] * pl->w[(size_t)k * step];
            y[k] = a + b;
            y[m + k] = a - b;
        }
        return;
    }
    if (p == 4) {
        const cplx jj = pl->sign * I;   /* w_4^1 = exp(sign*i*pi/2) */
        for (int k = 0; k < m; k++) {
            cplx a0 = y[k];
            cplx a1 = y[m + k] * pl->w[(size_t)k * step];
            cplx a2 = y[2 * m + k] * pl->w[(size_t)2 * k * step];
            cplx a3 = y[3 * m + k] * pl->w[(size_t)3 * k * step];
            cplx s02 = a0 + a2, d02 = a0 - a2, s13 = a1 + a3, d13 = (a1 - a3) * jj;
            y[k] = s02 + s13;
            y[m + k] = d02 + d13;
            y[2 * m + k] = s02 - s13;
            y[3 * m + k] = d02 - d13;
        }
        return;
    }
    cplx t[64];
    cplx *tp = (p <= 64) ? t : (cplx *)malloc(sizeof(cplx) * (size_t)p);
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) {
            long e = ((long)r * k) % n;
            tp[r] = y[(size_t)r * m + k] * pl->w[(e * step) % N];
        }
        for (int q = 0; q < p; q++) {
            cplx acc = 0.0;
            for (int r = 0; r < p; r++) {
                long e = ((long)r * q) % p;
                acc += tp[r] * pl->w[(e * m * step) % N];
            }
            y[(size_t)q * m + k] = acc;
        }
    }
    if (tp != t) free(tp);
}

static void fft_line(const fft_plan1 *pl, cplx *line, cplx *tmp)
{
    fft_rec(pl, pl->n, line, 1, tmp, 1);
    memcpy(line, tmp, sizeof(cplx) * (size_t)pl->n);
}

typedef struct {
    int n[3];
    fft_plan1 pl[3];
    cplx *buf, *tmp;
} fft_plan3;

static void plan3_init(fft_plan3 *P, int n0, int n1, int n2, int sign)
{
    P->n[0] = n0; P->n[1] = n1; P->n[2] = n2;
    int mx = 1;
    for (int a = 0; a < 3; a++) {
        plan1_init(&P->pl[a], P->n[a], sign);
        if (P->n[a] > mx) mx = P->n[a];
    }
    P->buf = (cplx *)malloc(sizeof(cplx) * (size_t)mx);
    P->tmp = (cplx *)malloc(sizeof(cplx) * (size_t)mx);
}

static void plan3_free(fft_plan3 *P)
{
    for (int a = 0; a < 3; a++) free(P->pl[a].w);
    free(P->buf);
    free(P->tmp);
}

static void plan3_exec(fft_plan3 *P, cplx *d)
{
    const int n0 = P->n[0], n1 = P->n[1], n2 = P->n[2];
    if (n2 > 1)
        for (size_t l = 0; l < (size_t)n0 * n1; l++)
            fft_line(&P->pl[2], d + l * n2, P->tmp);
    if (n1 > 1)
        for (int a = 0; a < n0; a++)
            for (int c = 0; c < n2; c++) {
                cplx *base = d + (size_t)a * n1 * n2 + c;
                for (int b = 0; b < n1; b++) P->buf[b] = base[(size_t)b * n2];
                fft_line(&P->pl[1], P->buf, P->tmp);
                for (int b = 0; b < n1; b++) base[(size_t)b * n2] = P->buf[b];
            }
    if (n0 > 1)
        for (size_t c = 0; c < (size_t)n1 * n2; c++) {
            cplx *base = d + c;
            size_t st = (size_t)n1 * n2;
            for (int a = 0; a < n0; a++) P->buf[a] = base[a * st];
            fft_line(&P->pl[0], P->buf, P->tmp);
            for (int a = 0; a < n0; a++) base[a * st] = P->buf[a];
        }
}

void orc_fft3d(int n0, int n1, int n2, double complex *data, int sign)
{
    fft_plan3 P;
    plan3_init(&P, n0, n1, n2, sign);
    plan3_exec(&P, data);
    plan3_free(&P);
}

void orc_dft_brute(int n, const double complex *in, double complex *out, int sign)
{
    for (int k = 0; k < n; k++) {
        long double re = 0.0L, im = 0.0L;
        for (int j = 0; j < n; j++) {
            long e = ((long)j * k) % n;
            long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)e / (long double)n;
            long double c = cosl(a), s = sign * sinl(a);
            long double xr = creal(in[j]), xi = cimag(in[j]);
            re += xr * c - xi * s;
            im += xr * s + xi * c;
        }
        out[k] = (double)re + (double)im * I;
    }
}

/* ------------------------------------------------------------------ */
/* State                                                               */
/* ------------------------------------------------------------------ */

struct orc_state {
    scft_params p;
    scft_derived d;
    int NX, NY, NZ, N;            /* N: z length of the transform grid */
    size_t V, Vext;
    double *w[13], *pf;
    double *mask[3];
    double *q[17], *qd[17];       /* [s][idx] */
    double *phi_x[4], *phi_y[4], *phi_h[10], *phi_s[3];
    double *tot[10], *phi_tot;
    double *expArg;               /* aExpArg, SCFT.c:983-1023 */
    fft_plan3 fwd, rev;
    unsigned long long x48;       /* drand48 state */
};

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

const scft_derived *orc_get_derived(const orc_state *st) { return &st->d; }

orc_state *orc_create(const scft_params *p)
{
    orc_state *st = (orc_state *)calloc(1, sizeof(*st));
    st->p = *p;
    if (orc_derive(p, &st->d) != 0) { free(st); return NULL; }
    st->NX = st->d.nx; st->NY = st->d.ny; st->NZ = st->d.nz; st->N = st->d.nzext;
    st->V = (size_t)st->NX * st->NY * st->NZ;
    st->Vext = (size_t)st->NX * st->NY * st->N;
    for (int f = 0; f < 13; f++) st->w[f] = dalloc(st->V);
    st->pf = dalloc(st->V);
    for (int s = 0; s < 3; s++) st->mask[s] = dalloc(st->V);
    for (int sp = 0; sp < 17; sp++) {
        int ns = st->d.ns_sp[sp];
        if (ns > 0) {
            st->q[sp] = dalloc(st->V * ns);
            st->qd[sp] = dalloc(st->V * ns);
            /* q[0] = 1, SCFT.c:881-981 */
            for (size_t i = 0; i < st->V; i++) { st->q[sp][i] = 1.0; st->qd[sp][i] = 1.0; }
        }
    }
    for (int b = 0; b < 4; b++) { st->phi_x[b] = dalloc(st->V); st->phi_y[b] = dalloc(st->V); }
    for (int h = 0; h < 10; h++) { st->phi_h[h] = dalloc(st->V); st->tot[h] = dalloc(st->V); }
    for (int s = 0; s < 3; s++) st->phi_s[s] = dalloc(st->V);
    st->phi_tot = dalloc(st->V);

    /* k-space operator, SCFT.c:983-1023 */
    const double DS = 1.0 / (p->ns - 1);
    const double DX = p->lx / st->NX, DY = p->ly / st->NY, DZ = p->lz / st->NZ;
    const double AX = DS / (DX * DX), AY = DS / (DY * DY), AZ = DS / (DZ * DZ);
    st->expArg = dalloc(st->Vext);
    const int NX = st->NX, NY = st->NY, NZN = st->N;
    for (int k = 0; k < NZN; k++)
        for (int j = 0; j < NY; j++)
            for (int i = 0; i < NX; i++) {
                size_t idx = k + (size_t)NZN * (j + (size_t)NY * i);
                double kx = 2.0 * M_PI * (NX / 2.0 - fabs(i - NX / 2.0)) / NX;
                double ky = 2.0 * M_PI * (NY / 2.0 - fabs(j - NY / 2.0)) / NY;
                double kz = 2.0 * M_PI * (NZN / 2.0 - fabs(k - NZN / 2.0)) / NZN;
                st->expArg[idx] = exp(-AX * pow(kx, 2) - AY * pow(ky, 2) - AZ * pow(kz, 2));
            }
    plan3_init(&st->fwd, NX, NY, NZN, -1);
    plan3_init(&st->rev, NX, NY, NZN, +1);
    st->x48 = 0x1234ABCD330EULL;   /* glibc's initial drand48 state */
    return st;
}

void orc_destroy(orc_state *st)
{
    if (!st) return;
    for (int f = 0; f < 13; f++) free(st->w[f]);
    free(st->pf);
    for (int s = 0; s < 3; s++) { free(st->mask[s]); free(st->phi_s[s]); }
    for (int sp = 0; sp < 17; sp++) { free(st->q[sp]); free(st->qd[sp]); }
    for (int b = 0; b < 4; b++) { free(st->phi_x[b]); free(st->phi_y[b]); }
    for (int h = 0; h < 10; h++) { free(st->phi_h[h]); free(st->tot[h]); }
    free(st->phi_tot);
    free(st->expArg);
    plan3_free(&st->fwd);
    plan3_free(&st->rev);
    free(st);
}

void orc_set_fields(orc_state *st, const double *w13, const double *pfield)
{
    for (int f = 0; f < 13; f++) memcpy(st->w[f], w13 + (size_t)f * st->V, st->V * sizeof(double));
    memcpy(st->pf, pfield, st->V * sizeof(double));
}

void orc_get_fields(const orc_state *st, double *w13, double *pfield)
{
    for (int f = 0; f < 13; f++) memcpy(w13 + (size_t)f * st->V, st->w[f], st->V * sizeof(double));
    if (pfield) memcpy(pfield, st->pf, st->V * sizeof(double));
}

void orc_set_mask(orc_state *st, const double *mask3)
{
    for (int s = 0; s < 3; s++) memcpy(st->mask[s], mask3 + (size_t)s * st->V, st->V * sizeof(double));
}

void orc_get_densities(const orc_state *st, double *phi13, double *phi_tot)
{
    for (int f = 0; f < 10; f++) memcpy(phi13 + (size_t)f * st->V, st->tot[f], st->V * sizeof(double));
    for (int s = 0; s < 3; s++) memcpy(phi13 + (size_t)(10 + s) * st->V, st->phi_s[s], st->V * sizeof(double));
    if (phi_tot) memcpy(phi_tot, st->phi_tot, st->V * sizeof(double));
}

int orc_get_propagator(const orc_state *st, int sp, int which, int s, double *out)
{
    if (sp < 0 || sp >= 17 || st->d.ns_sp[sp] <= 0 || s < 0 || s >= st->d.ns_sp[sp]) return SCFT_EINVAL;
    const double *src = (which ? st->qd[sp] : st->q[sp]) + (size_t)s * st->V;
    memcpy(out, src, st->V * sizeof(double));
    return 0;
}

/* ------------------------------------------------------------------ */
/* drand48 (POSIX 48-bit LCG) and funGaussDev (SCFT.c:1769-1782)        */
/* ------------------------------------------------------------------ */

void orc_seed48(orc_state *st, const unsigned short seed[3])
{
    st->x48 = ((unsigned long long)seed[2] << 32) | ((unsigned long long)seed[1] << 16) | seed[0];
}

void orc_get_seed(const orc_state *st, unsigned short seed[3])
{
    seed[0] = (unsigned short)(st->x48 & 0xffff);
    seed[1] = (unsigned short)((st->x48 >> 16) & 0xffff);
    seed[2] = (unsigned short)((st->x48 >> 32) & 0xffff);
}

double orc_drand48(orc_state *st)
{
    st->x48 = (0x5DEECE66DULL * st->x48 + 0xBULL) & 0xFFFFFFFFFFFFULL;
    return (double)st->x48 * (1.0 / 281474976710656.0);
}

double orc_gauss(orc_state *st)
{
    double x1, x2, w;
    do {
        x1 = 2.0 * orc_drand48(st) - 1.0;
        x2 = 2.0 * orc_drand48(st) - 1.0;
        w = x1 * x1 + x2 * x2;
    } while (w >= 1.0);
    w = sqrt((-2.0 * log(w)) / w);
    return x1 * w;
}

/* n successive deviates of the stream (what one iteration draws when n =
 * n_active * V), into out, or only advancing the state when out is NULL. */
void orc_gauss_n(orc_state *st, long n, double *out)
{
    for (long i = 0; i < n; i++) {
        const double g = orc_gauss(st);
        if (out) out[i] = g;
    }
}

/* ------------------------------------------------------------------ */
/* Propagator: funPDESolver_periodic (SCFT.c:2474-2525) and             */
/* funPDESolver_DirichletBC (SCFT.c:2527-2631).                         */
/* ------------------------------------------------------------------ */

static void step_periodic(orc_state *st, const cplx *E, cplx *in, cplx *work)
{
    const size_t V = st->V;
    for (size_t i = 0; i < V; i++) work[i] = in[i] * E[i];          /* 2479-2490 */
    plan3_exec(&st->fwd, work);                                       /* 2493 */
    for (size_t i = 0; i < V; i++) work[i] = work[i] * st->expArg[i];/* 2496-2507 */
    plan3_exec(&st->rev, work);                                       /* 2510 */
    for (size_t i = 0; i < V; i++) in[i] = work[i] * E[i];          /* 2513-2524 */
}

static void step_dirichlet(orc_state *st, const cplx *E, cplx *in, cplx *ext)
{
    const int NX = st->NX, NY = st->NY, NZ = st->NZ, N = st->N;
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++) {
            cplx *y = ext + (size_t)N * (j + (size_t)NY * i);
            const cplx *x = in + (size_t)NZ * (j + (size_t)NY * i);
            const cplx *e = E + (size_t)NZ * (j + (size_t)NY * i);
            y[0] = 0.0;                                               /* 2566-2567 */
            for (int k = 0; k < NZ; k++) y[k + 1] = x[k] * e[k];      /* 2555-2565 */
            y[NZ + 1] = 0.0;                                          /* 2570-2578 */
            for (int k = NZ + 2; k < N; k++)                          /* 2580-2592 */
                y[k] = -x[N - 1 - k] * e[N - 1 - k];
        }
    plan3_exec(&st->fwd, ext);                                        /* 2595 */
    for (size_t i = 0; i < st->Vext; i++) ext[i] = ext[i] * st->expArg[i]; /* 2598-2609 */
    plan3_exec(&st->rev, ext);                                        /* 2612 */
    for (int i = 0; i < NX; i++)                                      /* 2615-2630 */
        for (int j = 0; j < NY; j++) {
            const cplx *y = ext + (size_t)N * (j + (size_t)NY * i);
            cplx *x = in + (size_t)NZ * (j + (size_t)NY * i);
            const cplx *e = E + (size_t)NZ * (j + (size_t)NY * i);
            for (int k = 0; k < NZ; k++) x[k] = y[k + 1] * e[k];
        }
}

/* funPDESolver (SCFT.c:2310-2412) and funPDESolverBrush (2414-2472):
 * init 0 = continue from q[start], 1 = seedBool, 2 = brush mask. */
static void pde_block(orc_state *st, int start, int stop, int ns_eff, double *q,
                      const double *w, int init, const double *mask, unsigned long iter)
{
    const int NX = st->NX, NY = st->NY, NZ = st->NZ;
    const size_t V = st->V;
    const double DS = 1.0 / (st->p.ns - 1);
    const double dVol = st->p.mask ? (double)NX * NY * st->N : (double)NX * NY * NZ; /* 2325/2331, 2417 */
    const int D = st->p.seed_area_delta;
    cplx *E = (cplx *)malloc(sizeof(cplx) * V);
    cplx *in = (cplx *)malloc(sizeof(cplx) * V);
    cplx *work = (cplx *)malloc(sizeof(cplx) * (st->p.mask ? st->Vext : V));
    double *q0 = q + (size_t)start * V;
    for (int k = 0; k < NZ; k++)
        for (int j = 0; j < NY; j++)
            for (int i = 0; i < NX; i++) {
                size_t idx = k + (size_t)NZ * (j + (size_t)NY * i);
                E[idx] = cexp(DS * (-w[idx] / 2.0) * (1 + 0 * I));    /* 2344 */
                if (init == 2) {
                    q0[idx] = mask[idx];                             /* 2441 */
                } else if (init == 1 && iter <= (unsigned long)st->p.seed_steps) {
                    if (i > ((NX / 2) - D) && i < ((NX / 2) + D) && j > ((NY / 2) - D) && j < ((NY / 2) + D) &&
                        k > ((NZ / 4) - D) && k < ((NZ / 4) + D))
                        q0[idx] = 1.0;                               /* 2348-2356 */
                    else
                        q0[idx] = 0.0;                               /* 2358-2362 */
                } else if (init == 1) {
                    q0[idx] = 1.0;                                   /* 2364-2367 */
                }
                in[idx] = q0[idx] + 0 * I;
            }
    for (int s = start; s < stop; s++) {                            /* 2377-2402 */
        if (st->p.mask) step_dirichlet(st, E, in, work);
        else step_periodic(st, E, in, work);
        double *qs = q + (size_t)(s + 1) * V;
        for (size_t idx = 0; idx < V; idx++) {
            in[idx] = in[idx] / dVol;                                 /* 2397 */
            qs[idx] = creal(in[idx]);                                 /* 2398 */
        }
    }
    (void)ns_eff;
    free(E); free(in); free(work);
}

void orc_pde_block(orc_state *st, int start, int stop, int ns_eff, double *q,
                   const double *w, int seed_mode, unsigned long iter)
{
    pde_block(st, start, stop, ns_eff, q, w, seed_mode ? 1 : 0, NULL, iter);
}

/* funChemPotential, SCFT.c:1837-1872 */
static void chem_potential(const orc_state *st, const double *q, int ns_eff, double *chempot, double *qsum)
{
    const int NX = st->NX, NY = st->NY, NZ = st->NZ;
    const double *qe = q + (size_t)(ns_eff - 1) * st->V;
    double sum = 0.0;
    for (int k = 0; k < NZ; k++)
        for (int j = 0; j < NY; j++)
            for (int i = 0; i < NX; i++)
                sum = sum + qe[k + (size_t)NZ * (j + (size_t)NY * i)];
    sum = sum / (double)(NX * NY * NZ);
    double qmod = pow(sum, 2);
    *chempot = (qmod == 0) ? 0.0 : -log(qmod) / 2.0;
    *qsum = sum;
}

/* funMultiplyQs + funDensity, SCFT.c:2243-2308: Simpson over [start, stop]
 * of q[s]*qd[ns-1-s], times DS_EFF/Q with DS_EFF = 1/(NS_EFF-1). */
static void density(const orc_state *st, int start, int stop, int ns_eff, const double *q,
                    const double *qd, double qsum, double *phir)
{
    const size_t V = st->V;
    const double DS_EFF = 1.0 / (ns_eff - 1);
    for (size_t idx = 0; idx < V; idx++) {
#define QC(s) (q[(size_t)(s) * V + idx] * qd[(size_t)(ns_eff - 1 - (s)) * V + idx])
        double phi = QC(start) + QC(stop);
        for (int index = start + 1; index <= (stop - 1); index = index + 2) phi = phi + 4.0 * QC(index);
        for (int index = start + 2; index <= (stop - 2); index = index + 2) phi = phi + 2.0 * QC(index);
#undef QC
        phi = phi / 3.0;
        phir[idx] = phi * (DS_EFF / qsum);
    }
}

static void vol_frac(const orc_state *st, double frac, double *a)
{
    if (frac > 0.0)                                                    /* SCFT.c:1784-1803 */
        for (size_t i = 0; i < st->V; i++) a[i] = frac * a[i];
}

static const int kDbkFields[4][2] = {{0, 1}, {3, 4}, {5, 6}, {7, 8}};

/* funFreeEnergy, SCFT.c:1874-1917 (pair order of the reference) */
static double free_energy(const orc_state *st, double *const *phi)
{
    const int NX = st->NX, NY = st->NY, NZ = st->NZ;
    double (*chi)[13] = (double (*)[13])st->d.chi_eff;
    double H = 0.0;
    for (int k = 0; k < NZ; k++)
        for (int j = 0; j < NY; j++)
            for (int i = 0; i < NX; i++) {
                size_t idx = k + (size_t)NZ * (j + (size_t)NY * i);
                double h = 0.0;
                for (int f = 0; f < 10; f++) {
                    for (int g = f + 1; g < 10; g++) h += chi[f][g] * phi[f][idx] * phi[g][idx];
                    for (int s = 10; s < 13; s++) h += chi[s][f] * phi[f][idx] * phi[s][idx];
                }
                h += chi[10][11] * phi[10][idx] * phi[11][idx];
                h += chi[10][12] * phi[10][idx] * phi[12][idx];
                h += chi[11][12] * phi[11][idx] * phi[12][idx];
                double sphi = 1.0;
                for (int f = 0; f < 13; f++) h -= st->w[f][idx] * phi[f][idx];
                for (int f = 0; f < 13; f++) sphi -= phi[f][idx];
                h -= st->pf[idx] * sphi;
                H = H + h;
            }
    return H / (double)(NX * NY * NZ);
}

/* funUpdate, SCFT.c:1993-2124 */
static void update(orc_state *st, double dt, float MF, double *const *phi, double *dw2)
{
    const int NX = st->NX, NY = st->NY, NZ = st->NZ;
    const scft_derived *d = &st->d;
    const double G1 = st->p.g1, G2 = st->p.g1 / 6.0;
    double pAvg = 0.0;
    double sdw = 0.0;
    for (int k = 0; k < NZ; k++)
        for (int j = 0; j < NY; j++)
            for (int i = 0; i < NX; i++) {
                size_t idx = k + (size_t)NZ * (j + (size_t)NY * i);
                double wInit[13], r[13], delW[13];
                for (int f = 0; f < 13; f++)
                    if (d->active[f]) { wInit[f] = st->w[f][idx]; r[f] = phi[f][idx] - d->phi_avg[f]; }
                for (int f = 0; f < 13; f++) {
                    if (!d->active[f]) continue;
                    delW[f] = -wInit[f] + st->pf[idx];
                    for (int g = 0; g < 13; g++) {
                        if (!d->active[g] || f == g) continue;
                        delW[f] += d->chi_eff[f][g] * r[g];
                    }
                }
                double corr = 0.0;
                if (d->n_active > 1)
                    for (int f = 0; f < 13; f++) {
                        if (!d->active[f]) continue;
                        for (int g = f + 1; g < 13; g++) {
                            if (!d->active[g]) continue;
                            for (int h = 0; h < 13; h++)
                                if (d->active[h] && h != f && h != g) corr += d->chi_eff[f][g] * r[h];
                        }
                    }
                double psum = 0.0;
                for (int f = 0; f < 13; f++) {
                    if (!d->active[f]) continue;
                    double wn = wInit[f];
                    for (int g = 0; g < 13; g++)
                        if (d->active[g]) wn += ((f == g) ? G1 : G2) * delW[g];
                    wn += dt * MF * orc_gauss(st);
                    st->w[f][idx] = wn;
                    psum += wn;
                    sdw += (wn - wInit[f]) * (wn - wInit[f]);
                }
                st->pf[idx] = (psum + corr) / d->n_active;
                pAvg += st->pf[idx];
            }
    pAvg /= (double)(NX * NY * NZ);
    for (size_t idx = 0; idx < st->V; idx++) st->pf[idx] -= pAvg;
    *dw2 = sdw;
}

int orc_iterate(orc_state *st, unsigned long iter, scft_iter_out *out)
{
    const scft_derived *d = &st->d;
    const scft_params *p = &st->p;
    const size_t V = st->V;
    double qsum[17] = {0}, chem[17] = {0};

    /* homopolymers A..J, SCFT.c:1092-1240 */
    for (int h = 0; h < 10; h++) {
        int sp = 4 + h, ns = d->ns_sp[sp];
        if (ns <= 0) continue;
        pde_block(st, 0, ns - 1, ns, st->q[sp], st->w[h], 0, NULL, iter);
        chem_potential(st, st->q[sp], ns, &chem[sp], &qsum[sp]);
        density(st, 0, ns - 1, ns, st->q[sp], st->q[sp], qsum[sp], st->phi_h[h]);
    }
    /* diblocks 1..4, SCFT.c:1242-1410 */
    for (int b = 0; b < 4; b++) {
        if (!(p->vol_dbk[b] > 0)) continue;
        int ns = d->ns_sp[b], nsx = d->ns_x[b], nsy = d->ns_y[b];
        const double *wa = st->w[kDbkFields[b][0]], *wb = st->w[kDbkFields[b][1]];
        int seed = (b == 0) ? (p->enable_seed ? 1 : 0) : 0;
        pde_block(st, 0, nsx - 1, ns, st->q[b], wa, seed, NULL, iter);
        pde_block(st, nsx - 1, ns - 1, ns, st->q[b], wb, 0, NULL, iter);
        pde_block(st, 0, nsy, ns, st->qd[b], wb, 0, NULL, iter);
        pde_block(st, nsy, ns - 1, ns, st->qd[b], wa, 0, NULL, iter);
        /* dbk1/2 take Q from q-dagger, dbk3/4 from q (1271, 1314, 1357, 1399) */
        chem_potential(st, (b < 2) ? st->qd[b] : st->q[b], ns, &chem[b], &qsum[b]);
        density(st, 0, nsx - 1, ns, st->q[b], st->qd[b], qsum[b], st->phi_x[b]);
        density(st, nsx - 1, ns - 1, ns, st->q[b], st->qd[b], qsum[b], st->phi_y[b]);
    }
    /* brushes, SCFT.c:1412-1456 */
    for (int s = 0; s < 3; s++) {
        int sp = 14 + s, ns = d->ns_sp[sp];
        if (ns <= 0) continue;
        pde_block(st, 0, ns - 1, ns, st->q[sp], st->w[10 + s], 2, st->mask[s], iter);
        pde_block(st, 0, ns - 1, ns, st->qd[sp], st->w[10 + s], 0, NULL, iter);
        chem_potential(st, st->q[sp], ns, &chem[sp], &qsum[sp]);
        density(st, 0, ns - 1, ns, st->q[sp], st->qd[sp], qsum[sp], st->phi_s[s]);
    }
    /* funVolFrac, SCFT.c:1458-1482 */
    const double cv = d->coating_vol;
    for (int b = 0; b < 4; b++) {
        if (d->ns_sp[b] <= 0) continue;
        vol_frac(st, p->vol_dbk[b] * (1.0 - cv), st->phi_x[b]);
        vol_frac(st, p->vol_dbk[b] * (1.0 - cv), st->phi_y[b]);
    }
    for (int h = 0; h < 10; h++)
        if (d->ns_sp[4 + h] > 0) vol_frac(st, p->vol_homo[h] * (1.0 - cv), st->phi_h[h]);
    if (p->mask)
        for (int s = 0; s < 3; s++)
            if (d->ns_sp[14 + s] > 0) vol_frac(st, p->vol_s[s], st->phi_s[s]);

    /* totals, SCFT.c:1484-1513 (funAddArray leaves the sum untouched when
     * both operands are absent) */
    const double *dbk_part[10] = {0};
    for (int b = 0; b < 4; b++)
        if (d->ns_sp[b] > 0) {
            dbk_part[kDbkFields[b][0]] = st->phi_x[b];
            dbk_part[kDbkFields[b][1]] = st->phi_y[b];
        }
    for (int f = 0; f < 10; f++) {
        const double *a1 = dbk_part[f];
        const double *a2 = (d->ns_sp[4 + f] > 0) ? st->phi_h[f] : NULL;
        if (!a1 && !a2) continue;
        for (size_t i = 0; i < V; i++)
            st->tot[f][i] = !a1 ? a2[i] : (!a2 ? a1[i] : (1 * a1[i]) + (1 * a2[i]));
    }
    for (size_t i = 0; i < V; i++) {
        double t = st->tot[0][i] + st->tot[1][i];
        for (int f = 2; f < 10; f++) t = t + st->tot[f][i];
        if (p->mask)
            for (int s = 0; s < 3; s++) t = t + st->phi_s[s][i];
        st->phi_tot[i] = t;
    }

    float MF = (iter > p->relaxation) ? 0.0f : 1.0f;                  /* 1515-1522 */

    double *phi[13];
    for (int f = 0; f < 10; f++) phi[f] = st->tot[f];
    for (int s = 0; s < 3; s++) phi[10 + s] = st->phi_s[s];
    double H = free_energy(st, phi);                                 /* 1525-1527 */
    double frac = 0.0;                                               /* 1529-1555 */
    for (int b = 0; b < 4; b++)
        if (p->vol_dbk[b] > 0.0) frac += (p->vol_dbk[b] / d->alpha[b] * (1.0 - cv)) * log(qsum[b]);
    for (int h = 0; h < 10; h++)
        if (p->vol_homo[h] > 0.0) frac += (p->vol_homo[h] / d->alpha[4 + h] * (1.0 - cv)) * log(qsum[4 + h]);
    if (p->mask)
        for (int s = 0; s < 3; s++)
            if (p->vol_s[s] > 0.0) frac += p->vol_s[s] / d->alpha[14 + s] * log(qsum[14 + s]);
    H -= frac;

    double dw2 = 0.0;
    update(st, p->dt_noise, MF, phi, &dw2);                          /* 1558-1560 */

    double rms = 0.0;                                                /* 1564-1580 */
    for (int k = 0; k < st->NZ; k++)
        for (int j = 0; j < st->NY; j++)
            for (int i = 0; i < st->NX; i++) {
                size_t idx = k + (size_t)st->NZ * (j + (size_t)st->NY * i);
                rms = rms + pow((st->phi_tot[idx] - 1.0), 2);
            }
    rms = pow((rms / (double)(st->NX * st->NY * st->NZ)), 0.5);

    if (out) {
        memset(out, 0, sizeof(*out));
        out->free_energy = H;
        out->rmsdev = rms;
        out->dw_rms = (d->n_active > 0) ? sqrt(dw2 / ((double)d->n_active * (double)V)) : 0.0;
        for (int sp = 0; sp < 17; sp++) { out->qsum[sp] = qsum[sp]; out->chempot[sp] = chem[sp]; }
        out->mf = MF;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* main() with its file I/O (SCFT.c:84-101, 703-879, 1025-1085,         */
/* 1582-1664).                                                          */
/* ------------------------------------------------------------------ */

static FILE *open_in(const char *dir, const char *name, const char *mode)
{
    char path[4096];
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    return fopen(path, mode);
}

static void write_run_txt(FILE *fid, const scft_params *p, const scft_derived *d, unsigned long iIter)
{
    /* SCFT.c:1025-1085 */
    const double G1 = p->g1, G2 = p->g1 / 6.0;
    const double (*c)[13] = (const double (*)[13])d->chi_eff;
    const double DS = 1.0 / (p->ns - 1);
    fprintf(fid, "-----------------------------------------\n\n");
    fprintf(fid, "These are the run parameter values: \n\n");
    fprintf(fid, "Iteration  = %ld \n", (long)iIter);
    fprintf(fid, "NX, NY, NZ  = %d, %d, %d \n", d->nx, d->ny, d->nz);
    fprintf(fid, "NS  = %d \n\n", p->ns);
    for (int b = 0; b < 4; b++)
        fprintf(fid, "DBK_%d:: NS_DBK_%d, NS_DBK_%d_X, NS_DBK_%d_Y, LEN_DBK_%d, F_INPUT_DBK_%d, VOL_DBK_%d  = %d, %d, %d, %f, %f, %f \n",
                b + 1, b + 1, b + 1, b + 1, b + 1, b + 1, b + 1, d->ns_sp[b], p->vol_dbk[b] > 0.0 ? d->ns_x[b] : 0,
                p->vol_dbk[b] > 0.0 ? d->ns_y[b] : 0, p->len_dbk[b], p->f_dbk[b], p->vol_dbk[b]);
    for (int h = 0; h < 10; h++) {
        char L = (char)('A' + h);
        fprintf(fid, "HOMOPOLYMER %c:: NS_HOMO_%c, HOMOPOLYMER_LEN_%c, VOL_HOMOPOLYMER_%c  = %d, %f, %f \n%s", L, L, L, L,
                d->ns_sp[4 + h], p->len_homo[h], p->vol_homo[h], h == 9 ? "\n" : "");
    }
    fprintf(fid, "G1 = %f, G2 = %f, G3 = %f, G4 = %f, G5 = %f, G6 = %f, G7 = %f, G8 = %f, G9 = %f, G10 = %f, G11 = %f, G12 = %f, G13 = %f \n\n",
            G1, G2, G2, G2, G2, G2, G2, G2, G2, G2, G2, G2, G2);
    const double *pa = d->phi_avg;
    fprintf(fid, "phi_A_avg = %lf, phi_B_avg = %lf, phi_C_avg = %lf, phi_D_avg = %lf, phi_E_avg = %lf, phi_F_avg = %lf, phi_G_avg = %lf, phi_H_avg = %lf, phi_I_avg = %lf, phi_J_avg = %lf, phi_S1_avg = %lf, phi_S2_avg = %lf, phi_S3_avg = %lf \n\n",
            pa[0], pa[1], pa[2], pa[3], pa[4], pa[5], pa[6], pa[7], pa[8], pa[9], pa[10], pa[11], pa[12]);
    fprintf(fid, "Chi*N As, AB = %f, AC = %f, AD = %f, AE = %f, AF = %f, AG = %f, AH = %f, AI = %f, AJ = %f \n", c[0][1], c[0][2], c[0][3], c[0][4], c[0][5], c[0][6], c[0][7], c[0][8], c[0][9]);
    fprintf(fid, "Chi*N Bs, BC = %f, BD = %f, BE = %f, BF = %f, BG = %f, BH = %f, BI = %f, BJ = %f \n", c[1][2], c[1][3], c[1][4], c[1][5], c[1][6], c[1][7], c[1][8], c[1][9]);
    fprintf(fid, "Chi*N Cs, CD = %f, CE = %f, CF = %f, CG = %f, CH = %f, CI = %f, CJ = %f \n", c[2][3], c[2][4], c[2][5], c[2][6], c[2][7], c[2][8], c[2][9]);
    fprintf(fid, "Chi*N Ds, DE = %f, DF = %f, DG = %f, DH = %f, DI = %f, DJ = %f \n", c[3][4], c[3][5], c[3][6], c[3][7], c[3][8], c[3][9]);
    fprintf(fid, "Chi*N Es, EF = %f, EG = %f, EH = %f, EI = %f, EJ = %f \n", c[4][5], c[4][6], c[4][7], c[4][8], c[4][9]);
    fprintf(fid, "Chi*N Fs, FG = %f, FH = %f, FI = %f, FJ = %f \n", c[5][6], c[5][7], c[5][8], c[5][9]);
    fprintf(fid, "Chi*N Gs, GH = %f, GI = %f, GJ = %f \n", c[6][7], c[6][8], c[6][9]);
    fprintf(fid, "Chi*N Hs, HI = %f, HJ = %f \n", c[7][8], c[7][9]);
    fprintf(fid, "Chi*N Is, IJ = %f \n\n", c[8][9]);
    fprintf(fid, "Substrate settings: \n");
    fprintf(fid, "MASK = %d, vol S[1+2+3]: %f \n", p->mask, d->coating_vol);
    for (int s = 0; s < 3; s++)
        fprintf(fid, "S%d: NS_S%d, LEN_S%d, VOL_S%d  = %d, %f, %f \n%s", s + 1, s + 1, s + 1, s + 1, d->ns_sp[14 + s],
                p->len_s[s], p->vol_s[s], s == 2 ? "\n" : "");
    fprintf(fid, "CHI_N_S1_S2_EFF = %f, CHI_N_S1_S3_EFF = %f, CHI_N_S2_S3_EFF = %f \n", c[10][11], c[10][12], c[11][12]);
    for (int f = 0; f < 10; f++) {
        char L = (char)('A' + f);
        fprintf(fid, "CHI_N_S1_%c_EFF = %f, CHI_N_S2_%c_EFF = %f, CHI_N_S3_%c_EFF = %f \n%s", L, c[10][f], L, c[11][f], L,
                c[12][f], f == 9 ? "\n" : "");
    }
    fprintf(fid, "DS  = %f \n", DS);
    fprintf(fid, "LX, LY, LZ  = %f, %f, %f \n", p->lx, p->ly, p->lz);
    fprintf(fid, "DX, DY, DZ  = %f, %f, %f \n\n", p->lx / d->nx, p->ly / d->ny, p->lz / d->nz);
    fprintf(fid, "-----------------------------------------\n\n");
}

int orc_main(const scft_params *p, unsigned long iter_max, int d_wrt, int d_wrt_energies, const char *dir)
{
    scft_derived dd;
    if (orc_derive(p, &dd) == SCFT_EVOLUME) {                         /* 84-101 */
        double vsum = 0.0;
        for (int i = 0; i < 4; i++) vsum += p->vol_dbk[i];
        for (int i = 0; i < 10; i++) vsum += p->vol_homo[i];
        FILE *fe = open_in(dir, "DENSITY_CHECK_FAIL.error", "w");
        if (fe) { fprintf(fe, "%lf", vsum); fclose(fe); }
        return SCFT_EVOLUME;
    }
    orc_state *st = orc_create(p);
    const size_t V = st->V;
    const int NX = st->NX, NY = st->NY, NZ = st->NZ;
    unsigned long iIter;
    FILE *fid = open_in(dir, "counter.dat", "r");                     /* 703-716 */
    if (!fid) { orc_destroy(st); return SCFT_EINVAL; }
    if (fscanf(fid, "%lu", &iIter) != 1) { fclose(fid); orc_destroy(st); return SCFT_EINVAL; }
    fclose(fid);
    if (p->mask) {                                                    /* 756-781 */
        fid = open_in(dir, "wMask.dat", "r");
        if (!fid) { orc_destroy(st); return SCFT_EINVAL; }
        for (int k = 0; k < NZ; k++)
            for (int j = 0; j < NY; j++)
                for (int i = 0; i < NX; i++) {
                    size_t idx = k + (size_t)NZ * (j + (size_t)NY * i);
                    if (fscanf(fid, "%lf %lf %lf", &st->mask[0][idx], &st->mask[1][idx], &st->mask[2][idx]) != 3) break;
                }
        fclose(fid);
    }
    if (iIter == 0) {
        /* 785-843: time-seeded new run; fields start at zero */
        unsigned short seed[3] = {1, 2, 3};
        orc_seed48(st, seed);
    } else {                                                          /* 844-879 */
        fid = open_in(dir, "restart_fields.dat", "r");
        if (!fid) { orc_destroy(st); return SCFT_EINVAL; }
        for (int k = 0; k < NZ; k++)
            for (int j = 0; j < NY; j++)
                for (int i = 0; i < NX; i++) {
                    size_t idx = k + (size_t)NZ * (j + (size_t)NY * i);
                    double v[14];
                    if (fscanf(fid, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &v[0], &v[1], &v[2], &v[3], &v[4],
                               &v[5], &v[6], &v[7], &v[8], &v[9], &v[10], &v[11], &v[12], &v[13]) != 14)
                        break;
                    for (int f = 0; f < 13; f++) st->w[f][idx] = v[f];
                    st->pf[idx] = v[13];
                }
        fclose(fid);
        fid = open_in(dir, "restart_random.dat", "r");
        if (!fid) { orc_destroy(st); return SCFT_EINVAL; }
        unsigned short seed[3];
        if (fscanf(fid, "%hu %hu %hu", &seed[0], &seed[1], &seed[2]) != 3) { fclose(fid); orc_destroy(st); return SCFT_EINVAL; }
        fclose(fid);
        orc_seed48(st, seed);
    }
    fid = open_in(dir, "run.txt", "a");
    write_run_txt(fid, p, &st->d, iIter);
    fclose(fid);

    double *phi13 = dalloc(13 * V);
    while (iIter <= iter_max) {                                       /* 1088 */
        iIter++;
        scft_iter_out o;
        orc_iterate(st, iIter, &o);
        if (fmod((double)iIter, d_wrt_energies) == 0) {               /* 1583-1599 */
            fid = open_in(dir, "energies.dat", "a");
            fprintf(fid, "%lu %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf\n", iIter, o.chempot[0],
                    o.chempot[1], o.chempot[2], o.chempot[3], o.chempot[4], o.chempot[5], o.chempot[6], o.chempot[7],
                    o.chempot[8], o.chempot[9], o.chempot[10], o.chempot[11], o.chempot[12], o.chempot[13], o.chempot[14],
                    o.chempot[15], o.chempot[16]);
            fclose(fid);
            fid = open_in(dir, "free_energy.dat", "a");
            fprintf(fid, "%lu  %lf  %f \n", iIter, o.free_energy, o.mf);
            fclose(fid);
            fid = open_in(dir, "deviations.dat", "a");
            fprintf(fid, "%lu  %lf %lf %lf %lf \n", iIter, o.rmsdev, st->d.chi_eff[0][1], st->d.chi_eff[0][2],
                    st->d.chi_eff[1][2]);
            fclose(fid);
        }
        if (fmod((double)iIter, d_wrt) == 0 || iIter == 1) {          /* 1601-1664 */
            char name[64];
            snprintf(name, sizeof(name), "densities_dbk_%d.dat", (int)(100000000 + iIter));
            fid = open_in(dir, name, "w");
            orc_get_densities(st, phi13, NULL);
            for (int k = 0; k < NZ; k++)
                for (int j = 0; j < NY; j++)
                    for (int i = 0; i < NX; i++) {
                        size_t idx = k + (size_t)NZ * (j + (size_t)NY * i);
                        fprintf(fid, "%5d %5d %5d", i, j, k);
                        for (int f = 0; f < 13; f++) fprintf(fid, " %10.6lf", phi13[(size_t)f * V + idx]);
                        for (int f = 0; f < 13; f++) fprintf(fid, " %12.6lf", st->w[f][idx]);
                        fprintf(fid, " %12.6lf \n", st->pf[idx]);
                    }
            fclose(fid);
            fid = open_in(dir, "counter.dat", "w");
            fprintf(fid, "%lu", iIter);
            fclose(fid);
            fid = open_in(dir, "restart_fields.dat", "w");
            for (int k = 0; k < NZ; k++)
                for (int j = 0; j < NY; j++)
                    for (int i = 0; i < NX; i++) {
                        size_t idx = k + (size_t)NZ * (j + (size_t)NY * i);
                        for (int f = 0; f < 13; f++) fprintf(fid, "%lf ", st->w[f][idx]);
                        fprintf(fid, "%lf \n", st->pf[idx]);
                    }
            fclose(fid);
            unsigned short seed[3];
            orc_get_seed(st, seed);
            fid = open_in(dir, "restart_random.dat", "w");
            fprintf(fid, "%hu %hu %hu", seed[0], seed[1], seed[2]);
            fclose(fid);
        }
    }
    free(phi13);
    orc_destroy(st);
    return 0;
}
