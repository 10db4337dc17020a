/*
 * scft_run.c — drop-in driver for SCFT.c's main() on libscft.
 *
 * Built against the run's parameters.h (make -C csrc run PARAMS=...), it
 * reads the same inputs (counter.dat, wMask.dat, restart_fields.dat,
 * restart_random.dat) and writes the same outputs in the same formats
 * (run.txt, energies.dat, free_energy.dat, deviations.dat,
 * densities_dbk_<1e8+iter>.dat, counter.dat, restart_*.dat), with the
 * iteration loop running on the GPU through the C ABI of include/scft.h.
 * Snapshots are copied to host buffers at their iteration and written out
 * by a thread of their own while the GPU runs on.
 *
 * Replaces SCFT.c:82-1664.  Environment: SCFT_PRECISION=fp32 selects the
 * FP32-split propagators (fp32-nearest: slices rounded to nearest; default
 * FP64, the reference's arithmetic).
 * SCFT_EXACT_RESTART=1 also writes restart_fields.bin beside every
 * restart_fields.dat: the fields in full precision (the text file keeps the
 * reference's 6 decimals, SCFT.c:1645-1646).  A restart (counter != 0)
 * reads the binary file instead of the text when it exists and belongs to
 * the counter's iteration and this grid, so a resumed run continues
 * bit-exactly.  No counterpart in the reference.
 *
 * The reference draws one Gaussian deviate per active field and voxel every
 * iteration, also when MF = 0 (SCFT.c:2103), so restart_random.dat depends
 * on the whole stream.  By default the library replays that stream on the
 * GPU (scft_rng_seed / scft_rng_state).  SCFT_HOST_RNG=1 replays it here
 * instead: POSIX drand48 (48-bit LCG, a = 0x5DEECE66D, c = 0xB) and the
 * Marsaglia polar method; with MF = 1 the deviates are uploaded, with
 * MF = 0 only the generator state is advanced, on a host thread that
 * overlaps the GPU iteration.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "scft.h"
#include PARAMS_HEADER

/* ---------------------------------------------------------- parameters */

static void params_from_macros(scft_params *p)
{
    memset(p, 0, sizeof(*p));
    p->lx = LX; p->ly = LY; p->lz = LZ; p->dr_target = DR_TARGET;
    p->ns = NS; p->mask = MASK; p->g1 = G1; p->dt_noise = DT_NOISE;
    p->relaxation = RELAXATION;
    p->enable_seed = ENABLE_SEED; p->seed_steps = SEED_STEPS; p->seed_area_delta = SEED_AREA_DELTA;
    const double f[4] = {F_INPUT_DBK_1, F_INPUT_DBK_2, F_INPUT_DBK_3, F_INPUT_DBK_4};
    const double ld[4] = {LEN_DBK_1, LEN_DBK_2, LEN_DBK_3, LEN_DBK_4};
    const double vd[4] = {VOL_DBK_1, VOL_DBK_2, VOL_DBK_3, VOL_DBK_4};
    const double vh[10] = {VOL_HOMOPOLYMER_A, VOL_HOMOPOLYMER_B, VOL_HOMOPOLYMER_C, VOL_HOMOPOLYMER_D, VOL_HOMOPOLYMER_E,
                           VOL_HOMOPOLYMER_F, VOL_HOMOPOLYMER_G, VOL_HOMOPOLYMER_H, VOL_HOMOPOLYMER_I, VOL_HOMOPOLYMER_J};
    const double lh[10] = {LEN_HOMO_A, LEN_HOMO_B, LEN_HOMO_C, LEN_HOMO_D, LEN_HOMO_E,
                           LEN_HOMO_F, LEN_HOMO_G, LEN_HOMO_H, LEN_HOMO_I, LEN_HOMO_J};
    for (int i = 0; i < 4; i++) { p->f_dbk[i] = f[i]; p->len_dbk[i] = ld[i]; p->vol_dbk[i] = vd[i]; }
    for (int i = 0; i < 10; i++) { p->vol_homo[i] = vh[i]; p->len_homo[i] = lh[i]; }
    p->vol_s[0] = VOL_S1; p->vol_s[1] = VOL_S2; p->vol_s[2] = VOL_S3;
    p->len_s[0] = LEN_S1; p->len_s[1] = LEN_S2; p->len_s[2] = LEN_S3;
    /* chi over A..J, S1..S3: upper triangle from the CHI_N_* macros */
    const double cx[13][13] = {
        {0, CHI_N_AB, CHI_N_AC, CHI_N_AD, CHI_N_AE, CHI_N_AF, CHI_N_AG, CHI_N_AH, CHI_N_AI, CHI_N_AJ, CHI_N_S1_A, CHI_N_S2_A, CHI_N_S3_A},
        {0, 0, CHI_N_BC, CHI_N_BD, CHI_N_BE, CHI_N_BF, CHI_N_BG, CHI_N_BH, CHI_N_BI, CHI_N_BJ, CHI_N_S1_B, CHI_N_S2_B, CHI_N_S3_B},
        {0, 0, 0, CHI_N_CD, CHI_N_CE, CHI_N_CF, CHI_N_CG, CHI_N_CH, CHI_N_CI, CHI_N_CJ, CHI_N_S1_C, CHI_N_S2_C, CHI_N_S3_C},
        {0, 0, 0, 0, CHI_N_DE, CHI_N_DF, CHI_N_DG, CHI_N_DH, CHI_N_DI, CHI_N_DJ, CHI_N_S1_D, CHI_N_S2_D, CHI_N_S3_D},
        {0, 0, 0, 0, 0, CHI_N_EF, CHI_N_EG, CHI_N_EH, CHI_N_EI, CHI_N_EJ, CHI_N_S1_E, CHI_N_S2_E, CHI_N_S3_E},
        {0, 0, 0, 0, 0, 0, CHI_N_FG, CHI_N_FH, CHI_N_FI, CHI_N_FJ, CHI_N_S1_F, CHI_N_S2_F, CHI_N_S3_F},
        {0, 0, 0, 0, 0, 0, 0, CHI_N_GH, CHI_N_GI, CHI_N_GJ, CHI_N_S1_G, CHI_N_S2_G, CHI_N_S3_G},
        {0, 0, 0, 0, 0, 0, 0, 0, CHI_N_HI, CHI_N_HJ, CHI_N_S1_H, CHI_N_S2_H, CHI_N_S3_H},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, CHI_N_IJ, CHI_N_S1_I, CHI_N_S2_I, CHI_N_S3_I},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, CHI_N_S1_J, CHI_N_S2_J, CHI_N_S3_J},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, CHI_N_S1_S2, CHI_N_S1_S3},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, CHI_N_S2_S3},
        {0}};
    for (int a = 0; a < 13; a++)
        for (int b = a + 1; b < 13; b++) p->chi[a][b] = p->chi[b][a] = cx[a][b];
}

/* ------------------------------------------------------------- generator */

typedef struct {
    unsigned long long x;   /* 48-bit state */
} lcg48;

static void lcg_seed(lcg48 *g, const unsigned short s[3])
{
    g->x = ((unsigned long long)s[2] << 32) | ((unsigned long long)s[1] << 16) | s[0];
}

static void lcg_get(const lcg48 *g, unsigned short s[3])
{
    s[0] = (unsigned short)(g->x & 0xffff);
    s[1] = (unsigned short)((g->x >> 16) & 0xffff);
    s[2] = (unsigned short)((g->x >> 32) & 0xffff);
}

static inline double lcg_next(lcg48 *g)
{
    g->x = (0x5DEECE66DULL * g->x + 0xBULL) & 0xFFFFFFFFFFFFULL;
    return (double)g->x * (1.0 / 281474976710656.0);
}

/* funGaussDev (SCFT.c:1769-1782); value == NULL only advances the state */
static inline void gauss_dev(lcg48 *g, double *value)
{
    double x1, x2, w;
    do {
        x1 = 2.0 * lcg_next(g) - 1.0;
        x2 = 2.0 * lcg_next(g) - 1.0;
        w = x1 * x1 + x2 * x2;
    } while (w >= 1.0);
    if (value) *value = x1 * sqrt((-2.0 * log(w)) / w);
}

typedef struct {
    lcg48 *g;
    size_t n;
    double *out;
} draw_job;

static void *draw_thread(void *arg)
{
    draw_job *j = (draw_job *)arg;
    for (size_t i = 0; i < j->n; i++) gauss_dev(j->g, j->out ? &j->out[i] : NULL);
    return NULL;
}

/* ------------------------------------------------------------------ I/O */

static void die(const char *what)
{
    perror(what);
    exit(EXIT_FAILURE);
}

/* Full-precision restart sidecar: "SCFTRST1", u64 iteration, u64 V, then
 * w[13][V] and p[V] as FP64. */
#define EXACT_RESTART "restart_fields.bin"

static void write_exact_restart(unsigned long iter, size_t V, const double *w, const double *pf)
{
    FILE *f = fopen(EXACT_RESTART, "wb");
    if (!f) die(EXACT_RESTART);
    const unsigned long long hdr[2] = {iter, (unsigned long long)V};
    if (fwrite("SCFTRST1", 1, 8, f) != 8 || fwrite(hdr, sizeof(hdr), 1, f) != 1 ||
        fwrite(w, sizeof(double), 13 * V, f) != 13 * V || fwrite(pf, sizeof(double), V, f) != V)
        die(EXACT_RESTART);
    fclose(f);
}

/* 1 when the sidecar holds iteration `iter` of a V-point grid (fields read). */
static int read_exact_restart(unsigned long iter, size_t V, double *w, double *pf)
{
    FILE *f = fopen(EXACT_RESTART, "rb");
    if (!f) return 0;
    char magic[8];
    unsigned long long hdr[2];
    int ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "SCFTRST1", 8) == 0 && fread(hdr, sizeof(hdr), 1, f) == 1 &&
             hdr[0] == iter && hdr[1] == (unsigned long long)V;
    if (ok) ok = fread(w, sizeof(double), 13 * V, f) == 13 * V && fread(pf, sizeof(double), V, f) == V;
    fclose(f);
    return ok;
}

/* One snapshot (SCFT.c:1601-1664) written by a thread of its own while the
 * GPU runs the next iterations: the densities file, counter.dat,
 * restart_fields.dat (+ the full-precision sidecar) and restart_random.dat,
 * in the reference's order, from host copies taken at the snapshot
 * iteration. */
typedef struct {
    unsigned long iter;
    int nx, ny, nz, exact;
    double *w, *pf, *phi;      /* owned copies, 13V / V / 13V */
    unsigned short rng[3];
} snap_job;

static void *snap_thread(void *arg)
{
    const snap_job *j = (const snap_job *)arg;
    const size_t V = (size_t)j->nx * j->ny * j->nz;
    char name[64];
    snprintf(name, sizeof(name), "%s%d.dat", DENSITY, (int)(100000000 + j->iter));
    FILE *fid = fopen(name, "w");
    if (!fid) die(name);
    for (int k = 0; k < j->nz; k++)
        for (int jj = 0; jj < j->ny; jj++)
            for (int i = 0; i < j->nx; i++) {
                size_t idx = k + (size_t)j->nz * (jj + (size_t)j->ny * i);
                fprintf(fid, "%5d %5d %5d", i, jj, k);
                for (int f = 0; f < 13; f++) fprintf(fid, " %10.6lf", j->phi[(size_t)f * V + idx]);
                for (int f = 0; f < 13; f++) fprintf(fid, " %12.6lf", j->w[(size_t)f * V + idx]);
                fprintf(fid, " %12.6lf \n", j->pf[idx]);
            }
    fclose(fid);
    fid = fopen(COUNTER, "w");
    if (!fid) die(COUNTER);
    fprintf(fid, "%lu", j->iter);
    fclose(fid);
    fid = fopen(RESTART_FIELDS, "w");
    if (!fid) die(RESTART_FIELDS);
    for (int k = 0; k < j->nz; k++)
        for (int jj = 0; jj < j->ny; jj++)
            for (int i = 0; i < j->nx; i++) {
                size_t idx = k + (size_t)j->nz * (jj + (size_t)j->ny * i);
                for (int f = 0; f < 13; f++) fprintf(fid, "%lf ", j->w[(size_t)f * V + idx]);
                fprintf(fid, "%lf \n", j->pf[idx]);
            }
    fclose(fid);
    if (j->exact) write_exact_restart(j->iter, V, j->w, j->pf);
    fid = fopen(RESTART_RANDOM, "w");
    if (!fid) die(RESTART_RANDOM);
    fprintf(fid, "%hu %hu %hu", j->rng[0], j->rng[1], j->rng[2]);
    fclose(fid);
    return NULL;
}

static void check(int rc, const char *what)
{
    if (rc != SCFT_OK) {
        fprintf(stderr, "%s failed (%d): %s\n", what, rc, scft_last_error());
        exit(EXIT_FAILURE);
    }
}

static void write_run_txt(const scft_params *p, const scft_derived *d, unsigned long iIter)
{
    /* SCFT.c:1025-1085 */
    FILE *fid = fopen(RUN_PARAMS, "a");
    if (!fid) die("run.txt");
    const double (*c)[13] = (const double (*)[13])d->chi_eff;
    const double G2v = G1 / 6.0;
    fprintf(fid, "-----------------------------------------\n\n");
    fprintf(fid, "These are the run parameter values: \n\n");
    fprintf(fid, "Iteration  = %ld \n", (long)iIter);
    fprintf(fid, "NX, NY, NZ  = %d, %d, %d \n", d->nx, d->ny, d->nz);
    fprintf(fid, "NS  = %d \n\n", NS);
    for (int b = 0; b < 4; b++)
        fprintf(fid, "DBK_%d:: NS_DBK_%d, NS_DBK_%d_X, NS_DBK_%d_Y, LEN_DBK_%d, F_INPUT_DBK_%d, VOL_DBK_%d  = %d, %d, %d, %f, %f, %f \n",
                b + 1, b + 1, b + 1, b + 1, b + 1, b + 1, b + 1, d->ns_sp[b], d->ns_x[b], d->ns_y[b], p->len_dbk[b],
                p->f_dbk[b], p->vol_dbk[b]);
    for (int h = 0; h < 10; h++) {
        const char L = (char)('A' + h);
        fprintf(fid, "HOMOPOLYMER %c:: NS_HOMO_%c, HOMOPOLYMER_LEN_%c, VOL_HOMOPOLYMER_%c  = %d, %f, %f \n%s", L, L, L, L,
                d->ns_sp[4 + h], p->len_homo[h], p->vol_homo[h], h == 9 ? "\n" : "");
    }
    fprintf(fid, "G1 = %f, G2 = %f, G3 = %f, G4 = %f, G5 = %f, G6 = %f, G7 = %f, G8 = %f, G9 = %f, G10 = %f, G11 = %f, G12 = %f, G13 = %f \n\n",
            (double)G1, G2v, G2v, G2v, G2v, G2v, G2v, G2v, G2v, G2v, G2v, G2v, G2v);
    const double *pa = d->phi_avg;
    fprintf(fid, "phi_A_avg = %lf, phi_B_avg = %lf, phi_C_avg = %lf, phi_D_avg = %lf, phi_E_avg = %lf, phi_F_avg = %lf, phi_G_avg = %lf, phi_H_avg = %lf, phi_I_avg = %lf, phi_J_avg = %lf, phi_S1_avg = %lf, phi_S2_avg = %lf, phi_S3_avg = %lf \n\n",
            pa[0], pa[1], pa[2], pa[3], pa[4], pa[5], pa[6], pa[7], pa[8], pa[9], pa[10], pa[11], pa[12]);
    static const char *rows[9] = {"As", "Bs", "Cs", "Ds", "Es", "Fs", "Gs", "Hs", "Is"};
    for (int f = 0; f < 9; f++) {
        fprintf(fid, "Chi*N %s", rows[f]);
        for (int g = f + 1; g < 10; g++)
            fprintf(fid, ", %c%c = %f", 'A' + f, 'A' + g, c[f][g]);
        fprintf(fid, " \n%s", f == 8 ? "\n" : "");
    }
    fprintf(fid, "Substrate settings: \n");
    fprintf(fid, "MASK = %d, vol S[1+2+3]: %f \n", MASK, d->coating_vol);
    for (int s = 0; s < 3; s++)
        fprintf(fid, "S%d: NS_S%d, LEN_S%d, VOL_S%d  = %d, %f, %f \n%s", s + 1, s + 1, s + 1, s + 1, d->ns_sp[14 + s],
                p->len_s[s], p->vol_s[s], s == 2 ? "\n" : "");
    fprintf(fid, "CHI_N_S1_S2_EFF = %f, CHI_N_S1_S3_EFF = %f, CHI_N_S2_S3_EFF = %f \n", c[10][11], c[10][12], c[11][12]);
    for (int f = 0; f < 10; f++) {
        const char L = (char)('A' + f);
        fprintf(fid, "CHI_N_S1_%c_EFF = %f, CHI_N_S2_%c_EFF = %f, CHI_N_S3_%c_EFF = %f \n%s", L, c[10][f], L, c[11][f], L,
                c[12][f], f == 9 ? "\n" : "");
    }
    fprintf(fid, "DS  = %f \n", (double)DS);
    fprintf(fid, "LX, LY, LZ  = %f, %f, %f \n", (double)LX, (double)LY, (double)LZ);
    fprintf(fid, "DX, DY, DZ  = %f, %f, %f \n\n", (double)(LX / d->nx), (double)(LY / d->ny), (double)(LZ / d->nz));
    fprintf(fid, "-----------------------------------------\n\n");
    fclose(fid);
}

int main(void)
{
    scft_params p;
    params_from_macros(&p);
    /* SCFT_DUMP_PARAMS=path: write the parameter struct taken from the
     * compiled header and stop (lets tests compare header readers) */
    const char *dump = getenv("SCFT_DUMP_PARAMS");
    if (dump) {
        FILE *f = fopen(dump, "wb");
        if (!f || fwrite(&p, sizeof(p), 1, f) != 1) die(dump);
        fclose(f);
        return 0;
    }

    /* volume check, SCFT.c:84-101 */
    const double volumeCheck = VOL_DBK_1 + VOL_DBK_2 + VOL_DBK_3 + VOL_DBK_4 + VOL_HOMOPOLYMER_A + VOL_HOMOPOLYMER_B +
                               VOL_HOMOPOLYMER_C + VOL_HOMOPOLYMER_D + VOL_HOMOPOLYMER_E + VOL_HOMOPOLYMER_F +
                               VOL_HOMOPOLYMER_G + VOL_HOMOPOLYMER_H + VOL_HOMOPOLYMER_I + VOL_HOMOPOLYMER_J;
    if (fabs(1.0 - volumeCheck) >= 0.000001) {
        FILE *fe = fopen("DENSITY_CHECK_FAIL.error", "w");
        if (fe) { fprintf(fe, "%lf", volumeCheck); fclose(fe); }
        printf("Volumes do not sum up to 1.0, check the input parameters\n");
        exit(0);
    }
    scft_derived d;
    check(scft_derive(&p, &d), "scft_derive");
    const int NXv = d.nx, NYv = d.ny, NZv = d.nz;
    const size_t V = (size_t)NXv * NYv * NZv;

    unsigned long iIter;
    FILE *fid = fopen(COUNTER, "r");                                  /* SCFT.c:703-716 */
    if (!fid) die("Error opening counter file");
    if (fscanf(fid, "%lu", &iIter) != 1) die("Error reading the integer from counter file");
    fclose(fid);

    const char *prec = getenv("SCFT_PRECISION");
    scft_ctx *ctx;
    const int precision = !prec ? SCFT_FP64
                          : strcmp(prec, "fp32") == 0 ? SCFT_FP32_SPLIT
                          : strcmp(prec, "fp32-nearest") == 0 ? SCFT_FP32_NEAREST : SCFT_FP64;
    check(scft_create(&p, precision, &ctx), "scft_create");

    double *w = calloc(13 * V, sizeof(double)), *pf = calloc(V, sizeof(double));
    if (!w || !pf) die("calloc");
    if (MASK) {                                                       /* SCFT.c:756-781 */
        double *mask = calloc(3 * V, sizeof(double));
        fid = fopen(W_MASK, "r");
        if (!fid) die("Error opening the w field mask file");
        size_t got = 0;
        for (int k = 0; k < NZv; k++)
            for (int j = 0; j < NYv; j++)
                for (int i = 0; i < NXv; i++) {
                    size_t idx = k + (size_t)NZv * (j + (size_t)NYv * i);
                    if (fscanf(fid, "%lf %lf %lf", &mask[idx], &mask[V + idx], &mask[2 * V + idx]) != 3) goto mask_done;
                    got++;
                }
    mask_done:
        fclose(fid);
        if (got < V)   /* the reference reads on the same way; say so */
            fprintf(stderr, "warning: %s holds %zu of %zu voxel lines; the rest stay 0\n", W_MASK, got, V);
        check(scft_set_mask(ctx, mask), "scft_set_mask");
        free(mask);
    }
    lcg48 gen;
    unsigned short iSeed[3];
    if (iIter == 0) {                                                 /* SCFT.c:785-843 */
        struct timespec ts;
        clock_gettime(CLOCK_REALTIME, &ts);
        long t = ts.tv_nsec;
        iSeed[0] = (unsigned short)t;
        iSeed[1] = (unsigned short)(t + 1);
        iSeed[2] = (unsigned short)(t + 2);
        printf("iSeed new = %hu, %hu, %hu\n\n", iSeed[0], iSeed[1], iSeed[2]);
    } else if (read_exact_restart(iIter, V, w, pf)) {
        printf("restart fields: %s (full precision)\n", EXACT_RESTART);
        fid = fopen(RESTART_RANDOM, "r");
        if (!fid) die("Error opening restart random file");
        if (fscanf(fid, "%hu %hu %hu", &iSeed[0], &iSeed[1], &iSeed[2]) != 3) die("restart_random.dat");
        fclose(fid);
    } else {                                                          /* SCFT.c:844-879 */
        fid = fopen(RESTART_FIELDS, "r");
        if (!fid) die("Error opening restart fields file");
        size_t got = 0;
        for (int k = 0; k < NZv; k++)
            for (int j = 0; j < NYv; j++)
                for (int i = 0; i < NXv; i++) {
                    size_t idx = k + (size_t)NZv * (j + (size_t)NYv * i);
                    double v[14];
                    if (fscanf(fid, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &v[0], &v[1], &v[2], &v[3],
                               &v[4], &v[5], &v[6], &v[7], &v[8], &v[9], &v[10], &v[11], &v[12], &v[13]) != 14)
                        goto fields_done;
                    for (int f = 0; f < 13; f++) w[(size_t)f * V + idx] = v[f];
                    pf[idx] = v[13];
                    got++;
                }
    fields_done:
        fclose(fid);
        if (got < V)
            fprintf(stderr, "warning: %s holds %zu of %zu voxel lines; the rest stay 0\n", RESTART_FIELDS, got, V);
        fid = fopen(RESTART_RANDOM, "r");
        if (!fid) die("Error opening restart random file");
        if (fscanf(fid, "%hu %hu %hu", &iSeed[0], &iSeed[1], &iSeed[2]) != 3) die("restart_random.dat");
        fclose(fid);
    }
    lcg_seed(&gen, iSeed);
    const char *host_rng_env = getenv("SCFT_HOST_RNG");
    const int host_rng = host_rng_env && strcmp(host_rng_env, "0") != 0;
    if (!host_rng) check(scft_rng_seed(ctx, iSeed), "scft_rng_seed");
    check(scft_set_fields(ctx, w, pf), "scft_set_fields");
    write_run_txt(&p, &d, iIter);

    const char *exact_env = getenv("SCFT_EXACT_RESTART");
    const int exact = exact_env && strcmp(exact_env, "0") != 0;
    const size_t ndraw = (size_t)d.n_active * V;
    double *noise = host_rng ? calloc(ndraw ? ndraw : 1, sizeof(double)) : NULL;
    /* snapshots: two host buffers, each written out by its own thread */
    snap_job jobs[2];
    for (int b = 0; b < 2; b++) {
        jobs[b].nx = NXv;
        jobs[b].ny = NYv;
        jobs[b].nz = NZv;
        jobs[b].exact = exact;
        jobs[b].w = calloc(13 * V, sizeof(double));
        jobs[b].pf = calloc(V, sizeof(double));
        jobs[b].phi = calloc(13 * V, sizeof(double));
        if (!jobs[b].w || !jobs[b].pf || !jobs[b].phi) die("calloc");
    }
    pthread_t writer;
    int writing = 0;
    unsigned long nsnap = 0;
    while (iIter <= ITER) {                                           /* SCFT.c:1088 */
        iIter++;
        const float MF = (iIter > RELAXATION) ? 0.0f : 1.0f;           /* SCFT.c:1515-1522 */
        scft_iter_out o;
        draw_job job = {&gen, ndraw, NULL};
        if (!host_rng) {
            check(scft_iterate(ctx, iIter, MF, NULL, &o), "scft_iterate");
        } else if (MF != 0.0f) {
            job.out = noise;
            draw_thread(&job);
            check(scft_iterate(ctx, iIter, MF, noise, &o), "scft_iterate");
        } else {
            pthread_t th;
            const int threaded = pthread_create(&th, NULL, draw_thread, &job) == 0;
            if (!threaded) draw_thread(&job);
            check(scft_iterate(ctx, iIter, MF, NULL, &o), "scft_iterate");
            if (threaded) pthread_join(th, NULL);
        }
        if (fmod((double)iIter, D_WRT_ENERGIES) == 0) {               /* SCFT.c:1583-1599 */
            fid = fopen(ENERGY, "a");
            fprintf(fid, "%lu %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf\n", iIter, o.chempot[0],
                    o.chempot[1], o.chempot[2], o.chempot[3], o.chempot[4], o.chempot[5], o.chempot[6], o.chempot[7],
                    o.chempot[8], o.chempot[9], o.chempot[10], o.chempot[11], o.chempot[12], o.chempot[13], o.chempot[14],
                    o.chempot[15], o.chempot[16]);
            fclose(fid);
            fid = fopen(FREE_ENERGY, "a");
            fprintf(fid, "%lu  %lf  %f \n", iIter, o.free_energy, MF);
            fclose(fid);
            fid = fopen(DEVIATIONS, "a");
            fprintf(fid, "%lu  %lf %lf %lf %lf \n", iIter, o.rmsdev, d.chi_eff[0][1], d.chi_eff[0][2], d.chi_eff[1][2]);
            fclose(fid);
        }
        if (fmod((double)iIter, D_WRT) == 0 || iIter == 1) {          /* SCFT.c:1601-1664 */
            /* the previous snapshot's writer must be done with its buffers */
            if (writing) {
                pthread_join(writer, NULL);
                writing = 0;
            }
            snap_job *j = &jobs[nsnap++ & 1];
            j->iter = iIter;
            check(scft_get_fields(ctx, j->w, j->pf), "scft_get_fields");
            check(scft_get_densities(ctx, j->phi, NULL), "scft_get_densities");
            if (host_rng) lcg_get(&gen, j->rng);
            else check(scft_rng_state(ctx, j->rng), "scft_rng_state");
            if (pthread_create(&writer, NULL, snap_thread, j) == 0) writing = 1;
            else snap_thread(j);
        }
    }
    if (writing) pthread_join(writer, NULL);
    for (int b = 0; b < 2; b++) {
        free(jobs[b].w);
        free(jobs[b].pf);
        free(jobs[b].phi);
    }
    free(noise);
    free(w);
    free(pf);
    scft_destroy(ctx);
    return 0;
}
