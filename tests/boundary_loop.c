/*
 * boundary_loop.c — the reference's loop body bound to libscft through
 * include/scft.h exactly as a maintainer keeping SCFT.c's own main() would
 * bind it: the three calls scft_propagate -> scft_densities ->
 * scft_update_fields per iteration (SCFT.c:1088-1580), with the noise
 * deviates drawn on the host by drand48 + the polar method in the
 * reference's draw order (SCFT.c:1769-1782, 2103) and handed over.
 *
 * Test program (tests/test_gpu_boundary.py runs it; __graft_entry__.build()
 * compiles it).  Input file: scft_params (raw struct), int32 first
 * iteration, int32 iteration count, int32 rng flag, 3 x u16 seed48 state,
 * then w [13][V], p [V] and, for MASK 1, the brush masks [3][V] (doubles).
 * Output file: per iteration F, RMSDev, ||dw||, Q[17], -lnQ[17] from
 * scft_densities and -lnQ[17] from scft_update_fields, then the final w
 * and p.
 */
#define _DEFAULT_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "scft.h"

static double gauss_dev(void)
{
    double x1, x2, w;
    do {
        x1 = 2.0 * drand48() - 1.0;
        x2 = 2.0 * drand48() - 1.0;
        w = x1 * x1 + x2 * x2;
    } while (w >= 1.0);
    w = sqrt((-2.0 * log(w)) / w);
    return x1 * w;
}

static void die(const char *what, int rc)
{
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, scft_last_error());
    exit(1);
}

#define CALL(x) do { int rc_ = (x); if (rc_ != SCFT_OK) die(#x, rc_); } while (0)

static void read_all(FILE *f, void *dst, size_t n)
{
    if (fread(dst, 1, n, f) != n) {
        fprintf(stderr, "short input file\n");
        exit(1);
    }
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
        return 2;
    }
    FILE *in = fopen(argv[1], "rb");
    if (!in) {
        perror(argv[1]);
        return 1;
    }
    scft_params P;
    int32_t first, count, use_rng;
    unsigned short seed[3];
    read_all(in, &P, sizeof P);
    read_all(in, &first, sizeof first);
    read_all(in, &count, sizeof count);
    read_all(in, &use_rng, sizeof use_rng);
    read_all(in, seed, sizeof seed);

    scft_derived D;
    CALL(scft_derive(&P, &D));
    const size_t V = (size_t)D.nx * D.ny * D.nz;
    double *W = malloc(13 * V * sizeof(double));
    double *aP = malloc(V * sizeof(double));
    double *mask = malloc(3 * V * sizeof(double));
    double *noise = malloc(13 * V * sizeof(double));
    if (!W || !aP || !mask || !noise) {
        fprintf(stderr, "out of host memory\n");
        return 1;
    }
    read_all(in, W, 13 * V * sizeof(double));
    read_all(in, aP, V * sizeof(double));
    if (P.mask) read_all(in, mask, 3 * V * sizeof(double));
    fclose(in);

    scft_ctx *ctx;
    CALL(scft_create(&P, SCFT_FP64, &ctx));
    if (P.mask) CALL(scft_set_mask(ctx, mask));
    CALL(scft_set_fields(ctx, W, aP));
    if (use_rng) seed48(seed);

    FILE *out = fopen(argv[2], "wb");
    if (!out) {
        perror(argv[2]);
        return 1;
    }
    for (unsigned long iIter = (unsigned long)first; iIter < (unsigned long)first + (unsigned long)count; ++iIter) {
        const float MF = (iIter > P.relaxation) ? 0.0f : 1.0f;          /* SCFT.c:1515-1522 */
        if (use_rng)                                                    /* SCFT.c:2103, every iteration */
            for (size_t d = 0; d < (size_t)D.n_active * V; ++d) noise[d] = gauss_dev();
        double qsum[SCFT_NSPECIES], chempot[SCFT_NSPECIES];
        scft_iter_out o;
        CALL(scft_propagate(ctx, iIter));                               /* SCFT.c:1092-1456 */
        CALL(scft_densities(ctx, qsum, chempot));                       /* SCFT.c:1458-1513, 1837-1872, 2243-2308 */
        CALL(scft_update_fields(ctx, iIter, MF, use_rng ? noise : NULL, &o));  /* SCFT.c:1525-1580, 1874-2124 */
        double row[3] = {o.free_energy, o.rmsdev, o.dw_rms};
        fwrite(row, sizeof(double), 3, out);
        fwrite(qsum, sizeof(double), SCFT_NSPECIES, out);
        fwrite(chempot, sizeof(double), SCFT_NSPECIES, out);
        fwrite(o.chempot, sizeof(double), SCFT_NSPECIES, out);
    }
    CALL(scft_get_fields(ctx, W, aP));
    fwrite(W, sizeof(double), 13 * V, out);
    fwrite(aP, sizeof(double), V, out);
    fclose(out);
    CALL(scft_destroy(ctx));
    free(W);
    free(aP);
    free(mask);
    free(noise);
    return 0;
}
