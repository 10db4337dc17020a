/*
 * scft.h — C ABI of libscft, an MI355X (gfx950) implementation of the
 * iteration loop of the pseudo-spectral SCFT solver SCFT.c
 * (gahme-glitch/stimuli-responsive-nanodroplets).
 *
 * The reference has no library boundary: its loop body calls file-static
 * C functions from main().  Each entry point below names the reference
 * code it replaces.  Conventions of this ABI:
 *   - every function returns int: 0 = ok, < 0 = error (SCFT_E*);
 *   - plain pointers and sizes only, host memory on the host side;
 *   - a context owns its device memory and one HIP stream; contexts are
 *     independent, so two threads may drive two contexts;
 *   - grid arrays use the reference index idx = k + NZ*(j + NY*i)
 *     (z fastest), field arrays are [field][idx] with field order
 *     A..J, S1, S2, S3 (0..12) as in SCFT.c:2001-2015.
 */
#ifndef SCFT_H
#define SCFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCFT_NFIELDS   13   /* A..J, S1..S3 */
#define SCFT_NSPECIES  17   /* dbk1..4, homo A..J, S1..S3 (energies.dat order, SCFT.c:1586-1589) */

enum {
    SCFT_OK          = 0,
    SCFT_EINVAL      = -1,  /* bad argument */
    SCFT_EHIP        = -2,  /* a HIP runtime call failed */
    SCFT_ENOMEM      = -3,  /* device allocation failed */
    SCFT_EVOLUME     = -4,  /* free-chain volumes do not sum to 1 (SCFT.c:84-101) */
    SCFT_EUNSUPPORTED= -5,  /* grid/transform size outside what the kernels support */
    SCFT_ENONFINITE  = -6,  /* a NaN/Inf appeared in a reduction */
    SCFT_ERNG        = -7   /* device noise stream: too few accepted candidate pairs (see scft_rng_seed) */
};

enum {
    SCFT_FP64         = 0,  /* everything in double precision */
    SCFT_FP32_SPLIT   = 1,  /* FP32 propagator slices (stored with unbiased rounding), FP64 transforms,
                               Kahan-compensated FP64 contour sums, FP64 Q, F, update */
    SCFT_FP32_NEAREST = 2   /* as SCFT_FP32_SPLIT, slices rounded to nearest.  Rounding to nearest
                               erases perturbations below an FP32 step that FP64 carries, so an
                               unstable symmetric state that FP64 leaves can look stable */
};

/*
 * Run parameters: the macros of parameters.h:1-179 as a runtime struct.
 * chi is the symmetric CHI_N matrix over fields A..J,S1..S3 before the
 * zeroing of SCFT.c:303-396 (the diagonal is ignored).
 */
typedef struct scft_params {
    double lx, ly, lz;          /* LX, LY, LZ (R_g)                     parameters.h:2-4   */
    double dr_target;           /* DR_TARGET                            parameters.h:19    */
    int    ns;                  /* NS                                   parameters.h:18    */
    int    mask;                /* MASK: 0 periodic, 1 Dirichlet z      parameters.h:11    */
    double g1;                  /* G1                                   parameters.h:21    */
    double dt_noise;            /* DT_NOISE                             parameters.h:10    */
    unsigned long relaxation;   /* RELAXATION                           parameters.h:8     */
    int    enable_seed;         /* ENABLE_SEED                          parameters.h:14    */
    int    seed_steps;          /* SEED_STEPS                           parameters.h:15    */
    int    seed_area_delta;     /* SEED_AREA_DELTA                      parameters.h:16    */
    double chi[SCFT_NFIELDS][SCFT_NFIELDS];
    double f_dbk[4], len_dbk[4], vol_dbk[4];      /* parameters.h:77-87   */
    double vol_homo[10], len_homo[10];            /* parameters.h:89-109  */
    double vol_s[3], len_s[3];                    /* parameters.h:112-119 */
} scft_params;

/* Quantities the setup of SCFT.c:84-396, 718-734 derives from the parameters. */
typedef struct scft_derived {
    int    nx, ny, nz;                 /* SCFT.c:103-105 */
    int    nzext;                      /* N = 2(NZ+1) for MASK 1, else NZ (SCFT.c:680) */
    int    ns_sp[SCFT_NSPECIES];       /* NS per species, 0 = absent (SCFT.c:121-288) */
    int    ns_x[4], ns_y[4];           /* diblock block lengths */
    double alpha[SCFT_NSPECIES];       /* (NS_p - 1)*DS */
    double phi_avg[SCFT_NFIELDS];      /* phi_avgs[] after the coating rescale */
    double chi_eff[SCFT_NFIELDS][SCFT_NFIELDS];
    double coating_vol;
    int    active[SCFT_NFIELDS];       /* update-active fields (SCFT.c:2017-2042) */
    int    n_active;
} scft_derived;

/* Per-iteration scalars (all FP64). */
typedef struct scft_iter_out {
    double free_energy;                /* H_fe after the ln Q terms (SCFT.c:1525-1555) */
    double rmsdev;                     /* SCFT.c:1564-1580 */
    double dw_rms;                     /* sqrt(mean over active fields and voxels of (w_new - w_old)^2) */
    double qsum[SCFT_NSPECIES];        /* Q_p (0 for absent species) */
    double chempot[SCFT_NSPECIES];     /* -ln(Q_p^2)/2, 0 if absent (SCFT.c:1862-1869) */
    float  mf;                         /* MF (SCFT.c:1515-1522) */
} scft_iter_out;

/* Setup derivation only (host logic, no device). Replaces SCFT.c:84-396, 718-734. */
int scft_derive(const scft_params *p, scft_derived *out);

typedef struct scft_ctx scft_ctx;

/* Allocate device state for one SCFT instance on the current HIP device.
 * Replaces the allocation and planning of SCFT.c:433-701 and the
 * propagator start slices / k-space operator of SCFT.c:881-1023. */
int scft_create(const scft_params *p, int precision, scft_ctx **out);
int scft_destroy(scft_ctx *ctx);
int scft_get_derived(const scft_ctx *ctx, scft_derived *out);

/* Field transfer (restart / seeding / snapshots; SCFT.c:844-866, 1601-1650).
 * w: 13*V doubles [field][idx]; p: V doubles. */
int scft_set_fields(scft_ctx *ctx, const double *w, const double *p);
int scft_get_fields(scft_ctx *ctx, double *w, double *p);
/* Brush start masks, 3*V doubles [S1,S2,S3][idx] (wMask.dat, SCFT.c:756-781). */
int scft_set_mask(scft_ctx *ctx, const double *mask);
/* Total densities, 13*V doubles [A_tot..J_tot, S1, S2, S3][idx], and phi_tot (V). */
int scft_get_densities(scft_ctx *ctx, double *phi, double *phi_tot);
/* One propagator slice s of species sp (0..16); which = 0 q, 1 q-dagger (V doubles). */
int scft_get_propagator(scft_ctx *ctx, int sp, int which, int s, double *out);

/*
 * scft_propagate: every propagator of one iteration, in the species order
 * and block bounds of SCFT.c:1092-1456 (funPDESolver / funPDESolverBrush,
 * SCFT.c:2310-2472, with the periodic and Dirichlet steps of
 * SCFT.c:2474-2631).  iter is the iteration number after the increment of
 * SCFT.c:1090 (it selects the dbk1 seed start, SCFT.c:2346-2367).
 */
int scft_propagate(scft_ctx *ctx, unsigned long iter);

/*
 * scft_densities: Q_p and chemical potentials (funChemPotential,
 * SCFT.c:1837-1872), q*q-dagger contour integration (funMultiplyQs +
 * funDensity, SCFT.c:2243-2308), volume scaling and totals
 * (funVolFrac/funAddArray, SCFT.c:1458-1513).  qsum/chempot: 17 doubles
 * each, or NULL.
 */
int scft_densities(scft_ctx *ctx, double *qsum, double *chempot);

/*
 * scft_update_fields: free energy with pre-update fields (funFreeEnergy
 * and the ln Q terms, SCFT.c:1525-1555, 1874-1917), the field/pressure
 * update (funUpdate, SCFT.c:1993-2124) and RMSDev (SCFT.c:1564-1580).
 * noise: NULL, or the Gaussian deviates in the reference's draw order
 * (voxel k->j->i outer, active field inner), n_active*V doubles; it is
 * used only when mf != 0 (amplitude dt_noise*mf, SCFT.c:2103).
 */
int scft_update_fields(scft_ctx *ctx, unsigned long iter, float mf,
                       const double *noise, scft_iter_out *out);

/* One whole iteration: propagate + densities + update (SCFT.c:1088-1580). */
int scft_iterate(scft_ctx *ctx, unsigned long iter, float mf,
                 const double *noise, scft_iter_out *out);

/*
 * Device replay of the noise stream.  The reference draws one deviate per
 * active field and voxel every iteration, also when MF = 0 (funGaussDev,
 * SCFT.c:1769-1782, called at 2103), from drand48 seeded by seed48 or
 * restart_random.dat (SCFT.c:789-797, 870-879) and saved to
 * restart_random.dat (SCFT.c:1652-1663).  scft_rng_seed starts the same
 * stream on the GPU from a seed48 state (xsubi[0] = low 16 bits); from then
 * on scft_iterate / scft_update_fields take noise = NULL, draw the
 * iteration's deviates on the device in the reference's order and add
 * them when mf != 0.  scft_rng_state returns the state after the last
 * completed iteration (what restart_random.dat holds).  These act on slot
 * 0; scft_slot_rng_seed / scft_slot_rng_state below give every slot of a
 * batch its own stream.  Replaces the host-side drand48 loop of
 * SCFT.c:2103.  The candidate pairs drawn per iteration are sized for the
 * expected acceptance (pi/4) plus a margin of about 17 standard deviations;
 * if an iteration still finds too few accepted pairs the call returns
 * SCFT_ERNG and the generator state stays where it was, but that
 * iteration's field update has already added an incomplete set of
 * deviates: restore the fields from before the call before going on.
 */
int scft_rng_seed(scft_ctx *ctx, const unsigned short xsubi[3]);
int scft_rng_state(scft_ctx *ctx, unsigned short xsubi[3]);

/*
 * Batches.  A batch context holds nslots independent SCFT instances that
 * share one chain structure (species set, NS, lengths and block fractions,
 * MASK, LY, LZ, DR_TARGET, and whether LX == 1); chi, volumes, G1 and LX
 * (so NX <= max_nx) may differ per slot.  One scft_batch_iterate call
 * advances every active slot by one iteration; the propagator launches of
 * a contour step cover every chain of every active slot.  scft_create is a
 * batch of one slot, and the single-instance calls above act on slot 0.
 * Not in the reference (one process per SCFT run there); it serves the
 * phase-diagram workflow of README.md:59-128, whose trial solves are
 * independent SCFT runs.
 */
int scft_batch_create(const scft_params *proto, int nslots, int max_nx, int precision, scft_ctx **out);
int scft_batch_size(const scft_ctx *ctx);
/* (Re)initialise a slot for a new parameter set: fields and densities
 * zeroed, q[0] = 1; the slot is left inactive. */
int scft_slot_init(scft_ctx *ctx, int slot, const scft_params *p);
int scft_slot_activate(scft_ctx *ctx, int slot, int on);
int scft_slot_get_derived(const scft_ctx *ctx, int slot, scft_derived *out);
int scft_slot_set_fields(scft_ctx *ctx, int slot, const double *w, const double *p);
int scft_slot_get_fields(scft_ctx *ctx, int slot, double *w, double *p);
/* Device-to-device field copy (seed a trial solve from a converged slot). */
int scft_slot_copy_fields(scft_ctx *ctx, int dst_slot, int src_slot);
int scft_slot_set_mask(scft_ctx *ctx, int slot, const double *mask);
int scft_slot_get_densities(scft_ctx *ctx, int slot, double *phi, double *phi_tot);
int scft_slot_get_propagator(scft_ctx *ctx, int slot, int sp, int which, int s, double *out);
/* A slot's own noise stream (see scft_rng_seed; no counterpart in the
 * reference, which runs one instance per process).  A slot with a stream
 * draws n_active*V deviates every batch iteration and adds them while its
 * iteration is within its RELAXATION (MF = 1, SCFT.c:1515-1522).
 * scft_slot_init turns the stream off. */
int scft_slot_rng_seed(scft_ctx *ctx, int slot, const unsigned short xsubi[3]);
int scft_slot_rng_state(scft_ctx *ctx, int slot, unsigned short xsubi[3]);
/* One iteration of every active slot.  iters[slot] (or NULL: no seed box,
 * no noise) is the slot's iteration counter: it selects the dbk1 seed
 * start and, for slots with a noise stream, MF.  A slot whose iteration is
 * within its RELAXATION but that has no stream is an error (the reference
 * adds noise there).  outs[slot] is filled for active slots. */
int scft_batch_iterate(scft_ctx *ctx, const unsigned long *iters, scft_iter_out *outs);
/* scft_batch_iterate in two halves, so that the host can run other
 * contexts meanwhile: scft_batch_launch enqueues the iteration on the
 * context's stream and returns; scft_batch_complete waits for it and fills
 * outs.  Two contexts launched one after the other run on the GPU
 * concurrently (each has its own stream).  Between the two calls every
 * other call on this context except scft_sync, scft_destroy and the
 * read-only queries (scft_batch_size, scft_get_derived,
 * scft_slot_get_derived, scft_get_schedule) returns SCFT_EINVAL.  Not in
 * the reference: it serves solves of several chain structures at once. */
int scft_batch_launch(scft_ctx *ctx, const unsigned long *iters);
int scft_batch_complete(scft_ctx *ctx, scft_iter_out *outs);
/* Without waiting: 1 when the iteration started by scft_batch_launch has
 * finished on the GPU (scft_batch_complete then returns at once), 0 while
 * it runs, < 0 on error.  Lets a host loop advance each context at its own
 * pace. */
int scft_batch_query(scft_ctx *ctx);
/* Device bytes one slot of a batch with prototype p needs (fields,
 * propagators, scratch); the serving layer admits solves by this footprint. */
int scft_slot_bytes(const scft_params *p, int max_nx, int precision, size_t *bytes);
/* Free and total memory of the current HIP device. */
int scft_device_memory(size_t *free_bytes, size_t *total_bytes);
/* Measurement hook: run every launch of one kernel kind of the current
 * schedule (kind 0: z operator, 1: last axis) reps times between two HIP
 * events on the context's stream; returns the mean launch duration (us),
 * the bytes the kernel reads and writes (slices, scratch and Boltzmann
 * factors), the flops and the voxel-steps (grid points times chains) of
 * the mean launch.  Overwrites the chains' scratch and propagator slices,
 * so call it between solves.  No counterpart in the reference. */
int scft_time_launches(scft_ctx *ctx, int kind, int reps, double *avg_us, double *bytes, double *flops,
                       double *voxel_steps, int *nlaunch);
/* Human-readable schedule summary (z operator kind, launches per iteration). */
int scft_get_schedule(const scft_ctx *ctx, char *buf, int len);

/*
 * scft_step_noalloc: one contour step of one chain, q_out = E M(E q_in)
 * with E = exp(-DS w/2), on device arrays the caller owns (V doubles each,
 * idx = k + NZ*(j + NY*i)): w_dev the chain's field, q_in_dev the slice s,
 * q_out_dev the slice s + 1.  The same step as scft_propagate takes per
 * slice (M the periodic or, for MASK 1, Dirichlet z transform and the
 * periodic x and y transforms), without a context.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  It allocates nothing: its
 * tables, E and the scratch between axis operators live in one static
 * device array per GPU (about 4.1 MB, held until the process ends).
 * Limits: V <= 262144 voxels and transform lengths (NX, NY, and NZ, or
 * 2(NZ+1) for MASK 1) <= 1024, else SCFT_EUNSUPPORTED.  Calls on one GPU
 * share that array: calls that overlap on one GPU race, so callers on
 * several streams must order them.  stream: the hipStream_t to launch on
 * (NULL for the null stream); the call returns once the work is queued.
 */
int scft_step_noalloc(const scft_params *p, const double *w_dev, const double *q_in_dev, double *q_out_dev,
                      void *stream);

/* Wait for the context's stream. */
int scft_sync(scft_ctx *ctx);
/* Last error text of this thread (static storage). */
const char *scft_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* SCFT_H */
