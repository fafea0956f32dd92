/*
 * genpk_amd -- MI355X (gfx950) implementation of GenPK's power-spectrum path:
 * CIC deposit -> 3-D r2c FFT -> CIC-deconvolved log-binned P(k).
 *
 * Two groups of entry points, all C ABI, all returning 0 on success and a
 * nonzero status on failure (message from genpk_last_error(); no C++
 * exception crosses the boundary):
 *
 * 1. Device-resident plan API (stream ordered).  Pointers are device
 *    pointers; work is enqueued on the plan's HIP stream and is complete
 *    once that stream is synchronised.  A plan is not safe for concurrent
 *    use from several host threads; distinct plans are independent.
 *
 * 2. Drop-in replacements of the reference's host-pointer call surface
 *    (gen-pk.h), declared in genpk_compat.h.
 */
#ifndef GENPK_AMD_H
#define GENPK_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Precision of the grid and FFT: bytes per real.  Bin sums are FP64 in
 * both.  GENPK_FLOAT64 is the reference's precision (GENFLOAT = double,
 * gen-pk.h:31). */
#define GENPK_FLOAT32 4
#define GENPK_FLOAT64 8

typedef struct genpk_plan genpk_plan;

/* Message of the last failure on the calling thread ("" if none). */
const char *genpk_last_error(void);

/* Plan for a dims^3 grid.  dims: 2..4096 (the FFT needs an even dims whose
 * half factors into 2, 3 and 5; deposit-only plans accept any dims >= 1).
 * precision: GENPK_FLOAT32 or GENPK_FLOAT64.  stream: hipStream_t (NULL =
 * default stream).  Replaces the reference's grid allocation and FFTW plan
 * creation (gen-pk.cpp:169-193). */
int genpk_plan_create(genpk_plan **plan, int64_t dims, int precision, void *stream);
int genpk_plan_destroy(genpk_plan *plan);
int genpk_plan_set_stream(genpk_plan *plan, void *stream);
/* Bytes of the padded grid d*d*2(d/2+1) reals of the plan's precision. */
int genpk_plan_field_bytes(const genpk_plan *plan, size_t *bytes);
/* Device pointer of the plan's padded grid (real density or spectrum). */
int genpk_plan_field_ptr(genpk_plan *plan, void **ptr);
/* Make the plan work in a caller-owned device buffer of
 * genpk_plan_field_bytes() bytes (e.g. a framework tensor) instead of its
 * own grid; the caller keeps it alive until the plan is destroyed. */
int genpk_plan_attach_field(genpk_plan *plan, void *ptr);

/* Tuning options of a plan (no counterpart in the reference; the batched
 * form's autotuner picks them per grid size):
 *   tile_x, tile_y, tile_z: deposit tile edges in cells, powers of two
 *     (each clamped to the largest power of two <= dims; tile_z >= 2), whose
 *     halo'd tile of 64-bit accumulators, (tile_x+1)(tile_y+1)(tile_z+1) x 8
 *     bytes, fits in LDS (160 KB).  Default 16 x 32 x 32.
 *   fft_variant: radix ordering of the transforms: 0 = radix 8 first (the
 *     fewest stages; the FP32 default), 1 = the same radices with the
 *     leftover 4, 2, 3, 5 first, 2 = radix 4 first (no radix 8), 3 = one
 *     radix 4 first in the y and x passes, ordering 0 in the z pass (the
 *     FP64 default).
 * Changing the tiles drops the plan's deposits and its kept tile order. */
int genpk_plan_set_options(genpk_plan *plan, int tile_x, int tile_y, int tile_z, int fft_variant);
/* The plan's options in effect (tile edges after clamping to dims). */
int genpk_plan_get_options(const genpk_plan *plan, int *tile_x, int *tile_y, int *tile_z, int *fft_variant);
/* The radices, in stage order, of a length-n transform under fft_variant
 * (host only; the y and x passes use n = dims, the z passes n = dims/2,
 * with ordering 0 under variant 3).  fft_variant -1: the ordering with a
 * first radix 8, a last radix 2 and an even number of stages that the FP32
 * fused x pass + bin takes where the plan's own ordering is not of that
 * form (an error for lengths without one, such as 32 and 48). */
int genpk_radix_plan(int64_t n, int fft_variant, int *radices, int max, int *count);

/* Empty the density (the reference zeroes the grid, gen-pk.cpp:208). */
int genpk_clear(genpk_plan *plan);

/* CIC deposit of n particles, accumulating into the plan's density, as
 * fieldize(boxsize, dims, field, n, pos, masses, mass, 1) does
 * (fieldize.cpp:46-114).  pos: float [n][3]; masses: float [n] or NULL
 * (then every particle has mass `mass`). */
int genpk_deposit(genpk_plan *plan, double boxsize, int64_t n, const float *pos, const float *masses, double mass);

/* genpk_deposit that also keeps the particles' tile order (with each
 * particle's index) in the plan, so that a following genpk_adjoint_gather
 * of the same arrays (same pointers, n and boxsize, contents unchanged)
 * skips its own partition.  Any later deposit drops the kept order.  Kept
 * only when n <= 2^31.  No counterpart in the reference (which has no
 * gradients). */
int genpk_deposit_retain(genpk_plan *plan, double boxsize, int64_t n, const float *pos, const float *masses,
                         double mass);

/* Deposit again the particles whose tile order the last
 * genpk_deposit_retain (or genpk_adjoint_gather) kept, from the kept sorted
 * copy of their positions and masses (with the same boxsize and mass): the
 * tile accumulation runs without the partition.  This is how an adjoint
 * recomputes the forward density instead of storing the field.  Fails if no
 * order is kept.  No counterpart in the reference. */
int genpk_redeposit(genpk_plan *plan);

/* Forget the kept tile order, so that the next genpk_adjoint_gather
 * partitions the particles again (for callers whose particle arrays changed
 * in place since the retaining deposit). */
int genpk_drop_order(genpk_plan *plan);

/* Write the real-space density into out, a padded grid [d][d][out_rowlen]
 * of the plan's precision; accumulate != 0 adds to out.  Pad columns
 * (z >= d) are left untouched. */
int genpk_export_density(genpk_plan *plan, void *out, int64_t out_rowlen, int accumulate);

/* Replace the density by a caller grid [d][d][2(d/2+1)] of the plan's
 * precision (pads ignored); later deposits add to it. */
int genpk_import_density(genpk_plan *plan, const void *grid);

/* Replace the plan's content by a caller half spectrum [d][d][d/2+1]
 * (complex of the plan's precision), e.g. one saved by an earlier
 * genpk_r2c; the plan then holds a spectrum (for genpk_bin or the
 * adjoint). */
int genpk_import_spectrum(genpk_plan *plan, const void *spectrum);

/* In-place forward r2c of the density: the FFTW step of gen-pk.cpp:233.
 * Afterwards the plan's grid holds the complex half spectrum
 * [d][d][d/2+1] (unnormalised, sign -1). */
int genpk_r2c(genpk_plan *plan);

/* Unnormalised inverse (sign +1) of genpk_r2c on the plan's grid. */
int genpk_c2r(genpk_plan *plan);

/* Bin transformed spectra: powerspectrum(dims, Fa, Fb, nrbins, power,
 * count, keffs, ma, mb) of powerspectrum.c:35-110.  b = NULL: auto
 * spectrum of a into power_aa.  b != NULL: power_aa = P_aa (normalised by
 * ma*ma), power_bb = P_bb (mb*mb), power_ab = Re(Fa Fb*) (ma*mb); any of
 * the three may be NULL.  count (int [nrbins]) and keffs (double [nrbins])
 * may be NULL.  All outputs are device pointers. */
int genpk_bin(genpk_plan *a, genpk_plan *b, int nrbins, double *power_aa, double *power_bb, double *power_ab,
              int *count, double *keffs, double ma, double mb);

/* Fused forward: r2c of a (and b) whose last FFT pass bins the modes while
 * they are in LDS; the spectrum is not kept.  The plan's grid is left in
 * the state after the z and y passes: the plan then holds no density (a
 * deposit starts from zero), but genpk_adjoint_density can start from it.
 * Outputs as genpk_bin. */
int genpk_power(genpk_plan *a, genpk_plan *b, int nrbins, double *power_aa, double *power_bb, double *power_ab,
                int *count, double *keffs, double ma, double mb);

/* genpk_power that allocates no device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, work and results as genpk_power; the per-workgroup partial bins
 * live in a fixed 16 MiB device array of the library (one per GPU, held
 * until the process ends) instead of a plan buffer.  Requirements: the bin
 * tables of nrbins were built on plan a by an earlier genpk_power or
 * genpk_bin (else an error), and the partial bins (resident workgroups x
 * spectra x nrbins, e.g. 256 x 3 x 2730) fit in 2^21 doubles (else an error
 * before any work: use fewer bins).  Calls that overlap on one GPU share the
 * array and race.  (With stage timing enabled on the plan, the timer may
 * create hipEvents; it allocates no device memory either.) */
int genpk_power_noalloc(genpk_plan *a, genpk_plan *b, int nrbins, double *power_aa, double *power_bb,
                        double *power_ab, int *count, double *keffs, double ma, double mb);

/* Adjoint of the auto spectrum (no counterpart in the reference).  The
 * plan must hold one of: the forward spectrum (genpk_r2c or
 * genpk_import_spectrum), the state left by genpk_power with nothing
 * deposited since, or the forward density (deposits and/or an imported
 * density, as before genpk_power; the forward transform is then redone).
 * lambda: device double [nrbins] = dL/dP_b; m: the total mass the forward
 * normalised by.  Afterwards the plan's grid holds dL/drho on the padded
 * real layout. */
int genpk_adjoint_density(genpk_plan *plan, int nrbins, const double *lambda, double m);

/* Chain dL/drho to the particles: grad_pos float [n][3] = dL/dpos,
 * grad_mass float [n] = dL/dm_j (may be NULL).  pos, masses, mass, boxsize
 * as given to genpk_deposit.  Uses the tile order kept by
 * genpk_deposit_retain for the same arrays, else partitions the particles
 * itself (and keeps that order). */
int genpk_adjoint_gather(genpk_plan *plan, double boxsize, int64_t n, const float *pos, const float *masses,
                         double mass, float *grad_pos, float *grad_mass);

/* Batched forward (no counterpart in the reference, which runs one
 * realisation per process): many independent realisations with uneven
 * particle counts and grid sizes, captured into one hipGraph and launched
 * with one call.  Jobs run on `branches` parallel graph branches, each with
 * its own stream and plans.
 *   genpk_batch_create: stream = the stream launches are ordered on.
 *   genpk_batch_add: one job = genpk_clear + genpk_deposit + genpk_power of
 *     a dims^3 plan of `precision`: device pointers pos [n][3], masses [n] or
 *     NULL; outputs power/keffs double [nrbins], count int [nrbins] (count
 *     and keffs may be NULL); normalised by total_mass^2.  The buffers must
 *     stay valid while the batch exists.
 *   genpk_batch_build: assigns jobs to branches (branch_of_job[i] in
 *     [0, branches), or NULL for longest-processing-time-first balancing),
 *     runs every job once (sizing all buffers), captures and instantiates
 *     the graph.  Synchronises the batch stream first.
 *   genpk_batch_launch: one graph launch on the batch stream; the outputs are
 *     complete when that stream is synchronised.  Launches may be repeated
 *     (e.g. after new positions are written into the same buffers).
 *   genpk_batch_branch_of: the branch a job was assigned to.
 *   genpk_batch_plan_options: genpk_plan_set_options for the batch's plans
 *     of one dims and precision (before genpk_batch_build). */
typedef struct genpk_batch genpk_batch;
int genpk_batch_create(genpk_batch **batch, int branches, void *stream);
int genpk_batch_add(genpk_batch *batch, int64_t dims, int precision, double boxsize, int64_t n, const float *pos,
                    const float *masses, double mass, int nrbins, double total_mass, double *power, int *count,
                    double *keffs);
int genpk_batch_plan_options(genpk_batch *batch, int64_t dims, int precision, int tile_x, int tile_y, int tile_z,
                             int fft_variant);
int genpk_batch_build(genpk_batch *batch, const int *branch_of_job);
int genpk_batch_launch(genpk_batch *batch);
int genpk_batch_branch_of(const genpk_batch *batch, int64_t job, int *branch);
int genpk_batch_destroy(genpk_batch *batch);

/* Stage timing (no counterpart in the reference, which has no timers):
 * while enabled, hipEvents are recorded on the plan's stream around each
 * stage (sort passes, tile deposit, fold, FFT passes, binning, adjoint).
 * genpk_timing_read synchronises the stream, returns up to max entries of
 * (stage name, total milliseconds, launches) accumulated since the last
 * read and clears them; names are written NUL-separated into names. */
int genpk_timing_enable(genpk_plan *plan, int enable);
int genpk_timing_read(genpk_plan *plan, char *names, size_t names_len, double *ms, int *launches, int max,
                      int *count);

/* Host-side utilities of the reference (utils.cpp:35-42, 57-72). */
int genpk_nexttwo(int n);
/* Writes the file-name token of a particle type ("by", "DM", "nu", "st",
 * "xx") into out (>= 3 bytes). */
int genpk_type_str(int type, char *out, size_t len);
/* print_pk (utils.cpp:7-21): one "%e\t%e\t%d\n" line per bin with
 * count > 0.  Host pointers. */
int genpk_print_pk(const char *filename, int nrbins, const double *keffs, const double *power, const int *count);

#ifdef __cplusplus
}
#endif

#endif /* GENPK_AMD_H */
