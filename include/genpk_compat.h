/*
 * Drop-in replacements of the reference's power-spectrum call surface
 * (gen-pk.h).  Host pointers in and out; each call copies to the MI355X,
 * runs the HIP kernels of libgenpk_amd.so and copies back.  Same argument
 * meaning, accumulation and normalisation as the reference; a failure
 * returns nonzero (the reference always returns 0 and aborts on assertion)
 * and leaves the message in genpk_last_error().
 */
#ifndef GENPK_COMPAT_H
#define GENPK_COMPAT_H

#include <stdint.h>

/* fieldize.cpp:46 / gen-pk.h:93, C++ linkage (symbol _Z8fieldizediPdlPfS0_di):
 * CIC deposit of segment_particles particles onto out, a host FP64 grid
 * [dims][dims][2(dims/2+extra)], accumulating (+=).  masses may be NULL. */
#ifdef __cplusplus
int fieldize(double boxsize, int dims, double *out, int64_t segment_particles, float *positions, float *masses,
             double mass, int extra);
extern "C" {
#endif

/* fieldize.cpp:125 / gen-pk.h:104: (FP32 product of the three 1-D inverse
 * CIC window factors)^2; 0 for n = 0. */
double invwindow(int64_t kx, int64_t ky, int64_t kz, int64_t n);

/* powerspectrum.c:35 / gen-pk.h:119 with fftw_complex = double[2]: bins
 * Re(F1 F2*) of two host half spectra [dims][dims][dims/2+1] (may alias). */
int powerspectrum(int64_t dims, double (*outfield)[2], double (*outfield2)[2], int nrbins, double *power, int *count,
                  double *keffs, double total_mass, double total_mass2);

/* Replacement of the FFTW step the reference driver runs between the two
 * (fftw_plan_dft_r2c_3d(d,d,d,field,(fftw_complex*)field,FFTW_ESTIMATE) +
 * fftw_execute, gen-pk.cpp:193,233): in-place FP64 r2c of a host padded
 * grid [dims][dims][2(dims/2+1)], unnormalised, sign -1. */
int genpk_fft_r2c_3d_inplace(int64_t dims, double *field);

/* Resident mode of these shims (no counterpart in the reference; off by
 * default, on when the environment holds GENPK_COMPAT_RESIDENT=1 or after
 * genpk_compat_resident(1)).  It keeps the grid on the GPU across the
 * reference driver's sequence on one host array -- zero it, fieldize once
 * per chunk of particles (extra = 1), genpk_fft_r2c_3d_inplace, then
 * powerspectrum of that array (gen-pk.cpp:208-234, read_fieldize.cpp:
 * 43-94) -- so each forward moves the grid over PCIe once each way instead
 * of twice per fieldize call:
 *   - the first fieldize on an array loads it; later ones on the same array
 *     add to the device copy, and the host array lacks their deposits until
 *     the in-place FFT of that array (which writes the spectrum back, as
 *     eager mode does) or genpk_compat_flush;
 *   - powerspectrum of the array the last FFT wrote (as both fields) bins the
 *     device copy of that spectrum: the caller must not change the host
 *     array in between;
 *   - any call on another array (or powerspectrum of two arrays) first
 *     flushes, then runs as in eager mode.
 * Turning it off flushes. */
int genpk_compat_resident(int on);
/* Write a resident density back to its host array. */
int genpk_compat_flush(void);

#ifdef __cplusplus
}
#endif

#endif /* GENPK_COMPAT_H */
