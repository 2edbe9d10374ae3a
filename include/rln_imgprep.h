/* Particle image preparation on the GPU: the image part of MlOptimiser::getFourierTransformsAndCtfs
 * (src/ml_optimiser.cpp:5840-6531; GPU counterpart src/acc/acc_ml_optimiser_impl.h:11-1004) for 2-D particles
 * with one image each (no helical, multi-body or subtomogram branches, no aberration or MTF correction, no CTF
 * padding, no reconstruct images).
 *
 * Per particle, in the reference's order:
 *   1. norm correction: img *= avg_norm_correction / norm_correction            (src/ml_optimiser.cpp:6225-6237)
 *   2. integer re-centring by the rounded old offset, zero fill (selfTranslate DONT_WRAP, :6085, :6247)
 *   3. unmasked transform: FFT / N^2, window to current_size, CenterFFTbySign   (:6278-6289) -> fimg_nomask
 *   4. soft mask outside diameter/2 with a raised-cosine edge; background = average background value
 *      (zero_mask) or Gaussian noise of the model's noise spectrum              (:6291-6370, src/mask.cpp:43-96)
 *   5. masked transform, power spectrum of the full-size transform and the power beyond current_size
 *      (highres_Xi2), window, CenterFFTbySign                                    (:6379-6420) -> fimg
 *   6. CTF image at current_size, frequencies jp / (ori_size * pixel_size)      (:6483-6531, src/ctf.cpp:314-453)
 *
 * Outputs use the layout rln_estep_run consumes: interleaved complex float [n][s][s/2+1] (row i holds
 * frequency y = i for i <= s/2, i - s otherwise; column j holds x = j), CTF float [n][s][s/2+1].
 *
 * Background noise (zero_mask == 0): the reference draws it from its own random sequence seeded with
 * random_seed + part_id; this library uses a counter-based generator with the same seed, the same per-component
 * variance sigma2_fudge * sigma2_noise[ires] and Hermitian-consistent x = 0 and x = N/2 columns, so the noise
 * has the reference's distribution but not its instance. */
#ifndef RLN_IMGPREP_H
#define RLN_IMGPREP_H

#include "rln_estep.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rln_imgprep rln_imgprep;

typedef struct {
    int ori_size;                 /* image_full_size: real-space box, even */
    int current_size;             /* image_current_size: Fourier window, even, <= ori_size */
    double pixel_size;            /* Angstrom per pixel of the optics group */
    double particle_diameter;     /* Angstrom; mask radius diameter / (2 pixel_size); < 0: ori_size / 2 */
    double width_mask_edge;       /* raised-cosine edge width in pixels (reference default 5) */
    int zero_mask;                /* 1: background set to its average value; 0: Gaussian noise background */
    double sigma2_fudge;          /* noise variance multiplier of the background noise */
    unsigned long long random_seed;
    double avg_norm_correction;   /* mymodel.avg_norm_correction; <= 0 disables norm correction */
    int do_ctf_correction;        /* 0: Fctf = 1 */
    int ctf_phase_flipped;        /* |CTF| */
    int only_flip_phases;         /* sign(CTF) */
    int intact_ctf_first_peak;    /* CTF = 1 where |gamma| < pi/2 */
    int ctf_premultiplied;        /* Fctf = CTF^2 */
    double voltage;               /* kV */
    double cs;                    /* spherical aberration, mm */
    double q0;                    /* amplitude contrast */
} rln_imgprep_config;

typedef struct {
    double norm_correction;       /* METADATA_NORM */
    double old_offset_x;          /* METADATA_XOFF, pixels (rounded before use) */
    double old_offset_y;          /* METADATA_YOFF */
    double defocus_u;             /* Angstrom */
    double defocus_v;             /* Angstrom */
    double defocus_angle;         /* degrees */
    double bfactor;               /* Angstrom^2 */
    double kfactor;               /* CTF scale */
    double phase_shift;           /* degrees */
    long long part_id;            /* background noise seed offset */
} rln_image_meta;

/* Engine for one device. Replaces the per-thread FourierTransformer, CTF and mask objects of
 * getFourierTransformsAndCtfs (src/ml_optimiser.cpp:6255) and the acc path's hipFFT bundles
 * (src/acc/hip/hip_fft.h). */
int rln_imgprep_create(int device, const rln_imgprep_config *cfg, rln_imgprep **out);
void rln_imgprep_destroy(rln_imgprep *p);

/* Model noise spectrum for the noise background: sigma2[0 .. n-1] per resolution shell of the original box
 * (mymodel.sigma2_noise[optics_group], src/ml_optimiser.cpp:6309-6327 with equal image and model boxes). */
int rln_imgprep_set_sigma2_noise(rln_imgprep *p, const double *sigma2, int n);

/* Prepare n particles. images: real-space float [n][ori][ori], row-major, y rows (host, or device when
 * in_on_device). fimg, fimg_nomask: interleaved complex float [n][s][s/2+1]; fctf: float [n][s][s/2+1]; all three
 * on the host, or device memory when out_on_device. Host outputs (optional, may be NULL): highres_Xi2 [n]
 * (exp_highres_Xi2_img; 0 when current_size == ori_size), power [n][ori/2+1] (exp_power_img, filled also when
 * current_size == ori_size), old_offset [n][2] (the rounded offsets applied, exp_old_offset).
 * Errors: RLN_ERR_ARG for a norm correction <= 0 while norm correction is on (the reference checks the range
 * [0.001, 1000] only in debug builds, :6228-6235, and divides by it otherwise) and for n < 0. */
int rln_imgprep_run(rln_imgprep *p, int n, const float *images, int in_on_device, const rln_image_meta *meta,
                    float *fimg, float *fimg_nomask, float *fctf, int out_on_device, double *highres_Xi2,
                    double *power, double *old_offset);

/* Milliseconds of the last run per stage: [0] real-space kernel, [1] forward FFTs, [2] Fourier-space kernel,
 * [3] background noise (generation + inverse FFT), [4] whole run including copies. */
int rln_imgprep_timings(rln_imgprep *p, double *ms5);

#ifdef __cplusplus
}
#endif

#endif
