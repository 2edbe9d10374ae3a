/*
 * rln_proposal.h -- C ABI of the learned pose proposal (part of librln_estep.so).
 *
 * The proposal has no counterpart in the reference: the reference's coarse pass always enumerates the whole
 * HEALPix x psi grid (global search) or the window of the local priors (AccProjectorPlan::setup,
 * src/acc/acc_projector_plan_impl.h:99-410; HealpixSampling::selectOrientationsWithNonZeroPriorProbability,
 * src/healpix_sampling.cpp:695-1076). It is an addition that restricts that enumeration: a small convolutional
 * network reads a W x W image of the particle's low frequencies and ranks direction x psi bins; the top bins
 * become the particle's explicit coarse orientation list (rln_estep_set_orientation_sets), and particles whose
 * top-1 probability is low keep the exhaustive search.
 *
 * Network: input W x W (the particle's Fourier window |kx|, |ky| <= W/2, inverse transformed and normalised to
 * zero mean and unit variance); six 3 x 3 convolutions with padding 1, strides 1, 2, 1, 2, 1, 2, C channels
 * each and ReLU; one fully connected layer to n_dir_bins * n_psi_bins logits (bin = idir * n_psi_bins + ipsi,
 * idir a HEALPix NEST pixel at healpix_order, ipsi = round(psi / (360 / n_psi_bins)) mod n_psi_bins).
 * Products run in BF16 with FP32 accumulation on MFMA; master weights are FP32 and trained with Adam.
 *
 * Parameter layout (FP32, rln_proposal_get_params): for each convolution l = 0..5, W_l [C][Kp_l] with
 * k = (3*ky + kx) * Cin_l + c (Kp_0 = 16, zero padded past 9; Kp_l = 9C otherwise) then b_l [C]; then the
 * fully connected W [NB][F] with F = C * (W/8)^2 in (y, x, c) order, then b [NB].
 *
 * Every call returns RLN_OK (0) or a negative status of rln_estep.h; rln_last_error() describes failures.
 */
#ifndef RLN_PROPOSAL_H
#define RLN_PROPOSAL_H

#include "rln_estep.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rln_proposal rln_proposal;

typedef struct rln_proposal_config {
    int window;              /* W: even, 16..64, divisible by 8 */
    int channels;            /* C: multiple of 8 */
    int healpix_order;       /* order of the direction bins (2: 192 directions) */
    int n_psi_bins;          /* psi bins over 360 degrees (24) */
    unsigned long long seed; /* weight initialisation and sample shuffling */
} rln_proposal_config;

/* Training projections: the reference projected at uniform random poses (unit quaternions), shifted
 * uniformly within +-offset_max pixels, multiplied by a CTF with defocus uniform in [defocus_min,
 * defocus_max] and astigmatism up to max_astigmatism, plus Gaussian noise of per-component variance
 * sigma2_noise[ires] (the model's noise spectrum, ires = round(|k|) in the original box). */
typedef struct rln_proposal_gen {
    int ori_size;
    double pixel_size;
    double voltage, cs, q0;
    double defocus_min, defocus_max, max_astigmatism; /* Angstrom */
    double offset_max;                                 /* pixels */
    const double *sigma2_noise;                        /* [ori_size/2 + 1] */
} rln_proposal_gen;

int rln_proposal_create(int device, const rln_proposal_config *cfg, rln_proposal **out);
int rln_proposal_destroy(rln_proposal *p);
int rln_proposal_num_params(const rln_proposal *p, long long *n);
int rln_proposal_num_bins(const rln_proposal *p, int *n);
int rln_proposal_get_params(rln_proposal *p, float *out);
int rln_proposal_set_params(rln_proposal *p, const float *in);

/* Training set: generated projections (poses, if not null, receive rot, tilt, psi per sample) and particle
 * images labelled with poses (e.g. the previous iteration's max-likelihood assignments). fimg: n FFTW
 * half-complex images of current_size (host, or device if on_device); fctf: their CTFs in the same layout, or
 * null. The network sees phase-flipped windows: particle windows are multiplied by the sign of fctf (when
 * given), generated windows by the sign of their own CTF. */
int rln_proposal_clear_training(rln_proposal *p);
int rln_proposal_generate(rln_proposal *p, const rln_projector *reference, const rln_proposal_gen *gen, int n,
                          long long first, double *poses);
int rln_proposal_add_images(rln_proposal *p, int n, const float *fimg, const float *fctf, int current_size,
                            int on_device, const double *poses);
int rln_proposal_training_size(const rln_proposal *p, long long *n);

/* Adam over the training set in shuffled minibatches; mean_loss receives the mean cross entropy of the last
 * epoch (may be null). */
int rln_proposal_train(rln_proposal *p, int epochs, int batch, double lr, double *mean_loss);

/* Top-k bins and their softmax probabilities per particle (bins[n][topk], probs[n][topk], descending). */
int rln_proposal_predict(rln_proposal *p, int n, const float *fimg, const float *fctf, int current_size,
                         int on_device, int topk, int *bins, float *probs);
/* The same in two halves, so that a caller can overlap the network with other GPU or host work (the next batch's
 * inference with the current batch's E-step): launch enqueues the inference and the copy of the top-k into pinned
 * host memory on the proposal's stream and returns; finish waits for them and fills bins and probs. One prediction
 * may be in flight per proposal (a second launch before the finish is RLN_ERR_STATE); host inputs (on_device = 0)
 * must stay valid until the finish. Other calls on the proposal stay ordered after the prediction on its stream. */
int rln_proposal_predict_launch(rln_proposal *p, int n, const float *fimg, const float *fctf, int current_size,
                                int on_device, int topk);
int rln_proposal_predict_finish(rln_proposal *p, int *bins, float *probs);

/* The bin of a pose, and the coarse grid orientations (idir * n_psi_grid + ipsi at grid_order >=
 * healpix_order, grid psi step 360 / n_psi_grid) that a bin covers: the NEST children of its direction times
 * the grid psi values within half a bin of its centre, boundaries included. */
int rln_proposal_pose_bin(const rln_proposal *p, double rot, double tilt, double psi, int *bin);
int rln_proposal_bin_orientations(const rln_proposal *p, int grid_order, int n_psi_grid, int bin, int *out,
                                  int capacity, int *count);

/* Coarse orientation lists of n particles from their top-k bins (rln_proposal_predict output), in the format of
 * rln_estep_set_orientation_sets: the sorted union of the bins' grid orientations, or an empty list (exhaustive
 * search) when the particle's top-1 probability is below min_top1. set_offsets[n+1]; set_orients holds at most
 * capacity entries (n * topk * the orientations of one bin always suffices). */
int rln_proposal_orientation_sets(const rln_proposal *p, int n, const int *bins, const float *probs, int topk,
                                  int grid_order, int n_psi_grid, double min_top1, int *set_offsets,
                                  int *set_orients, long long capacity);

/* Check hooks: the network input images (n x W x W FP32) of particle images, and one forward + backward
 * pass over given input images (FP32, rounded to BF16) with labels: loss = mean cross entropy, grads in the
 * parameter layout (FP32), logits [n][NB]. The weights are not updated. */
int rln_proposal_images(rln_proposal *p, int n, const float *fimg, const float *fctf, int current_size, int on_device,
                        float *out);
int rln_proposal_forward_backward(rln_proposal *p, int n, const float *images, const int *labels, double *loss,
                                  float *grads, float *logits);

#ifdef __cplusplus
}
#endif

#endif /* RLN_PROPOSAL_H */
