/*
 * hst_closure.h -- C ABI of the learned subgrid closure of libhstgpu.so.
 *
 * The closure has no counterpart in the reference (hst is a DNS code): a
 * pointwise network maps the 9 filtered velocity gradients du_i/dx_j to the 6
 * subgrid stresses tau_ij = filt(u_i u_j) - filt(u_i) filt(u_j) (order uu vv
 * ww uv vw uw), and a coarse run adds tau_ij to its six physical products
 * before the forward transform, the place where the reference forms the
 * products (build_products, src/hst_transforms.f90:97-121): the right-hand
 * sides then take -div(u_i u_j + tau_ij) (src/hst_equations.f90:143-150).
 *
 * Pipeline: a DNS (the live hst_gpu state) is filtered onto a coarse grid
 * (hst_closure_filter), samples of (gradients, stresses) at the coarse grid
 * points are appended to a device-resident data set (hst_closure_sample),
 * the network is trained by the library's own forward/backward kernels on
 * the matrix cores (bfloat16 operands, FP32 accumulation) with Adam
 * (hst_closure_train), and a coarse run attaches it (hst_closure_attach).
 *
 * Filter: the Fourier columns |ix| <= nx_c, |iz| <= nz_c are kept (a sharp
 * spectral box in x and z; alfa0, beta0 and ly are shared), and in y a
 * top-hat of width ly/ny_c centred on each coarse row is applied to the
 * piecewise-linear interpolant of the fine rows.
 *
 * Network: 9 -> 64 -> 64 -> 6, tanh hidden units, inputs and targets
 * standardised by norm = [mean_x(9) std_x(9) mean_y(6) std_y(6)].  The
 * parameters are one FP32 array of HST_MLP_NPARAM entries:
 * W1[64][9] b1[64] W2[64][64] b2[64] W3[6][64] b3[6] (row = output unit).
 * Loss: the mean over samples and outputs of the squared standardised error.
 *
 * Every function returns 0 or a status (text from hst_gpu_last_error).  One
 * context at a time may be attached; calls are not re-entrant.
 */
#ifndef HST_CLOSURE_H
#define HST_CLOSURE_H

#include <stdint.h>

#include "hst_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HST_MLP_NPARAM 5190
#define HST_CLOSURE_NORM 30

/* Fields of a context, reference layout (like hst_gpu_upload_field):
 * 0..2 the filtered velocity, 3..8 the filtered products (uu vv ww uv vw uw),
 * 9..14 the modelled stress spectra of the last hst_closure_stress. */
#define HST_CLOSURE_U 0
#define HST_CLOSURE_PBAR 3
#define HST_CLOSURE_TAU 9

typedef struct hst_closure hst_closure;

/* Host only: the y filter from ny_fine to ny_coarse rows.  For coarse row k,
 * first_row[k] is the first fine row (may be -2 or -1: ghost rows) and
 * weights[8k .. 8k+7] the weights of rows first_row[k] .. first_row[k]+7. */
int hst_closure_filter_weights(int ny_fine, int ny_coarse, int32_t *first_row, double *weights);

/* A context on the coarse grid of `coarse` (nx, ny, nz, alfa0, beta0, ly,
 * S and the S2 clock are used) holding up to `capacity` samples. */
int hst_closure_create(const hst_deck *coarse, int64_t capacity, hst_closure **out);
int hst_closure_free(hst_closure *c);

/* The live state's velocity and, after a product pass (hst_gpu_nonlinear or
 * a time step), its six products, filtered onto the context's grid; ghost
 * rows at the shifts of `time`.  The state's grid must be finer in every
 * direction, with the same alfa0, beta0 and ly. */
int hst_closure_filter(hst_closure *c, double time);
int hst_closure_get_field(hst_closure *c, int which, double *host);
int hst_closure_set_field(hst_closure *c, int which, const double *host);

/* Appends one sample per coarse grid point of the interior rows
 * (ny * nzd * 2 nxd of them, index ((iy nzd + zp) 2 nxd + x)): inputs the 9
 * gradients of fields 0..2, targets the 6 stresses from fields 0..8.  The
 * velocity's ghost rows must hold the shifts of `time`.  *count: samples held. */
int hst_closure_sample(hst_closure *c, double time, int64_t *count);
int hst_closure_set_dataset(hst_closure *c, int64_t n, const float *x, const float *y);
int hst_closure_get_dataset(hst_closure *c, int64_t first, int64_t n, float *x, float *y);
int hst_closure_count(hst_closure *c, int64_t *count);

/* norm[30] from the data set (device reduction), stored and returned. */
int hst_closure_normalise(hst_closure *c, double *norm);
int hst_closure_set_norm(hst_closure *c, const double *norm);
int hst_closure_get_norm(hst_closure *c, double *norm);

/* Glorot-uniform weights, zero biases, from a seeded generator; Adam state reset. */
int hst_closure_init_params(hst_closure *c, uint64_t seed);
int hst_closure_set_params(hst_closure *c, const float *params);
int hst_closure_get_params(hst_closure *c, float *params);

/* Gradient of the loss over samples [first, first + n) of the data set and the loss. */
int hst_closure_grad(hst_closure *c, int64_t first, int64_t n, float *grad, double *loss);

/* `epochs` passes of Adam (beta 0.9, 0.999, eps 1e-8, rate lr) over the data
 * set in batches of `batch` samples (a multiple of 128, at most the data set),
 * reshuffled every epoch from `seed`; loss[e]: the mean batch loss of epoch e.
 * Samples [n_train, count) are held out: val_loss[e] (may be NULL) is their
 * loss after epoch e.  n_train <= 0: all samples train. */
int hst_closure_train(hst_closure *c, int epochs, int batch, double lr, uint64_t seed, int64_t n_train,
                      double *loss, double *val_loss);

/* The network on host inputs: y[n][6] stresses (de-standardised) of x[n][9]. */
int hst_closure_predict(hst_closure *c, int64_t n, const float *x, float *y);

/* Modelled stress spectra (fields 9..14, ghosts at the shifts of `time`) of
 * the context's velocity (fields 0..2, ghosts valid). */
int hst_closure_stress(hst_closure *c, double time);

/* on = 1: every substep of hst_gpu_timestep adds the modelled stresses of the
 * live state's velocity to its products (the state's grid must be the
 * context's); on = 0 detaches, and the run is the plain method again. */
int hst_closure_attach(hst_closure *c, int on);

#ifdef __cplusplus
}
#endif

#endif
