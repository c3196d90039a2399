/* mdft_wsnet.h -- C-ABI of the learned warm start: a periodic 3D
 * convolutional network with hand-written HIP forward, backward and Adam
 * kernels (FP32), and the entry points that couple it to an MDFT context.
 *
 * None of these has a counterpart in the reference (src/ has no learned
 * component).  They serve callers that want a better starting density for
 * the minimisation driven by src/energy_minimization.f90:282; the
 * minimisation itself and its answer are unchanged.
 *
 * Network: a_0 = input [C_0][nz][ny][nx] (x fastest, FP32),
 *   z_l = conv3d(a_{l-1}; W_l, kernel k_l in {1, 3}, dilation d_l) + b_l
 *         (periodic box; a 3x3x3 layer has <= 32 channels in and out, a
 *         pointwise layer <= 256),
 *   a_l = leaky_relu(z_l) = (z_l > 0 ? z_l : 0.1 z_l) for l < L, output a_L = z_L.
 * Parameters: for each layer W_l[co][ci][k][k][k] then b_l[co], one FP32 vector.
 * Loss: sum(mask (a_L - target)^2) / (sum(mask) C_L), mask per grid point.
 * All field pointers are device pointers; the caller makes them ready before
 * the call (the network runs on its own stream and returns synchronised).
 */
#ifndef MDFT_WSNET_H
#define MDFT_WSNET_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdft_wsnet mdft_wsnet;
struct mdft_ctx;

/* channels[0..nlayers], ksize[0..nlayers-1] (1 or 3), dilations[0..nlayers-1]
 * (>= 1); He-normal weights from `seed`, zero biases. */
int mdft_wsnet_create(int nlayers, const int *channels, const int *ksize, const int *dilations,
                      unsigned long long seed, int device, mdft_wsnet **out);
void mdft_wsnet_destroy(mdft_wsnet *net);
const char *mdft_wsnet_last_error(const mdft_wsnet *net); /* NULL: last create error */
long long mdft_wsnet_num_params(const mdft_wsnet *net);
int mdft_wsnet_get_params(mdft_wsnet *net, float *host);
int mdft_wsnet_set_params(mdft_wsnet *net, const float *host); /* also resets Adam's moments */
int mdft_wsnet_get_grads(mdft_wsnet *net, float *host);

/* d_out [C_L][N] = network(d_in [C_0][N]) */
int mdft_wsnet_forward(mdft_wsnet *net, int nx, int ny, int nz, const float *d_in, float *d_out);

/* Forward, loss and backward: stores dLoss/dparams (read with
 * mdft_wsnet_get_grads, applied by mdft_wsnet_adam); d_mask [N] and d_gin
 * ([C_0][N], dLoss/dinput) may be NULL. */
int mdft_wsnet_loss_grad(mdft_wsnet *net, int nx, int ny, int nz, const float *d_in, const float *d_target,
                         const float *d_mask, float *d_gin, double *loss);

/* Backward pass from a given gradient of the output, d_gout [C_L][N] (device):
 * accumulates nothing -- the stored parameter gradients are overwritten, as by
 * mdft_wsnet_loss_grad -- and writes dL/d(input) to d_gin when not NULL.  Must
 * follow mdft_wsnet_forward on the same d_in and grid (it uses the activations
 * that call kept). */
int mdft_wsnet_backward(mdft_wsnet *net, int nx, int ny, int nz, const float *d_in, const float *d_gout,
                        float *d_gin);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) with the stored gradients. */
int mdft_wsnet_adam(mdft_wsnet *net, float lr);

/* ---- coupling to an MDFT context (its grid, orientations, vext and c(k)) ----
 * Network-side fields are the context's angular projections as 2 np real
 * channels [2 np][N] (FP32): channel 2 ip is the real part of projection ip
 * (the reference's projection order), 2 ip + 1 the imaginary part.
 * Molecular symmetry order 2 only. */

/* d_feat [2 np][N]: projections of clip(-beta vext / 2, -4, 4) / 4 (the bounded log of the
 * square-root Boltzmann factor), channel k multiplied by h_scale[k] (host array of 2 np
 * floats; NULL: no scaling) -- the network's input normalisation. */
int mdft_ws_features(struct mdft_ctx *ctx, float *d_feat, const float *h_scale);

/* For the density xi (device): d_gamma [N][no] (FP64, may be NULL) receives
 * gamma(r, Omega) = c (*) drho in units of ln(rho/rho0), i.e.
 * ln(rho/rho0) = -beta vext + gamma at the minimum; d_target [2 np][N] (FP32,
 * may be NULL) receives the projections of gamma. */
int mdft_ws_gamma(struct mdft_ctx *ctx, const double *d_xi, double *d_gamma, float *d_target);

/* d_target [2 np][N] (FP32): projections of rho/rho0 - 1 for the density xi
 * (device): 0 in the bulk, -1 where the solvent is excluded. */
int mdft_ws_density(struct mdft_ctx *ctx, const double *d_xi, float *d_target);

enum { MDFT_WS_GAMMA = 0, MDFT_WS_DENSITY = 1, MDFT_WS_BOLTZMANN = 2 };

/* Starting density from predicted projections d_pred [2 np][N]:
 * kind MDFT_WS_DENSITY: d_pred are projections of rho/rho0 - 1 (see mdft_ws_density),
 *   d_xi = 0 where the cold start (below) is 0, else sqrt(max(1 + pred, 0.01)) -- the
 *   floor keeps every point the minimiser may fill away from xi = 0, where the
 *   gradient vanishes;
 * kind MDFT_WS_GAMMA: d_pred are projections of gamma, d_xi = exp((-beta vext + gamma) / 2),
 * exponent clipped to [-200, 7].
 * With d_pred NULL: the reference's cold start (src/module_density.f90:267-284),
 * xi = 0 where vext >= (-ln eps - 0.05) kT, else 1; or, kind MDFT_WS_BOLTZMANN, the
 * Boltzmann factor capped at rho/rho0 = e^3, xi = exp(-max(beta vext, -3) / 2) (kind
 * MDFT_WS_BOLTZMANN with a prediction is MDFT_EINVAL).  All on the device. */
int mdft_ws_guess(struct mdft_ctx *ctx, const float *d_pred, int kind, double *d_xi);

/* Gradient of F at a density guess with respect to the prediction it came from:
 * d_xi is mdft_ws_guess(d_pred, MDFT_WS_DENSITY) and d_df the gradient of
 * mdft_energy_and_gradient at d_xi (device, [N][no], dF = (1/dv) dF/dxi);
 * d_gpred [2 np][N] (FP32) receives dF/d(pred), zero through the floor and the
 * excluded points.  With mdft_wsnet_backward this trains the network on F
 * itself. */
int mdft_ws_start_gradient(struct mdft_ctx *ctx, const double *d_xi, const double *d_df, float *d_gpred);

/* d_out = sqrt((1 - lam) xa^2 + lam xb^2): xi of the density mixture (1 - lam) rho_a + lam rho_b. */
int mdft_ws_mix(struct mdft_ctx *ctx, double lam, const double *d_xa, const double *d_xb, double *d_out);

#ifdef __cplusplus
}
#endif

#endif
