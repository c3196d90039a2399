/* mdft_wsnet.h -- C-ABI of the learned warm start: a periodic dilated 3D
 * convolutional network with hand-written HIP forward, backward and Adam
 * kernels (FP32), and the entry points that couple it to an MDFT context.
 *
 * None of these has a counterpart in the reference (src/ has no learned
 * component).  They serve callers that want a better starting density for
 * the minimisation driven by src/energy_minimization.f90:282; the
 * minimisation itself and its answer are unchanged.
 *
 * Network: a_0 = input [C_0][nz][ny][nx] (x fastest, FP32),
 *   z_l = conv3d(a_{l-1}; W_l, dilation d_l) + b_l (3x3x3 taps, periodic box),
 *   a_l = relu(z_l) for l < L, output a_L = z_L.
 * Parameters: for each layer W_l[co][ci][3][3][3] then b_l[co], one FP32 vector.
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

/* channels[0..nlayers] (1..32 each), dilations[0..nlayers-1] (>= 1); He-normal
 * weights from `seed`, zero biases. */
int mdft_wsnet_create(int nlayers, const int *channels, const int *dilations, unsigned long long seed, int device,
                      mdft_wsnet **out);
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

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) with the stored gradients. */
int mdft_wsnet_adam(mdft_wsnet *net, float lr);

/* ---- coupling to an MDFT context (its grid, orientations and vext) ----
 * Low angular basis B_k(Omega), k = 0..3: 1 and the lab-frame components
 * (x, y, z) of the molecule's z axis, i.e. the m <= 1 projections for
 * molecular symmetry order 2; projections c_k(r) = sum_io w_io f(r,io) B_k(io)
 * / sum_io w_io B_k(io)^2. */

/* d_feat [8][N]: projections of exp(-max(beta vext, -8) / 2) - 1 (channels
 * 0..3) and of tanh(-beta vext / 8) (channels 4..7), from the context's vext. */
int mdft_ws_features(struct mdft_ctx *ctx, float *d_feat);

/* d_gamma [N][no] (FP64): gamma(r, Omega) = c (*) drho in units of ln(rho/rho0)
 * for the density xi (device), i.e. ln(rho/rho0) = -beta vext + gamma at the
 * minimum.  d_proj4 [4][N] (FP32, may be NULL): its projections c_k(r). */
int mdft_ws_gamma(struct mdft_ctx *ctx, const double *d_xi, double *d_gamma, float *d_proj4);

/* d_xi [N][no] = exp((-beta vext + sum_k pred_k(r) B_k(Omega)) / 2), exponent
 * clipped to [-200, 7].  With d_pred NULL: the reference's cold start
 * (src/module_density.f90:267-284), xi = 0 where vext >= (-ln eps - 0.05) kT,
 * else 1, computed on the device. */
int mdft_ws_guess(struct mdft_ctx *ctx, const float *d_pred, double *d_xi);

#ifdef __cplusplus
}
#endif

#endif
