/* C-ABI of the coarse-grid neural operator that predicts the initial
 * scattered field e0 of a forward solve from (sigma, omega).
 *
 * No counterpart in the reference: its forward solve starts from a zero
 * scattered field (src/forward/forward_solver_fv.cpp:353), which stays the
 * fallback whenever the predicted e0 does not reduce the initial residual.
 *
 * Network: four 3x3x3 convolutions on a 32^3 grid (dilations d0..d3),
 * channels 2 -> 32 -> 32 -> 32 -> 12, ReLU between layers, BF16 storage and
 * FP32 accumulation on the matrix cores, Adam on FP32 master weights.
 * Input per voxel: (log10 sigma, log10 f) as FP32 [batch][32^3][2] in z, y, x
 * order; output per voxel: E_sec of both polarisations (x, y, z components,
 * re / im) as FP32 [batch][32^3][12].  Data pointers are device memory; calls
 * are ordered on the context stream and return after it drains.
 */
#ifndef M3D_NET_H
#define M3D_NET_H

#include "m3d_device.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct m3d_net m3d_net;

const char* m3d_net_last_error(void);
int m3d_net_create(m3d_context* ctx, int max_batch, const int* dilations, unsigned seed, m3d_net** out);
int m3d_net_destroy(m3d_net* net);
/* Weights 27*32*(8 + 3*32) floats, each layer [tap][co][ci]; biases 4*32. */
int m3d_net_num_params(long long* n_weights, long long* n_biases);
int m3d_net_get_params(m3d_net* net, float* w, float* b, float* grad_w, float* grad_b);
int m3d_net_set_params(m3d_net* net, const float* w, const float* b);
int m3d_net_forward(m3d_net* net, int batch, const float* d_in, float* d_out);
/* Mean-squared-error loss and parameter gradients, without an update. */
int m3d_net_gradients(m3d_net* net, int batch, const float* d_in, const float* d_target, float* loss);
/* Global gradient-norm clipping for m3d_net_train_step: gradients are scaled
 * by min(1, max_norm / ||g||) before the Adam update (0 disables; default). */
int m3d_net_set_clip(m3d_net* net, float max_norm);
/* One Adam step (beta 0.9 / 0.999) at learning rate lr; returns the loss. */
int m3d_net_train_step(m3d_net* net, int batch, const float* d_in, const float* d_target, float lr,
                       float* loss);

/* e0 on edges from the network output: trilinear interpolation of the voxel
 * field (voxel centres of box = x0,x1,y0,y1,z0,z1) at edge midpoints
 * (d_edge_xyz [n][3] FP32, d_edge_dir [n] 0/1/2).  d_x0 receives complex128
 * [2*batch][n]: system 2b is polarisation 1 of sample b, 2b+1 polarisation 2. */
int m3d_net_prolong(m3d_net* net, int batch, const float* d_field, int n_edges,
                    const float* d_edge_xyz, const unsigned char* d_edge_dir, const double* box,
                    double* d_x0);

#ifdef __cplusplus
}
#endif

#endif /* M3D_NET_H */
