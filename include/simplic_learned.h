/*
 * Learned interface normal for the MI355X SimPLIC library (libsimplic_amd.so).
 *
 * No counterpart in the reference.  It adds the orientation method
 * "learned" next to the selector map of reconstruction.C:60-68: a small
 * network (27 -> 128 -> 128 -> 3, GELU; BF16 weights and activations, FP32
 * accumulation on MFMA) maps the 3x3x3 volume-fraction stencil of an
 * interface cell of a uniform hex mesh to its normal.  Where the volume
 * residual of that normal exceeds learnedResidualTol, and on cells without a
 * full uniform stencil, the least-squares normal (reconstruction.C:85-141) is
 * used instead.  The residual is the mean |fraction - alpha| over the 26
 * neighbour cubes of the plane with the network normal that holds the cell's
 * own alpha in its cube.  The unchanged FP64 findSignedDistance
 * (cutCell.C:611-799) then places every plane.
 *
 * The network is trained on the GPU by this library: training pairs are cut
 * from random planes, spheres and ellipsoids (curvature radius 1.5-50 cells)
 * by exact plane volumes and adaptive subdivision, and forward, backward and
 * Adam run as HIP kernels.
 *
 * Parameter order of the packed FP32 vectors (SVP_NM_N_PARAMS values):
 * W1[128][27], b1[128], W2[128][128], b2[128], W3[3][128], b3[3]; the input
 * is 2*alpha - 1 of the stencil in the canonical frame (axes reflected so the
 * 1-2-1 central-difference normal estimate is non-negative and ordered by
 * decreasing component), lattice index 9(dx+1) + 3(dy+1) + (dz+1).
 */
#ifndef SIMPLIC_LEARNED_H
#define SIMPLIC_LEARNED_H

#include <stdint.h>

#include "simplic_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SVP_NM_N_PARAMS 20483
#define SVP_NM_SHAPE_DOUBLES 19 /* kind, c[3], a[3], Q[9], normal[3] */

typedef struct svp_normal_model svp_normal_model;

/* New model on `device` with seeded uniform(-1/sqrt(fan_in), 1/sqrt(fan_in))
 * weights and biases. */
int svp_nm_create(int device, uint64_t seed, svp_normal_model** out);
int svp_nm_destroy(svp_normal_model* m);
int svp_nm_set_params(svp_normal_model* m, const float* params, int32_t n);
int svp_nm_get_params(svp_normal_model* m, float* params, int32_t n);

/* Training set of n samples on the device (replaces the previous one).
 * Sample i uses shape kind i % 3 (plane, sphere, ellipsoid); curved shapes
 * are integrated with `depth` (1..7) octree levels per cell. */
int svp_nm_generate(svp_normal_model* m, int64_t n, uint64_t seed, int32_t depth);

/* Copies of samples [first, first+count) of the training set; any output
 * may be NULL.  alpha27: mesh-frame stencil (lattice order); shape: 19
 * doubles per sample; x: 32 BF16 network inputs; y: 4 floats (canonical
 * target normal, 0). */
int svp_nm_dataset_get(svp_normal_model* m, int64_t first, int64_t count, double* alpha27, double* shape,
                       uint16_t* x, float* y);

/* Adam training over the set: `epochs` passes, `batch` samples per step (a
 * multiple of 64), learning rate cosine from lr0 to lr1, epoch order a
 * seeded affine permutation.  loss_per_epoch (may be NULL): mean squared
 * error per sample of each epoch.  seconds (may be NULL): device time. */
int svp_nm_train(svp_normal_model* m, int32_t epochs, int32_t batch, float lr0, float lr1, uint64_t seed,
                 float* loss_per_epoch, double* seconds);

/* Network outputs (3 floats per sample, canonical frame, not normalised)
 * for n BF16 input rows of 32 values. */
int svp_nm_forward(svp_normal_model* m, const uint16_t* x, int64_t n, float* out, int where);

/* Gradient of the mean squared error over n samples (n a multiple of 64) in
 * packed parameter order, and the loss (sum of squared errors / n). */
int svp_nm_gradient(svp_normal_model* m, const uint16_t* x, const float* y, int64_t n, float* grad, double* loss,
                    int where);

/* Liquid fraction of the cube [lo, lo+h]^3 on the side n.x + d < 0 (the
 * closed form the residual test uses), for k planes on the device: in[k*8] =
 * nx, ny, nz, d, lox, loy, loz, h. */
int svp_nm_cube_fractions(const double* in, int64_t k, double* out, int device);

/* Use the model's current parameters for a handle whose orientationMethod is
 * "learned" (copied: the model may be destroyed afterwards).  Builds the
 * uniform-lattice stencil map of the mesh. */
int svp_set_normal_model(svp_handle* h, const svp_normal_model* m);

/* Counts of the last reconstruct of a "learned" handle: [0] interface cells,
 * [1] with a learned normal kept, [2] fallen back by the residual test,
 * [3] without a full uniform stencil or with a degenerate network output. */
int svp_learned_stats(svp_handle* h, int64_t* out4);

/* Residuals of the last reconstruct, per interface-cell position; -1 where no
 * network normal was tried.  SVP_E_ARG when learnedResidualTol >= 1 (the test
 * is not run). */
int svp_learned_residuals(svp_handle* h, double* out, int where);

#ifdef __cplusplus
}
#endif

#endif
