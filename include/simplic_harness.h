/*
 * Harness entry points of libsimplic_amd.so: the deformation test driver of
 * the reference (applications/test/plicVofAdvectionFoam) on the GPU.  Not part
 * of the solveVofEqu drop-in; they set the flow that a solver would provide.
 */
#ifndef SIMPLIC_HARNESS_H
#define SIMPLIC_HARNESS_H

#include "simplic_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Unscaled deformation velocity at cell centres (3*n_cells) and face flux
 * U(face centre) . Sf (n_faces) on the handle's geometry, evaluated on the
 * host with the C library sin (updateU.H:66-112).  Host output arrays. */
int svp_harness_deformation(svp_handle* h, double* U0, double* phi0);

/* Solid-body rotation U = omega x (x - origin) at cell centres and faces
 * (linear, so face-centre fluxes are exactly divergence free). */
int svp_harness_rotation(svp_handle* h, const double* omega, const double* origin, double* U0,
                         double* phi0);

/* Upload unscaled fields once (host arrays); the handle keeps them on the GPU. */
int svp_harness_upload_base_flow(svp_handle* h, const double* U0, const double* phi0,
                                 const double* U0_boundary);

/* phi = phi0*factor, U = U0*factor, U_boundary = U0_boundary*factor on the
 * GPU (updateU.H:117-127). */
int svp_harness_scale_flow(svp_handle* h, double factor);

/* Initial alpha of a union of n_shapes ellipsoids (spheres: equal semi-axes,
 * identity rotation), each given by 15 doubles: centre (3), semi-axes (3),
 * rotation rows (9).  Cells are split into tets that are refined 1:8 near the
 * surface down to max_depth levels, where the linear interpolant of the shape
 * function is integrated exactly.  Writes the handle's alpha. */
int svp_harness_init_alpha(svp_handle* h, const double* shapes, int n_shapes, int max_depth);

/* Initial alpha from a closed, outward-oriented triangulated surface (the
 * setVofField utility's exact cut, setAlphaField.H:41-62, without CGAL):
 * tris holds n_tris triangles of 3 points of 3 coordinates.  The inside
 * volume of each cell is exact: the signed sum over the triangles of the
 * cone from the cell centre clipped to the cell's (centre, face centre,
 * edge) tets.  mode 0 sets alpha to it, 1 adds it, 2 subtracts it (the
 * utility's ADD / SUBTRACT modes); the result is clipped to [0, 1]. */
int svp_harness_init_alpha_surface(svp_handle* h, const double* tris, int32_t n_tris, int32_t mode);

/* max over cells of sum|phi|/V, and of the same restricted to
 * 0.01 <= alpha <= 0.99 (CourantNo.H, alphaCourantNo.H:95-103), from the
 * current phi and alpha; multiply by 0.5*dt for the Courant numbers.
 * Synchronises with the handle's stream. */
int svp_harness_courant(svp_handle* h, double* max_sum_phi_by_v, double* max_alpha_sum_phi_by_v);
/* The same two maxima for every case of a batch (n_cases each), one launch. */
int svp_batch_courant(svp_batch* b, double* max_sum_phi_by_v, double* max_alpha_sum_phi_by_v);
/* Every case's uploaded base flow times its own factor (n_cases), one launch. */
int svp_batch_scale_flow(svp_batch* b, const double* factors);


#ifdef __cplusplus
}
#endif
#endif
