/*
 * C-ABI of the MI355X SimPLIC alpha-transport library (libsimplic_amd.so).
 *
 * One handle = one alpha field on one mesh = one reference `solveVofEqu`
 * object (src/SimPLIC/solveVofEqu/solveVofEqu.H:67-216): a reconstruction
 * (src/SimPLIC/reconstruction/reconstruction.C) and an advection
 * (src/SimPLIC/advection/advection.C) sharing the fvSolution "alpha.*" dict.
 * Fields live in GPU memory; the caller hands over host or device pointers
 * (`where` = SVP_HOST or SVP_DEVICE).  Every entry point returns 0 on success
 * or a negative SVP_E_* code; svp_last_error() gives the message.  The
 * reference aborts the process (FatalError) where this ABI returns an error.
 *
 * Threading: a handle is not reentrant (the reference objects hold mutable
 * scratch state as well); distinct handles may be used from distinct threads.
 * All work of a handle is ordered on its HIP stream.
 */
#ifndef SIMPLIC_AMD_H
#define SIMPLIC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVP_HOST 0
#define SVP_DEVICE 1

#define SVP_OK 0
#define SVP_E_ARG (-1)          /* invalid argument or size */
#define SVP_E_DICT (-2)         /* invalid dictionary entry (reference: FatalError) */
#define SVP_E_UNSUPPORTED (-3)  /* valid reference option not provided by this library */
#define SVP_E_HIP (-4)          /* HIP runtime failure */
#define SVP_E_MESH (-5)         /* mesh exceeds the device cut limits */
#define SVP_E_DEVICE (-6)       /* a kernel reported an error (see svp_last_error) */
#define SVP_E_NAN (-7)          /* non-finite input */

#define SVP_PATCH_WALL 0
#define SVP_PATCH_EMPTY 1

typedef struct svp_handle svp_handle;

/* Raw OpenFOAM polyMesh: owner-outward face point lists, internal faces in
 * upper-triangular order followed by boundary faces patch by patch. */
typedef struct {
    int32_t n_points, n_faces, n_internal_faces, n_cells, n_patches;
    const double* points;          /* 3*n_points */
    const int32_t* face_offsets;   /* n_faces+1 */
    const int32_t* face_points;    /* face_offsets[n_faces] */
    const int32_t* owner;          /* n_faces */
    const int32_t* neighbour;      /* n_internal_faces */
    const int32_t* patch_start;    /* n_patches */
    const int32_t* patch_size;     /* n_patches */
    const int32_t* patch_type;     /* n_patches: SVP_PATCH_* */
} svp_mesh;

/* The fvSolution solvers{"alpha.*"{...}} keys the reference reads
 * (reconstruction.C:502-516, advection.C:455-457). */
typedef struct {
    char orientation_method[32];   /* "isoAlphaGrad" (default), "LS", "alphaGrad", "isoRDF", "RDF" or "learned" (simplic_learned.h) */
    double mixed_cell_tol;         /* 1e-8 */
    int32_t n_alpha_bounds;        /* 10 */
    double snap_tol;               /* 0 */
    int32_t clip;                  /* 1 */
    int32_t split_warped_face;     /* 0 */
    int32_t write_plic_fields;     /* 0 */
    double learned_residual_tol;   /* learnedResidualTol, "learned" method only: 0.25 (>= 1: no fallback test) */
    double rdf_tol;                /* tol, isoRDF only: 1e-6 (reconstruction.C:513) */
    double rdf_rel_tol;            /* relTol, isoRDF only: 0.1 (reconstruction.C:514) */
    int32_t rdf_iterations;        /* iterations, isoRDF only: 5 (reconstruction.C:515) */
    char grad_scheme[32];          /* alphaGrad only: the fvSchemes gradSchemes entry grad(alpha1) that
                                      fvc::grad(alpha1_, "grad(alpha1)") reads (reconstruction.C:78):
                                      "Gauss linear" (default) or "Gauss pointLinear" */
} svp_dict;

/* Defaults of every key (reconstruction.C:502-516, advection.C:455-457). */
void svp_dict_defaults(svp_dict* d);

/* Parse the body of an fvSolution "alpha.*" dictionary ("key value;" lines,
 * C/C++ comments allowed).  Unknown keys are ignored, as OpenFOAM does.
 * `reconstructionScheme` is accepted as an alias of `orientationMethod`.
 * Returns SVP_E_DICT for an unknown orientation method
 * (reconstruction.C:610-625).  alphaGrad is the Gauss gradient normal
 * (reconstruction.C:74-82) with the scheme of key "grad(alpha1)" (or
 * "gradScheme"): Gauss linear or Gauss pointLinear, others SVP_E_UNSUPPORTED;
 * isoRDF / RDF the reconstructed-distance-function
 * normal (reconstruction.C:194-405) with keys tol, relTol and iterations;
 * contact-angle boundaries (updateContactAngle) are not part of this mesh
 * description, so its boundary correction is not applied.
 * "learned" (alias "MLP") selects the learned normal of simplic_learned.h. */
int svp_dict_parse(const char* text, svp_dict* out);

/* solveVofEqu constructor (solveVofEqu.C:55-91): uploads the mesh, computes
 * its geometry and face flatness on the GPU (updateFaceFlatness,
 * reconstruction.C:408-473).  `device` selects the HIP device. */
int svp_create(const svp_mesh* mesh, const svp_dict* dict, int device, svp_handle** out);
int svp_destroy(svp_handle* h);

/* Fields.  alpha: n_cells.  phi: n_faces (internal then boundary faces).
 * U: 3*n_cells cell values.  U_boundary: 3*(n_faces-n_internal_faces) boundary
 * face values of U (NULL = zero), used by the cell-point interpolation. */
int svp_set_alpha(svp_handle* h, const double* alpha, int where);
int svp_set_flow(svp_handle* h, const double* phi, const double* U, const double* U_boundary,
                 int where);

/* solveVofEqu::reconstruct (solveVofEqu.C:96-99 -> reconstruction.C:680-722). */
int svp_reconstruct(svp_handle* h);

/* solveVofEqu::advect<Sp,Su> (solveVofEquTemplates.C:35-43 ->
 * advectionTemplates.C:352-418) over time step dt.  Sp/Su: n_cells implicit /
 * explicit sources or NULL for zeroField. */
int svp_advect(svp_handle* h, double dt, const double* Sp, const double* Su, int where);

/* Outputs. */
int svp_get_alpha(svp_handle* h, double* alpha, int where);
/* alphaPhi = dVf/dt of the last advect (advectionTemplates.C:417). */
int svp_get_alpha_phi(svp_handle* h, double* alpha_phi, int where);
/* getRhoPhi = (rho1 - rho2) alphaPhi + rho2 phi (advection.H:329-343). */
int svp_get_rho_phi(svp_handle* h, double rho1, double rho2, double* rho_phi, int where);
/* The alpha sub-step of interPlicFoam-class solvers (alphaEqnSubCycle.H:30-60
 * with alphaEqn.H:9-20; nAlphaSubCycles from alphaControls.H:3): n_sub > 1
 * runs n_sub reconstruct + advect steps of dt/n_sub and returns
 * rhoPhi = sum over them of (dt/n_sub)/dt * getRhoPhi(rho1, rho2); n_sub == 1
 * is one reconstruct + advect whose getRhoPhi is returned.  Sp/Su (n_cells
 * or NULL) are held for every sub-step; a caller whose sources change with
 * alpha calls svp_advect per sub-step instead.  rho_phi (n_faces) may be
 * NULL.  The PIMPLE outer-corrector averaging of phi and U
 * (alphaEqnSubCycle.H:1-28, 62-67) stays with the caller (svp_set_flow,
 * svp_get_alpha / svp_set_alpha). */
int svp_advect_sub_cycles(svp_handle* h, double dt, int32_t n_sub, const double* Sp, const double* Su,
                          double rho1, double rho2, double* rho_phi, int where);
/* solveVofEqu::mapAlphaField (solveVofEqu.C:102-105 -> reconstruction.C:725-784)
 * after a mesh refinement: every cell whose alpha lies in
 * [lower_refine_level, upper_refine_level] (dynamicRefineFvMeshCoeffs) takes
 * the volume fraction of the plane n.x + d = 0 in it, cut through its mesh
 * faces.  n (3 per cell) and d (per cell) are interfaceN / interfaceD as the
 * caller's mesh mapper carried them onto this mesh; the reference's own
 * guards (dynamicRefineFvMesh, mesh changing, mapAlphaField true) are the
 * caller's. */
int svp_map_alpha_field(svp_handle* h, const double* n, const double* d, double lower_refine_level,
                        double upper_refine_level, int where);
/* Iterations the last isoRDF / RDF reconstruct ran (the loop of
 * reconstruction.C:236-400); no counterpart in the reference, which does not
 * report it.  0 for the other orientation methods. */
int svp_rdf_iterations(svp_handle* h, int32_t* n);
/* Number of mixed cells of the last reconstruct (reconstruction.C:675-676). */
int svp_n_mixed(svp_handle* h, int32_t* n_mixed);
/* Interface data of the last reconstruct per mixed cell (ascending cells):
 * cell label, status (-1/0/1), normal n (3), distance d, centre C (3), area
 * vector S (3).  Any output pointer may be NULL. */
int svp_get_interface(svp_handle* h, int32_t* cells, int32_t* status, double* n, double* d,
                      double* C, double* S, int where);
/* PLIC surface of the last reconstruct (reconstruction::interface,
 * reconstruction.C:787-835, returned as a plicSurface, plicSurface.H:58-100):
 * every mixed cell is cut again by its plane through its mesh faces, and each
 * cell the plane cuts gives one polygon, cells ascending.  Its points are the
 * plane crossings of the cut faces ordered by angle about the interface
 * centre, near-equal angles (1e-8) merged (cutCell::interfacePoints,
 * cutCell.C:545-608).  n_faces / n_points are always written; face_cells
 * (n_faces), face_offsets (n_faces + 1, polygon f is points
 * [face_offsets[f], face_offsets[f+1])) and points (3 n_points) are filled
 * when all three are non-NULL (all NULL: size query). */
int svp_get_plic_surface(svp_handle* h, int32_t* n_faces, int32_t* n_points, int32_t* face_cells,
                         int32_t* face_offsets, double* points, int where);
/* No counterpart in the reference.  The same export for callers that cannot
 * allocate device memory (code between HIP graph captures, tight serving
 * loops, harnesses that forbid hipMalloc): face_cells, face_offsets and points
 * are device buffers of the caller (sized by a prior size query), and the
 * working space is a static device array of the library, one per GPU, so the
 * call makes no device or pinned allocation.  Results equal
 * svp_get_plic_surface's.  At most SVP_PLIC_NOALLOC_MAX_CELLS mixed cells
 * (SVP_E_ARG beyond).  Calls that overlap on one GPU share that array and
 * race; n_faces / n_points are host pointers. */
#define SVP_PLIC_NOALLOC_MAX_CELLS (1 << 20)
int svp_get_plic_surface_device(svp_handle* h, int32_t* n_faces, int32_t* n_points, int32_t* face_cells,
                                int32_t* face_offsets, double* points);
/* Face-cut classification of the final plane cut per mixed cell: 2 bits per
 * cell face ((status+1) << 2*k, cells() face order) and the number of plane
 * crossings (cut vertices) on its cut faces. */
int svp_get_cut_class(svp_handle* h, uint64_t* face_class, int32_t* n_cut_points, int where);
/* Interface normal velocity Un0 of each mixed cell of the last advect. */
int svp_get_un0(svp_handle* h, double* un0, int where);
/* Mesh geometry computed at create time: face centres/areas (3*n_faces),
 * |Sf|, flatness (n_faces), cell centres (3*n_cells), volumes (n_cells). */
int svp_get_geometry(svp_handle* h, double* Cf, double* Sf, double* magSf, double* flatness,
                     double* C, double* V);
/* Cumulative GPU seconds spent in reconstruct and advect (the reference's
 * reconstructionTime/advectionTime, solveVofEqu.H:195-211) and the running sum
 * of mixed-cell counts over all reconstruct calls. */
int svp_timers(svp_handle* h, double* reconstruction_s, double* advection_s, uint64_t* n_ic_total);
/* solveVofEqu::alphaMappingTime (solveVofEqu.H:208-210 ->
 * reconstruction.H:396-399): cumulative GPU seconds of svp_map_alpha_field. */
int svp_alpha_mapping_time(svp_handle* h, double* seconds);
/* The reference's full-mesh interface fields after the last reconstruct
 * (reconstruction.C:520-570; zeroed every step by initialize,
 * reconstruction.C:638-641, and filled in the mixed cells): interfaceN
 * (3 n_cells), interfaceD (n_cells), interfaceC (3 n_cells), interfaceS
 * (3 n_cells), and the face flatness zeta (n_faces, reconstruction.C:448-471).
 * The reference writes them to disk when writePlicFields is on
 * (svp_dict.write_plic_fields); here the caller writes them
 * (geometricvofext_amd.plic_io.write_plic_fields).  Any output may be NULL. */
int svp_get_plic_fields(svp_handle* h, double* interface_n, double* interface_d, double* interface_c,
                        double* interface_s, double* zeta, int where);

/* Device pointers of the resident fields, for zero-copy drivers. */
typedef struct {
    double* alpha;
    double* phi;
    double* U;
    double* U_boundary;
    double* dVf;
    const double* Cf;
    const double* Sf;
    const double* C;
    const double* V;
    void* stream;   /* hipStream_t */
} svp_device_view;
int svp_device_view_get(svp_handle* h, svp_device_view* out);

/* Per-stage GPU timing with HIP events on the handle's stream (no
 * counterpart in the reference, which has only the two cumulative CPU timers).
 * Stages: 0 detect, 1 normal, 2 findSignedDistance, 3 upwind, 4 interface
 * flux, 5 update, 6 bounding, 7 snap/clip; bit i of mask enables stage i.
 * svp_profile_get fills seconds[8] and counts[8] (either may be NULL) and
 * optionally resets them. */
#define SVP_N_STAGES 8
/* Statistics of the bounding passes since the last call, 4 ints per pass:
 * candidates, cells that wrote corrections, dependency rounds, corrected
 * faces.  Zero entries after the last pass that ran. */
int svp_bound_stats(svp_handle* h, int32_t* out, int cap);
/* Running sums over every bounding pass that had candidates since the
 * handle was created: candidates, active candidates, dependency rounds,
 * corrected faces, passes. */
int svp_bound_totals(svp_handle* h, uint64_t* out5);
int svp_profile_set(svp_handle* h, int mask);
int svp_profile_get(svp_handle* h, double* seconds, int64_t* counts, int reset);

/* Wait for all queued work of the handle and report the device error bits
 * its kernels set (SVP_E_DEVICE; the bits are cleared).  svp_get_alpha,
 * svp_get_alpha_phi, svp_get_rho_phi and svp_advect_sub_cycles report them
 * too when their output is a host buffer (they wait for the stream anyway);
 * with device outputs they do not wait, and svp_synchronize reports. */
int svp_synchronize(svp_handle* h);

/* Test hook, no counterpart in the reference: sets device error bits as a
 * failing kernel would, so the error path of every reporting call can be
 * exercised without making a kernel fail. */
int svp_test_set_error_bits(svp_handle* h, int32_t bits);

/* Batch of independent cases advanced together (no counterpart in the
 * reference, which runs one case per process: applications/test/
 * plicVofAdvectionFoam/plicVof.H:13-57 is the loop each case follows).  The
 * handles keep their own meshes and fields; svp_batch_step runs
 * reconstruct + advect of every case with its own time step, and each stage
 * of the step is one launch over all cases, each case running the kernels of
 * a single handle on its own data, so every case's results equal those of its
 * handle stepping alone.  The cases must share the alpha dictionary, use the
 * least-squares normal and the same cell type (all hexahedra or all
 * tetrahedra, at most 8 faces and 16 points, no splitWarpedFace); otherwise
 * SVP_E_UNSUPPORTED / SVP_E_ARG.  While the batch exists its cases work on
 * the batch's stream, so a handle may still be read (svp_get_alpha,
 * svp_get_interface, ...) between steps; destroy the batch before using the
 * handles on their own again.  A handle belongs to at most one batch. */
typedef struct svp_batch svp_batch;
int svp_batch_create(svp_handle* const* cases, int32_t n_cases, svp_batch** out);
int svp_batch_destroy(svp_batch* b);
/* One step of every case k with dt[k] > 0 (host array of n_cases); a case
 * with dt[k] <= 0 is left as it is (a finished case).  The step's launches
 * are captured once as a graph and replayed (the time steps travel in device
 * memory); environment SVP_BATCH_GRAPH=0 launches them one by one. */
int svp_batch_step(svp_batch* b, const double* dt);
/* Launch shapes: the x extent per case of the stages whose kernels stride
 * over their work (least-squares normal, plane placement, flux list,
 * interface flux, bounding records, dataflow bounding pass, bounding apply),
 * shapes[7] in that order.  svp_batch_tune times each over a few values on one step from the
 * current fields (the fields are restored after every trial), keeps the
 * fastest and writes them to shapes (may be NULL); results do not depend on
 * the shapes.  svp_batch_set_shapes sets them directly. */
int svp_batch_tune(svp_batch* b, const double* dt, int32_t* shapes);
int svp_batch_set_shapes(svp_batch* b, const int32_t* shapes);
/* GPU seconds of the batch's steps and the summed mixed-cell counts of the
 * cases' reconstructs (their running totals, svp_timers' n_ic_total). */
int svp_batch_timers(svp_batch* b, double* seconds, uint64_t* n_ic_total);
/* Per-stage timing of batch steps (the stage numbering of svp_profile_set). */
int svp_batch_profile_set(svp_batch* b, int mask);
int svp_batch_profile_get(svp_batch* b, double* seconds, int64_t* counts, int reset);

/* Message of the last failing call on this thread. */

const char* svp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* SIMPLIC_AMD_H */
