// Device-resident data of one SimPLIC alpha-transport instance and the
// launchers of its kernels.  Shared by the kernels and the host facade.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svp {

// Static mesh in HBM.  Index arrays are int32, values FP64, 3-vectors AoS.
struct DevMesh {
    int nP, nF, nIF, nC;
    const double* pts;      // 3*nP
    const int* fOff;        // nF+1, face -> point CSR
    const int* fPts;
    const int* own;         // nF
    const int* nbr;         // nF, -1 on boundary faces
    const unsigned char* fEmpty;  // nF, 1 for faces of empty patches
    // cell -> face CSR in primitiveMesh::cells() order (owned faces ascending,
    // then neighbour faces ascending)
    const int* cOff;
    const int* cFaces;
    // same cells, faces in ascending face order (fvc::surfaceIntegrate order),
    // neighbour faces stored as ~f
    const int* cFacesAsc;
    // cell -> point CSR (unique points, ascending labels)
    const int* cpOff;
    const int* cPts;
    int maxCellPoints;
    int triQuadFaces;       // 1 when every face has 3 or 4 points
    // cell-point-cell stencil: the cell first, then the other cells ascending
    const int* stOff;
    const int* st;
    int maxStencil;         // largest stencil (cells)
    // volPointInterpolation sources per point: cell c >= 0 or boundary face ~f
    const int* psOff;
    const int* psSrc;
    // geometry (computed on the device at create time)
    double* Cf;      // 3*nF
    double* Sf;      // 3*nF
    double* magSf;   // nF
    double* flat;    // nF
    double* fullCS;  // 6*nF: face::centre and face::areaNormal of the stored face (a fully submerged sub-face)
    double* C;       // 3*nC
    double* V;       // nC
    double* psW;     // point weights, psOff CSR
    double* linW;    // nIF, linear interpolation weight of the owner value
};

// Per-step state.  Interface data are compact arrays indexed by the position
// of a cell in the ascending mixed-cell list.
struct DevState {
    double* alpha;      // nC, current alpha
    double* alphaOld;   // nC
    double* phi;        // nF
    double* U;          // 3*nC
    double* Ub;         // 3*(nF-nIF), boundary values of U
    double* dVf;        // nF
    double* alphaPhi;   // nF
    double* corr;       // nF, bounding corrections (kept zero between passes)
    int* clearAtDetect; // [2] or null: counters the detection kernel zeroes (learned-normal counts)
    int* nbTag;         // nC: number of the last step that found the cell mixed or next to a mixed cell
    int epoch;          // number of the current reconstruct (1, 2, ...)
    int* nMixed;        // [1]
    int* mixed;         // capacity nC
    int* status;
    int* nCutPts;
    unsigned long long* faceClass;
    double* ifN;        // 3*cap
    double* ifD;
    double* ifC;        // 3*cap
    double* ifS;        // 3*cap
    double* un0;
    // bounding bookkeeping (limitFlux/boundFlux)
    int* oobCount;      // cells with alpha-1 > aTol or alpha < -aTol, spread over OOB_SLOTS counters
    int* fluxList;      // (cut cell, downwind face) pairs of the interface flux, cell-face CSR size
    int* fluxCount;     // [1]
    int* candCount;     // [1] candidates of the last pass (diagnostics)
    unsigned long long* candBits;  // nC bits: alpha outside [-aTol,1+aTol] and needBounding
    int* cand;          // 2*nC: ascending candidates of a pass, halves alternating by pass
    unsigned char* active;  // nC, per candidate: writes corrections this pass
    int* pend;          // nC, per candidate: active upwind predecessors not yet done
    int* queue;         // 2*nC, Kahn frontiers
    int* actIdx;        // nC cells -> candidate index of active cells, else -1
    int* corrPos;       // nF faces -> order key in correctedFaces (cell * max faces + rank), else -1
    int* touched;       // nC, cells whose alpha gets corrections
    int* touchedFlag;   // nC
    int* boundStats;    // [4*maxPasses]: candidates, active, rounds, corrected faces
    unsigned long long* boundTotals;  // [5] running sums: candidates, active, rounds, corrected faces, passes
    int* bcount;        // [32] per-pass counters of the bounding kernels
    unsigned long long* bndLook;  // per candidate-list block: epoch, look-back flag and count
    unsigned long long* detLook;  // per detection block: the same
    unsigned long long* detTicket;  // detection block tickets (grows by the block count per launch)
    unsigned long long* bndTicket;  // candidate-list block tickets (the same)
    int un0Fused;       // plane placement by 8-lane groups, which also computes Un0
    int fluxFan;        // face flux with four lanes per face (most faces warped)
    int fsdKeep;        // 1: where findSignedDistance writes no d/C/S they keep their values (isoRDF)
    int gridRounds;     // bounding rounds run as grid-wide kernels before the one-workgroup tail
    int bndGrid;        // workgroups of the grid-stride bounding kernels
    int bndFlow;        // workgroups of the one-launch dataflow pass, 0 = rounds and tail
    int bndListWords;   // bitmap words per block of the candidate list (64..4096)
    void* bndRec;       // nC records BndRec<MF> of the active bounding candidates
    int maxCellFaces;
    double* minmax;     // [4]: min, max before bounding, min, max after
    int* errFlag;       // [1] device error bits
    unsigned long long* nIcTotal;  // [1] running sum of mixed-cell counts
};

// The other half of the bounding double buffers (cand, active, corr, corrPos,
// actIdx): what the previous pass left there is cleared by the next pass's
// candidate-list kernel.
struct BndPrev {
    const int* cand;
    const unsigned char* active;
    double* corr;
    int* corrPos;
    int* actIdx;
};


struct Params {
    double mixedCellTol;
    int nAlphaBounds;
    double snapTol;
    int clip;
    int splitWarpedFace;
};

// error bits written to DevState::errFlag
enum : int {
    ERR_CELL_TOO_LARGE = 1,
    ERR_TOO_MANY_CROSSINGS = 2,
    ERR_BOUND_SCHEDULE = 4,
    ERR_BOUND_STALL = 16,
    ERR_DETECT_STALL = 32,
};
constexpr int MAX_BOUND_PASSES = 64;
constexpr int MAX_GRID_ROUNDS = 8;
constexpr int OOB_SLOTS = 64;   // counters of the out-of-bounds count, one per 64-byte line
constexpr int OOB_STRIDE = 16;

constexpr int DETECT_BLOCK = 256;

// Face data of an active bounding candidate, gathered once per pass with the
// first iteration of its redistribution loop done, so that the dependent
// part reads one contiguous record.  flags: 4 bits per face (owner,
// downwind, writable, upwind writer with a lower label).
template <int MF>
struct BndRec {
    double phf[MF], dvf[MF];
    double cor0[MF];    // downwind corrections after iteration 0
    int fid[MF];
    int oth[MF];        // cell on the other side of the face, -1 on boundary faces
    int succ[MF];       // Kahn schedule: candidate index of the active downwind neighbour with a higher label, else -1
    unsigned int flg[(MF + 7) / 8];
    int nf;
    unsigned int first;  // face slots of the first redistribution loop, in correctedFaces order
    int nFaces0;         // faces still able to take fluid after iteration 0
    double Vi;
    double num0;         // alpha_old/dt (+Su) - sum(dVf)/V/dt, the part of the new alpha's numerator without corrections
    double den;          // 1/dt (-Sp)
};

// One case of a batch of independent handles advanced together (one launch
// per stage, blockIdx.y = case).  st holds the four views of the case's state
// a step can need: bit 1 of the index swaps alpha and alphaOld (the double
// buffer), bit 0 takes the other half of the candidate list.  View 0 is the
// handle's state when the batch took it.
struct BatchCase {
    DevMesh m;
    DevState st[4];
    int detBlocks;        // blocks of the one-launch detection
    int listBlocks;       // blocks of the one-launch candidate list
    int listWords;        // candidate-bitmap words
    int nB;               // boundary faces
    // harness flow of the case (base fields scaled per step)
    const double* U0;
    const double* phi0;
    const double* U0b;
};

// Per case and step: its time step (<= 0: the case is not advanced this
// step) and the parities of its alpha buffers and candidate-list halves
// relative to view 0.
struct BatchStepInfo {
    double dt;
    int aPar;
    int hPar;
    int epoch;   // the case's reconstruct number of this step
};

// grid x dimension of each batched stage (the y dimension is the case count)
struct BatchDims {
    int ls, fsd, fluxList, fluxFaces, prep, flow, apply;
};

// launchers (simplic_kernels.hip)
void launch_geometry(const DevMesh& m, hipStream_t s);
void launch_detect(const DevMesh& m, const DevState& st, const Params& p, hipStream_t s);
// wl/wn: optional device list of mixed-cell positions and its length
void launch_ls_normal(const DevMesh& m, const DevState& st, hipStream_t s, const int* wl = nullptr,
                      const int* wn = nullptr);
// alphaGrad normal: minus the Gauss-linear gradient of alpha, normalised
void launch_grad_normal(const DevMesh& m, const DevState& st, bool pointLinear, hipStream_t s);

// isoRDF work buffers (reconstruction.C:194-405); zero between reconstructs
enum { RDF_NNEXT = 0, RDF_DONE = 1, RDF_ITERS = 2, RDF_WN = 3, RDF_WN2 = 4, RDF_CTL = 8 };
struct RdfBufs {
    int* next;                 // nC: nextToInterface (mixed cells and the cells sharing a point with one)
    int* nextList;             // nC: those cells, ctl[RDF_NNEXT] of them
    unsigned char* tooCoarse;  // nC
    double* rdf;               // nC: reconstructed distance function
    double* cellN;             // 3*nC: interfaceN per cell
    double* cellC;             // 3*nC: interfaceC per cell
    double* g;                 // 3*nC: per mixed cell, the least-squares gradient (interfaceNormal)
    double* res;               // 2*nC: per mixed cell, its normalRes and normRes terms
    int* counted;              // nC: per mixed cell, 1 when it enters the residual averages
    int* wl;                   // nC: mixed cells placed this iteration (not too coarse)
    int* wl2;                  // nC: mixed cells too coarse at the end
    int* ctl;                  // [RDF_CTL]: next count, done flag, iterations run, wl and wl2 counts
};
// (ctl cleared by the caller) mark nextToInterface and list it, zero d/C/S/n
// of the mixed cells, g = least-squares gradient of alpha
void launch_rdf_begin(const DevMesh& m, const DevState& st, const RdfBufs& r, hipStream_t s);
// one iteration: normals, plane placement of the cells that are not too
// coarse, distance function, its gradient, residuals and the stopping rule
// (a no-op once the rule has stopped the loop)
void launch_rdf_iteration(const DevMesh& m, const DevState& st, const Params& p, const RdfBufs& r, int iter,
                          int iterations, double tol, double relTol, hipStream_t s);
// the too-coarse work list, then clear the buffers for the next reconstruct
void launch_rdf_end(const DevMesh& m, const DevState& st, const RdfBufs& r, hipStream_t s);
void launch_find_signed_distance(const DevMesh& m, const DevState& st, const Params& p, hipStream_t s,
                                 const int* wl = nullptr, const int* wn = nullptr);
// pipeline stages of one step, in launch order (also the profiler's slots)
enum Stage : int {
    STAGE_DETECT = 0,
    STAGE_NORMAL = 1,
    STAGE_FSD = 2,
    STAGE_UPWIND = 3,
    STAGE_IFLUX = 4,
    STAGE_UPDATE = 5,
    STAGE_BOUND = 6,
    STAGE_SNAPCLIP = 7,
    N_STAGES = 8,
};
// passSeq: the handle's running count of bounding passes (selects the
// candidate-list half of each pass)
void launch_advect_stage(int stage, const DevMesh& m, const DevState& st, const Params& p, double dt,
                         const double* Sp, const double* Su, int* passSeq, hipStream_t s);
void launch_alpha_phi(const DevMesh& m, const DevState& st, double dt, hipStream_t s);
void launch_map_alpha(const DevMesh& m, const DevState& st, const double* n, const double* d, double lower,
                      double upper, hipStream_t s);
// pass 0: polygon sizes and their scans (totals[0] points, totals[1] faces);
// pass 1: the polygons
void launch_plic_surface(const DevMesh& m, const DevState& st, int nMixed, int pass, int* ptOff, int* faceIdx,
                         int* totals, double* pts, int* faceCells, int* faceOff, hipStream_t s);
void launch_plic_close(int* faceOff, const int* totals, hipStream_t s);
// full-mesh interfaceN / interfaceD / interfaceC / interfaceS of the last
// reconstruct (zero outside the mixed cells); null outputs are skipped
void launch_plic_fields(const DevMesh& m, const DevState& st, double* N, double* D, double* C, double* S,
                        hipStream_t s);
void launch_rho_phi(int nF, const double* ap, const double* phi, double rho1, double rho2, double w,
                    double* out, int mode, hipStream_t s);
int detect_blocks(int nC);
int detect_list_blocks(int nC);   // blocks of the one-launch detection
// face slots of the bounding records and kernels for cells of at most
// maxCellFaces faces (4, 6, 8 or 24), and the bytes of one record
int bound_faces(int maxCellFaces);
size_t bound_record_bytes(int maxCellFaces);
void launch_count_warped(const DevMesh& m, int* count, hipStream_t s);  // faces below flatness 1 - tSmall
// the state of bounding pass number seq of a handle: the half (seq & 1) of
// the double-buffered bounding arrays
DevState bound_pass_view(const DevMesh& m, const DevState& st, int seq);
// batched stages (same order and kernels as one handle's reconstruct and
// advect); G is the plane-placement lane group and MF the bounding record
// size shared by all cases
void launch_batch_stage(int stage, const BatchCase* bc, const BatchStepInfo* info, int nCases, int maxC, int maxF,
                        int maxDetBlocks, int maxListBlocks, const BatchDims& d, const Params& p, int G, int MF,
                        int triQuad, hipStream_t s);
void launch_batch_courant(const BatchCase* bc, const BatchStepInfo* info, int nCases, int maxC,
                          unsigned long long* out, hipStream_t s);
void launch_batch_scale_flow(const BatchCase* bc, const BatchStepInfo* info, const double* factors, int nCases,
                             long maxN, hipStream_t s);

}  // namespace svp
