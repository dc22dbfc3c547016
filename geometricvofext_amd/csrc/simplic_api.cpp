// Host side of libsimplic_amd.so: the solveVofEqu facade over the HIP kernels
// and its C-ABI (include/simplic_amd.h, include/simplic_harness.h).
//
// Reference interface mirrored: src/SimPLIC/solveVofEqu/solveVofEqu.H:67-216
// (constructor, reconstruct(), advect(Sp, Su), alphaPhi(), getRhoPhi(),
// reconstructionTime(), advectionTime()).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "api_common.hpp"
#include "simplic_amd.h"
#include "simplic_harness.h"
#include "simplic_device.hpp"
#include "simplic_learned.h"
#include "simplic_types.hpp"
#include "learned_kernels.hpp"
#include "learned_model.hpp"

namespace svp {
void launch_scale_flow(int nC, int nF, int nB, const double* U0, const double* phi0, const double* U0b,
                       double factor, double* U, double* phi, double* Ub, hipStream_t s);
void launch_courant(const DevMesh& m, const double* phi, const double* alpha, unsigned long long* out,
                    hipStream_t s);
void launch_count_nonfinite(const double* x, long n, int* count, hipStream_t s);
void launch_init_alpha(const DevMesh& m, const void* shapes, int nShapes, int maxDepth, double* alpha,
                       hipStream_t s);
void launch_init_alpha_surface(const DevMesh& m, const double* tris, int nT, int mode, double* frac, int* cutList,
                               int* nCut, double* alpha, hipStream_t s);
}  // namespace svp

using namespace svp::api;

// Scratch of svp_get_plic_surface_device, the PLIC export that allocates
// nothing: the point count and polygon index of every mixed cell (scanned in
// place into offsets), then the two totals.  A file-scope device array, so
// every GPU the library runs on holds its own copy until the process ends.
// In host code its name is only a stand-in: kernels get the address that
// hipGetSymbolAddress returns for the current device, taken once per device
// and kept.  Every call on a GPU shares that GPU's copy (calls that overlap on
// one GPU race); each call writes all it reads, so nothing is cleared.
static __device__ __align__(256) int scratch[2 * SVP_PLIC_NOALLOC_MAX_CELLS + 2];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
int* g_scratchAddr[kMaxDevices] = {};

// address of this device's scratch (the caller has made `device` current)
int* scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) throw SvpError(SVP_E_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (!g_scratchAddr[device]) {
        void* p = nullptr;
        HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw SvpError(SVP_E_HIP, "hipGetSymbolAddress returned no address for the PLIC scratch");
        g_scratchAddr[device] = static_cast<int*>(p);
    }
    return g_scratchAddr[device];
}

// grad(alpha1) scheme name with single spaces; "" when not one of the two
// the alphaGrad normal provides
std::string grad_scheme_name(const std::string& v) {
    std::string out, tok;
    for (char ch : v + " ") {
        if (ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r') {
            if (!tok.empty()) out += (out.empty() ? "" : " ") + tok;
            tok.clear();
        } else {
            tok += ch;
        }
    }
    return (out == "Gauss linear" || out == "Gauss pointLinear") ? out : std::string();
}
}  // namespace

namespace gve {
namespace SimPLIC {

using namespace svp;

// Host mirror of the reference's solveVofEqu: owns the device mesh, the
// per-step state and the stream all kernels of this alpha field run on.
class solveVofEqu {
public:
    solveVofEqu(const svp_mesh& mesh, const svp_dict& dict, int device) : dict_(dict), device_(device) {
        HIPCHECK(hipSetDevice(device));
        HIPCHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        params_.mixedCellTol = dict.mixed_cell_tol;
        params_.nAlphaBounds = dict.n_alpha_bounds;
        params_.snapTol = dict.snap_tol;
        params_.clip = dict.clip;
        params_.splitWarpedFace = dict.split_warped_face;
        learned_.active = std::string(dict.orientation_method) == "learned";
        gaussNormal_ = std::string(dict.orientation_method) == "alphaGrad";
        {
            const std::string g = grad_scheme_name(std::string(dict.grad_scheme, strnlen(dict.grad_scheme, sizeof dict.grad_scheme)));
            if (g.empty())
                throw SvpError(SVP_E_UNSUPPORTED, "grad_scheme: this library provides Gauss linear and Gauss pointLinear");
            pointLinear_ = g == "Gauss pointLinear";
        }
        rdfNormal_ = std::string(dict.orientation_method) == "isoRDF" || std::string(dict.orientation_method) == "RDF";
        learned_.tol = dict.learned_residual_tol;
        buildMesh(mesh);
        allocState();
        launch_geometry(m_, stream_);
        HIPCHECK(hipGetLastError());
        // face flux with four lanes per face where most faces are warped
        // (fan-integrated); SVP_FLUX_FAN=0/1 forces it off/on
        if (m_.triQuadFaces && st_.un0Fused) {
            int nWarped = 0;
            launch_count_warped(m_, nonfinite_, stream_);
            HIPCHECK(hipMemcpyAsync(&nWarped, nonfinite_, sizeof(int), hipMemcpyDeviceToHost, stream_));
            HIPCHECK(hipStreamSynchronize(stream_));
            st_.fluxFan = 2L * nWarped > (long)m_.nF ? 1 : 0;
            if (const char* e = std::getenv("SVP_FLUX_FAN")) st_.fluxFan = std::atoi(e) != 0 ? 1 : 0;
        }
        HIPCHECK(hipStreamSynchronize(stream_));
    }

    ~solveVofEqu() {
        (void)hipSetDevice(device_);
        if (stream_) (void)hipStreamSynchronize(stream_);
        for (auto& e : events_) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    // reconstruction::reconstruct (reconstruction.C:680-722)
    void reconstruct() {
        hipEvent_t a = event(), b = event();
        HIPCHECK(hipEventRecord(a, stream_));
        st_.epoch = nextEpoch();
        stage(STAGE_DETECT, [&] { launch_detect(m_, st_, params_, stream_); });
        if (learned_.active) {
            reconstructLearned();
        } else {
            stage(STAGE_NORMAL, [&] {
                if (rdfNormal_) rdfNormals();
                else if (gaussNormal_) launch_grad_normal(m_, st_, pointLinear_, stream_);
                else launch_ls_normal(m_, st_, stream_);
            });
            stage(STAGE_FSD, [&] { placePlanes(); });
        }
        HIPCHECK(hipEventRecord(b, stream_));
        HIPCHECK(hipGetLastError());
        pendingRecon_.push_back({a, b});
        reconstructed_ = true;
    }

    // advection::advect (advectionTemplates.C:352-418); alpha1.oldTime() is
    // the field at the call, i.e. the double buffer swaps here.
    void advect(double dt, const double* Sp, const double* Su) {
        if (!(dt > 0.0) || !std::isfinite(dt)) throw SvpError(SVP_E_ARG, "advect: dt must be positive");
        if (!reconstructed_) reconstruct();
        hipEvent_t a = event(), b = event();
        HIPCHECK(hipEventRecord(a, stream_));
        std::swap(st_.alpha, st_.alphaOld);
        for (int sg = STAGE_UPWIND; sg <= STAGE_SNAPCLIP; ++sg)
            stage(sg, [&] { launch_advect_stage(sg, m_, st_, params_, dt, Sp, Su, &bndPassSeq_, stream_); });
        HIPCHECK(hipEventRecord(b, stream_));
        HIPCHECK(hipGetLastError());
        pendingAdvect_.push_back({a, b});
        lastDt_ = dt;
        alphaPhiValid_ = false;
        reconstructed_ = false;
    }

    // orientationMethod learned: network normals on cells with a uniform
    // 3x3x3 stencil; the residual test of each network normal's cube plane;
    // least-squares normals on the other cells and on those that fail it; one
    // plane placement for all.
    void reconstructLearned() {
        if (!learned_.ready)
            throw SvpError(SVP_E_ARG, "orientationMethod learned needs a model (svp_set_normal_model)");
        int* nLs = learned_.counts;
        int* nFb = learned_.counts + 1;
        stage(STAGE_NORMAL, [&] {
            // the counts were zeroed by the detection kernel (clearAtDetect);
            // the residual test runs inside the inference kernel (tol < 1)
            nm::launch_infer(learned_.lat, learned_.stride, st_.alpha, st_.mixed, st_.nMixed, m_.C, m_.V, learned_.wb, learned_.pq,
                             learned_.tol, st_.ifN, learned_.learned, learned_.resid, learned_.lsList, nLs, nFb,
                             m_.nC, stream_);
            launch_ls_normal(m_, st_, stream_, learned_.lsList, nLs);
        });
        stage(STAGE_FSD, [&] { launch_find_signed_distance(m_, st_, params_, stream_); });
    }

    void setNormalModel(const svp_normal_model& nmod) {
        if (!learned_.active)
            throw SvpError(SVP_E_ARG, "svp_set_normal_model: the handle's orientationMethod is not learned");
        if (nmod.device != device_) throw SvpError(SVP_E_ARG, "svp_set_normal_model: model on another device");
        HIPCHECK(hipStreamSynchronize(nmod.stream));
        if (!learned_.lat) {
            learned_.lat = dalloc<int>(27L * m_.nC);
            learned_.pq = dalloc<float>(nm::N_PADDED);
            learned_.wb = dalloc<uint16_t>(nm::N_BF16);
            learned_.learned = dalloc<unsigned char>(m_.nC);
            learned_.lsList = dalloc<int>(m_.nC);
            learned_.counts = dalloc<int>(2);
            st_.clearAtDetect = learned_.counts;
            learned_.resid = dalloc<double>(m_.nC);
            DBuf cube;
            cube.alloc<unsigned char>(m_.nC);
            nm::launch_lattice(m_, (unsigned char*)cube.p, learned_.lat, stream_);
            DBuf probe;
            probe.alloc<int>(2 + 27);
            nm::launch_lattice_probe(learned_.lat, m_.nC, (int*)probe.p, stream_);
            HIPCHECK(hipGetLastError());
            int pr[2 + 27];
            HIPCHECK(hipMemcpyAsync(pr, probe.p, sizeof pr, hipMemcpyDeviceToHost, stream_));
            HIPCHECK(hipStreamSynchronize(stream_));
            learned_.stride.affine = pr[1];
            for (int L = 0; L < 27; ++L) learned_.stride.off[L] = pr[2 + L];
            if (const char* e = std::getenv("SVP_LATTICE_AFFINE"))
                if (std::atoi(e) == 0) learned_.stride.affine = 0;
        }
        HIPCHECK(hipMemcpyAsync(learned_.pq, nmod.pq.p, sizeof(float) * nm::N_PADDED, hipMemcpyDeviceToDevice,
                                stream_));
        HIPCHECK(hipMemcpyAsync(learned_.wb, nmod.wb.p, sizeof(uint16_t) * nm::N_BF16, hipMemcpyDeviceToDevice,
                                stream_));
        HIPCHECK(hipStreamSynchronize(stream_));
        learned_.ready = true;
        reconstructed_ = false;
    }

    void learnedStats(int64_t* out4) {
        if (!learned_.ready) throw SvpError(SVP_E_ARG, "svp_learned_stats: no learned reconstruction");
        int c[2] = {0, 0};
        const int n = nMixed();
        HIPCHECK(hipMemcpy(c, learned_.counts, sizeof(c), hipMemcpyDeviceToHost));
        out4[0] = n;
        out4[1] = n - c[0];
        out4[2] = c[1];
        out4[3] = c[0] - c[1];
    }

    void learnedResiduals(double* out, int where) {
        if (!learned_.ready) throw SvpError(SVP_E_ARG, "svp_learned_residuals: no learned reconstruction");
        if (learned_.tol >= 1.0) throw SvpError(SVP_E_ARG, "svp_learned_residuals: the test is off (tolerance >= 1)");
        copyOut(out, learned_.resid, nMixed(), where);
    }

    // Per-stage HIP-event timing on this handle's stream (bit i of mask =
    // stage i); accumulated on readout.
    void setProfile(int mask) {
        drainProfile();
        profMask_ = mask;
    }
    void profile(double* seconds, int64_t* counts) {
        HIPCHECK(hipStreamSynchronize(stream_));
        drainProfile();
        for (int i = 0; i < N_STAGES; ++i) {
            if (seconds) seconds[i] = profSec_[i];
            if (counts) counts[i] = profCnt_[i];
        }
    }
    void resetProfile() {
        HIPCHECK(hipStreamSynchronize(stream_));
        drainProfile();
        for (int i = 0; i < N_STAGES; ++i) { profSec_[i] = 0.0; profCnt_[i] = 0; }
    }

    void checkDeviceErrors() {
        int err = 0;
        HIPCHECK(hipMemcpyAsync(&err, st_.errFlag, sizeof(int), hipMemcpyDeviceToHost, stream_));
        HIPCHECK(hipStreamSynchronize(stream_));
        if (err) {
            HIPCHECK(hipMemsetAsync(st_.errFlag, 0, sizeof(int), stream_));
            char buf[256];
            std::snprintf(buf, sizeof buf,
                          "device error bits 0x%x (1: cell exceeds cut limits, 2: more than %d plane "
                          "crossings on a face, 4: bounding candidate overflow, 8: bounding work overflow, "
                          "16: bounding wait without progress, 32: mixed-cell detection wait without "
                          "progress)",
                          err, MAX_IF_POINTS);
            throw SvpError(SVP_E_DEVICE, buf);
        }
    }

    void setAlpha(const double* a, int where) {
        copyIn(st_.alpha, a, m_.nC, where);
        checkFinite(st_.alpha, m_.nC, "alpha");
        reconstructed_ = false;
    }
    void setFlow(const double* phi, const double* U, const double* Ub, int where) {
        copyIn(st_.phi, phi, m_.nF, where);
        copyIn(st_.U, U, 3L * m_.nC, where);
        if (Ub) copyIn(st_.Ub, Ub, 3L * nB_, where);
        else HIPCHECK(hipMemsetAsync(st_.Ub, 0, sizeof(double) * 3 * std::max(nB_, 1), stream_));
        checkFinite(st_.phi, m_.nF, "phi");
        checkFinite(st_.U, 3L * m_.nC, "U");
    }

    // Host outputs wait for the stream anyway, so their calls also report
    // the device error bits of the work before them; with device outputs the
    // caller's svp_synchronize does.
    void getAlpha(double* out, int where) {
        copyOut(out, st_.alpha, m_.nC, where);
        if (where != SVP_DEVICE) checkDeviceErrors();
    }

    void setErrorBits(int bits) {
        HIPCHECK(hipMemcpyAsync(st_.errFlag, &bits, sizeof(int), hipMemcpyHostToDevice, stream_));
        HIPCHECK(hipStreamSynchronize(stream_));
    }

    const double* alphaPhi() {
        if (!alphaPhiValid_) {
            if (lastDt_ <= 0.0) throw SvpError(SVP_E_ARG, "alphaPhi: advect has not run");
            launch_alpha_phi(m_, st_, lastDt_, stream_);
            alphaPhiValid_ = true;
        }
        return st_.alphaPhi;
    }

    // calcInterfaceNFromIsoRDF (reconstruction.C:194-405): every iteration
    // runs on the device, the stopping rule included, so the host queues
    // `iterations` of them without waiting; those after the stop do nothing.
    // The planes of the cells that are not too coarse are placed by the
    // findSignedDistance kernels over a work list; the final placement of
    // reconstruct then only has the too-coarse cells left (the others would
    // get the same planes again from the same normals).
    void rdfNormals() {
        const long nC = m_.nC;
        if (!rdf_.next) {
            rdf_.next = dalloc<int>(nC);
            rdf_.nextList = dalloc<int>(nC);
            rdf_.tooCoarse = dalloc<unsigned char>(nC);
            rdf_.rdf = dalloc<double>(nC);
            rdf_.cellN = dalloc<double>(3 * nC);
            rdf_.cellC = dalloc<double>(3 * nC);
            rdf_.g = dalloc<double>(3 * nC);
            rdf_.res = dalloc<double>(2 * nC);
            rdf_.counted = dalloc<int>(nC);
            rdf_.wl = dalloc<int>(nC);
            rdf_.wl2 = dalloc<int>(nC);
            rdf_.ctl = dalloc<int>(RDF_CTL);
        }
        HIPCHECK(hipMemsetAsync(rdf_.ctl, 0, sizeof(int) * RDF_CTL, stream_));
        launch_rdf_begin(m_, st_, rdf_, stream_);
        const int iterations = dict_.rdf_iterations;
        for (int iter = 0; iter < iterations; ++iter)
            launch_rdf_iteration(m_, st_, params_, rdf_, iter, iterations, dict_.rdf_tol, dict_.rdf_rel_tol,
                                 stream_);
        launch_rdf_end(m_, st_, rdf_, stream_);
        HIPCHECK(hipGetLastError());
    }

    // final plane placement of reconstruct (reconstruction.C:711-716)
    void placePlanes() {
        if (rdfNormal_ && dict_.rdf_iterations > 0)
            launch_find_signed_distance(m_, st_, params_, stream_, rdf_.wl2, rdf_.ctl + RDF_WN2);
        else
            launch_find_signed_distance(m_, st_, params_, stream_);
    }

    // getRhoPhi with constant densities (advection.H:329-343) of the last
    // advect into the device buffer out.
    void rhoPhi(double rho1, double rho2, double* out) {
        launch_rho_phi(m_.nF, alphaPhi(), st_.phi, rho1, rho2, 0.0, out, 0, stream_);
    }

    // The alpha sub-step of interPlicFoam (alphaEqnSubCycle.H:30-60 with
    // alphaEqn.H:9-20): nSub > 1 runs nSub reconstruct + advect steps of
    // deltaT/nSub (Time::subCycle divides deltaT by nSub) and sums
    // (subDeltaT/totalDeltaT) rhoPhi over them; nSub == 1 is one step whose
    // rhoPhi is returned as is.  Sp/Su are held for every sub-step.  Returns
    // the device buffer holding rhoPhi (n_faces).
    const double* advectSubCycles(double dt, int nSub, const double* Sp, const double* Su, double rho1,
                                  double rho2) {
        if (nSub < 1) throw SvpError(SVP_E_ARG, "nAlphaSubCycles must be at least 1");
        if (!(dt > 0.0) || !std::isfinite(dt)) throw SvpError(SVP_E_ARG, "advect: dt must be positive");
        if (!std::isfinite(rho1) || !std::isfinite(rho2)) throw SvpError(SVP_E_ARG, "rho1/rho2 must be finite");
        double* sum = rhoPhiSum_.p ? (double*)rhoPhiSum_.p : rhoPhiSum_.alloc<double>(m_.nF);
        if (nSub == 1) {
            reconstruct();
            advect(dt, Sp, Su);
            rhoPhi(rho1, rho2, sum);
            return sum;
        }
        const double subDt = dt / nSub;
        const double w = subDt / dt;
        for (int k = 0; k < nSub; ++k) {
            reconstruct();
            advect(subDt, Sp, Su);
            launch_rho_phi(m_.nF, alphaPhi(), st_.phi, rho1, rho2, w, sum, k == 0 ? 1 : 2, stream_);
        }
        return sum;
    }

    // reconstruction::interface (reconstruction.C:787-835) of the last
    // reconstruct: sizes always; the polygons when all three outputs are given.
    void plicSurface(int32_t* nFaces, int32_t* nPts, int32_t* cells, int32_t* offs, double* pts, int where) {
        if (!nFaces || !nPts) throw SvpError(SVP_E_ARG, "svp_get_plic_surface: null size output");
        const bool want = cells || offs || pts;
        if (want && !(cells && offs && pts))
            throw SvpError(SVP_E_ARG, "svp_get_plic_surface: give face_cells, face_offsets and points, or none");
        checkDeviceErrors();
        const int k = nMixed();
        int tot[2] = {0, 0};
        if (k > 0) {
            if (plicIdx_.bytes < sizeof(int) * (2L * k + 2)) plicIdx_.alloc<int>(2L * k + 2);
            int* ptOff = (int*)plicIdx_.p;
            int* faceIdx = ptOff + k;
            int* totals = faceIdx + k;
            launch_plic_surface(m_, st_, k, 0, ptOff, faceIdx, totals, nullptr, nullptr, nullptr, stream_);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipMemcpyAsync(tot, totals, sizeof(tot), hipMemcpyDeviceToHost, stream_));
            HIPCHECK(hipStreamSynchronize(stream_));
            checkDeviceErrors();
        }
        *nPts = tot[0];
        *nFaces = tot[1];
        if (!want) return;
        int32_t* dCells = cells;
        int32_t* dOffs = offs;
        double* dPts = pts;
        if (where != SVP_DEVICE) {
            const size_t need = sizeof(double) * 3 * std::max(tot[0], 1) + sizeof(int) * (2L * tot[1] + 1) + 64;
            if (plicOut_.bytes < need) plicOut_.alloc<char>(need);
            dPts = (double*)plicOut_.p;
            dCells = (int32_t*)(dPts + 3L * std::max(tot[0], 1));
            dOffs = dCells + tot[1];
        }
        if (k > 0) {
            int* ptOff = (int*)plicIdx_.p;
            launch_plic_surface(m_, st_, k, 1, ptOff, ptOff + k, nullptr, dPts, dCells, dOffs, stream_);
            HIPCHECK(hipGetLastError());
        }
        // the closing offset: face_offsets[n_faces] = n_points
        HIPCHECK(hipMemcpyAsync(dOffs + tot[1], &tot[0], sizeof(int), hipMemcpyHostToDevice, stream_));
        HIPCHECK(hipStreamSynchronize(stream_));
        checkDeviceErrors();
        if (where != SVP_DEVICE) {
            copyOut(pts, dPts, 3L * tot[0], SVP_HOST);
            copyOutT(cells, dCells, tot[1], SVP_HOST);
            copyOutT(offs, dOffs, tot[1] + 1L, SVP_HOST);
        }
    }

    // The same export into the caller's device buffers, with the static
    // scratch in place of plicIdx_ and the closing offset written by a kernel:
    // no device or pinned allocation.  Mixed-cell counts above
    // SVP_PLIC_NOALLOC_MAX_CELLS are refused.
    void plicSurfaceNoAlloc(int32_t* nFaces, int32_t* nPts, int32_t* cells, int32_t* offs, double* pts) {
        if (!nFaces || !nPts) throw SvpError(SVP_E_ARG, "svp_get_plic_surface_device: null size output");
        const bool want = cells || offs || pts;
        if (want && !(cells && offs && pts))
            throw SvpError(SVP_E_ARG,
                           "svp_get_plic_surface_device: give face_cells, face_offsets and points, or none");
        int* ptOff = scratch_address(device_);
        checkDeviceErrors();
        const int k = nMixed();
        if (k > SVP_PLIC_NOALLOC_MAX_CELLS)
            throw SvpError(SVP_E_ARG, "svp_get_plic_surface_device: " + std::to_string(k) +
                                          " mixed cells exceed the scratch of " +
                                          std::to_string(SVP_PLIC_NOALLOC_MAX_CELLS) +
                                          " cells (svp_get_plic_surface has no limit)");
        int* faceIdx = ptOff + k;
        int* totals = faceIdx + k;
        int tot[2] = {0, 0};
        if (k > 0) {
            launch_plic_surface(m_, st_, k, 0, ptOff, faceIdx, totals, nullptr, nullptr, nullptr, stream_);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipMemcpyAsync(tot, totals, sizeof(tot), hipMemcpyDeviceToHost, stream_));
            HIPCHECK(hipStreamSynchronize(stream_));
            checkDeviceErrors();
        }
        *nPts = tot[0];
        *nFaces = tot[1];
        if (!want) return;
        if (k > 0) launch_plic_surface(m_, st_, k, 1, ptOff, faceIdx, nullptr, pts, cells, offs, stream_);
        launch_plic_close(offs, k > 0 ? totals : nullptr, stream_);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(stream_));
        checkDeviceErrors();
    }

    int rdfIterationsRun() {
        if (!rdfNormal_ || !rdf_.ctl) return 0;
        int n = 0;
        HIPCHECK(hipMemcpyAsync(&n, rdf_.ctl + RDF_ITERS, sizeof(int), hipMemcpyDeviceToHost, stream_));
        HIPCHECK(hipStreamSynchronize(stream_));
        return n;
    }

    // reconstruction::mapAlphaField (reconstruction.C:725-784) after a mesh
    // change: n (3 per cell) and d (per cell) are the interface fields the
    // caller mapped onto this mesh.
    void mapAlphaField(const double* n, const double* d, double lower, double upper, int where) {
        if (!n || !d) throw SvpError(SVP_E_ARG, "svp_map_alpha_field: null n or d");
        if (!std::isfinite(lower) || !std::isfinite(upper)) throw SvpError(SVP_E_ARG, "refine levels must be finite");
        const double* dn = n;
        const double* dd = d;
        if (where != SVP_DEVICE) {
            if (mapBuf_.bytes < sizeof(double) * 4 * (size_t)m_.nC) mapBuf_.alloc<double>(4L * m_.nC);
            double* b = (double*)mapBuf_.p;
            copyIn(b, n, 3L * m_.nC, SVP_HOST);
            copyIn(b + 3L * m_.nC, d, m_.nC, SVP_HOST);
            dn = b;
            dd = b + 3L * m_.nC;
        }
        hipEvent_t a = event(), b = event();
        HIPCHECK(hipEventRecord(a, stream_));
        launch_map_alpha(m_, st_, dn, dd, lower, upper, stream_);
        HIPCHECK(hipEventRecord(b, stream_));
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(stream_));
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, a, b));
        mapTime_ += 1e-3 * ms;
        freeEvents_.push_back(a);
        freeEvents_.push_back(b);
        checkDeviceErrors();
        reconstructed_ = false;
    }

    // reconstruction::alphaMappingTime (reconstruction.H:396-399): GPU seconds
    // spent in mapAlphaField
    double alphaMappingTime() const { return mapTime_; }

    // The full-mesh interface fields and the face flatness zeta of the
    // reference (reconstruction.C:448-471, 520-584; written by the reference
    // when writePlicFields is on).  Any output may be null.
    void plicFields(double* N, double* D, double* C, double* S, double* zeta, int where) {
        checkDeviceErrors();
        const long nC = m_.nC;
        if (N || D || C || S) {
            double *dN = N, *dD = D, *dC = C, *dS = S;
            if (where != SVP_DEVICE) {
                if (plicFields_.bytes < sizeof(double) * 10 * (size_t)nC) plicFields_.alloc<double>(10L * nC);
                double* b = (double*)plicFields_.p;
                dN = N ? b : nullptr;
                dD = D ? b + 3 * nC : nullptr;
                dC = C ? b + 4 * nC : nullptr;
                dS = S ? b + 7 * nC : nullptr;
            }
            launch_plic_fields(m_, st_, dN, dD, dC, dS, stream_);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipStreamSynchronize(stream_));
            if (where != SVP_DEVICE) {
                if (N) copyOut(N, dN, 3 * nC, SVP_HOST);
                if (D) copyOut(D, dD, nC, SVP_HOST);
                if (C) copyOut(C, dC, 3 * nC, SVP_HOST);
                if (S) copyOut(S, dS, 3 * nC, SVP_HOST);
            }
        }
        if (zeta) copyOut(zeta, m_.flat, m_.nF, where);
        if (where == SVP_DEVICE) HIPCHECK(hipStreamSynchronize(stream_));
    }

    // Host sources are staged in device buffers kept by the handle.
    void stageSources(const double*& Sp, const double*& Su, int where) {
        if (where != SVP_HOST) return;
        if (Sp) {
            if (!harnessBufs_[1].p) allocHarness(harnessBufs_[1], m_.nC);
            copyIn((double*)harnessBufs_[1].p, Sp, m_.nC, SVP_HOST);
            Sp = (const double*)harnessBufs_[1].p;
        }
        if (Su) {
            if (!harnessBufs_[2].p) allocHarness(harnessBufs_[2], m_.nC);
            copyIn((double*)harnessBufs_[2].p, Su, m_.nC, SVP_HOST);
            Su = (const double*)harnessBufs_[2].p;
        }
    }

    int nMixed() {
        int n = 0;
        HIPCHECK(hipMemcpyAsync(&n, st_.nMixed, sizeof(int), hipMemcpyDeviceToHost, stream_));
        HIPCHECK(hipStreamSynchronize(stream_));
        return n;
    }

    void timers(double& rec, double& adv, uint64_t& nic) {
        HIPCHECK(hipStreamSynchronize(stream_));
        auto drain = [&](std::vector<std::pair<hipEvent_t, hipEvent_t>>& v, double& acc) {
            for (auto& e : v) {
                float ms = 0.f;
                HIPCHECK(hipEventElapsedTime(&ms, e.first, e.second));
                acc += 1e-3 * ms;
                freeEvents_.push_back(e.first);
                freeEvents_.push_back(e.second);
            }
            v.clear();
        };
        drain(pendingRecon_, reconTime_);
        drain(pendingAdvect_, advectTime_);
        rec = reconTime_;
        adv = advectTime_;
        unsigned long long n = 0;
        HIPCHECK(hipMemcpy(&n, st_.nIcTotal, sizeof(n), hipMemcpyDeviceToHost));
        nic = n;
    }

    // Batched stepping (class caseBatch): whether this handle's step is the
    // plain path a batch runs (least-squares normal, fast plane placement, no
    // warped-face split), its view of the state, and the bookkeeping a batch
    // step does for it.
    bool batchable(std::string& why) const {
        if (learned_.active || rdfNormal_ || gaussNormal_) {
            why = "a batch runs the least-squares normal (orientationMethod LS / isoAlphaGrad)";
            return false;
        }
        if (params_.splitWarpedFace || !st_.un0Fused) {
            why = "a batch runs cells of at most 8 faces and 16 points without splitWarpedFace";
            return false;
        }
        return true;
    }
    int batchGroup() const { return st_.maxCellFaces <= 4 ? 4 : (st_.maxCellFaces <= 6 ? 6 : 8); }
    int passSeq() const { return bndPassSeq_; }
    // number of the next reconstruct (tags of needBounding); the tags are
    // cleared before the count wraps
    int nextEpoch() {
        if (st_.epoch == 0x7ffffffe) {
            HIPCHECK(hipMemsetAsync(st_.nbTag, 0, sizeof(int) * m_.nC, stream_));
            st_.epoch = 0;
        }
        return st_.epoch + 1;
    }

    void afterBatchStep(double dt, int nPasses) {
        st_.epoch = nextEpoch();
        std::swap(st_.alpha, st_.alphaOld);
        bndPassSeq_ += nPasses;
        lastDt_ = dt;
        alphaPhiValid_ = false;
        reconstructed_ = false;
    }

    // accessors used by the C-ABI
    DevMesh m_{};
    DevState st_{};
    Params params_{};
    svp_dict dict_;
    int device_ = 0;
    int nB_ = 0;
    hipStream_t stream_ = nullptr;
    double lastDt_ = 0.0;
    // harness base flow
    double* U0_ = nullptr;
    double* phi0_ = nullptr;
    double* U0b_ = nullptr;
    unsigned long long* courant_ = nullptr;
    int* nonfinite_ = nullptr;
    std::vector<double> hostC_, hostCf_, hostSf_;
    // learned orientation method
    struct {
        bool active = false, ready = false;
        double tol = 0.0;
        int* lat = nullptr;          // 27 per cell, -1 without a uniform stencil
        nm::LatStride stride{};      // the map as c + off[L] where that holds for every cell
        float* pq = nullptr;         // network parameters (padded FP32)
        uint16_t* wb = nullptr;      // BF16 weights
        unsigned char* learned = nullptr;
        int* lsList = nullptr;       // positions that take the least-squares normal
        int* counts = nullptr;       // [2]: lsList length, of which residual fallbacks
        double* resid = nullptr;
    } learned_;

    void copyIn(double* dst, const double* src, long n, int where) {
        if (!src) throw SvpError(SVP_E_ARG, "null input pointer");
        HIPCHECK(hipMemcpyAsync(dst, src, sizeof(double) * n,
                                where == SVP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                stream_));
        if (where != SVP_DEVICE) HIPCHECK(hipStreamSynchronize(stream_));
    }
    void copyOut(double* dst, const double* src, long n, int where) {
        if (!dst) return;
        HIPCHECK(hipMemcpyAsync(dst, src, sizeof(double) * n,
                                where == SVP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                stream_));
        if (where != SVP_DEVICE) HIPCHECK(hipStreamSynchronize(stream_));
    }
    template <class T>
    void copyOutT(T* dst, const T* src, long n, int where) {
        if (!dst || n <= 0) return;
        HIPCHECK(hipMemcpyAsync(dst, src, sizeof(T) * n,
                                where == SVP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                stream_));
        if (where != SVP_DEVICE) HIPCHECK(hipStreamSynchronize(stream_));
    }
    void checkFinite(const double* x, long n, const char* what) {
        HIPCHECK(hipMemsetAsync(nonfinite_, 0, sizeof(int), stream_));
        launch_count_nonfinite(x, n, nonfinite_, stream_);
        int bad = 0;
        HIPCHECK(hipMemcpyAsync(&bad, nonfinite_, sizeof(int), hipMemcpyDeviceToHost, stream_));
        HIPCHECK(hipStreamSynchronize(stream_));
        if (bad) throw SvpError(SVP_E_NAN, std::string(what) + ": " + std::to_string(bad) + " non-finite values");
    }
    void ensureHostGeometry() {
        if (!hostC_.empty()) return;
        hostC_.resize(3L * m_.nC);
        hostCf_.resize(3L * m_.nF);
        hostSf_.resize(3L * m_.nF);
        HIPCHECK(hipMemcpy(hostC_.data(), m_.C, sizeof(double) * hostC_.size(), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(hostCf_.data(), m_.Cf, sizeof(double) * hostCf_.size(), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(hostSf_.data(), m_.Sf, sizeof(double) * hostSf_.size(), hipMemcpyDeviceToHost));
    }
    double* allocHarness(DBuf& b, long n) { return b.alloc<double>(n); }
    DBuf harnessBufs_[3];
    DBuf rhoPhiSum_;
    DBuf plicIdx_, plicOut_, mapBuf_, plicFields_;

private:
    std::vector<std::unique_ptr<DBuf>> bufs_;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pendingRecon_, pendingAdvect_;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events_;
    std::vector<hipEvent_t> freeEvents_;
    double reconTime_ = 0.0, advectTime_ = 0.0, mapTime_ = 0.0;
    bool reconstructed_ = false;
    bool gaussNormal_ = false;  // orientationMethod alphaGrad
    bool pointLinear_ = false;  // alphaGrad with grad(alpha1) Gauss pointLinear
    bool rdfNormal_ = false;    // orientationMethod isoRDF / RDF
    RdfBufs rdf_{};
    bool alphaPhiValid_ = false;
    int maxCellFaces_ = 0;
    long cellFaceEntries_ = 0;  // size of the cell -> face CSR
    int bndPassSeq_ = 0;        // bounding passes launched so far (candidate-list half)
    int profMask_ = 0;
    double profSec_[N_STAGES] = {};
    int64_t profCnt_[N_STAGES] = {};
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> profPending_;

    template <class F>
    void stage(int sg, F&& launch) {
        if (!(profMask_ & (1 << sg))) {
            launch();
            return;
        }
        hipEvent_t a = event(), b = event();
        HIPCHECK(hipEventRecord(a, stream_));
        launch();
        HIPCHECK(hipEventRecord(b, stream_));
        profPending_.push_back({sg, {a, b}});
    }
    void drainProfile() {
        if (profPending_.empty()) return;
        HIPCHECK(hipStreamSynchronize(stream_));
        for (auto& e : profPending_) {
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, e.second.first, e.second.second));
            profSec_[e.first] += 1e-3 * ms;
            profCnt_[e.first] += 1;
            freeEvents_.push_back(e.second.first);
            freeEvents_.push_back(e.second.second);
        }
        profPending_.clear();
    }

    template <class T>
    T* dalloc(size_t n) {
        bufs_.emplace_back(new DBuf());
        return bufs_.back()->alloc<T>(n);
    }
    template <class T>
    T* dupload(const std::vector<T>& v) {
        bufs_.emplace_back(new DBuf());
        return bufs_.back()->upload<T>(v);
    }
    hipEvent_t event() {
        if (!freeEvents_.empty()) {
            hipEvent_t e = freeEvents_.back();
            freeEvents_.pop_back();
            return e;
        }
        // timing only (read after a stream synchronize): without the
        // system-scope fence a record does not write back and invalidate the
        // L2 between the kernels of a step
        hipEvent_t e;
        HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        events_.push_back({e, e});
        return e;
    }

    // Host-built addressing of the raw polyMesh (primitiveMesh::cells(),
    // cell-point-cell stencil, volPointInterpolation sources).
    void buildMesh(const svp_mesh& M) {
        const int nP = M.n_points, nF = M.n_faces, nIF = M.n_internal_faces, nC = M.n_cells;
        if (nP <= 0 || nF <= 0 || nC <= 0 || nIF < 0 || nIF > nF || !M.points || !M.face_offsets ||
            !M.face_points || !M.owner || (nIF > 0 && !M.neighbour))
            throw SvpError(SVP_E_ARG, "svp_create: incomplete mesh");
        nB_ = nF - nIF;
        m_.nP = nP; m_.nF = nF; m_.nIF = nIF; m_.nC = nC;
        std::vector<int> fOff(M.face_offsets, M.face_offsets + nF + 1);
        if (fOff[0] != 0) throw SvpError(SVP_E_ARG, "face_offsets[0] != 0");
        for (int f = 0; f < nF; ++f) {
            const int n = fOff[f + 1] - fOff[f];
            if (n < 3 || n > MAX_FACE_POINTS)
                throw SvpError(SVP_E_MESH, "face " + std::to_string(f) + " has " + std::to_string(n) +
                                               " points (supported 3.." + std::to_string(MAX_FACE_POINTS) + ")");
        }
        std::vector<int> fPts(M.face_points, M.face_points + fOff[nF]);
        for (int p : fPts)
            if (p < 0 || p >= nP) throw SvpError(SVP_E_ARG, "face point label out of range");
        std::vector<int> own(M.owner, M.owner + nF), nbr(nF, -1);
        for (int f = 0; f < nIF; ++f) nbr[f] = M.neighbour[f];
        for (int f = 0; f < nF; ++f) {
            if (own[f] < 0 || own[f] >= nC) throw SvpError(SVP_E_ARG, "owner out of range");
            if (f < nIF && (nbr[f] < 0 || nbr[f] >= nC || nbr[f] <= own[f]))
                throw SvpError(SVP_E_ARG, "neighbour must exceed owner on every internal face");
        }
        std::vector<unsigned char> fEmpty(nF, 0);
        for (int p = 0; p < M.n_patches; ++p) {
            const int s = M.patch_start[p], n = M.patch_size[p];
            if (s < nIF || s + n > nF) throw SvpError(SVP_E_ARG, "patch range outside boundary faces");
            if (M.patch_type && M.patch_type[p] == SVP_PATCH_EMPTY)
                for (int k = 0; k < n; ++k) fEmpty[s + k] = 1;
        }
        // cells(): owned faces ascending then neighbour faces ascending
        std::vector<int> cnt(nC + 1, 0);
        for (int f = 0; f < nF; ++f) cnt[own[f] + 1]++;
        for (int f = 0; f < nIF; ++f) cnt[nbr[f] + 1]++;
        std::vector<int> cOff(nC + 1, 0);
        for (int c = 0; c < nC; ++c) cOff[c + 1] = cOff[c] + cnt[c + 1];
        cellFaceEntries_ = cOff[nC];
        std::vector<int> fill(cOff.begin(), cOff.end() - 1), cFaces(cOff[nC]), cFacesAsc(cOff[nC]);
        for (int f = 0; f < nF; ++f) cFaces[fill[own[f]]++] = f;
        for (int f = 0; f < nIF; ++f) cFaces[fill[nbr[f]]++] = f;
        int maxCellFaces = 0;
        for (int c = 0; c < nC; ++c) {
            const int b = cOff[c], e = cOff[c + 1];
            maxCellFaces = std::max(maxCellFaces, e - b);
            std::vector<int> srt(cFaces.begin() + b, cFaces.begin() + e);
            std::sort(srt.begin(), srt.end());
            for (int k = 0; k < e - b; ++k) cFacesAsc[b + k] = (own[srt[k]] == c) ? srt[k] : ~srt[k];
        }
        maxCellFaces_ = maxCellFaces;
        const int faceLimit = dict_.split_warped_face ? MAX_CELL_FACES : GEN_CELL_FACES;
        if (maxCellFaces > faceLimit)
            throw SvpError(SVP_E_MESH, "a cell has " + std::to_string(maxCellFaces) + " faces (supported " +
                                           std::to_string(faceLimit) + ")");
        // point -> cells (ascending) and cell points
        std::vector<std::vector<int>> cellPts(nC);
        std::vector<int> pcCnt(nP + 1, 0);
        int maxCellPts = 0;
        for (int c = 0; c < nC; ++c) {
            auto& v = cellPts[c];
            for (int k = cOff[c]; k < cOff[c + 1]; ++k)
                for (int q = fOff[cFaces[k]]; q < fOff[cFaces[k] + 1]; ++q) v.push_back(fPts[q]);
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            maxCellPts = std::max(maxCellPts, (int)v.size());
            for (int p : v) pcCnt[p + 1]++;
        }
        if (maxCellPts > (dict_.split_warped_face ? MAX_CELL_POINTS - 1 : GEN_CELL_POINTS))
            throw SvpError(SVP_E_MESH, "a cell has " + std::to_string(maxCellPts) + " points");
        std::vector<int> pcOff(nP + 1, 0);
        for (int p = 0; p < nP; ++p) pcOff[p + 1] = pcOff[p] + pcCnt[p + 1];
        std::vector<int> pcFill(pcOff.begin(), pcOff.end() - 1), pCells(pcOff[nP]);
        for (int c = 0; c < nC; ++c)
            for (int p : cellPts[c]) pCells[pcFill[p]++] = c;
        // cell-point-cell stencil: cell first, then the others ascending
        std::vector<int> stOff(nC + 1, 0), st;
        st.reserve((size_t)nC * 27);
        std::vector<int> tmp;
        for (int c = 0; c < nC; ++c) {
            tmp.clear();
            for (int p : cellPts[c])
                for (int k = pcOff[p]; k < pcOff[p + 1]; ++k)
                    if (pCells[k] != c) tmp.push_back(pCells[k]);
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            st.push_back(c);
            st.insert(st.end(), tmp.begin(), tmp.end());
            stOff[c + 1] = (int)st.size();
            m_.maxStencil = std::max(m_.maxStencil, stOff[c + 1] - stOff[c]);
        }
        std::vector<int> cpOff(nC + 1, 0), cPtsFlat;
        cPtsFlat.reserve((size_t)nC * 8);
        for (int c = 0; c < nC; ++c) {
            cPtsFlat.insert(cPtsFlat.end(), cellPts[c].begin(), cellPts[c].end());
            cpOff[c + 1] = (int)cPtsFlat.size();
        }
        cellPts.clear();
        cellPts.shrink_to_fit();
        m_.cpOff = dupload(cpOff);
        m_.cPts = dupload(cPtsFlat);
        m_.maxCellPoints = maxCellPts;
        m_.triQuadFaces = 1;
        for (int f = 0; f < nF; ++f)
            if (fOff[f + 1] - fOff[f] > 4) m_.triQuadFaces = 0;
        // point interpolation sources: boundary points -> their non-empty
        // boundary faces (ascending); other points -> their cells (ascending)
        std::vector<std::vector<int>> bFaces(nP);
        std::vector<unsigned char> isPatch(nP, 0);
        for (int f = nIF; f < nF; ++f) {
            if (fEmpty[f]) continue;
            for (int q = fOff[f]; q < fOff[f + 1]; ++q) {
                isPatch[fPts[q]] = 1;
                bFaces[fPts[q]].push_back(f);
            }
        }
        std::vector<int> psOff(nP + 1, 0), psSrc;
        psSrc.reserve(pCells.size());
        for (int p = 0; p < nP; ++p) {
            if (isPatch[p]) {
                for (int f : bFaces[p]) psSrc.push_back(~f);
            } else {
                for (int k = pcOff[p]; k < pcOff[p + 1]; ++k) psSrc.push_back(pCells[k]);
            }
            psOff[p + 1] = (int)psSrc.size();
        }
        std::vector<double> pts(M.points, M.points + 3L * nP);
        m_.pts = dupload(pts);
        m_.fOff = dupload(fOff);
        m_.fPts = dupload(fPts);
        m_.own = dupload(own);
        m_.nbr = dupload(nbr);
        m_.fEmpty = dupload(fEmpty);
        m_.cOff = dupload(cOff);
        m_.cFaces = dupload(cFaces);
        m_.cFacesAsc = dupload(cFacesAsc);
        m_.stOff = dupload(stOff);
        m_.st = dupload(st);
        m_.psOff = dupload(psOff);
        m_.psSrc = dupload(psSrc);
        m_.Cf = dalloc<double>(3L * nF);
        m_.Sf = dalloc<double>(3L * nF);
        m_.magSf = dalloc<double>(nF);
        m_.flat = dalloc<double>(nF);
        m_.fullCS = dalloc<double>(6L * nF);
        m_.C = dalloc<double>(3L * nC);
        m_.V = dalloc<double>(nC);
        m_.psW = dalloc<double>(psSrc.size());
        m_.linW = dalloc<double>(std::max(nIF, 1));
    }

    void allocState() {
        const int nC = m_.nC, nF = m_.nF;
        st_.alpha = dalloc<double>(nC);
        st_.alphaOld = dalloc<double>(nC);
        st_.phi = dalloc<double>(nF);
        st_.U = dalloc<double>(3L * nC);
        st_.Ub = dalloc<double>(3L * std::max(nB_, 1));
        st_.dVf = dalloc<double>(nF);
        st_.alphaPhi = dalloc<double>(nF);
        // bounding arrays double-buffered by pass parity: corrections, active
        // flags, writer indices and correctedFaces keys
        st_.corr = dalloc<double>(2L * nF);
        st_.nbTag = dalloc<int>(nC);
        st_.nMixed = dalloc<int>(1);
        st_.mixed = dalloc<int>(nC);
        st_.status = dalloc<int>(nC);
        st_.nCutPts = dalloc<int>(nC);
        st_.faceClass = dalloc<unsigned long long>(nC);
        st_.ifN = dalloc<double>(3L * nC);
        st_.ifD = dalloc<double>(nC);
        st_.ifC = dalloc<double>(3L * nC);
        st_.ifS = dalloc<double>(3L * nC);
        st_.un0 = dalloc<double>(nC);
        st_.oobCount = dalloc<int>(OOB_SLOTS * OOB_STRIDE);
        st_.fluxList = dalloc<int>(cellFaceEntries_);
        st_.fluxCount = dalloc<int>(1);
        st_.candCount = dalloc<int>(1);
        st_.candBits = dalloc<unsigned long long>((nC + 63) / 64);
        st_.cand = dalloc<int>(2L * nC);
        st_.active = dalloc<unsigned char>(2L * nC);
        st_.pend = dalloc<int>(nC);
        st_.queue = dalloc<int>(2L * nC);
        st_.actIdx = dalloc<int>(2L * nC);
        HIPCHECK(hipMemsetAsync(st_.actIdx, 0xFF, sizeof(int) * 2 * nC, stream_));
        // correctedFaces order keys are cell * (max faces) + rank in 31 bits
        if ((long long)nC * bound_faces(maxCellFaces_) >= 0x7FFFFFFFll)
            throw SvpError(SVP_E_UNSUPPORTED, "mesh too large for 31-bit bounding order keys");
        st_.corrPos = dalloc<int>(2L * nF);
        HIPCHECK(hipMemsetAsync(st_.corrPos, 0xFF, sizeof(int) * 2 * nF, stream_));
        st_.touched = dalloc<int>(nC);
        st_.touchedFlag = dalloc<int>(nC);
        st_.boundStats = dalloc<int>(4 * MAX_BOUND_PASSES);
        st_.boundTotals = dalloc<unsigned long long>(5);
        st_.bcount = dalloc<int>(32);
        // one word per candidate-list block (blocks of at least 64 bitmap words)
        st_.bndLook = dalloc<unsigned long long>((nC + 63) / 64 / 64 + 1);
        st_.detLook = dalloc<unsigned long long>(detect_blocks(nC));
        st_.detTicket = dalloc<unsigned long long>(1);
        st_.bndTicket = dalloc<unsigned long long>(1);
        HIPCHECK(hipMemsetAsync(st_.bcount, 0, 32 * sizeof(int), stream_));
        st_.maxCellFaces = maxCellFaces_;
        // bounding rounds run grid-wide before the one-workgroup tail (tuning
        // knob SVP_BND_GRID_ROUNDS, 0..MAX_GRID_ROUNDS; results do not depend on it)
        st_.gridRounds = 4;
        if (const char* e = std::getenv("SVP_BND_GRID_ROUNDS")) {
            const int r = std::atoi(e);
            if (r >= 0 && r <= MAX_GRID_ROUNDS) st_.gridRounds = r;
        }
        // bitmap words per candidate-list block (tuning knob SVP_BND_LIST_WORDS;
        // results do not depend on it)
        st_.bndListWords = 256;
        if (const char* e = std::getenv("SVP_BND_LIST_WORDS")) {
            const int w = std::atoi(e);
            if (w >= 64 && w <= 4096) st_.bndListWords = w;
        }
        st_.bndGrid = 128;
        if (const char* e = std::getenv("SVP_BND_GRID")) {
            const int g = std::atoi(e);
            if (g >= 1 && g <= 4096) st_.bndGrid = g;
        }
        // dependent part of a bounding pass: one dataflow launch of this many
        // workgroups (knob SVP_BND_FLOW, 0 = grid rounds and the one-workgroup
        // tail; results do not depend on it)
        st_.bndFlow = 128;
        if (const char* e = std::getenv("SVP_BND_FLOW")) {
            const int g = std::atoi(e);
            if (g >= 0 && g <= 4096) st_.bndFlow = g;
        }
        st_.bndRec = (void*)dalloc<unsigned char>(bound_record_bytes(maxCellFaces_) * (size_t)nC);
        st_.un0Fused = (!dict_.split_warped_face && maxCellFaces_ <= 8 && m_.maxCellPoints <= NS_CELL_POINTS &&
                        m_.triQuadFaces) ? 1 : 0;
        st_.fsdKeep = rdfNormal_ ? 1 : 0;
        st_.minmax = dalloc<double>(4);
        st_.errFlag = dalloc<int>(1);
        st_.nIcTotal = dalloc<unsigned long long>(1);
        courant_ = dalloc<unsigned long long>(2);
        nonfinite_ = dalloc<int>(1);
    }
};

}  // namespace SimPLIC
}  // namespace gve

struct svp_handle {
    std::unique_ptr<gve::SimPLIC::solveVofEqu> s;
};

using gve::SimPLIC::solveVofEqu;
using namespace svp;

// A batch of handles advanced together: every stage of a step is one launch
// over all cases (blockIdx.y = case), each case running the single-case
// kernel body on its own mesh and state.  While the batch exists its cases
// work on the batch's stream, so calls on a case handle stay ordered with the
// batch's steps.
class caseBatch {
public:
    caseBatch(svp_handle* const* hs, int n) {
        if (n < 1 || n > 65535 || !hs) throw SvpError(SVP_E_ARG, "svp_batch_create: 1..65535 cases");
        for (int k = 0; k < n; ++k) {
            if (!hs[k] || !hs[k]->s) throw SvpError(SVP_E_ARG, "svp_batch_create: null handle");
            for (int j = 0; j < k; ++j)
                if (hs[j] == hs[k]) throw SvpError(SVP_E_ARG, "svp_batch_create: a handle is given twice");
            c_.push_back(hs[k]->s.get());
        }
        solveVofEqu& f = *c_[0];
        device_ = f.device_;
        HIPCHECK(hipSetDevice(device_));
        params_ = f.params_;
        G_ = f.batchGroup();
        MF_ = bound_faces(f.st_.maxCellFaces);
        triQuad_ = f.m_.triQuadFaces;
        for (solveVofEqu* h : c_) {
            std::string why;
            if (h->device_ != device_) throw SvpError(SVP_E_ARG, "svp_batch_create: cases on different devices");
            if (!h->batchable(why)) throw SvpError(SVP_E_UNSUPPORTED, "svp_batch_create: " + why);
            const Params& q = h->params_;
            if (q.mixedCellTol != params_.mixedCellTol || q.nAlphaBounds != params_.nAlphaBounds ||
                q.snapTol != params_.snapTol || q.clip != params_.clip)
                throw SvpError(SVP_E_ARG, "svp_batch_create: the cases' alpha dictionaries differ");
            if (h->batchGroup() != G_ || bound_faces(h->st_.maxCellFaces) != MF_ ||
                h->m_.triQuadFaces != triQuad_)
                throw SvpError(SVP_E_ARG, "svp_batch_create: the cases' cell types differ (hexahedra and "
                                          "tetrahedra go to separate batches)");
        }
        HIPCHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        // the cases' earlier work is done; from now on they use the batch's stream
        std::vector<BatchCase> host(n);
        for (int k = 0; k < n; ++k) {
            solveVofEqu& h = *c_[k];
            HIPCHECK(hipStreamSynchronize(h.stream_));
            ownStream_.push_back(h.stream_);
            h.stream_ = stream_;
            BatchCase& b = host[k];
            std::memset(&b, 0, sizeof b);
            b.m = h.m_;
            const int nC = h.m_.nC, S = h.passSeq();
            for (int a = 0; a < 2; ++a)
                for (int half = 0; half < 2; ++half) {
                    DevState v = h.st_;
                    if (a) std::swap(v.alpha, v.alphaOld);
                    v = bound_pass_view(h.m_, v, S + half);
                    b.st[(a << 1) | half] = v;
                }
            b.detBlocks = detect_list_blocks(nC);
            b.listWords = (nC + 63) / 64;
            b.listBlocks = (b.listWords + h.st_.bndListWords - 1) / h.st_.bndListWords;
            b.nB = h.m_.nF - h.m_.nIF;
            b.U0 = h.U0_;
            b.phi0 = h.phi0_;
            b.U0b = h.U0b_;
            maxC_ = std::max(maxC_, nC);
            maxF_ = std::max(maxF_, h.m_.nF);
            maxDet_ = std::max(maxDet_, b.detBlocks);
            maxList_ = std::max(maxList_, b.listBlocks);
            maxN_ = std::max(maxN_, std::max<long>(h.m_.nF, 3L * std::max(nC, b.nB)));
            haveFlow_ = haveFlow_ && h.U0_ != nullptr;
        }
        bc_.upload(host);
        info_.alloc<BatchStepInfo>(n);
        factors_.alloc<double>(n);
        courant_.alloc<unsigned long long>(2L * n);
        for (int r = 0; r < RING; ++r) {
            HIPCHECK(hipHostMalloc((void**)&ring_[r], sizeof(BatchStepInfo) * n + sizeof(double) * n, 0));
            HIPCHECK(hipEventCreateWithFlags(&ringEv_[r], hipEventDisableTiming));
            HIPCHECK(hipEventRecord(ringEv_[r], stream_));
        }
        aPar_.assign(n, 0);
        hPar_.assign(n, 0);
        // launch shapes of the grid-stride stages (x per case; svp_batch_tune)
        dims_.ls = 32;
        dims_.fsd = 64;
        dims_.fluxList = 8;
        dims_.fluxFaces = 64;
        dims_.prep = 8;
        dims_.flow = 4;
        dims_.apply = 8;
        if (const char* e = std::getenv("SVP_BATCH_GRAPH")) useGraph_ = std::atoi(e) != 0;
        HIPCHECK(hipStreamSynchronize(stream_));
    }

    ~caseBatch() {
        (void)hipSetDevice(device_);
        (void)hipStreamSynchronize(stream_);
        for (size_t k = 0; k < c_.size(); ++k) c_[k]->stream_ = ownStream_[k];
        dropGraph();
        for (hipEvent_t e : allEvents_) (void)hipEventDestroy(e);
        for (int r = 0; r < RING; ++r) {
            if (ring_[r]) (void)hipHostFree(ring_[r]);
            if (ringEv_[r]) (void)hipEventDestroy(ringEv_[r]);
        }
        (void)hipStreamDestroy(stream_);
    }

    int size() const { return (int)c_.size(); }

    // reconstruct + advect of every case with dt[k] > 0 (dt <= 0: the case is
    // left as it is), one launch per stage for all of them
    void step(const double* dt) {
        if (!dt) throw SvpError(SVP_E_ARG, "svp_batch_step: null dt");
        const int n = size();
        for (int k = 0; k < n; ++k)
            if (!std::isfinite(dt[k])) throw SvpError(SVP_E_ARG, "svp_batch_step: non-finite dt");
        upload(dt, nullptr);
        hipEvent_t a = event(), b = event();
        HIPCHECK(hipEventRecord(a, stream_));
        if (profMask_ || !useGraph_) {
            for (int sg = STAGE_DETECT; sg <= STAGE_SNAPCLIP; ++sg)
                profiled(sg, [&] { launchStage(sg); });
        } else {
            // the step's launches do not depend on dt or the buffer
            // parities (both are read from info_ on the device), so one
            // captured graph replays every step
            if (!graph_) {
                hipGraph_t g = nullptr;
                HIPCHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
                for (int sg = STAGE_DETECT; sg <= STAGE_SNAPCLIP; ++sg) launchStage(sg);
                const hipError_t launched = hipGetLastError();
                HIPCHECK(hipStreamEndCapture(stream_, &g));
                HIPCHECK(launched);
                const hipError_t made = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIPCHECK(made);
            }
            HIPCHECK(hipGraphLaunch(graph_, stream_));
        }
        HIPCHECK(hipEventRecord(b, stream_));
        HIPCHECK(hipGetLastError());
        pending_.push_back({a, b});
        const int nPasses = std::min(params_.nAlphaBounds, MAX_BOUND_PASSES);
        for (int k = 0; k < n; ++k) {
            if (!(dt[k] > 0.0)) continue;
            aPar_[k] ^= 1;
            hPar_[k] ^= (nPasses & 1);
            c_[k]->afterBatchStep(dt[k], nPasses);
        }
        steps_++;
    }

    void launchStage(int sg) {
        launch_batch_stage(sg, (const BatchCase*)bc_.p, (const BatchStepInfo*)info_.p, size(), maxC_, maxF_,
                           maxDet_, maxList_, dims_, params_, G_, MF_, triQuad_, stream_);
    }

    void dropGraph() {
        if (graph_) (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }

    // harness: max over each case's cells of sum|phi|/V and of the same for
    // 0.01 <= alpha <= 0.99 (svp_harness_courant per case)
    void courant(double* co, double* aco) {
        const int n = size();
        upload(nullptr, nullptr);
        launch_batch_courant((const BatchCase*)bc_.p, (const BatchStepInfo*)info_.p, n, maxC_,
                             (unsigned long long*)courant_.p, stream_);
        HIPCHECK(hipGetLastError());
        std::vector<unsigned long long> r(2L * n);
        HIPCHECK(hipMemcpyAsync(r.data(), courant_.p, sizeof(unsigned long long) * r.size(), hipMemcpyDeviceToHost,
                                stream_));
        HIPCHECK(hipStreamSynchronize(stream_));
        for (int k = 0; k < n; ++k) {
            double v[2];
            std::memcpy(v, &r[2L * k], sizeof v);
            if (co) co[k] = v[0];
            if (aco) aco[k] = v[1];
        }
    }

    // harness: each case's base flow times its own factor
    void scaleFlow(const double* factors) {
        if (!factors) throw SvpError(SVP_E_ARG, "svp_batch_scale_flow: null factors");
        if (!haveFlow_) throw SvpError(SVP_E_ARG, "svp_batch_scale_flow: a case has no base flow uploaded");
        upload(nullptr, factors);
        launch_batch_scale_flow((const BatchCase*)bc_.p, (const BatchStepInfo*)info_.p, (const double*)factors_.p,
                                size(), maxN_, stream_);
        HIPCHECK(hipGetLastError());
    }

    // GPU seconds of the batch steps so far and the summed mixed-cell counts
    // of the cases' reconstructs (their running totals)
    void timers(double& sec, uint64_t& nic) {
        HIPCHECK(hipStreamSynchronize(stream_));
        for (auto& e : pending_) {
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, e.first, e.second));
            stepSec_ += 1e-3 * ms;
            free_.push_back(e.first);
            free_.push_back(e.second);
        }
        pending_.clear();
        sec = stepSec_;
        nic = 0;
        for (solveVofEqu* h : c_) {
            unsigned long long v = 0;
            HIPCHECK(hipMemcpy(&v, h->st_.nIcTotal, sizeof v, hipMemcpyDeviceToHost));
            nic += v;
        }
    }

    // per-stage GPU time (the handle's stage numbering) of the steps since
    // profiling was switched on
    void profileSet(int mask) {
        drainProfile();
        profMask_ = mask;
    }
    void profileGet(double* sec, int64_t* cnt, bool reset) {
        drainProfile();
        for (int k = 0; k < N_STAGES; ++k) {
            if (sec) sec[k] = profSec_[k];
            if (cnt) cnt[k] = profCnt_[k];
        }
        if (reset) {
            std::fill(profSec_, profSec_ + N_STAGES, 0.0);
            std::fill(profCnt_, profCnt_ + N_STAGES, (int64_t)0);
        }
    }

    // Launch-shape tuner: the x extent of each grid-stride stage is tried over
    // a few values on the batch's current step (each trial restarts from the
    // same fields), and the fastest kept.  Results do not depend on these
    // shapes; only the stage times do.
    void tune(const double* dt, int32_t* chosen) {
        if (!dt) throw SvpError(SVP_E_ARG, "svp_batch_tune: null dt");
        dropGraph();
        const int n = size();
        // the fields every trial starts from
        std::vector<DBuf> saved(n);
        for (int k = 0; k < n; ++k) {
            saved[k].alloc<double>(c_[k]->m_.nC);
            HIPCHECK(hipMemcpyAsync(saved[k].p, c_[k]->st_.alpha, sizeof(double) * c_[k]->m_.nC,
                                    hipMemcpyDeviceToDevice, stream_));
        }
        auto restore = [&] {
            for (int k = 0; k < n; ++k)
                HIPCHECK(hipMemcpyAsync(c_[k]->st_.alpha, saved[k].p, sizeof(double) * c_[k]->m_.nC,
                                        hipMemcpyDeviceToDevice, stream_));
        };
        const int oldMask = profMask_;
        auto trial = [&](int stage) {
            profileSet(1 << stage);
            step(dt);
            double sec[N_STAGES];
            profileGet(sec, nullptr, true);
            restore();
            return sec[stage];
        };
        struct Knob { int stage; int* v; std::vector<int> cand; };
        std::vector<Knob> knobs = {{STAGE_NORMAL, &dims_.ls, {8, 16, 32, 64, 128}},
                                   {STAGE_FSD, &dims_.fsd, {16, 32, 64, 128, 256}},
                                   {STAGE_IFLUX, &dims_.fluxFaces, {16, 32, 64, 128, 256}},
                                   {STAGE_IFLUX, &dims_.fluxList, {2, 4, 8, 16, 32}},
                                   {STAGE_BOUND, &dims_.flow, {1, 2, 4, 8}},
                                   {STAGE_BOUND, &dims_.prep, {2, 4, 8, 16, 32}},
                                   {STAGE_BOUND, &dims_.apply, {2, 4, 8, 16, 32}}};
        trial(STAGE_DETECT);  // warm-up
        for (Knob& kb : knobs) {
            int best = *kb.v;
            double bestT = 1e30;
            for (int v : kb.cand) {
                *kb.v = v;
                double t = 0.0;
                for (int r = 0; r < 2; ++r) t += trial(kb.stage);
                if (t < bestT) {
                    bestT = t;
                    best = v;
                }
            }
            *kb.v = best;
        }
        profileSet(oldMask);
        HIPCHECK(hipStreamSynchronize(stream_));
        if (chosen) {
            chosen[0] = dims_.ls;
            chosen[1] = dims_.fsd;
            chosen[2] = dims_.fluxList;
            chosen[3] = dims_.fluxFaces;
            chosen[4] = dims_.prep;
            chosen[5] = dims_.flow;
            chosen[6] = dims_.apply;
        }
    }

    void setDims(const int32_t* d) {
        if (!d) throw SvpError(SVP_E_ARG, "svp_batch_set_shapes: null");
        for (int k = 0; k < 7; ++k)
            if (d[k] < 1 || d[k] > 65535) throw SvpError(SVP_E_ARG, "svp_batch_set_shapes: 1..65535 per stage");
        dropGraph();
        dims_.ls = d[0];
        dims_.fsd = d[1];
        dims_.fluxList = d[2];
        dims_.fluxFaces = d[3];
        dims_.prep = d[4];
        dims_.flow = d[5];
        dims_.apply = d[6];
    }

    int device_ = 0;

private:
    static constexpr int RING = 8;
    std::vector<solveVofEqu*> c_;
    std::vector<hipStream_t> ownStream_;
    hipStream_t stream_ = nullptr;
    Params params_{};
    int G_ = 6, MF_ = NS_CELL_FACES, triQuad_ = 1;
    int maxC_ = 0, maxF_ = 0, maxDet_ = 0, maxList_ = 0;
    long maxN_ = 0;
    bool haveFlow_ = true;
    BatchDims dims_{};
    // the step's launches as one graph (SVP_BATCH_GRAPH=0: launched one by
    // one); rebuilt after the launch shapes change
    bool useGraph_ = true;
    hipGraphExec_t graph_ = nullptr;
    DBuf bc_, info_, factors_, courant_;
    char* ring_[RING] = {};
    hipEvent_t ringEv_[RING] = {};
    int ringPos_ = 0;
    std::vector<int> aPar_, hPar_;
    long steps_ = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending_;
    std::vector<hipEvent_t> free_, allEvents_;
    double stepSec_ = 0.0;
    int profMask_ = 0;
    double profSec_[N_STAGES] = {};
    int64_t profCnt_[N_STAGES] = {};
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> profPending_;

    // the per-case step records (dt, buffer parities) and, when given, the
    // flow factors, through a ring of pinned buffers: a buffer is refilled
    // only after the copy that last read it has run
    void upload(const double* dt, const double* factors) {
        const int n = size();
        const int r = ringPos_;
        ringPos_ = (ringPos_ + 1) % RING;
        HIPCHECK(hipEventSynchronize(ringEv_[r]));
        BatchStepInfo* inf = (BatchStepInfo*)ring_[r];
        double* fac = (double*)(ring_[r] + sizeof(BatchStepInfo) * n);
        for (int k = 0; k < n; ++k) {
            inf[k].dt = dt ? (dt[k] > 0.0 ? dt[k] : 0.0) : 0.0;
            inf[k].aPar = aPar_[k];
            inf[k].hPar = hPar_[k];
            inf[k].epoch = c_[k]->nextEpoch();
            if (factors) fac[k] = factors[k];
        }
        HIPCHECK(hipMemcpyAsync(info_.p, inf, sizeof(BatchStepInfo) * n, hipMemcpyHostToDevice, stream_));
        if (factors) HIPCHECK(hipMemcpyAsync(factors_.p, fac, sizeof(double) * n, hipMemcpyHostToDevice, stream_));
        HIPCHECK(hipEventRecord(ringEv_[r], stream_));
    }

    hipEvent_t event() {
        if (!free_.empty()) {
            hipEvent_t e = free_.back();
            free_.pop_back();
            return e;
        }
        hipEvent_t e;  // timing only, as the handle's step events
        HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        allEvents_.push_back(e);
        return e;
    }

    template <class F>
    void profiled(int sg, F&& launch) {
        if (!(profMask_ & (1 << sg))) {
            launch();
            return;
        }
        hipEvent_t a = event(), b = event();
        HIPCHECK(hipEventRecord(a, stream_));
        launch();
        HIPCHECK(hipEventRecord(b, stream_));
        profPending_.push_back({sg, {a, b}});
    }

    void drainProfile() {
        if (profPending_.empty()) return;
        HIPCHECK(hipStreamSynchronize(stream_));
        for (auto& e : profPending_) {
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, e.second.first, e.second.second));
            profSec_[e.first] += 1e-3 * ms;
            profCnt_[e.first] += 1;
            free_.push_back(e.second.first);
            free_.push_back(e.second.second);
        }
        profPending_.clear();
    }
};

struct svp_batch {
    std::unique_ptr<caseBatch> b;
};

static caseBatch& B(svp_batch* b) {
    if (!b || !b->b) throw SvpError(SVP_E_ARG, "null batch");
    HIPCHECK(hipSetDevice(b->b->device_));
    return *b->b;
}

static solveVofEqu& H(svp_handle* h) {
    if (!h || !h->s) throw SvpError(SVP_E_ARG, "null handle");
    HIPCHECK(hipSetDevice(h->s->device_));
    return *h->s;
}

extern "C" {

void svp_dict_defaults(svp_dict* d) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    std::snprintf(d->orientation_method, sizeof d->orientation_method, "isoAlphaGrad");
    d->mixed_cell_tol = 1e-8;
    d->n_alpha_bounds = 10;
    d->snap_tol = 0.0;
    d->clip = 1;
    d->split_warped_face = 0;
    d->write_plic_fields = 0;
    d->learned_residual_tol = 0.25;
    d->rdf_tol = 1e-6;
    d->rdf_rel_tol = 0.1;
    d->rdf_iterations = 5;
    std::snprintf(d->grad_scheme, sizeof d->grad_scheme, "Gauss linear");
}


// OpenFOAM dictionary entries "key value;" with comments; sub-dictionaries
// are skipped.
int svp_dict_parse(const char* text, svp_dict* out) {
    return guarded([&] {
        if (!text || !out) throw SvpError(SVP_E_ARG, "svp_dict_parse: null argument");
        svp_dict d;
        svp_dict_defaults(&d);
        std::string s(text), clean;
        for (size_t i = 0; i < s.size(); ++i) {
            if (s[i] == '/' && i + 1 < s.size() && s[i + 1] == '/') {
                while (i < s.size() && s[i] != '\n') ++i;
                clean += '\n';
            } else if (s[i] == '/' && i + 1 < s.size() && s[i + 1] == '*') {
                i += 2;
                while (i + 1 < s.size() && !(s[i] == '*' && s[i + 1] == '/')) ++i;
                ++i;
                clean += ' ';
            } else {
                clean += s[i];
            }
        }
        std::string method;
        size_t pos = 0;
        int depth = 0;
        while (pos < clean.size()) {
            const size_t semi = clean.find_first_of(";{}", pos);
            if (semi == std::string::npos) break;
            const char delim = clean[semi];
            std::string stmt = clean.substr(pos, semi - pos);
            pos = semi + 1;
            if (delim == '{') { depth++; continue; }
            if (delim == '}') { depth--; continue; }
            if (depth != 0) continue;
            size_t a = stmt.find_first_not_of(" \t\r\n");
            if (a == std::string::npos) continue;
            size_t b = stmt.find_first_of(" \t\r\n", a);
            if (b == std::string::npos) continue;
            std::string key = stmt.substr(a, b - a);
            size_t c = stmt.find_first_not_of(" \t\r\n", b);
            size_t e = stmt.find_last_not_of(" \t\r\n");
            if (c == std::string::npos) continue;
            std::string val = stmt.substr(c, e - c + 1);
            auto toBool = [&](const std::string& v) {
                if (v == "true" || v == "on" || v == "yes" || v == "1") return 1;
                if (v == "false" || v == "off" || v == "no" || v == "0" || v == "none") return 0;
                throw SvpError(SVP_E_DICT, "bad bool '" + v + "' for " + key);
            };
            if (key == "orientationMethod" || (key == "reconstructionScheme" && method.empty())) method = val;
            else if (key == "mixedCellTol") d.mixed_cell_tol = std::stod(val);
            else if (key == "nAlphaBounds") d.n_alpha_bounds = std::stoi(val);
            else if (key == "snapTol") d.snap_tol = std::stod(val);
            else if (key == "clip") d.clip = toBool(val);
            else if (key == "splitWarpedFace") d.split_warped_face = toBool(val);
            else if (key == "writePlicFields") d.write_plic_fields = toBool(val);
            else if (key == "learnedResidualTol") d.learned_residual_tol = std::stod(val);
            else if (key == "tol") d.rdf_tol = std::stod(val);
            else if (key == "relTol") d.rdf_rel_tol = std::stod(val);
            else if (key == "iterations") d.rdf_iterations = std::stoi(val);
            else if (key == "grad(alpha1)" || key == "gradScheme") {
                const std::string g = grad_scheme_name(val);
                if (g.empty())
                    throw SvpError(SVP_E_UNSUPPORTED, "grad(alpha1) scheme '" + val +
                                                          "': this library provides Gauss linear and Gauss pointLinear");
                std::snprintf(d.grad_scheme, sizeof d.grad_scheme, "%s", g.c_str());
            }
        }
        if (!method.empty()) {
            if (method == "LS" || method == "isoAlphaGrad" || method == "alphaGrad" || method == "learned" ||
                method == "MLP" || method == "isoRDF" || method == "RDF") {
                std::snprintf(d.orientation_method, sizeof d.orientation_method, "%s",
                              method == "MLP" ? "learned" : method.c_str());
            } else {
                throw SvpError(SVP_E_DICT, "Invalid orientationMethod " + method +
                                               ". Valid methods are: alphaGrad isoAlphaGrad isoRDF LS RDF learned");
            }
        }
        *out = d;
    });
}

int svp_create(const svp_mesh* mesh, const svp_dict* dict, int device, svp_handle** out) {
    return guarded([&] {
        if (!mesh || !out) throw SvpError(SVP_E_ARG, "svp_create: null argument");
        svp_dict d;
        if (dict) d = *dict;
        else svp_dict_defaults(&d);
        const std::string m(d.orientation_method);
        if (m != "LS" && m != "isoAlphaGrad" && m != "alphaGrad" && m != "learned" && m != "isoRDF" && m != "RDF")
            throw SvpError(SVP_E_DICT, "Invalid orientationMethod " + m);
        if (d.n_alpha_bounds < 0 || !(d.mixed_cell_tol >= 0.0) || d.snap_tol < 0.0 || !(d.learned_residual_tol >= 0.0))
            throw SvpError(SVP_E_DICT, "invalid tolerance or bound count");
        if ((m == "isoRDF" || m == "RDF") && (!(d.rdf_tol >= 0.0) || !(d.rdf_rel_tol >= 0.0) || d.rdf_iterations < 0))
            throw SvpError(SVP_E_DICT, "invalid isoRDF tol, relTol or iterations");
        auto h = std::make_unique<svp_handle>();
        h->s = std::make_unique<solveVofEqu>(*mesh, d, device);
        *out = h.release();
    });
}

int svp_destroy(svp_handle* h) {
    return guarded([&] { delete h; });
}

int svp_set_alpha(svp_handle* h, const double* alpha, int where) {
    return guarded([&] { H(h).setAlpha(alpha, where); });
}

int svp_set_flow(svp_handle* h, const double* phi, const double* U, const double* Ub, int where) {
    return guarded([&] { H(h).setFlow(phi, U, Ub, where); });
}

int svp_reconstruct(svp_handle* h) {
    return guarded([&] { H(h).reconstruct(); });
}

int svp_advect(svp_handle* h, double dt, const double* Sp, const double* Su, int where) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        s.stageSources(Sp, Su, where);
        s.advect(dt, Sp, Su);
    });
}

int svp_advect_sub_cycles(svp_handle* h, double dt, int32_t n_sub, const double* Sp, const double* Su,
                          double rho1, double rho2, double* rho_phi, int where) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        s.stageSources(Sp, Su, where);
        const double* sum = s.advectSubCycles(dt, n_sub, Sp, Su, rho1, rho2);
        s.copyOut(rho_phi, sum, s.m_.nF, where);
        if (where != SVP_DEVICE) s.checkDeviceErrors();
    });
}

int svp_get_alpha(svp_handle* h, double* alpha, int where) {
    return guarded([&] { H(h).getAlpha(alpha, where); });
}

int svp_get_alpha_phi(svp_handle* h, double* ap, int where) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        s.copyOut(ap, s.alphaPhi(), s.m_.nF, where);
        if (where != SVP_DEVICE) s.checkDeviceErrors();
    });
}

int svp_get_rho_phi(svp_handle* h, double rho1, double rho2, double* out, int where) {
    return guarded([&] {
        if (!out) throw SvpError(SVP_E_ARG, "null output");
        solveVofEqu& s = H(h);
        if (where == SVP_DEVICE) {
            s.rhoPhi(rho1, rho2, out);
        } else {
            if (!s.rhoPhiSum_.p) s.rhoPhiSum_.alloc<double>(s.m_.nF);
            s.rhoPhi(rho1, rho2, (double*)s.rhoPhiSum_.p);
            s.copyOut(out, (const double*)s.rhoPhiSum_.p, s.m_.nF, SVP_HOST);
            s.checkDeviceErrors();
        }
    });
}

int svp_map_alpha_field(svp_handle* h, const double* n, const double* d, double lower_refine_level,
                        double upper_refine_level, int where) {
    return guarded([&] { H(h).mapAlphaField(n, d, lower_refine_level, upper_refine_level, where); });
}

int svp_rdf_iterations(svp_handle* h, int32_t* n) {
    return guarded([&] {
        if (!n) throw SvpError(SVP_E_ARG, "null output");
        *n = H(h).rdfIterationsRun();
    });
}

int svp_get_plic_surface(svp_handle* h, int32_t* n_faces, int32_t* n_points, int32_t* face_cells,
                         int32_t* face_offsets, double* points, int where) {
    return guarded([&] { H(h).plicSurface(n_faces, n_points, face_cells, face_offsets, points, where); });
}

int svp_get_plic_surface_device(svp_handle* h, int32_t* n_faces, int32_t* n_points, int32_t* face_cells,
                                int32_t* face_offsets, double* points) {
    return guarded([&] { H(h).plicSurfaceNoAlloc(n_faces, n_points, face_cells, face_offsets, points); });
}

int svp_n_mixed(svp_handle* h, int32_t* n) {
    return guarded([&] {
        if (!n) throw SvpError(SVP_E_ARG, "null output");
        solveVofEqu& s = H(h);
        s.checkDeviceErrors();
        *n = s.nMixed();
    });
}

int svp_get_interface(svp_handle* h, int32_t* cells, int32_t* status, double* n, double* d, double* C,
                      double* S, int where) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        s.checkDeviceErrors();
        const int k = s.nMixed();
        s.copyOutT(cells, s.st_.mixed, k, where);
        s.copyOutT(status, s.st_.status, k, where);
        s.copyOut(n, s.st_.ifN, 3L * k, where);
        s.copyOut(d, s.st_.ifD, k, where);
        s.copyOut(C, s.st_.ifC, 3L * k, where);
        s.copyOut(S, s.st_.ifS, 3L * k, where);
    });
}

int svp_get_cut_class(svp_handle* h, uint64_t* fc, int32_t* ncp, int where) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        const int k = s.nMixed();
        s.copyOutT((unsigned long long*)fc, s.st_.faceClass, k, where);
        s.copyOutT(ncp, s.st_.nCutPts, k, where);
    });
}

int svp_get_un0(svp_handle* h, double* un0, int where) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        s.copyOut(un0, s.st_.un0, s.nMixed(), where);
    });
}

int svp_get_geometry(svp_handle* h, double* Cf, double* Sf, double* magSf, double* flat, double* C, double* V) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        s.copyOut(Cf, s.m_.Cf, 3L * s.m_.nF, SVP_HOST);
        s.copyOut(Sf, s.m_.Sf, 3L * s.m_.nF, SVP_HOST);
        s.copyOut(magSf, s.m_.magSf, s.m_.nF, SVP_HOST);
        s.copyOut(flat, s.m_.flat, s.m_.nF, SVP_HOST);
        s.copyOut(C, s.m_.C, 3L * s.m_.nC, SVP_HOST);
        s.copyOut(V, s.m_.V, s.m_.nC, SVP_HOST);
    });
}

int svp_alpha_mapping_time(svp_handle* h, double* seconds) {
    return guarded([&] {
        if (!seconds) throw SvpError(SVP_E_ARG, "svp_alpha_mapping_time: null output");
        *seconds = H(h).alphaMappingTime();
    });
}

int svp_get_plic_fields(svp_handle* h, double* interface_n, double* interface_d, double* interface_c,
                        double* interface_s, double* zeta, int where) {
    return guarded([&] { H(h).plicFields(interface_n, interface_d, interface_c, interface_s, zeta, where); });
}

int svp_timers(svp_handle* h, double* rec, double* adv, uint64_t* nic) {
    return guarded([&] {
        double r, a;
        uint64_t n;
        H(h).timers(r, a, n);
        if (rec) *rec = r;
        if (adv) *adv = a;
        if (nic) *nic = n;
    });
}

int svp_device_view_get(svp_handle* h, svp_device_view* out) {
    return guarded([&] {
        if (!out) throw SvpError(SVP_E_ARG, "null output");
        solveVofEqu& s = H(h);
        out->alpha = s.st_.alpha;
        out->phi = s.st_.phi;
        out->U = s.st_.U;
        out->U_boundary = s.st_.Ub;
        out->dVf = s.st_.dVf;
        out->Cf = s.m_.Cf;
        out->Sf = s.m_.Sf;
        out->C = s.m_.C;
        out->V = s.m_.V;
        out->stream = (void*)s.stream_;
    });
}

int svp_bound_totals(svp_handle* h, uint64_t* out5) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        if (!out5) throw SvpError(SVP_E_ARG, "svp_bound_totals: null output");
        s.copyOutT((unsigned long long*)out5, s.st_.boundTotals, 5, SVP_HOST);
    });
}

int svp_bound_stats(svp_handle* h, int32_t* out, int cap) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        std::vector<int> v(4 * svp::MAX_BOUND_PASSES);
        s.copyOutT(v.data(), s.st_.boundStats, (long)v.size(), SVP_HOST);
        for (int i = 0; i < cap && i < (int)v.size(); ++i) out[i] = v[i];
        HIPCHECK(hipMemsetAsync(s.st_.boundStats, 0, sizeof(int) * v.size(), s.stream_));
    });
}

int svp_profile_set(svp_handle* h, int mask) {
    return guarded([&] { H(h).setProfile(mask); });
}

int svp_profile_get(svp_handle* h, double* seconds, int64_t* counts, int reset) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        s.profile(seconds, counts);
        if (reset) s.resetProfile();
    });
}

int svp_synchronize(svp_handle* h) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        HIPCHECK(hipStreamSynchronize(s.stream_));
        s.checkDeviceErrors();
    });
}


int svp_set_normal_model(svp_handle* h, const svp_normal_model* m) {
    return guarded([&] {
        if (!m) throw SvpError(SVP_E_ARG, "svp_set_normal_model: null model");
        H(h).setNormalModel(*m);
    });
}

int svp_learned_stats(svp_handle* h, int64_t* out4) {
    return guarded([&] {
        if (!out4) throw SvpError(SVP_E_ARG, "svp_learned_stats: null output");
        H(h).learnedStats(out4);
    });
}

int svp_learned_residuals(svp_handle* h, double* out, int where) {
    return guarded([&] { H(h).learnedResiduals(out, where); });
}

int svp_test_set_error_bits(svp_handle* h, int32_t bits) {
    return guarded([&] { H(h).setErrorBits(bits); });
}

int svp_batch_create(svp_handle* const* cases, int32_t n_cases, svp_batch** out) {
    return guarded([&] {
        if (!out) throw SvpError(SVP_E_ARG, "svp_batch_create: null output");
        *out = nullptr;
        auto b = std::make_unique<svp_batch>();
        b->b = std::make_unique<caseBatch>(cases, n_cases);
        *out = b.release();
    });
}

int svp_batch_destroy(svp_batch* b) {
    return guarded([&] {
        if (b && b->b) (void)hipSetDevice(b->b->device_);
        delete b;
    });
}

int svp_batch_step(svp_batch* b, const double* dt) {
    return guarded([&] { B(b).step(dt); });
}

int svp_batch_tune(svp_batch* b, const double* dt, int32_t* shapes) {
    return guarded([&] { B(b).tune(dt, shapes); });
}

int svp_batch_set_shapes(svp_batch* b, const int32_t* shapes) {
    return guarded([&] { B(b).setDims(shapes); });
}

int svp_batch_timers(svp_batch* b, double* seconds, uint64_t* n_ic_total) {
    return guarded([&] {
        double s = 0.0;
        uint64_t n = 0;
        B(b).timers(s, n);
        if (seconds) *seconds = s;
        if (n_ic_total) *n_ic_total = n;
    });
}

int svp_batch_profile_set(svp_batch* b, int mask) {
    return guarded([&] { B(b).profileSet(mask); });
}

int svp_batch_profile_get(svp_batch* b, double* seconds, int64_t* counts, int reset) {
    return guarded([&] { B(b).profileGet(seconds, counts, reset != 0); });
}

int svp_batch_courant(svp_batch* b, double* max_sum_phi_by_v, double* max_alpha_sum_phi_by_v) {
    return guarded([&] { B(b).courant(max_sum_phi_by_v, max_alpha_sum_phi_by_v); });
}

int svp_batch_scale_flow(svp_batch* b, const double* factors) {
    return guarded([&] { B(b).scaleFlow(factors); });
}

const char* svp_last_error(void) { return g_err.c_str(); }

// ------------------------------- harness -----------------------------------

int svp_harness_deformation(svp_handle* h, double* U0, double* phi0) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        s.ensureHostGeometry();
        const double PI = M_PI;
        auto U = [&](double x, double y, double z, double* u) {
            u[0] = 2 * ((std::sin(PI * x)) * (std::sin(PI * x))) * std::sin(2 * PI * y) * std::sin(2 * PI * z);
            u[1] = -std::sin(2 * PI * x) * ((std::sin(PI * y)) * (std::sin(PI * y))) * std::sin(2 * PI * z);
            u[2] = -std::sin(2 * PI * x) * std::sin(2 * PI * y) * ((std::sin(PI * z)) * (std::sin(PI * z)));
        };
        for (int c = 0; c < s.m_.nC; ++c) U(s.hostC_[3 * c], s.hostC_[3 * c + 1], s.hostC_[3 * c + 2], U0 + 3 * c);
        for (int f = 0; f < s.m_.nF; ++f) {
            double u[3];
            U(s.hostCf_[3 * f], s.hostCf_[3 * f + 1], s.hostCf_[3 * f + 2], u);
            phi0[f] = u[0] * s.hostSf_[3 * f] + u[1] * s.hostSf_[3 * f + 1] + u[2] * s.hostSf_[3 * f + 2];
        }
    });
}

int svp_harness_rotation(svp_handle* h, const double* w, const double* o, double* U0, double* phi0) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        s.ensureHostGeometry();
        auto U = [&](const double* x, double* u) {
            const double r0 = x[0] - o[0], r1 = x[1] - o[1], r2 = x[2] - o[2];
            u[0] = w[1] * r2 - w[2] * r1;
            u[1] = w[2] * r0 - w[0] * r2;
            u[2] = w[0] * r1 - w[1] * r0;
        };
        for (int c = 0; c < s.m_.nC; ++c) U(&s.hostC_[3 * c], U0 + 3 * c);
        for (int f = 0; f < s.m_.nF; ++f) {
            double u[3];
            U(&s.hostCf_[3 * f], u);
            phi0[f] = u[0] * s.hostSf_[3 * f] + u[1] * s.hostSf_[3 * f + 1] + u[2] * s.hostSf_[3 * f + 2];
        }
    });
}

int svp_harness_upload_base_flow(svp_handle* h, const double* U0, const double* phi0, const double* U0b) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        if (!s.U0_) {
            s.U0_ = s.allocHarness(s.harnessBufs_[0], 3L * s.m_.nC + s.m_.nF + 3L * std::max(s.nB_, 1));
            s.phi0_ = s.U0_ + 3L * s.m_.nC;
            s.U0b_ = s.phi0_ + s.m_.nF;
        }
        s.copyIn(s.U0_, U0, 3L * s.m_.nC, SVP_HOST);
        s.copyIn(s.phi0_, phi0, s.m_.nF, SVP_HOST);
        if (U0b) s.copyIn(s.U0b_, U0b, 3L * s.nB_, SVP_HOST);
        else HIPCHECK(hipMemsetAsync(s.U0b_, 0, sizeof(double) * 3 * std::max(s.nB_, 1), s.stream_));
    });
}

int svp_harness_scale_flow(svp_handle* h, double factor) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        if (!s.U0_) throw SvpError(SVP_E_ARG, "base flow not uploaded");
        svp::launch_scale_flow(s.m_.nC, s.m_.nF, s.nB_, s.U0_, s.phi0_, s.U0b_, factor, s.st_.U, s.st_.phi,
                               s.st_.Ub, s.stream_);
        HIPCHECK(hipGetLastError());
    });
}

int svp_harness_init_alpha_surface(svp_handle* h, const double* tris, int32_t n_tris, int32_t mode) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        if (!tris || n_tris <= 0 || mode < 0 || mode > 2)
            throw SvpError(SVP_E_ARG, "init_alpha_surface: need triangles and mode 0 (set), 1 (add) or 2 (subtract)");
        for (long k = 0; k < 9L * n_tris; ++k)
            if (!std::isfinite(tris[k])) throw SvpError(SVP_E_NAN, "init_alpha_surface: non-finite point");
        const int nC = s.m_.nC;
        DBuf dt, df, dl, dn;
        double* dtr = dt.alloc<double>(9L * n_tris);
        double* frac = df.alloc<double>(nC);
        int* cut = dl.alloc<int>(nC);
        int* nCut = dn.alloc<int>(1);
        HIPCHECK(hipMemcpy(dtr, tris, sizeof(double) * 9 * n_tris, hipMemcpyHostToDevice));
        svp::launch_init_alpha_surface(s.m_, dtr, n_tris, mode, frac, cut, nCut, s.st_.alpha, s.stream_);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(s.stream_));
    });
}

int svp_harness_init_alpha(svp_handle* h, const double* shapes, int nShapes, int maxDepth) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        if (!shapes || nShapes <= 0 || maxDepth < 0 || maxDepth > 8)
            throw SvpError(SVP_E_ARG, "init_alpha: need shapes and 0 <= max_depth <= 8");
        DBuf d;
        double* ds = d.alloc<double>(15L * nShapes);
        HIPCHECK(hipMemcpy(ds, shapes, sizeof(double) * 15 * nShapes, hipMemcpyHostToDevice));
        svp::launch_init_alpha(s.m_, ds, nShapes, maxDepth, s.st_.alpha, s.stream_);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(s.stream_));
    });
}

int svp_harness_courant(svp_handle* h, double* co, double* aco) {
    return guarded([&] {
        solveVofEqu& s = H(h);
        svp::launch_courant(s.m_, s.st_.phi, s.st_.alpha, s.courant_, s.stream_);
        unsigned long long r[2];
        HIPCHECK(hipMemcpyAsync(r, s.courant_, sizeof r, hipMemcpyDeviceToHost, s.stream_));
        HIPCHECK(hipStreamSynchronize(s.stream_));
        double v[2];
        std::memcpy(v, r, sizeof v);
        if (co) *co = v[0];
        if (aco) *aco = v[1];
    });
}

}  // extern "C"
