// HIP kernels of the SimPLIC per-time-step alpha transport on gfx950.
//
// Step = reconstruct (detection -> LS normal -> plane placement) + advect
// (upwind face volumes -> interface face volumes -> conservative update ->
// bounding).  All arithmetic is FP64 in the reference's operation order
// (compiled with -ffp-contract=off); each kernel cites the reference lines it
// follows.  Control flow that the reference runs on the host (mixed-cell count,
// bounding passes) stays on the device, so a step never waits for the host.
#include "simplic_device.hpp"
#include "simplic_types.hpp"

namespace svp {

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// Mesh geometry (OpenFOAM primitiveMeshTools), computed once at create time
// ---------------------------------------------------------------------------
__global__ void k_face_geom(DevMesh m) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nF) return;
    const int s = m.fOff[f];
    const int n = m.fOff[f + 1] - s;
    V3 p[MAX_FACE_POINTS];
    for (int k = 0; k < n; ++k) p[k] = ld3(m.pts, m.fPts[s + k]);
    V3 fc, fa;
    primitiveFaceGeom(p, n, fc, fa);
    st3(m.Cf, f, fc);
    st3(m.Sf, f, fa);
    // the sub-face of a fully submerged face (cutFace::calcSubFace)
    const V3 c0 = faceCentre(p, n), a0 = faceAreaNormal(p, n);
    m.fullCS[6L * f] = c0.x;
    m.fullCS[6L * f + 1] = c0.y;
    m.fullCS[6L * f + 2] = c0.z;
    m.fullCS[6L * f + 3] = a0.x;
    m.fullCS[6L * f + 4] = a0.y;
    m.fullCS[6L * f + 5] = a0.z;
    const double ms = mag(fa);
    m.magSf[f] = ms;
    // reconstruction::updateFaceFlatness (reconstruction.C:408-440)
    double fl = 1.0;
    if (n > 3 && ms > SVP_ROOTVSMALL) {
        double sumA = 0.0;
        for (int i = 0; i < n; ++i) {
            const V3 th = p[i], nx = p[(i + 1) % n];
            sumA += mag(0.5 * cross(nx - th, fc - th));
        }
        fl = ms / (sumA + SVP_ROOTVSMALL);
    }
    m.flat[f] = fl;
}

// primitiveMeshTools::makeCellCentresAndVols; cells() face order equals the
// owner-loop-then-neighbour-loop order of the reference implementation.
__global__ void k_cell_geom(DevMesh m) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const int b = m.cOff[c], e = m.cOff[c + 1];
    V3 cEst = VZERO;
    for (int k = b; k < e; ++k) cEst = cEst + ld3(m.Cf, m.cFaces[k]);
    cEst = cEst / (double)(e - b);
    V3 ctr = VZERO;
    double vol = 0.0;
    for (int k = b; k < e; ++k) {
        const int f = m.cFaces[k];
        const V3 fc = ld3(m.Cf, f), fa = ld3(m.Sf, f);
        const double pyr3Vol = (m.own[f] == c) ? dot(fa, fc - cEst) : dot(fa, cEst - fc);
        const V3 pc = (3.0 / 4.0) * fc + (1.0 / 4.0) * cEst;
        ctr = ctr + pyr3Vol * pc;
        vol += pyr3Vol;
    }
    if (fabs(vol) > SVP_VSMALL) ctr = ctr / vol;
    else ctr = cEst;
    st3(m.C, c, ctr);
    m.V[c] = vol * (1.0 / 3.0);
}

// volPointInterpolation weights: inverse distance to the point's cells, or for
// boundary points to its boundary faces, normalised by their sum.
__global__ void k_point_weights(DevMesh m) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m.nP) return;
    const V3 x = ld3(m.pts, p);
    double sumW = 0.0;
    for (int k = m.psOff[p]; k < m.psOff[p + 1]; ++k) {
        const int s = m.psSrc[k];
        const V3 y = (s >= 0) ? ld3(m.C, s) : ld3(m.Cf, ~s);
        const double w = 1.0 / mag(x - y);
        m.psW[k] = w;
        sumW += w;
    }
    for (int k = m.psOff[p]; k < m.psOff[p + 1]; ++k) m.psW[k] = m.psW[k] / sumW;
}

// surfaceInterpolation::makeWeights (OpenFOAM v2312): owner weight of the
// linear interpolation, |Sf.(C_N - Cf)| / (|Sf.(Cf - C_P)| + |Sf.(C_N - Cf)|)
__global__ void k_lin_weights(DevMesh m) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nIF) return;
    const V3 Sf = ld3(m.Sf, f), Cf = ld3(m.Cf, f);
    const double sfdOwn = fabs(dot(Sf, Cf - ld3(m.C, m.own[f])));
    const double sfdNei = fabs(dot(Sf, ld3(m.C, m.nbr[f]) - Cf));
    m.linW[f] = sfdNei / (sfdOwn + sfdNei);
}

void launch_geometry(const DevMesh& m, hipStream_t s) {
    k_face_geom<<<cdiv(m.nF, 256), 256, 0, s>>>(m);
    k_cell_geom<<<cdiv(m.nC, 256), 256, 0, s>>>(m);
    k_point_weights<<<cdiv(m.nP, 256), 256, 0, s>>>(m);
    if (m.nIF > 0) k_lin_weights<<<cdiv(m.nIF, 256), 256, 0, s>>>(m);
}

// ---------------------------------------------------------------------------
// Ordered compaction helpers (block-count / scan / write)
// ---------------------------------------------------------------------------
int detect_blocks(int nC) { return cdiv(nC, DETECT_BLOCK); }

// Ordered compaction in one launch (decoupled look-back): every block takes a
// ticket in the order blocks start, publishes its total, and sums the totals
// of the blocks before it until it meets a published prefix.  A block only
// waits for blocks that took earlier tickets, which are already running, and
// each of those waits only on its own predecessors.  The published words
// carry the launch's epoch, so a word of an earlier launch reads as not yet
// published.  Counts stay below 2^30.
constexpr unsigned long long LOOK_AGG = 1ull << 30, LOOK_PREFIX = 2ull << 30;
constexpr int LOOK_SPIN_LIMIT = 1 << 22;

// where a gathered empty-patch flag of face f is read: f itself for a
// boundary face; for an interior face (never empty, and its flag unused) the
// first boundary face's, clamped into the array when there is none
__device__ inline int empty_slot(const DevMesh& m, int f) { return min(max(f, m.nIF), m.nF - 1); }

// The block's ticket (thread 0 only).  The counter only grows, by nBlocks per
// launch (every launch of a kernel on one counter has the same grid, and every
// block takes one ticket), so ticket / nBlocks numbers the launch and
// ticket % nBlocks is the block's place in it.  The launch number gives the
// look-back epoch (1..2^32-1, never the 0 of cleared words) on the device, so
// a launch replayed from a captured graph gets a fresh epoch as well.
__device__ inline int take_ticket(unsigned long long* counter, int nBlocks, unsigned& epoch) {
    const unsigned long long t = atomicAdd(counter, 1ull);
    epoch = (unsigned)((t / (unsigned long long)nBlocks) % 0xFFFFFFFFull) + 1u;
    return (int)(t % (unsigned long long)nBlocks);
}

// Wave 0 of the block (all 64 lanes): publish this block's total and return
// the sum of the totals of blocks 0..b-1 (the same value in every lane).  The
// lanes read 64 predecessors at once; the walk stops at the nearest published
// prefix and waits only while a word before it is unpublished.  A wait
// without progress for about 0.1 s sets errBit.
__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ inline unsigned long long look_back(unsigned long long* look, int b, unsigned epoch,
                                               unsigned long long total, int* errFlag, int errBit) {
    const int lane = threadIdx.x & 63;
    const unsigned long long tag = (unsigned long long)epoch << 32;
    unsigned long long excl = 0;
    if (b > 0) {
        if (lane == 0)
            __hip_atomic_store(&look[b], tag | LOOK_AGG | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int j0 = b - 1, spins = 0;
        while (true) {
            const int j = j0 - lane;
            // before block 0: a published prefix of zero
            const unsigned long long v =
                j >= 0 ? __hip_atomic_load(&look[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : (tag | LOOK_PREFIX);
            const bool ready = (v >> 32) == epoch;
            const unsigned long long notReady = __ballot(!ready);
            const unsigned long long isPrefix = __ballot(ready && (v & LOOK_PREFIX));
            const int p = isPrefix ? __ffsll((long long)isPrefix) - 1 : 64;
            const unsigned long long need = p >= 63 ? ~0ull : ((1ull << (p + 1)) - 1ull);
            if (notReady & need) {
                if (++spins > LOOK_SPIN_LIMIT) {
                    if (lane == 0) atomicOr(errFlag, errBit);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            excl += wave_sum_u64(lane <= p ? (v & (LOOK_AGG - 1)) : 0ull);
            if (p < 64) break;
            j0 -= 64;
        }
    }
    if (lane == 0)
        __hip_atomic_store(&look[b], tag | LOOK_PREFIX | (excl + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return excl;
}

// workgroup exclusive scan of one int per thread (blockDim.x == 1024)
// exclusive scan over a workgroup of 1024 threads: wave scans by shuffles,
// then a scan of the 16 wave totals; lds needs 16 ints
__device__ inline int wg_scan1024(int v, int* lds, int& total) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[w] = x;
    __syncthreads();
    if (w == 0) {
        int s2 = lane < 16 ? lds[lane] : 0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const int y = __shfl_up(s2, o);
            if (lane >= o) s2 += y;
        }
        if (lane < 16) lds[lane] = s2;
    }
    __syncthreads();
    const int pre = w > 0 ? lds[w - 1] : 0;
    total = lds[15];
    __syncthreads();
    return pre + x - v;
}

// single-workgroup exclusive scan of counts[0..nb) in place; total -> *out.
// Tiles of 8192: eight consecutive counts per thread, one workgroup scan.
__global__ void __launch_bounds__(1024) k_scan_counts(int* counts, int nb, int* out, unsigned long long* accum) {
    __shared__ int lds[16];
    constexpr int PT = 8;
    const int t = threadIdx.x;
    int carry = 0;
    for (int base = 0; base < nb; base += 1024 * PT) {
        int v[PT];
        const int i = base + PT * t;
#pragma unroll
        for (int j = 0; j < PT; ++j) v[j] = (i + j < nb) ? counts[i + j] : 0;
        int sum = 0;
#pragma unroll
        for (int j = 0; j < PT; ++j) sum += v[j];
        int total;
        int off = carry + wg_scan1024(sum, lds, total);
#pragma unroll
        for (int j = 0; j < PT; ++j) {
            if (i + j < nb) counts[i + j] = off;
            off += v[j];
        }
        carry += total;
    }
    if (t == 0) {
        *out = carry;
        if (accum) *accum += (unsigned long long)carry;
    }
}

// ---------------------------------------------------------------------------
// reconstruction::initialize mixed-cell detection (reconstruction.C:634-677,
// isAMixedCell reconstruction.H:281-288): ascending list of mixed cells
// ---------------------------------------------------------------------------
// Detection in one launch: flags, a look-back over the blocks, and the
// ascending mixed-cell list; the last ticket writes the count.  A block of
// 1024 threads takes 8 chunks of 1024 cells (thread t holds cell t of every
// chunk, so the reads are coalesced and the flags stay in registers); few,
// large blocks keep the look-back chain short.
constexpr int DETECT_LIST_THREADS = 1024, DETECT_LIST_CHUNKS = 8;
constexpr int DETECT_LIST_CELLS = DETECT_LIST_THREADS * DETECT_LIST_CHUNKS;

int detect_list_blocks(int nC) { return cdiv(nC, DETECT_LIST_CELLS); }

__device__ __attribute__((always_inline)) inline void k_detect_list_body(const DevMesh& m, const DevState& st, double tol, int nBlocks) {
    constexpr int NW = DETECT_LIST_THREADS / 64;
    __shared__ int cnt[DETECT_LIST_CHUNKS][NW];
    __shared__ int sTicket, sBase;
    __shared__ unsigned sEpoch;
    if (threadIdx.x == 0) sTicket = take_ticket(st.detTicket, nBlocks, sEpoch);
    __syncthreads();
    const int b = sTicket;
    const unsigned epoch = sEpoch;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c0 = b * DETECT_LIST_CELLS + threadIdx.x;
    if (b == 0 && threadIdx.x < 2 && st.clearAtDetect) st.clearAtDetect[threadIdx.x] = 0;
    double a[DETECT_LIST_CHUNKS];
#pragma unroll
    for (int k = 0; k < DETECT_LIST_CHUNKS; ++k) {
        const int c = c0 + k * DETECT_LIST_THREADS;
        a[k] = c < m.nC ? st.alpha[c] : 0.0;
    }
    unsigned flags = 0;
    int inWave[DETECT_LIST_CHUNKS];
#pragma unroll
    for (int k = 0; k < DETECT_LIST_CHUNKS; ++k) {
        const int c = c0 + k * DETECT_LIST_THREADS;
        bool flag = false;
        if (c < m.nC) flag = (tol < a[k]) && (a[k] < 1.0 - tol);
        flags |= flag ? (1u << k) : 0u;
        const unsigned long long bal = __ballot(flag);
        inWave[k] = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) cnt[k][wave] = __popcll(bal);
    }
    __syncthreads();
    int total = 0;
    for (int k = 0; k < DETECT_LIST_CHUNKS; ++k)
        for (int w = 0; w < NW; ++w) total += cnt[k][w];
    if (threadIdx.x < 64) {
        const unsigned long long excl =
            look_back(st.detLook, b, epoch, (unsigned long long)total, st.errFlag, ERR_DETECT_STALL);
        if (threadIdx.x == 0) {
            if (b == nBlocks - 1) {
                *st.nMixed = (int)(excl + total);
                *st.nIcTotal += excl + total;
            }
            sBase = (int)excl;
        }
    }
    __syncthreads();
    int off = sBase;
#pragma unroll
    for (int k = 0; k < DETECT_LIST_CHUNKS; ++k) {
        int pre = 0, tk = 0;
        for (int w = 0; w < NW; ++w) {
            pre += w < wave ? cnt[k][w] : 0;
            tk += cnt[k][w];
        }
        if (flags & (1u << k)) st.mixed[off + pre + inWave[k]] = c0 + k * DETECT_LIST_THREADS;
        off += tk;
    }
}

__global__ void __launch_bounds__(DETECT_LIST_THREADS) k_detect_list(DevMesh m, DevState st, double tol, int nBlocks) {
    k_detect_list_body(m, st, tol, nBlocks);
}

void launch_detect(const DevMesh& m, const DevState& st, const Params& p, hipStream_t s) {
    const int nb = cdiv(m.nC, DETECT_LIST_CELLS);
    k_detect_list<<<nb, DETECT_LIST_THREADS, 0, s>>>(m, st, p.mixedCellTol, nb);
}

// ---------------------------------------------------------------------------
// LS normal (reconstruction.C:85-141): n = -grad(alpha)/(|grad| + SMALL) with
// the polyDegree1 least-squares gradient over the cell-point-cell stencil
// (normal equations, Cholesky decomposition and inverse of OpenFOAM's
// scalarMatrices, x = inv(A) B).
// ---------------------------------------------------------------------------
template <int SB = 8>
__device__ inline V3 ls_gradient(const DevMesh& m, const double* alpha, int c) {
    double A[4][4] = {};
    double B[4] = {};
    const V3 cc = ld3(m.C, c);
    // stencil cells in batches of SB: the labels, then their centres and
    // alphas, are loaded together; the sums still run in stencil order
    const int b = m.stOff[c], e = m.stOff[c + 1];
    for (int k0 = b; k0 < e; k0 += SB) {
        // every load is unconditional (past the end: the last entry again)
        // and entries past the end leave the sums unchanged by selection,
        // so nothing splits the loads into dependent groups
        int g[SB];
#pragma unroll
        for (int j = 0; j < SB; ++j) g[j] = m.st[min(k0 + j, e - 1)];
        V3 pc[SB];
        double av[SB];
#pragma unroll
        for (int j = 0; j < SB; ++j) {
            pc[j] = ld3(m.C, g[j]);
            av[j] = alpha[g[j]];
        }
        // keep the scheduler from sinking those loads into the sums
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < SB; ++j) {
            const bool in = k0 + j < e;
            const V3 d = pc[j] - cc;
            const double t[4] = {1.0, d.x, d.y, d.z};
            const double v = av[j];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double a = A[r][q] + t[r] * t[q];
                    A[r][q] = in ? a : A[r][q];
                }
                const double bb = B[r] + t[r] * v;
                B[r] = in ? bb : B[r];
            }
        }
    }
    // LUDecompose(scalarSymmetricSquareMatrix)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int k = j + 1; k < 4; k++) A[j][k] = 0.0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        double d = 0.0;
#pragma unroll
        for (int k = 0; k < j; k++) {
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < k; i++) s += A[i][k] * A[i][j];
            s = (A[j][k] - s) / A[k][k];
            A[k][j] = s;
            A[j][k] = s;
            d += s * s;
        }
        d = A[j][j] - d;
        A[j][j] = sqrt(d);
    }
    // invDecomposed
    double inv[4][4] = {};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        inv[i][i] = 1.0 / A[i][i];
#pragma unroll
        for (int j = 0; j < i; ++j) {
            double sum = 0.0;
#pragma unroll
            for (int k = j; k < i; k++) sum -= A[i][k] * inv[k][j];
            inv[i][j] = sum / A[i][i];
        }
    }
    double R[4][4] = {};
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int i = 0; i <= k; i++)
#pragma unroll
            for (int j = 0; j <= k; j++) R[i][j] += inv[k][i] * inv[k][j];
    double x[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) x[i] += R[i][j] * B[j];
    return v3(x[1], x[2], x[3]);
}

// wl: optional list of mixed-cell positions (*wn entries) instead of all of
// them.  SB: stencil cells loaded per batch (one wave per 64 cells has the
// registers to spare): 28 takes a 3x3x3 hexahedral stencil in one batch, 36
// the 15-71 cells of a tetrahedral one in two.
template <int SB = 28>
__device__ __attribute__((always_inline)) inline void k_ls_normal_body(const DevMesh& m, const DevState& st, const int* wl, const int* wn) {
    const int nWork = wl ? *wn : *st.nMixed;
    for (int wi = blockIdx.x * blockDim.x + threadIdx.x; wi < nWork; wi += gridDim.x * blockDim.x) {
        const int i = wl ? wl[wi] : wi;
        const int c = st.mixed[i];
        V3 n = -ls_gradient<SB>(m, st.alpha, c);
        n = n / (mag(n) + SVP_SMALL);
        st3(st.ifN, i, n);
    }
}

template <int SB>
__global__ void __launch_bounds__(64, 1) k_ls_normal(DevMesh m, DevState st, const int* wl, const int* wn) {
    k_ls_normal_body<SB>(m, st, wl, wn);
}

void launch_ls_normal(const DevMesh& m, const DevState& st, hipStream_t s, const int* wl, const int* wn) {
    if (m.maxStencil > 28) k_ls_normal<36><<<1024, 64, 0, s>>>(m, st, wl, wn);
    else k_ls_normal<28><<<1024, 64, 0, s>>>(m, st, wl, wn);
}

// calcInterfaceNFromRegAlphaGrad (reconstruction.C:74-82): n = -grad(alpha) /
// (|grad| + SMALL) with fvc::grad = Gauss linear, i.e. OpenFOAM's
// gaussGrad::gradf: the faces of the cell in ascending order add Sf*ssf
// (owner) or subtract it (neighbour), boundary faces of non-empty patches
// add Sf*alpha (zeroGradient), then the sum is divided by V.  Only mixed
// cells need the normal.  Faces in batches of 8 with all loads in flight.
// volPointInterpolation of alpha at point p: internal points over their
// cells, boundary points over their boundary faces, whose zeroGradient value
// is the owner's alpha (the point sources and weights of the U interpolation)
__device__ inline double point_alpha(const DevMesh& m, const double* alpha, int p) {
    double t = 0.0;
    for (int k = m.psOff[p]; k < m.psOff[p + 1]; ++k) {
        const int s = m.psSrc[k];
        t += m.psW[k] * alpha[s >= 0 ? s : m.own[~s]];
    }
    return t;
}

// Gauss pointLinear face value of internal face f (OpenFOAM v2312
// pointLinear::correction, selected by the NAG case's fvSchemes,
// tutorials/test/plicVofOrientationFoam/NAG/system/fvSchemes:35): the fan of
// triangles (pi, p_k, p_k-1), closed by (pi, p_0, p_last), around
// pi = w C_own + (1 - w) C_nei with w the face weight at the owner cell's
// index (as pointLinear indexes it); the area average of
// (linear value + the two point values) / 3, returned as
// linear + (average - linear).
__device__ inline double point_linear_face_value(const DevMesh& m, const double* alpha, int f, double lin) {
    const int o = m.own[f], n = m.nbr[f];
    const double wq = m.linW[o];
    const V3 pi = wq * ld3(m.C, o) + (1.0 - wq) * ld3(m.C, n);
    const int b = m.fOff[f], np = m.fOff[f + 1] - b;
    const int p0 = m.fPts[b], pl = m.fPts[b + np - 1];
    double at = mag(0.5 * cross(ld3(m.pts, p0) - pi, ld3(m.pts, pl) - pi));
    double sumAt = at;
    double sumPsip = at * (1.0 / 3.0) * (lin + point_alpha(m, alpha, p0) + point_alpha(m, alpha, pl));
    for (int k = 1; k < np; ++k) {
        const int pk = m.fPts[b + k], pk1 = m.fPts[b + k - 1];
        at = mag(0.5 * cross(ld3(m.pts, pk) - pi, ld3(m.pts, pk1) - pi));
        sumAt += at;
        sumPsip += at * (1.0 / 3.0) * (lin + point_alpha(m, alpha, pk) + point_alpha(m, alpha, pk1));
    }
    const double corr = sumPsip / sumAt - lin;
    return lin + corr;
}

template <bool POINT_LINEAR>
__global__ void __launch_bounds__(64, 1) k_grad_normal(DevMesh m, DevState st) {
    const int nMixed = *st.nMixed;
    constexpr int FB = 8;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nMixed; i += gridDim.x * blockDim.x) {
        const int c = st.mixed[i];
        const double ac = st.alpha[c];
        const double Vc = m.V[c];
        V3 g = VZERO;
        const int b = m.cOff[c], e = m.cOff[c + 1];
        for (int k0 = b; k0 < e; k0 += FB) {
            int ev[FB];
#pragma unroll
            for (int j = 0; j < FB; ++j) ev[j] = m.cFacesAsc[min(k0 + j, e - 1)];
            V3 sf[FB];
            double w[FB], ao[FB];
            unsigned char em[FB];
#pragma unroll
            for (int j = 0; j < FB; ++j) {
                const int f = ev[j] >= 0 ? ev[j] : ~ev[j];
                const bool in = f < m.nIF;
                sf[j] = ld3(m.Sf, f);
                w[j] = m.linW[in ? f : 0];
                em[j] = m.fEmpty[empty_slot(m, f)];
                // the other cell's alpha (the cell's own on boundary faces)
                const int o = ev[j] >= 0 ? (in ? m.nbr[f] : c) : m.own[f];
                ao[j] = st.alpha[o];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < FB; ++j) {
                if (k0 + j >= e) break;
                const int ef = ev[j];
                if (ef >= 0) {
                    if (ef < m.nIF) {
                        double ssf = w[j] * (ac - ao[j]) + ao[j];
                        if (POINT_LINEAR) ssf = point_linear_face_value(m, st.alpha, ef, ssf);
                        g = g + ssf * sf[j];
                    } else if (!em[j]) {
                        g = g + ac * sf[j];
                    }
                } else {
                    double ssf = w[j] * (ao[j] - ac) + ac;
                    if (POINT_LINEAR) ssf = point_linear_face_value(m, st.alpha, ~ef, ssf);
                    g = g - ssf * sf[j];
                }
            }
        }
        const V3 n = -(g / Vc);
        st3(st.ifN, i, n / (mag(n) + SVP_SMALL));
    }
}

void launch_grad_normal(const DevMesh& m, const DevState& st, bool pointLinear, hipStream_t s) {
    if (pointLinear) k_grad_normal<true><<<1024, 64, 0, s>>>(m, st);
    else k_grad_normal<false><<<1024, 64, 0, s>>>(m, st);
}

// ---------------------------------------------------------------------------
// isoRDF / RDF normals (reconstruction.C:194-405) with OpenFOAM's
// reconstructedDistanceFunction (markCellsNearSurf with one layer,
// constructRDF); one lane per mixed cell or per cell
// ---------------------------------------------------------------------------
// Vector::normalise(tol) of OpenFOAM
__device__ inline V3 normalise_tol(V3 v, double tol) {
    const double s = mag(v);
    return s < tol ? VZERO : v / s;
}

// The buffers are clean (zero) between reconstructs: the first allocation
// zeroes them and k_rdf_end clears what one reconstruct touched, all of it
// inside the nextToInterface cells.  ctl holds the small counters.
__global__ void k_rdf_begin(DevMesh m, DevState st, RdfBufs r) {
    const int nMixed = *st.nMixed;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nMixed; i += gridDim.x * blockDim.x) {
        const int c = st.mixed[i];
        // level 0 and level 1 of markCellsNearSurf: the stencil holds the cell
        // and every cell sharing a point with it
        for (int k = m.stOff[c]; k < m.stOff[c + 1]; ++k) {
            const int o = m.st[k];
            if (atomicExch(&r.next[o], 1) == 0) r.nextList[atomicAdd(&r.ctl[RDF_NNEXT], 1)] = o;
        }
        st.ifD[i] = 0.0;
        st3(st.ifN, i, VZERO);
        st3(st.ifC, i, VZERO);
        st3(st.ifS, i, VZERO);
        st3(r.g, i, ls_gradient(m, st.alpha, c));
    }
}

__global__ void k_rdf_normals(DevState st, RdfBufs r) {
    if (r.ctl[RDF_DONE]) return;
    const int nMixed = *st.nMixed;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nMixed; i += gridDim.x * blockDim.x) {
        const int c = st.mixed[i];
        const V3 gi = normalise_tol(ld3(r.g, i), SVP_SMALL);
        st3(r.g, i, gi);
        const V3 n = -gi;
        st3(st.ifN, i, n);
        st3(r.cellN, c, n);
        if (!r.tooCoarse[c]) r.wl[atomicAdd(&r.ctl[RDF_WN], 1)] = i;
    }
}

__global__ void k_rdf_scatter_c(DevState st, RdfBufs r) {
    if (r.ctl[RDF_DONE]) return;
    const int nMixed = *st.nMixed;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nMixed; i += gridDim.x * blockDim.x)
        st3(r.cellC, st.mixed[i], ld3(st.ifC, i));
}

// constructRDF(nextToInterface, interfaceC, -interfaceN) over the
// nextToInterface cells (the others hold 0): the signed distance to the
// cell's own plane in an interface cell, else the average over the stencil's
// planes weighted by the squared cosine between the plane normal and the
// direction to the plane centre; a cell with no weight keeps its value.
__global__ void __launch_bounds__(64) k_rdf_construct(DevMesh m, RdfBufs r) {
    if (r.ctl[RDF_DONE]) return;
    const int nNext = r.ctl[RDF_NNEXT];
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nNext; j += gridDim.x * blockDim.x) {
        const int c = r.nextList[j];
        const V3 cc = ld3(m.C, c);
        const V3 nrm = -ld3(r.cellN, c);
        if (mag(nrm) != 0.0) {
            const V3 n = (-nrm) / mag(nrm);
            r.rdf[c] = dot(ld3(r.cellC, c) - cc, n);
            continue;
        }
        double averageWeight = 0.0, avgDist = 0.0;
        // stencil entries in batches of 8 with every load issued first (past
        // the end: the last entry again, left out of the sums); the sums run
        // in stencil order
        constexpr int SB = 8;
        const int kb = m.stOff[c], ke = m.stOff[c + 1];
        for (int k0 = kb; k0 < ke; k0 += SB) {
            int o[SB];
#pragma unroll
            for (int j = 0; j < SB; ++j) o[j] = m.st[min(k0 + j, ke - 1)];
            V3 nb[SB], cb[SB];
#pragma unroll
            for (int j = 0; j < SB; ++j) {
                nb[j] = ld3(r.cellN, o[j]);
                cb[j] = ld3(r.cellC, o[j]);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < SB; ++j) {
                V3 n = -(-nb[j]);
                if (k0 + j < ke && mag(n) != 0.0) {
                    n = n / mag(n);
                    V3 dv = cb[j] - cc;
                    const double distToSurf = dot(dv, n);
                    double weight;
                    if (mag(dv) != 0.0) {
                        dv = dv / mag(dv);
                        const double w = fabs(dot(dv, n));
                        weight = w * w;
                    } else {
                        weight = 1.0;
                    }
                    averageWeight += weight;
                    avgDist += distToSurf * weight;
                }
            }
        }
        if (averageWeight != 0.0) r.rdf[c] = avgDist / averageWeight;
    }
}

__global__ void k_rdf_grad(DevMesh m, DevState st, RdfBufs r) {
    if (r.ctl[RDF_DONE]) return;
    const int nMixed = *st.nMixed;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nMixed; i += gridDim.x * blockDim.x)
        st3(r.g, i, ls_gradient(m, r.rdf, st.mixed[i]));
}

// Residual of one mixed cell (reconstruction.C:276-330) and its share of
// the averages (:333-357).  g is normalised in place as the reference does:
// once inside the zero-normal test when the cell's normal is not zero, once
// for the new normal.
__global__ void __launch_bounds__(64) k_rdf_residual(DevMesh m, DevState st, RdfBufs r, int iter, double tol) {
    if (r.ctl[RDF_DONE]) return;
    const int nMixed = *st.nMixed;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nMixed; i += gridDim.x * blockDim.x) {
        const int c = st.mixed[i];
        const V3 cellNormal = ld3(r.cellN, c);
        V3 gi = ld3(r.g, i);
        if (!(mag(cellNormal) < SVP_TSMALL)) gi = normalise_tol(gi, SVP_SMALL);
        double avgDiffNormal = 0.0, weight = 0.0;
        // the stencil past the cell itself, in batches of 8 with the loads first
        constexpr int SB = 8;
        const int kb = m.stOff[c] + 1, ke = m.stOff[c + 1];
        for (int k0 = kb; k0 < ke; k0 += SB) {
            V3 nb[SB];
#pragma unroll
            for (int j = 0; j < SB; ++j) nb[j] = ld3(r.cellN, m.st[min(k0 + j, ke - 1)]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < SB; ++j) {
                const V3 normal = nb[j];
                if (k0 + j < ke && mag(normal) >= SVP_TSMALL) {
                    const V3 n = normal / mag(normal);
                    const double cosAngle = fmax(fmin(dot(cellNormal, n), 1.0), -1.0);
                    avgDiffNormal += acos(cosAngle) * mag(normal);
                    weight += mag(normal);
                }
            }
        }
        avgDiffNormal = weight != 0.0 ? avgDiffNormal / weight : 0.0;
        gi = normalise_tol(gi, SVP_SMALL);
        st3(r.g, i, gi);
        const double normalRes = 1.0 - dot(cellNormal, -gi);
        if (avgDiffNormal > 0.26 && iter > 0) {
            r.tooCoarse[c] = 1;
            r.res[2 * i] = 0.0;
            r.res[2 * i + 1] = 0.0;
            r.counted[i] = 0;
        } else {
            const double discreteError = 0.01 * (avgDiffNormal * avgDiffNormal);
            const double normRes = discreteError != 0.0 ? normalRes / fmax(discreteError, tol) : normalRes / tol;
            r.res[2 * i] = normalRes;
            r.res[2 * i + 1] = normRes;
            r.counted[i] = 1;
        }
    }
}

// Sums of the residual terms in one workgroup (a fixed order: the sums only
// feed the stopping test), then the stopping rule of reconstruction.C:359-399
// on the device: it sets the done flag that turns the remaining iterations'
// kernels into no-ops.  Also counts the iteration and clears the work-list
// count for the next one.
__global__ void __launch_bounds__(1024) k_rdf_reduce(DevState st, RdfBufs r, int iter, int iterations, double tol,
                                                     double relTol) {
    if (r.ctl[RDF_DONE]) return;
    __shared__ double s0[1024], s1[1024], s2[1024];
    const int nMixed = *st.nMixed;
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0, n = 0.0;
    for (int i = t; i < nMixed; i += 1024) {
        a += r.res[2 * i];
        b += r.res[2 * i + 1];
        n += (double)r.counted[i];
    }
    s0[t] = a;
    s1[t] = b;
    s2[t] = n;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) {
            s0[t] += s0[t + w];
            s1[t] += s1[t + w];
            s2[t] += s2[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        double avgRes = s0[0], avgNormRes = s1[0];
        long long resCounter = (long long)s2[0];
        if (resCounter == 0) {
            resCounter = 1;
            avgRes = 0.0;
            avgNormRes = 0.0;
        }
        const bool stop = ((avgNormRes / resCounter < relTol || avgRes / resCounter < tol) && iter >= 1) ||
                          iter + 1 == iterations;
        r.ctl[RDF_ITERS] = iter + 1;
        r.ctl[RDF_WN] = 0;
        if (stop) r.ctl[RDF_DONE] = 1;
    }
}

// after the iterations: the work list of the cells that are (or became) too
// coarse, whose planes the final placement still has to compute
__global__ void k_rdf_coarse_list(DevState st, RdfBufs r) {
    const int nMixed = *st.nMixed;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nMixed; i += gridDim.x * blockDim.x)
        if (r.tooCoarse[st.mixed[i]]) r.wl2[atomicAdd(&r.ctl[RDF_WN2], 1)] = i;
}

// clear what this reconstruct touched (all inside the nextToInterface cells)
__global__ void k_rdf_end(RdfBufs r) {
    const int nNext = r.ctl[RDF_NNEXT];
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nNext; j += gridDim.x * blockDim.x) {
        const int c = r.nextList[j];
        r.next[c] = 0;
        r.tooCoarse[c] = 0;
        r.rdf[c] = 0.0;
        st3(r.cellN, c, VZERO);
        st3(r.cellC, c, VZERO);
    }
}

void launch_rdf_begin(const DevMesh& m, const DevState& st, const RdfBufs& r, hipStream_t s) {
    k_rdf_begin<<<1024, 64, 0, s>>>(m, st, r);
}

void launch_rdf_iteration(const DevMesh& m, const DevState& st, const Params& p, const RdfBufs& r, int iter,
                          int iterations, double tol, double relTol, hipStream_t s) {
    k_rdf_normals<<<1024, 256, 0, s>>>(st, r);
    launch_find_signed_distance(m, st, p, s, r.wl, r.ctl + RDF_WN);
    k_rdf_scatter_c<<<1024, 256, 0, s>>>(st, r);
    k_rdf_construct<<<1024, 64, 0, s>>>(m, r);
    k_rdf_grad<<<1024, 64, 0, s>>>(m, st, r);
    k_rdf_residual<<<1024, 64, 0, s>>>(m, st, r, iter, tol);
    k_rdf_reduce<<<1, 1024, 0, s>>>(st, r, iter, iterations, tol, relTol);
}

void launch_rdf_end(const DevMesh& m, const DevState& st, const RdfBufs& r, hipStream_t s) {
    (void)m;
    k_rdf_coarse_list<<<1024, 256, 0, s>>>(st, r);
    k_rdf_end<<<1024, 256, 0, s>>>(r);
}

// ---------------------------------------------------------------------------
// Local polyhedron of a cell (cutCell.C:140-236)
// ---------------------------------------------------------------------------
template <class Poly>
__device__ inline bool build_cell_poly(const DevMesh& m, int c, bool split, Poly& P) {
    constexpr int MP = Poly::maxPts, MF = Poly::maxFaces, MI = Poly::maxIdx;
    P.nPts = 0;
    P.nFaces = 0;
    P.fStart[0] = 0;
    int gl[MP];    // global labels of P.pts in first-appearance order
    int nIdx = 0;
    bool ok = true;
    const int b = m.cOff[c], e = m.cOff[c + 1];
    // c.labels(faces): unique points in order of first appearance
    for (int k = b; k < e; ++k) {
        const int f = m.cFaces[k];
        for (int q = m.fOff[f]; q < m.fOff[f + 1]; ++q) {
            const int p = m.fPts[q];
            bool found = false;
            for (int r = 0; r < P.nPts; ++r) found |= (gl[r] == p);
            if (!found) {
                if (P.nPts < MP) {
                    gl[P.nPts] = p;
                    P.pts[P.nPts] = ld3(m.pts, p);
                }
                P.nPts++;
            }
        }
    }
    if (P.nPts > MP) return false;
    auto local = [&](int p) {
        int r = 0;
        while (gl[r] != p) ++r;
        return r;
    };
    for (int k = b; k < e; ++k) {
        const int f = m.cFaces[k];
        const int fs = m.fOff[f], n = m.fOff[f + 1] - fs;
        const bool rev = split && (m.own[f] != c);
        if (!split || m.flat[f] > (1.0 - SVP_TSMALL)) {
            if (P.nFaces >= MF || nIdx + n > MI) { ok = false; break; }
            for (int q = 0; q < n; ++q) {
                // face::reverseFace keeps the first point
                const int qq = rev ? (q == 0 ? 0 : n - q) : q;
                P.fIdx[nIdx++] = (unsigned char)local(m.fPts[fs + qq]);
            }
            P.fStart[++P.nFaces] = nIdx;
        } else {
            if (P.nPts >= MP) { ok = false; break; }
            P.pts[P.nPts] = ld3(m.Cf, f);
            const int centre = P.nPts++;
            for (int q = 0; q < n; ++q) {
                if (P.nFaces >= MF || nIdx + 3 > MI) { ok = false; break; }
                const int a0 = centre, a1 = local(m.fPts[fs + q]), a2 = local(m.fPts[fs + (q + 1) % n]);
                P.fIdx[nIdx++] = (unsigned char)a0;
                P.fIdx[nIdx++] = (unsigned char)(rev ? a2 : a1);
                P.fIdx[nIdx++] = (unsigned char)(rev ? a1 : a2);
                P.fStart[++P.nFaces] = nIdx;
            }
            if (!ok) break;
        }
    }
    return ok;
}

// ---------------------------------------------------------------------------
// cutCell::findSignedDistance (cutCell.C:611-799), one lane per mixed cell
// ---------------------------------------------------------------------------
struct FsdOut {
    int status, nCutPts;
    unsigned long long faceClass;
    double d;
    V3 C, S;
    bool wrote;
};

__device__ inline void classify(const FaceCut* fc, int nF, FsdOut& o) {
    o.faceClass = 0;
    o.nCutPts = 0;
    for (int f = 0; f < nF; ++f) {
        o.faceClass |= (unsigned long long)(fc[f].status + 1) << (2 * f);
        if (fc[f].status == 0) o.nCutPts += fc[f].nIf;
    }
}

template <class Poly>
__device__ inline void find_signed_distance(const Poly& P, V3 Cc, double Vc, double alphaI,
                                            V3 normalI, FaceCut* fc, FsdOut& o, bool& ok) {
    o.wrote = false;
    o.faceClass = 0;
    o.nCutPts = 0;
    if (mag(normalI) < SVP_TSMALL) {
        o.status = (int)sgn(0.5 - alphaI);
        return;
    }
    double vd[Poly::maxPts];
    const int n = P.nPts;
    for (int k = 0; k < n; ++k) vd[k] = -dot(normalI, P.pts[k]);
    // descending sort of the vertex levels (only the sorted values are used)
    for (int k = 1; k < n; ++k) {
        const double v = vd[k];
        int j = k - 1;
        while (j >= 0 && vd[j] < v) {
            vd[j + 1] = vd[j];
            --j;
        }
        vd[j + 1] = v;
    }
    double lowDistance = vd[0], upDistance = vd[n - 1];
    int lowLabel = 0, upLabel = n - 1;
    double lowAlpha = 0.0, upAlpha = 1.0;
    CellCut cut;
    while ((upLabel - lowLabel) > 1) {
        const int midLabel = (int)round(0.5 * (double)(upLabel + lowLabel));
        const double midDistance = vd[midLabel];
        ok &= calcSubCell(P, Cc, Vc, normalI, midDistance, fc, cut);
        const double midAlpha = cut.VOF;
        if (fabs(midAlpha - alphaI) < SVP_TSMALL) {
            o.d = midDistance;
            o.C = cut.ifCentre;
            o.S = cut.ifArea;
            o.status = cut.status;
            o.wrote = true;
            classify(fc, P.nFaces, o);
            return;
        }
        if (midAlpha > alphaI) {
            upLabel = midLabel;
            upDistance = midDistance;
            upAlpha = midAlpha;
        } else {
            lowLabel = midLabel;
            lowDistance = midDistance;
            lowAlpha = midAlpha;
        }
    }
    if (fabs(lowDistance - upDistance) < SVP_TSMALL) {
        const double midDistance = 0.5 * (lowDistance + upDistance);
        ok &= calcSubCell(P, Cc, Vc, normalI, midDistance, fc, cut);
        o.d = midDistance;
        o.C = cut.ifCentre;
        o.S = cut.ifArea;
        o.status = cut.status;
        o.wrote = true;
        classify(fc, P.nFaces, o);
        return;
    }
    const double alphaPrismatoid = upAlpha - lowAlpha;
    const double deltaDistance = (upDistance - lowDistance) / 3.0;
    const double distanceOneThird = lowDistance + deltaDistance;
    ok &= calcSubCell(P, Cc, Vc, normalI, distanceOneThird, fc, cut);
    const double alphaOneThird = cut.VOF - lowAlpha;
    const double distanceTwoThirds = lowDistance + 2.0 * deltaDistance;
    ok &= calcSubCell(P, Cc, Vc, normalI, distanceTwoThirds, fc, cut);
    const double alphaTwoThirds = cut.VOF - lowAlpha;
    const double a = 13.5 * alphaOneThird - 13.5 * alphaTwoThirds + 4.5 * alphaPrismatoid;
    const double b = -22.5 * alphaOneThird + 18.0 * alphaTwoThirds - 4.5 * alphaPrismatoid;
    const double c = 9.0 * alphaOneThird - 4.5 * alphaTwoThirds + 1.0 * alphaPrismatoid;
    const double d = lowAlpha - alphaI;
    double lambda = 0.5;
    for (int iter = 0; iter < 100; iter++) {
        const double func = a * (lambda * lambda * lambda) + b * (lambda * lambda) + c * lambda + d;
        const double funcPrime = 3.0 * a * (lambda * lambda) + 2.0 * b * lambda + c;
        const double lambdaNew = lambda - (func / funcPrime);
        if (fabs(lambdaNew - lambda) < SVP_TSMALL) break;
        lambda = lambdaNew;
    }
    const double distance0 = lowDistance - lambda * (lowDistance - upDistance);
    ok &= calcSubCell(P, Cc, Vc, normalI, distance0, fc, cut);
    o.d = distance0;
    o.C = cut.ifCentre;
    o.S = cut.ifArea;
    o.status = cut.status;
    o.wrote = true;
    classify(fc, P.nFaces, o);
}

template <class Poly>
__global__ void __launch_bounds__(64) k_find_signed_distance(DevMesh m, DevState st, int split, const int* wl,
                                                             const int* wn) {
    const int nWork = wl ? *wn : *st.nMixed;
    for (int wi = blockIdx.x * blockDim.x + threadIdx.x; wi < nWork; wi += gridDim.x * blockDim.x) {
        const int i = wl ? wl[wi] : wi;
        const int c = st.mixed[i];
        Poly P;
        FaceCut fc[Poly::maxFaces];
        FsdOut o;
        bool ok = build_cell_poly(m, c, split != 0, P);
        if (!ok) {
            atomicOr(st.errFlag, ERR_CELL_TOO_LARGE);
            st.status[i] = -100;
            continue;
        }
        find_signed_distance(P, ld3(m.C, c), m.V[c], st.alpha[c], ld3(st.ifN, i), fc, o, ok);
        if (!ok) atomicOr(st.errFlag, ERR_TOO_MANY_CROSSINGS);
        st.status[i] = o.status;
        st.nCutPts[i] = o.nCutPts;
        st.faceClass[i] = o.faceClass;
        // interfaceD/C/S stay as they were where the reference returns before
        // writing: zero after initialize, the last value inside the isoRDF loop
        if (o.wrote || !st.fsdKeep) {
            st.ifD[i] = o.wrote ? o.d : 0.0;
            st3(st.ifC, i, o.wrote ? o.C : VZERO);
            st3(st.ifS, i, o.wrote ? o.S : VZERO);
        }
    }
}

// findSignedDistance with one G-lane group per mixed cell and one lane per
// cell face (cells of at most G faces and 16 points, faces as stored; G is
// the mesh's largest face count: 4 for tetrahedra (16 cells per wave), 6 for
// hexahedra (10 cells per wave, 4 lanes idle), else 8).  A
// plane cut = every lane cuts its face (points in registers), writes the face
// record to LDS, and every lane of the group combines the records in face
// order, so all lanes hold the same VOF and follow the same bracketing and
// Newton steps as the one-lane version.  The vertex levels are sorted by rank
// in LDS.  One wave per workgroup: LDS traffic of a wave is in order.

// volPointInterpolation value at a point: the weighted sum of its source
// values in source order.  A point of up to PU_SLOTS sources (every interior
// point of a hexahedral mesh) loads all source labels and weights, then all
// their values, before the sum, instead of waiting on two loads per source.
constexpr int PU_SLOTS = 8;
__device__ inline V3 point_u(const DevMesh& m, const DevState& st, int p) {
    const int o0 = m.psOff[p], n = m.psOff[p + 1] - o0;
    V3 t = VZERO;
    if (n >= 1 && n <= PU_SLOTS) {
        int src[PU_SLOTS];
        double w[PU_SLOTS];
        V3 u[PU_SLOTS];
#pragma unroll
        for (int k = 0; k < PU_SLOTS; ++k) {
            src[k] = m.psSrc[o0 + min(k, n - 1)];
            w[k] = m.psW[o0 + min(k, n - 1)];
        }
#pragma unroll
        for (int k = 0; k < PU_SLOTS; ++k) {
            // a source is a cell (>= 0) or a boundary face (~face)
            const double* base = src[k] >= 0 ? st.U : st.Ub;
            const int idx = src[k] >= 0 ? src[k] : (~src[k]) - m.nIF;
            u[k] = ld3(base, idx);
        }
#pragma unroll
        for (int k = 0; k < PU_SLOTS; ++k)
            if (k < n) t = t + w[k] * u[k];
        return t;
    }
    for (int k = o0; k < o0 + n; ++k) {
        const int s = m.psSrc[k];
        const V3 u = (s >= 0) ? ld3(st.U, s) : ld3(st.Ub, (~s) - m.nIF);
        t = t + m.psW[k] * u;
    }
    return t;
}

__device__ __attribute__((always_inline)) inline void fsd_face_cut(int fsz, const V3 (&q4)[4], const V3 (&q3)[3], V3 n, double dist,
                                    FaceCut& fc, bool& ok, const double* full, const double (&dots)[4]) {
    if (fsz == 4) ok &= calcSubFaceN<4, true, true>(q4, n, dist, fc, full, dots);
    else ok &= calcSubFaceN<3, true, true>(q3, n, dist, fc, full, dots);
}

// waves per SIMD the register allocation of the interface-cell kernels
// targets (build-time tuning; results do not depend on it)
#ifndef SVP_FSD_WAVES
#define SVP_FSD_WAVES 2
#endif
#ifndef SVP_FLUX_WAVES
#define SVP_FLUX_WAVES 1
#endif

template <int G>
__device__ __attribute__((always_inline)) inline void k_fsd_group_body(const DevMesh& m, const DevState& st, const int* wl, const int* wn) {
    static_assert(G >= 3 && G <= 8, "groups of 3 to 8 lanes");
    constexpr int NG = 64 / G;                         // groups per wave; lanes past NG * G idle
    constexpr int PPL = (NS_CELL_POINTS + G - 1) / G;  // cell points per lane
    __shared__ FaceCut rec[NG][G];
    __shared__ double full[NG][G][6];   // the lane's face as a fully submerged sub-face
    __shared__ SegTerm seg[NG][G];      // per-face terms of the cut combination
    __shared__ double pyr[NG][G];
    __shared__ double vraw[NG][NS_CELL_POINTS];
    __shared__ double vsrt[NG][NS_CELL_POINTS];
    const int g = threadIdx.x / G, l = threadIdx.x % G;
    if (g >= NG) return;
    const int nWork = wl ? *wn : *st.nMixed;
    for (int wi = blockIdx.x * NG + g; wi < nWork; wi += gridDim.x * NG) {
        const int i = wl ? wl[wi] : wi;
        const int c = st.mixed[i];
        const int fb = m.cOff[c], nF = m.cOff[c + 1] - fb;
        const int pb = m.cpOff[c], np = m.cpOff[c + 1] - pb;
        const V3 normalI = ld3(st.ifN, i);
        const double alphaI = st.alpha[c];
        const V3 Cc = ld3(m.C, c);
        const double Vc = m.V[c];
        if (mag(normalI) < SVP_TSMALL) {
            if (l == 0) {
                st.status[i] = (int)sgn(0.5 - alphaI);
                st.nCutPts[i] = 0;
                st.faceClass[i] = 0ull;
                if (!st.fsdKeep) {
                    st.ifD[i] = 0.0;
                    st3(st.ifC, i, VZERO);
                    st3(st.ifS, i, VZERO);
                }
                st.un0[i] = 0.0;
            }
            continue;
        }
        // The lane's face (faces of 3 or 4 points) and its two cell points,
        // every load unconditional (slots past the cell's faces or points,
        // or past a triangle's third point, repeat the last one): the face
        // chain and the point chain then go out together, where loads under
        // a lane condition would each wait for all the loads before them.
        const int fL = m.cFaces[fb + min(l, nF - 1)];
        const int cpA = m.cPts[pb + min(l, np - 1)], cpB = m.cPts[pb + min(l + G, np - 1)];
        const int s0 = m.fOff[fL], fszL = m.fOff[fL + 1] - s0;
        int plL[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) plL[k] = m.fPts[s0 + min(k, fszL - 1)];
        V3 qq[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) qq[k] = ld3(m.pts, plL[k]);
        const V3 ptA = ld3(m.pts, cpA), ptB = ld3(m.pts, cpB);
        double fcs[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) fcs[k] = m.fullCS[6L * fL + k];
        int fsz = 0, myFace = -1;
        V3 q4[4], q3[3];
        int pl[4] = {0, 0, 0, 0};
        if (l < nF) {
            myFace = fL;
            fsz = fszL;
#pragma unroll
            for (int k = 0; k < 4; ++k) pl[k] = k < fsz ? plL[k] : 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) q4[k] = qq[k];
#pragma unroll
            for (int k = 0; k < 3; ++k) q3[k] = qq[k];
#pragma unroll
            for (int k = 0; k < 6; ++k) full[g][l][k] = fcs[k];
        }
        // the plane's normal is fixed for the cell: dot(point, n) once
        double dots[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) dots[k] = (fsz == 4) ? dot(q4[k], normalI) : (k < 3 ? dot(q3[k], normalI) : 0.0);
        // vertex levels -n.x of the cell points, sorted descending by rank
        // (points past PPL per lane cannot occur: np <= NS_CELL_POINTS = PPL * G)
        double va = 0.0, vb = 0.0;
        if (l < np) va = -dot(normalI, ptA);
        if (l + G < np) vb = -dot(normalI, ptB);
        if (l < np) vraw[g][l] = va;
        if (l + G < np) vraw[g][l + G] = vb;
        if constexpr (PPL > 2) {
            for (int k = l + 2 * G; k < np; k += G) vraw[g][k] = -dot(normalI, ld3(m.pts, m.cPts[pb + k]));
        }
        __builtin_amdgcn_wave_barrier();
        {
            int ra = 0, rb = 0;
            for (int j = 0; j < np; ++j) {
                const double v = vraw[g][j];
                ra += (v > va) || (v == va && j < l);
                rb += (v > vb) || (v == vb && j < l + G);
            }
            if (l < np) vsrt[g][ra] = va;
            if (l + G < np) vsrt[g][rb] = vb;
            if constexpr (PPL > 2) {
                for (int k = l + 2 * G; k < np; k += G) {
                    const double vk = vraw[g][k];
                    int rk = 0;
                    for (int j = 0; j < np; ++j) {
                        const double v = vraw[g][j];
                        rk += (v > vk) || (v == vk && j < k);
                    }
                    vsrt[g][rk] = vk;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        bool ok = true;
        auto cut = [&](double dist, CellCut& cc) {
            if (l < nF) fsd_face_cut(fsz, q4, q3, normalI, dist, rec[g][l], ok, &full[g][l][0], dots);
            __builtin_amdgcn_wave_barrier();
            combineSubCellGroup(&rec[g][0], &seg[g][0], &pyr[g][0], nF, l, Vc, cc);
            __builtin_amdgcn_wave_barrier();
        };
        // cutCell::findSignedDistance (cutCell.C:611-799) as one loop around a
        // single cut, so the cut's code exists once (the kernel's instruction
        // footprint decides its speed): phase 0 bisects over the vertex
        // levels, 1 and 2 cut at a third and two thirds of the last bracket,
        // 3 is the final cut at the plane found
        double lowDistance = vsrt[g][0], upDistance = vsrt[g][np - 1];
        int lowLabel = 0, upLabel = np - 1;
        double lowAlpha = 0.0, upAlpha = 1.0;
        CellCut cc;
        double dOut = 0.0, deltaDistance = 0.0, alphaOneThird = 0.0;
        int phase = 0, midLabel = 0;
        while (true) {
            double dist;
            if (phase == 0) {
                if ((upLabel - lowLabel) > 1) {
                    midLabel = (int)round(0.5 * (double)(upLabel + lowLabel));
                    dist = vsrt[g][midLabel];
                } else if (fabs(lowDistance - upDistance) < SVP_TSMALL) {
                    dOut = 0.5 * (lowDistance + upDistance);
                    dist = dOut;
                    phase = 3;
                } else {
                    deltaDistance = (upDistance - lowDistance) / 3.0;
                    dist = lowDistance + deltaDistance;
                    phase = 1;
                }
            } else if (phase == 2) {
                dist = lowDistance + 2.0 * deltaDistance;
            } else {
                dist = dOut;
            }
            cut(dist, cc);
            if (phase == 0) {
                const double midDistance = dist;
                const double midAlpha = cc.VOF;
                if (fabs(midAlpha - alphaI) < SVP_TSMALL) {
                    dOut = midDistance;
                    break;
                }
                if (midAlpha > alphaI) {
                    upLabel = midLabel;
                    upDistance = midDistance;
                    upAlpha = midAlpha;
                } else {
                    lowLabel = midLabel;
                    lowDistance = midDistance;
                    lowAlpha = midAlpha;
                }
            } else if (phase == 1) {
                alphaOneThird = cc.VOF - lowAlpha;
                phase = 2;
            } else if (phase == 2) {
                const double alphaPrismatoid = upAlpha - lowAlpha;
                const double alphaTwoThirds = cc.VOF - lowAlpha;
                const double a = 13.5 * alphaOneThird - 13.5 * alphaTwoThirds + 4.5 * alphaPrismatoid;
                const double b = -22.5 * alphaOneThird + 18.0 * alphaTwoThirds - 4.5 * alphaPrismatoid;
                const double cq = 9.0 * alphaOneThird - 4.5 * alphaTwoThirds + 1.0 * alphaPrismatoid;
                const double d = lowAlpha - alphaI;
                double lambda = 0.5;
                for (int iter = 0; iter < 100; iter++) {
                    const double func = a * (lambda * lambda * lambda) + b * (lambda * lambda) + cq * lambda + d;
                    const double funcPrime = 3.0 * a * (lambda * lambda) + 2.0 * b * lambda + cq;
                    const double lambdaNew = lambda - (func / funcPrime);
                    if (fabs(lambdaNew - lambda) < SVP_TSMALL) break;
                    lambda = lambdaNew;
                }
                dOut = lowDistance - lambda * (lowDistance - upDistance);
                phase = 3;
            } else {
                break;
            }
        }
        if (l == 0) {
            unsigned long long fcls = 0ull;
            int ncp = 0;
            for (int k = 0; k < nF; ++k) {
                const int s2 = rec[g][k].status;
                fcls |= (unsigned long long)(s2 + 1) << (2 * k);
                if (s2 == 0) ncp += rec[g][k].nIf;
            }
            st.status[i] = cc.status;
            st.nCutPts[i] = ncp;
            st.faceClass[i] = fcls;
            st.ifD[i] = dOut;
            st3(st.ifC, i, cc.ifCentre);
            st3(st.ifS, i, cc.ifArea);
        }
        if (!ok) atomicOr(st.errFlag, ERR_TOO_MANY_CROSSINGS);
        __builtin_amdgcn_wave_barrier();
        // Un0 = U(interface centre) . n for a cut cell (advection.C:126):
        // interpolationCellPoint's tet search with one lane per face
        double un0 = 0.0;
        if (cc.status == 0) {
            const double tol = SVP_SMALL;
            const V3 pos = cc.ifCentre;
            bool found = false;
            double fw[4] = {0, 0, 0, 0}, bw[4] = {0, 0, 0, 0}, bestMin = -1e300;
            int fp[3] = {0, 0, 0}, bp[3] = {0, 0, 0};
            if (l < nF) {
                const bool ownFace = (m.own[myFace] == c);
#pragma unroll
                for (int t = 1; t < 3; ++t) {
                    if (t >= fsz - 1 || found) continue;
                    int a0 = pl[0], a1 = (t == 1) ? pl[1] : pl[2], a2 = (t == 1) ? pl[2] : pl[3];
                    V3 b0 = (fsz == 4) ? q4[0] : q3[0];
                    V3 b1 = (fsz == 4) ? ((t == 1) ? q4[1] : q4[2]) : q3[1];
                    V3 b2 = (fsz == 4) ? ((t == 1) ? q4[2] : q4[3]) : q3[2];
                    if (!ownFace) {
                        const int ti = a1; a1 = a2; a2 = ti;
                        const V3 tv = b1; b1 = b2; b2 = tv;
                    }
                    const V3 e0 = Cc - b2, e1 = b0 - b2, e2 = b1 - b2;
                    const double detT = dot(e0, cross(e1, e2));
                    double w[4];
                    if (fabs(detT) < SVP_SMALL) {
                        w[0] = w[1] = w[2] = w[3] = -1.0;
                    } else {
                        const V3 r = pos - b2;
                        w[0] = dot(r, cross(e1, e2)) / detT;
                        w[1] = dot(e0, cross(r, e2)) / detT;
                        w[2] = dot(e0, cross(e1, r)) / detT;
                        w[3] = 1.0 - w[0] - w[1] - w[2];
                    }
                    const double mn = fmin(fmin(w[0], w[1]), fmin(w[2], w[3]));
                    if (mn + tol > 0.0) {
                        found = true;
                        fw[0] = w[0]; fw[1] = w[1]; fw[2] = w[2]; fw[3] = w[3];
                        fp[0] = a0; fp[1] = a1; fp[2] = a2;
                    } else if (mn > bestMin) {
                        bestMin = mn;
                        bw[0] = w[0]; bw[1] = w[1]; bw[2] = w[2]; bw[3] = w[3];
                        bp[0] = a0; bp[1] = a1; bp[2] = a2;
                    }
                }
            }
            // first face (in cell-face order) with a containing tet, else the
            // tet of largest minimum weight, earliest on ties
            const unsigned long long grpMask = ((1ull << G) - 1ull) << (g * G);
            const unsigned long long foundMask = __ballot(found && l < nF) & grpMask;
            int src = -1;
            if (foundMask) {
                src = __ffsll((long long)foundMask) - 1 - g * G;
            } else {
                double best = -1e300;
                for (int k = 0; k < nF; ++k) {
                    const double v = __shfl(bestMin, g * G + k);
                    if (v > best) {
                        best = v;
                        src = k;
                    }
                }
            }
            double w[4];
            int pp[3];
#pragma unroll
            for (int q = 0; q < 4; ++q) w[q] = __shfl(foundMask ? fw[q] : bw[q], g * G + src);
#pragma unroll
            for (int q = 0; q < 3; ++q) pp[q] = __shfl(foundMask ? fp[q] : bp[q], g * G + src);
            // point values: lane q of the group evaluates point q
            V3 pu = VZERO;
            if (l < 3) pu = point_u(m, st, (l == 0) ? pp[0] : (l == 1 ? pp[1] : pp[2]));
            V3 pv[3];
#pragma unroll
            for (int q = 0; q < 3; ++q)
                pv[q] = v3(__shfl(pu.x, g * G + q), __shfl(pu.y, g * G + q), __shfl(pu.z, g * G + q));
            const V3 Ui = w[0] * ld3(st.U, c) + w[1] * pv[0] + w[2] * pv[1] + w[3] * pv[2];
            un0 = dot(Ui, normalI);
        }
        if (l == 0) st.un0[i] = un0;
        __builtin_amdgcn_wave_barrier();
    }
}

template <int G>
__global__ void __launch_bounds__(64, SVP_FSD_WAVES) k_fsd_group(DevMesh m, DevState st, const int* wl, const int* wn) {
    k_fsd_group_body<G>(m, st, wl, wn);
}

void launch_find_signed_distance(const DevMesh& m, const DevState& st, const Params& p, hipStream_t s,
                                 const int* wl, const int* wn) {
    if (p.splitWarpedFace) k_find_signed_distance<CellPolySP><<<1024, 64, 0, s>>>(m, st, 1, wl, wn);
    else if (st.un0Fused && st.maxCellFaces <= 4) k_fsd_group<4><<<2048, 64, 0, s>>>(m, st, wl, wn);
    else if (st.un0Fused && st.maxCellFaces <= 6) k_fsd_group<6><<<2048, 64, 0, s>>>(m, st, wl, wn);
    else if (st.un0Fused) k_fsd_group<8><<<2048, 64, 0, s>>>(m, st, wl, wn);
    else if (st.maxCellFaces <= NS_CELL_FACES && m.maxCellPoints <= NS_CELL_POINTS)
        k_find_signed_distance<CellPolyNS><<<1024, 64, 0, s>>>(m, st, 0, wl, wn);
    else k_find_signed_distance<CellPolyGen><<<1024, 64, 0, s>>>(m, st, 0, wl, wn);
}

// ---------------------------------------------------------------------------
// PLIC surface: reconstruction::interface (reconstruction.C:787-835) with
// cutCell::interfacePoints (cutCell.C:545-608), one lane per mixed cell
// ---------------------------------------------------------------------------
// Vector::normalise of OpenFOAM: zero below ROOTVSMALL, else divided by its
// magnitude component by component
__device__ inline V3 normalised(V3 v) {
    const double s = mag(v);
    return s < SVP_ROOTVSMALL ? VZERO : v / s;
}

// The plane crossings of the cut faces (face order, crossing order) sorted
// by their angle about the interface centre, in a frame with z along the
// interface area vector and x towards the first crossing; a crossing whose
// angle lies within 1e-8 of the previous one in that order is dropped.  The
// insertion sort is stable, like sortedOrder.  Returns the polygon size.
template <int K>
__device__ inline int interface_points(const FaceCut* fc, int nF, const CellCut& cut, V3 (&q)[K]) {
    V3 raw[K];
    double ang[K];
    int ord[K];
    int n = 0;
    for (int f = 0; f < nF; ++f) {
        if (fc[f].status != 0) continue;
        for (int k = 0; k < fc[f].nIf && n < K; ++k) raw[n++] = fc[f].ifPts[k];
    }
    if (n == 0) return 0;
    const V3 C = cut.ifCentre;
    const V3 zhat = cut.ifArea / mag(cut.ifArea);
    V3 xhat = raw[0] - C;
    xhat = xhat - dot(xhat, zhat) * zhat;
    xhat = normalised(xhat);
    const V3 yhat = normalised(cross(zhat, xhat));
    for (int k = 0; k < n; ++k) {
        const V3 r = raw[k] - C;
        ang[k] = atan2(dot(r, yhat), dot(r, xhat));
        ord[k] = k;
    }
    for (int a = 1; a < n; ++a) {
        const int t = ord[a];
        const double ta = ang[t];
        int b = a - 1;
        while (b >= 0 && ang[ord[b]] > ta) {
            ord[b + 1] = ord[b];
            --b;
        }
        ord[b + 1] = t;
    }
    int out = 0;
    q[out++] = raw[ord[0]];
    for (int k = 1; k < n; ++k)
        if (fabs(ang[ord[k]] - ang[ord[k - 1]]) > 1e-8) q[out++] = raw[ord[k]];
    return out;
}

// Count mode (pts == nullptr): ptOff[i] = polygon size of mixed cell i,
// faceIdx[i] = 1 if it has a polygon.  Write mode: ptOff and faceIdx hold the
// exclusive scans of those; the polygon goes to pts[ptOff[i]..], its cell and
// first point to faceCells / faceOff at faceIdx[i].  Every mixed cell is cut
// again through its mesh faces (no warped-face split), as the reference does.
template <class Poly>
__global__ void __launch_bounds__(64) k_plic_polygons(DevMesh m, DevState st, int nMixed, int* ptOff, int* faceIdx,
                                                      double* pts, int* faceCells, int* faceOff) {
    constexpr int K = Poly::maxFaces * MAX_IF_POINTS;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nMixed; i += gridDim.x * blockDim.x) {
        const int c = st.mixed[i];
        Poly P;
        FaceCut fc[Poly::maxFaces];
        CellCut cut;
        V3 q[K];
        int n = 0;
        if (!build_cell_poly(m, c, false, P)) {
            atomicOr(st.errFlag, ERR_CELL_TOO_LARGE);
        } else {
            if (!calcSubCell(P, ld3(m.C, c), m.V[c], ld3(st.ifN, i), st.ifD[i], fc, cut))
                atomicOr(st.errFlag, ERR_TOO_MANY_CROSSINGS);
            if (cut.status == 0) n = interface_points(fc, P.nFaces, cut, q);
        }
        if (!pts) {
            ptOff[i] = n;
            faceIdx[i] = n > 0 ? 1 : 0;
            continue;
        }
        if (n == 0) continue;
        const int o = ptOff[i], fi = faceIdx[i];
        faceCells[fi] = c;
        faceOff[fi] = o;
        for (int k = 0; k < n; ++k) st3(pts, o + k, q[k]);
    }
}

// reconstruction::mapAlphaField (reconstruction.C:725-784): every cell whose
// alpha lies in [lower, upper] takes the volume fraction of its plane (n, d)
// cut through its mesh faces (no warped-face split); n and d are per cell.
template <class Poly>
__global__ void __launch_bounds__(64) k_map_alpha(DevMesh m, DevState st, const double* __restrict__ n,
                                                  const double* __restrict__ d, double lower, double upper) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < m.nC; c += gridDim.x * blockDim.x) {
        const double a = st.alpha[c];
        if (!(a >= lower && a <= upper)) continue;
        Poly P;
        FaceCut fc[Poly::maxFaces];
        CellCut cut;
        if (!build_cell_poly(m, c, false, P)) {
            atomicOr(st.errFlag, ERR_CELL_TOO_LARGE);
            continue;
        }
        if (!calcSubCell(P, ld3(m.C, c), m.V[c], ld3(n, c), d[c], fc, cut))
            atomicOr(st.errFlag, ERR_TOO_MANY_CROSSINGS);
        st.alpha[c] = cut.VOF;
    }
}

void launch_map_alpha(const DevMesh& m, const DevState& st, const double* n, const double* d, double lower,
                      double upper, hipStream_t s) {
    if (m.nC <= 0) return;
    if (st.maxCellFaces <= NS_CELL_FACES && m.maxCellPoints <= NS_CELL_POINTS)
        k_map_alpha<CellPolyNS><<<std::min(4096, cdiv(m.nC, 64)), 64, 0, s>>>(m, st, n, d, lower, upper);
    else
        k_map_alpha<CellPolyGen><<<std::min(4096, cdiv(m.nC, 64)), 64, 0, s>>>(m, st, n, d, lower, upper);
}

void launch_plic_surface(const DevMesh& m, const DevState& st, int nMixed, int pass, int* ptOff, int* faceIdx,
                         int* totals, double* pts, int* faceCells, int* faceOff, hipStream_t s) {
    if (nMixed <= 0) return;
    const int grid = std::min(1024, cdiv(nMixed, 64));
    const bool gen = st.maxCellFaces > NS_CELL_FACES || m.maxCellPoints > NS_CELL_POINTS;
    if (pass == 0) {
        if (gen) k_plic_polygons<CellPolyGen><<<grid, 64, 0, s>>>(m, st, nMixed, ptOff, faceIdx, nullptr, nullptr, nullptr);
        else k_plic_polygons<CellPolyNS><<<grid, 64, 0, s>>>(m, st, nMixed, ptOff, faceIdx, nullptr, nullptr, nullptr);
        k_scan_counts<<<1, 1024, 0, s>>>(ptOff, nMixed, totals, nullptr);
        k_scan_counts<<<1, 1024, 0, s>>>(faceIdx, nMixed, totals + 1, nullptr);
    } else {
        if (gen) k_plic_polygons<CellPolyGen><<<grid, 64, 0, s>>>(m, st, nMixed, ptOff, faceIdx, pts, faceCells, faceOff);
        else k_plic_polygons<CellPolyNS><<<grid, 64, 0, s>>>(m, st, nMixed, ptOff, faceIdx, pts, faceCells, faceOff);
    }
}

// The full-mesh interface fields of the reference (interfaceN/D/C/S,
// reconstruction.C:520-570): zero in every cell (initialize,
// reconstruction.C:638-641), then the values of the mixed cells from the
// compact arrays.  Any output may be null.
__global__ void k_plic_fields_zero(int nC, double* N, double* D, double* C, double* S) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < 3L * nC; i += (long)gridDim.x * blockDim.x) {
        if (N) N[i] = 0.0;
        if (C) C[i] = 0.0;
        if (S) S[i] = 0.0;
        if (D && i < nC) D[i] = 0.0;
    }
}

__global__ void k_plic_fields_scatter(DevState st, double* N, double* D, double* C, double* S) {
    const int k = *st.nMixed;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) {
        const int c = st.mixed[i];
        if (N) st3(N, c, ld3(st.ifN, i));
        if (C) st3(C, c, ld3(st.ifC, i));
        if (S) st3(S, c, ld3(st.ifS, i));
        if (D) D[c] = st.ifD[i];
    }
}

void launch_plic_fields(const DevMesh& m, const DevState& st, double* N, double* D, double* C, double* S,
                        hipStream_t s) {
    k_plic_fields_zero<<<std::min(4096, cdiv(3L * m.nC, 256)), 256, 0, s>>>(m.nC, N, D, C, S);
    k_plic_fields_scatter<<<256, 256, 0, s>>>(st, N, D, C, S);
}

// the closing polygon offset face_offsets[n_faces] = n_points, from the
// totals of the size pass (totals[0] points, totals[1] polygons); without
// totals (no mixed cell) the one offset face_offsets[0] = 0
__global__ void k_plic_close(int* faceOff, const int* totals) {
    if (threadIdx.x != 0) return;
    if (totals) faceOff[totals[1]] = totals[0];
    else faceOff[0] = 0;
}

void launch_plic_close(int* faceOff, const int* totals, hipStream_t s) {
    k_plic_close<<<1, 64, 0, s>>>(faceOff, totals);
}

// ---------------------------------------------------------------------------
// Advection
// ---------------------------------------------------------------------------


// dVf = upwind(phi).flux(alpha)*dt (advectionTemplates.C:371): face value
// lambda*(aP - aN) + aN with lambda = pos0(phi); zeroGradient boundary value
// is the owner value; empty patches carry nothing.
__device__ __attribute__((always_inline)) inline void k_upwind_body(const DevMesh& m, const DevState& st, double dtStep) {
    // the out-of-bounds counters of the update and the flux work-list count
    // of the interface flux that follow start at zero
    if (blockIdx.x == 0) {
        for (int k = threadIdx.x; k < OOB_SLOTS * OOB_STRIDE; k += blockDim.x) st.oobCount[k] = 0;
        if (threadIdx.x == 0) *st.fluxCount = 0;
    }
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nF) return;
    const double dt = dtStep;
    const double ph = st.phi[f];
    double v;
    if (f < m.nIF) {
        const double aP = st.alphaOld[m.own[f]], aN = st.alphaOld[m.nbr[f]];
        const double af = pos0(ph) * (aP - aN) + aN;
        v = (ph * af) * dt;
    } else if (m.fEmpty[f]) {
        v = 0.0;
    } else {
        v = (ph * st.alphaOld[m.own[f]]) * dt;
    }
    st.dVf[f] = v;
}

__global__ void __launch_bounds__(256) k_upwind(DevMesh m, DevState st, double dtStep) {
    k_upwind_body(m, st, dtStep);
}


// interpolationCellPoint<vector>::interpolate(position, cell) (advection.C:91,126):
// barycentric weights in the tets (cell centre, face point 0, fan triangle).
__device__ inline V3 interpolate_u(const DevMesh& m, const DevState& st, V3 pos, int c) {
    const double tol = SVP_SMALL;
    const V3 cc = ld3(m.C, c);
    double bestMin = -1e300;
    double bw[4] = {0.0, 0.0, 0.0, 0.0};
    int bp[3] = {0, 0, 0};
    for (int k = m.cOff[c]; k < m.cOff[c + 1]; ++k) {
        const int f = m.cFaces[k];
        const int fs = m.fOff[f], n = m.fOff[f + 1] - fs;
        for (int t = 1; t < n - 1; ++t) {
            int p0 = m.fPts[fs], p1 = m.fPts[fs + t], p2 = m.fPts[fs + t + 1];
            if (m.own[f] != c) {
                const int tmp = p1;
                p1 = p2;
                p2 = tmp;
            }
            const V3 b = ld3(m.pts, p0), cq = ld3(m.pts, p1), d = ld3(m.pts, p2);
            const V3 e0 = cc - d, e1 = b - d, e2 = cq - d;
            const double detT = dot(e0, cross(e1, e2));
            double w[4];
            if (fabs(detT) < SVP_SMALL) {
                w[0] = w[1] = w[2] = w[3] = -1.0;
            } else {
                const V3 r = pos - d;
                w[0] = dot(r, cross(e1, e2)) / detT;
                w[1] = dot(e0, cross(r, e2)) / detT;
                w[2] = dot(e0, cross(e1, r)) / detT;
                w[3] = 1.0 - w[0] - w[1] - w[2];
            }
            const double mn = fmin(fmin(w[0], w[1]), fmin(w[2], w[3]));
            if (mn + tol > 0.0) {
                return w[0] * ld3(st.U, c) + w[1] * point_u(m, st, p0) + w[2] * point_u(m, st, p1) +
                       w[3] * point_u(m, st, p2);
            }
            if (mn > bestMin) {
                bestMin = mn;
                bw[0] = w[0]; bw[1] = w[1]; bw[2] = w[2]; bw[3] = w[3];
                bp[0] = p0; bp[1] = p1; bp[2] = p2;
            }
        }
    }
    return bw[0] * ld3(st.U, c) + bw[1] * point_u(m, st, bp[0]) + bw[2] * point_u(m, st, bp[1]) +
           bw[3] * point_u(m, st, bp[2]);
}

// cutFace::timeIntegratedArea (cutFace.C:392-506)
__device__ inline double time_integrated_area(const V3* fPts, int n, V3 normal, double distance,
                                              const double* pTimesIn, double Un0, double dt,
                                              double magSf, bool& ok) {
    double pt[MAX_FACE_POINTS];
    for (int k = 0; k < n; ++k) pt[k] = pTimesIn[k];
    for (int k = 1; k < n; ++k) {
        const double v = pt[k];
        int j = k - 1;
        while (j >= 0 && pt[j] > v) {
            pt[j + 1] = pt[j];
            --j;
        }
        pt[j + 1] = v;
    }
    const double firstTime = pt[0], lastTime = pt[n - 1];
    if (lastTime <= 0.0) return magSf * dt * pos0(Un0);
    if (firstTime >= dt) return magSf * dt * (1.0 - pos0(Un0));
    double tIntArea = 0.0;
    double sortedTimes[MAX_FACE_POINTS + 2];
    int nSorted = 0;
    double prevTime;
    double subAreaOld;
    FaceCut fc;
    if (firstTime > 0.0) {
        subAreaOld = magSf * (1.0 - pos0(Un0));
        tIntArea = subAreaOld * firstTime;
        sortedTimes[nSorted++] = firstTime;
        prevTime = firstTime;
    } else {
        sortedTimes[nSorted++] = 0.0;
        prevTime = 0.0;
        ok &= calcSubFace(fPts, n, normal, distance, fc);
        subAreaOld = mag(fc.area);
    }
    const double smallTime = fmax(SVP_TSMALL / fabs(Un0), SVP_TSMALL);
    for (int ti = 0; ti < n; ++ti) {
        const double timeI = pt[ti];
        if (timeI > (prevTime + smallTime) && timeI < dt) {
            sortedTimes[nSorted++] = timeI;
            prevTime = timeI;
        }
    }
    if (lastTime > dt) sortedTimes[nSorted++] = dt;
    else tIntArea += magSf * (dt - lastTime) * pos0(Un0);
    for (int k = 0; k < nSorted - 1; k++) {
        const double tauOld = sortedTimes[k], tauNew = sortedTimes[k + 1];
        const double deltaTau = 0.5 * (tauNew - tauOld);
        ok &= calcSubFace(fPts, n, normal, distance - tauNew * Un0, fc);
        const double subAreaNew = mag(fc.area);
        ok &= calcSubFace(fPts, n, normal, distance - (tauOld + deltaTau) * Un0, fc);
        const double subAreaMid = mag(fc.area);
        tIntArea += (deltaTau / 3.0) * (subAreaOld + 4.0 * subAreaMid + subAreaNew);
        subAreaOld = subAreaNew;
    }
    return tIntArea;
}

// cutFace::timeIntegratedFaceFlux (cutFace.C:262-389)
__device__ inline double time_integrated_face_flux(const DevMesh& m, int f, V3 normal, double distance,
                                                   double Un0, double dt, double phi, double magSf,
                                                   bool& ok) {
    if (fabs(phi) <= SVP_TSMALL) return 0.0;
    const int fs = m.fOff[f], n = m.fOff[f + 1] - fs;
    V3 p[MAX_FACE_POINTS];
    for (int k = 0; k < n; ++k) p[k] = ld3(m.pts, m.fPts[fs + k]);
    const bool flat = m.flat[f] > (1.0 - SVP_TSMALL);
    if (fabs(Un0 * dt) > SVP_TSMALL) {
        double pTimes[MAX_FACE_POINTS];
        for (int i = 0; i < n; ++i) {
            const double t = (dot(p[i], normal) + distance) / Un0;
            pTimes[i] = fabs(t) < SVP_TSMALL ? 0.0 : t;
        }
        if (flat) {
            return phi / magSf * time_integrated_area(p, n, normal, distance, pTimes, Un0, dt, magSf, ok);
        }
        double dVf = 0.0;
        V3 tri[3];
        double tt[3];
        tri[0] = ld3(m.Cf, f);
        const double t0 = (dot(tri[0], normal) + distance) / Un0;
        tt[0] = fabs(t0) < SVP_TSMALL ? 0.0 : t0;
        for (int pi = 0; pi < n; ++pi) {
            tri[1] = p[pi];
            tt[1] = pTimes[pi];
            tri[2] = p[(pi + 1) % n];
            tt[2] = pTimes[(pi + 1) % n];
            const double magSfTri = mag(triAreaNormal(tri[0], tri[1], tri[2]));
            const double phiTri = phi * magSfTri / magSf;
            dVf += phiTri / magSfTri *
                   time_integrated_area(tri, 3, normal, distance, tt, Un0, dt, magSfTri, ok);
        }
        return dVf;
    }
    FaceCut fc;
    if (flat) {
        ok &= calcSubFace(p, n, normal, distance, fc);
        const double alphaf = mag(fc.area) / magSf;
        return (phi * dt * alphaf);
    }
    V3 tri[3];
    tri[0] = ld3(m.Cf, f);
    double dVf = 0.0;
    for (int pi = 0; pi < n; ++pi) {
        tri[1] = p[pi];
        tri[2] = p[(pi + 1) % n];
        const double magSfTri = mag(triAreaNormal(tri[0], tri[1], tri[2]));
        const double phiTri = phi * magSfTri / magSf;
        ok &= calcSubFace(tri, 3, normal, distance, fc);
        const double alphafTri = mag(fc.area) / magSfTri;
        dVf += (phiTri * dt * alphafTri);
    }
    return dVf;
}

// timeIntegratedArea for a face of N points in registers: the sorted arrival
// times are walked once and every accepted sub-interval gets its Simpson term
// right away; the additions to tIntArea keep the reference's order (leading
// uncut segment, trailing uncut segment, then the Simpson terms in time order).
// FULL: full points to the face's fully submerged sub-face (fullCS)
template <int N, bool FULL = false>
__device__ inline double tia_regs(const V3 (&P)[N], const double (&ptIn)[N], V3 normal, double distance,
                                  double Un0, double dt, double magSf, bool& ok, const double* full = nullptr) {
    // the plane normal is fixed over the time step: dot(point, n) once
    double dots[N];
#pragma unroll
    for (int k = 0; k < N; ++k) dots[k] = dot(P[k], normal);
    double pt[N];
#pragma unroll
    for (int k = 0; k < N; ++k) pt[k] = ptIn[k];
    // ascending sort (values only): odd-even transposition, static indices
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
        for (int k = r & 1; k + 1 < N; k += 2) {
            const double a = pt[k], b = pt[k + 1];
            pt[k] = fmin(a, b);
            pt[k + 1] = fmax(a, b);
        }
    const double firstTime = pt[0], lastTime = pt[N - 1];
    if (lastTime <= 0.0) return magSf * dt * pos0(Un0);
    if (firstTime >= dt) return magSf * dt * (1.0 - pos0(Un0));
    double tIntArea = 0.0;
    double prevTime, tauOld, subAreaOld;
    FaceCut fc;
    if (firstTime > 0.0) {
        subAreaOld = magSf * (1.0 - pos0(Un0));
        tIntArea = subAreaOld * firstTime;
        prevTime = firstTime;
        tauOld = firstTime;
    } else {
        prevTime = 0.0;
        tauOld = 0.0;
        ok &= calcSubFaceN<N, false, FULL, true>(P, normal, distance, fc, full, dots);
        subAreaOld = mag(fc.area);
    }
    if (!(lastTime > dt)) tIntArea += magSf * (dt - lastTime) * pos0(Un0);
    const double smallTime = fmax(SVP_TSMALL / fabs(Un0), SVP_TSMALL);
    auto simpson = [&](double tauNew) {
        const double deltaTau = 0.5 * (tauNew - tauOld);
        ok &= calcSubFaceN<N, false, FULL, true>(P, normal, distance - tauNew * Un0, fc, full, dots);
        const double subAreaNew = mag(fc.area);
        ok &= calcSubFaceN<N, false, FULL, true>(P, normal, distance - (tauOld + deltaTau) * Un0, fc, full, dots);
        const double subAreaMid = mag(fc.area);
        tIntArea += (deltaTau / 3.0) * (subAreaOld + 4.0 * subAreaMid + subAreaNew);
        subAreaOld = subAreaNew;
        tauOld = tauNew;
    };
#pragma unroll
    for (int ti = 0; ti < N; ++ti) {
        const double timeI = pt[ti];
        if (timeI > (prevTime + smallTime) && timeI < dt) {
            simpson(timeI);
            prevTime = timeI;
        }
    }
    if (lastTime > dt) simpson(dt);
    return tIntArea;
}

// timeIntegratedFaceFlux (cutFace.C:262-389) for a face of N points in registers
template <int N>
__device__ __attribute__((always_inline)) inline double tiff_regs(const DevMesh& m, int f, const V3 (&p)[N], V3 normal, double distance,
                                   double Un0, double dt, double phi, double magSf, bool& ok) {
    if (fabs(phi) <= SVP_TSMALL) return 0.0;
    const bool flat = m.flat[f] > (1.0 - SVP_TSMALL);
    if (fabs(Un0 * dt) > SVP_TSMALL) {
        double pTimes[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const double t = (dot(p[i], normal) + distance) / Un0;
            pTimes[i] = fabs(t) < SVP_TSMALL ? 0.0 : t;
        }
        if (flat) return phi / magSf * tia_regs<N, true>(p, pTimes, normal, distance, Un0, dt, magSf, ok, &m.fullCS[6L * f]);
        double dVf = 0.0;
        const V3 fc = ld3(m.Cf, f);
        const double t0 = (dot(fc, normal) + distance) / Un0;
        const double tc = fabs(t0) < SVP_TSMALL ? 0.0 : t0;
        // fan triangle pi uses points pi and pi+1: the arrays rotate one slot
        // per triangle so that those are always slots 0 and 1
        V3 rp[N];
        double rt[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            rp[i] = p[i];
            rt[i] = pTimes[i];
        }
#pragma unroll 1
        for (int pi = 0; pi < N; ++pi) {
            const V3 tri[3] = {fc, rp[0], rp[1]};
            const double tt[3] = {tc, rt[0], rt[1]};
            rotate_left(rp, 1);
            rotate_left(rt, 1);
            const double magSfTri = mag(triAreaNormal(tri[0], tri[1], tri[2]));
            const double phiTri = phi * magSfTri / magSf;
            dVf += phiTri / magSfTri * tia_regs<3>(tri, tt, normal, distance, Un0, dt, magSfTri, ok);
        }
        return dVf;
    }
    FaceCut fcut;
    if (flat) {
        ok &= calcSubFaceN<N, false>(p, normal, distance, fcut);
        const double alphaf = mag(fcut.area) / magSf;
        return (phi * dt * alphaf);
    }
    const V3 fc = ld3(m.Cf, f);
    double dVf = 0.0;
#pragma unroll
    for (int pi = 0; pi < N; ++pi) {
        const V3 tri[3] = {fc, p[pi], p[(pi + 1) % N]};
        const double magSfTri = mag(triAreaNormal(tri[0], tri[1], tri[2]));
        const double phiTri = phi * magSfTri / magSf;
        ok &= calcSubFaceN<3, false>(tri, normal, distance, fcut);
        const double alphafTri = mag(fcut.area) / magSfTri;
        dVf += (phiTri * dt * alphafTri);
    }
    return dVf;
}

// timeIntegratedFlux of the listed (cut cell, downwind face) pairs with four
// lanes per face (faces of 3 or 4 points).  On a warped face, which the
// reference integrates as the fan of triangles about the face centre
// (cutFace.C:316-345 with a time step, 369-386 without), lane j integrates
// triangle j (points j and j+1) and the group's first lane adds the
// triangle volumes in fan order, the reference's sum; a flat face (one
// timeIntegratedArea of the whole face) runs on the first lane alone.  Every
// triangle's volume is the expression tiff_regs evaluates, so the face
// volumes are identical to k_interface_flux_faces.
constexpr int FAN_LANES = 4;

__device__ __attribute__((always_inline)) inline void k_flux_fan_body(const DevMesh& m, const DevState& st, double dtStep, int slots) {
    const long nWork = (long)*st.fluxCount;
    const int lane = threadIdx.x & 63, j = lane & (FAN_LANES - 1), base = lane & ~(FAN_LANES - 1);
    constexpr int PER_WAVE = 64 / FAN_LANES;
    // one wave per workgroup; the loop bound is uniform across the wave, so
    // every lane reaches the shuffles
    for (long w0 = (long)blockIdx.x * PER_WAVE; w0 < nWork; w0 += (long)gridDim.x * PER_WAVE) {
        const long wi = w0 + lane / FAN_LANES;
        const bool live = wi < nWork;
        int f = 0, i = 0, fsz = 3, s0 = 0;
        double ph = 0.0;
        if (live) {
            const int w = st.fluxList[wi];
            i = w / slots;
            const int k = w % slots;
            f = m.cFaces[m.cOff[st.mixed[i]] + k];
            ph = st.phi[f];
            s0 = m.fOff[f];
            fsz = m.fOff[f + 1] - s0;
        }
        const double dt = dtStep;
        const bool zero = !live || fabs(ph) <= SVP_TSMALL;
        const bool flat = live && m.flat[f] > (1.0 - SVP_TSMALL);
        const bool fan = !zero && !flat;
        double term = 0.0;
        bool ok = true;
        if (!zero) {
            const V3 n = ld3(st.ifN, i);
            const double d = st.ifD[i], Un0 = st.un0[i];
            const double magSf = m.magSf[f];
            if (fan && j < fsz) {
                const V3 fc = ld3(m.Cf, f);
                const V3 pa = ld3(m.pts, m.fPts[s0 + j]);
                const V3 pb = ld3(m.pts, m.fPts[s0 + (j + 1 == fsz ? 0 : j + 1)]);
                const V3 tri[3] = {fc, pa, pb};
                const double magSfTri = mag(triAreaNormal(tri[0], tri[1], tri[2]));
                const double phiTri = ph * magSfTri / magSf;
                if (fabs(Un0 * dt) > SVP_TSMALL) {
                    auto arrival = [&](V3 q) {
                        const double t = (dot(q, n) + d) / Un0;
                        return fabs(t) < SVP_TSMALL ? 0.0 : t;
                    };
                    const double tt[3] = {arrival(fc), arrival(pa), arrival(pb)};
                    term = phiTri / magSfTri * tia_regs<3>(tri, tt, n, d, Un0, dt, magSfTri, ok);
                } else {
                    FaceCut fcut;
                    ok &= calcSubFaceN<3, false>(tri, n, d, fcut);
                    const double alphafTri = mag(fcut.area) / magSfTri;
                    term = (phiTri * dt * alphafTri);
                }
            } else if (!fan && j == 0) {
                if (fsz == 4) {
                    V3 p[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) p[q] = ld3(m.pts, m.fPts[s0 + q]);
                    term = tiff_regs<4>(m, f, p, n, d, Un0, dt, ph, magSf, ok);
                } else {
                    V3 p[3];
#pragma unroll
                    for (int q = 0; q < 3; ++q) p[q] = ld3(m.pts, m.fPts[s0 + q]);
                    term = tiff_regs<3>(m, f, p, n, d, Un0, dt, ph, magSf, ok);
                }
            }
        }
        double v = term;
        if (fan) {
            v = 0.0;
#pragma unroll
            for (int t = 0; t < FAN_LANES; ++t) {
                const double x = __shfl(term, base + t);
                if (t < fsz) v += x;
            }
        } else {
#pragma unroll
            for (int t = 0; t < FAN_LANES; ++t) (void)__shfl(term, base + t);
        }
        if (live && j == 0) st.dVf[f] = zero ? 0.0 : v;
        if (!ok) atomicOr(st.errFlag, ERR_TOO_MANY_CROSSINGS);
    }
}

__global__ void __launch_bounds__(64, SVP_FLUX_WAVES) k_flux_fan(DevMesh m, DevState st, double dtStep, int slots) {
    k_flux_fan_body(m, st, dtStep, slots);
}

// faces whose flatness is below 1 - tSmall (the fan-integrated ones)
__global__ void __launch_bounds__(256) k_count_warped(DevMesh m, int* count) {
    const int f = blockIdx.x * 256 + threadIdx.x;
    const bool w = f < m.nF && !(m.flat[f] > (1.0 - SVP_TSMALL));
    const unsigned long long b = __ballot(w);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(count, __popcll(b));
}

__global__ void k_zero_one(int* p) {
    if (threadIdx.x == 0) *p = 0;
}

void launch_count_warped(const DevMesh& m, int* count, hipStream_t s) {
    k_zero_one<<<1, 64, 0, s>>>(count);
    k_count_warped<<<cdiv(m.nF, 256), 256, 0, s>>>(m, count);
}

// Un0 = U(interface centre) . n of every cut cell (advection.C:126)
__global__ void __launch_bounds__(64) k_un0(DevMesh m, DevState st) {
    const int nMixed = *st.nMixed;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nMixed; i += gridDim.x * blockDim.x) {
        double u = 0.0;
        if (st.status[i] == 0) u = dot(interpolate_u(m, st, ld3(st.ifC, i), st.mixed[i]), ld3(st.ifN, i));
        st.un0[i] = u;
    }
}

// The (cut cell, cell face) pairs whose face is downwind of the cell, packed
// as i * slots + k: about half of a cut cell's faces, so the flux kernel
// below runs on full waves.  Block-aggregated appends (the order does not
// matter: each face is downwind of one cell and gets one volume), one atomic
// per block and round (a single counter serialises its atomics).  The count
// is cleared by the upwind kernel.
// needBounding (advection.C:54-82): each mixed cell of this step's reconstruct
// tags itself (slot 0) and the neighbour across each internal face (one slot
// per face) with the step's number, so the update and the bounding apply test
// it with one load; mixed cells of every status are tagged (here, or by the
// unlisted flux kernel on tetrahedra).

__device__ __attribute__((always_inline)) inline void k_flux_list_body(const DevMesh& m, const DevState& st, int slots, int epoch) {
    const int nMixed = *st.nMixed;
    const long total = (long)nMixed * slots;
    const int lane = threadIdx.x & 63;
    const long stride = (long)gridDim.x * blockDim.x;
    __shared__ int wcnt[4], sBase;
    const int wave = threadIdx.x >> 6;
    // block-uniform trip count: every wave of the block takes part in the
    // block's single atomic per round
    for (long b0 = (long)blockIdx.x * blockDim.x; b0 < total; b0 += stride) {
        const long w = b0 + threadIdx.x;
        bool take = false;
        if (w < total) {
            // the pair's cell, face and the face's cells loaded once (the
            // needBounding tags are stores, after which reloads would wait)
            const int i = (int)(w / slots), k = (int)(w % slots);
            const int c = st.mixed[i];
            const int status = st.status[i];
            const int fb = m.cOff[c], nf = m.cOff[c + 1] - fb;
            if (k < nf) {
                const int f = m.cFaces[fb + k];
                const int fo = m.own[f];
                const int fn = f < m.nIF ? m.nbr[f] : -1;
                const double ph = st.phi[f];
                const unsigned char em = m.fEmpty[empty_slot(m, f)];
                // needBounding: the cell (slot 0) and its neighbour across
                // each internal face
                if (k == 0) st.nbTag[c] = epoch;
                if (f < m.nIF) st.nbTag[fo == c ? fn : fo] = epoch;
                if (status == 0) {
                    if (f < m.nIF) take = (fo == c) ? (ph >= 0.0) : (ph < 0.0);
                    else take = !em && (ph >= 0.0);
                }
            }
        }
        const unsigned long long bal = __ballot(take);
        if (lane == 0) wcnt[wave] = __popcll(bal);
        __syncthreads();
        if (threadIdx.x == 0) {
            const int tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
            sBase = tot ? atomicAdd(st.fluxCount, tot) : 0;
        }
        __syncthreads();
        int pre = sBase;
        for (int k = 0; k < wave; ++k) pre += wcnt[k];
        if (take) st.fluxList[pre + __popcll(bal & ((1ull << lane) - 1ull))] = (int)w;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_flux_list(DevMesh m, DevState st, int slots) {
    k_flux_list_body(m, st, slots, st.epoch);
}

// timeIntegratedFlux with one thread per (cut cell, downwind face): each
// downwind face's volume is independent of the others.  LISTED takes the
// pairs from the work list; otherwise every (cut cell, face slot) pair is a
// thread and the upwind faces drop out (cheaper for tetrahedra, whose light
// faces do not repay the list).
template <bool TRIQUAD, bool LISTED>
__device__ __attribute__((always_inline)) inline void k_interface_flux_faces_body(const DevMesh& m, const DevState& st, double dtStep, int slots,
                                                                                  int epoch) {
    const long nWork = LISTED ? (long)*st.fluxCount : (long)*st.nMixed * slots;
    for (long wi = (long)blockIdx.x * blockDim.x + threadIdx.x; wi < nWork; wi += (long)gridDim.x * blockDim.x) {
        const int w = LISTED ? st.fluxList[wi] : (int)wi;
        const int i = w / slots, k = w % slots;
        // the pair's cell and face loaded once, before the needBounding
        // stores (after which reloads of them would wait)
        const int c = st.mixed[i];
        const int fb = m.cOff[c];
        if (!LISTED && k >= m.cOff[c + 1] - fb) continue;
        const double dt = dtStep;
        const int f = m.cFaces[fb + k];
        const double ph = st.phi[f];
        if (!LISTED) {
            const int status = st.status[i];
            const int fo = m.own[f];
            const int fn = f < m.nIF ? m.nbr[f] : -1;
            const unsigned char em = m.fEmpty[empty_slot(m, f)];
            // needBounding: the cell (slot 0) and its neighbour across each
            // internal face, for every status
            if (k == 0) st.nbTag[c] = epoch;
            if (f < m.nIF) st.nbTag[fo == c ? fn : fo] = epoch;
            if (status != 0) continue;
            bool down;
            if (f < m.nIF) down = (fo == c) ? (ph >= 0.0) : (ph < 0.0);
            else down = !em && (ph >= 0.0);
            if (!down) continue;
        }
        const V3 n = ld3(st.ifN, i);
        const double d = st.ifD[i], Un0 = st.un0[i];
        const int s0 = m.fOff[f], fsz = m.fOff[f + 1] - s0;
        bool ok = true;
        double v;
        if (fsz == 4) {
            V3 p[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) p[q] = ld3(m.pts, m.fPts[s0 + q]);
            v = tiff_regs<4>(m, f, p, n, d, Un0, dt, ph, m.magSf[f], ok);
        } else if (TRIQUAD || fsz == 3) {
            V3 p[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) p[q] = ld3(m.pts, m.fPts[s0 + q]);
            v = tiff_regs<3>(m, f, p, n, d, Un0, dt, ph, m.magSf[f], ok);
        } else {
            v = time_integrated_face_flux(m, f, n, d, Un0, dt, ph, m.magSf[f], ok);
        }
        st.dVf[f] = v;
        if (!ok) atomicOr(st.errFlag, ERR_TOO_MANY_CROSSINGS);
    }
}

template <bool TRIQUAD, bool LISTED>
__global__ void __launch_bounds__(64, SVP_FLUX_WAVES) k_interface_flux_faces(DevMesh m, DevState st, double dtStep, int slots) {
    k_interface_flux_faces_body<TRIQUAD, LISTED>(m, st, dtStep, slots, st.epoch);
}

__device__ inline bool loop_oob(double a, double aTol) { return (a - 1.0 > aTol) || (a < -aTol); }
__device__ inline bool cell_oob(double a, double aTol) { return (a < -aTol) || (a > 1.0 + aTol); }

// needBounding = mixed cells and their face neighbours (advection.C:54-82):
// tagged with the reconstruct's number by the detection kernel
__device__ inline bool need_bounding(const DevState& st, int c, int epoch) { return st.nbTag[c] == epoch; }

// Conservative update (advectionTemplates.C:399-404) with fvc::surfaceIntegrate
// summing the cell's faces in ascending face order.  Also counts cells outside
// [-aTol, 1+aTol] for the bounding loop and writes the bounding candidates
// (outside the bounds and next to the interface) as a bitmap, one 64-bit word
// per wave.
constexpr int UPDATE_BLOCK = 256;

// fvc::surfaceIntegrate's face sum of dVf for cell c, in ascending face
// order (neighbour faces subtract, empty-patch faces carry nothing).  With
// MF > 0, a cell of at most MF faces has all its face labels loaded at once
// and then all their volumes, so a thread waits on two gathers instead of two
// per face; larger cells take the face-by-face loop.
template <int MF>
__device__ __attribute__((always_inline)) inline double cell_face_sum(const DevMesh& m, const DevState& st, int c) {
    const int b = m.cOff[c], nf = m.cOff[c + 1] - b;
    double sI = 0.0;
    if (MF > 0 && nf <= MF) {
        constexpr int N = MF > 0 ? MF : 1;
        int e[N];
        double v[N];
        unsigned char em[N];
#pragma unroll
        for (int k = 0; k < N; ++k) e[k] = m.cFacesAsc[b + min(k, nf - 1)];
        // (unconditional loads: a load under a branch would wait for all the
        // loads before it).  Only boundary faces can be empty: an interior
        // face's flag load reads the first boundary face's byte, one line
        // that every thread shares, instead of a line per face.
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const int f = e[k] >= 0 ? e[k] : ~e[k];
            v[k] = st.dVf[f];
            em[k] = m.fEmpty[empty_slot(m, f)];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < N; ++k) {
            if (k >= nf) break;
            if (e[k] >= 0) {
                if (e[k] >= m.nIF && em[k]) continue;
                sI = sI + v[k];
            } else {
                sI = sI - v[k];
            }
        }
        return sI;
    }
    for (int k = b; k < b + nf; ++k) {
        const int e = m.cFacesAsc[k];
        if (e >= 0) {
            if (e >= m.nIF && m.fEmpty[e]) continue;
            sI = sI + st.dVf[e];
        } else {
            sI = sI - st.dVf[~e];
        }
    }
    return sI;
}

template <int MF>
__device__ __attribute__((always_inline)) inline void k_update_body(const DevMesh& m, const DevState& st, double dtStep, const double* Sp, const double* Su,
                                                                    int epoch) {
    const double aTol = 100.0 * SVP_SMALL;
    const int c = blockIdx.x * UPDATE_BLOCK + threadIdx.x;
    bool cand = false, oob = false;
    const double dt = dtStep;
    if (c < m.nC) {
        const double rDeltaT = 1.0 / dt;
        double sI = cell_face_sum<MF>(m, st, c);
        sI = sI / m.V[c];
        double num = st.alphaOld[c] * rDeltaT;
        if (Su) num = num + Su[c];
        num = num - sI * rDeltaT;
        const double den = Sp ? (rDeltaT - Sp[c]) : rDeltaT;
        const double a = num / den;
        st.alpha[c] = a;
        oob = loop_oob(a, aTol);
        cand = cell_oob(a, aTol) && need_bounding(st, c, epoch);
    }
    const unsigned long long bits = __ballot(cand);
    const unsigned long long oobBits = __ballot(oob);
    const int lane = threadIdx.x & 63;
    if (lane == 0) {
        const int word = c >> 6;
        if (word < (m.nC + 63) / 64) st.candBits[word] = bits;
        if (oobBits) atomicAdd(&st.oobCount[OOB_STRIDE * (blockIdx.x & (OOB_SLOTS - 1))], __popcll(oobBits));
    }
}

template <int MF>
__global__ void __launch_bounds__(UPDATE_BLOCK) k_update(DevMesh m, DevState st, double dtStep, const double* Sp, const double* Su) {
    k_update_body<MF>(m, st, dtStep, Sp, Su, st.epoch);
}

// face slots of the update's gathered form: the mesh's largest cell when it
// has at most 8 faces, else the face-by-face loop (0)
static int update_faces(int maxCellFaces) {
    return maxCellFaces <= 4 ? 4 : (maxCellFaces <= 6 ? 6 : (maxCellFaces <= 8 ? 8 : 0));
}

static void launch_update(const DevMesh& m, const DevState& st, double dt, const double* Sp, const double* Su,
                          hipStream_t s) {
    const dim3 grid(cdiv(m.nC, UPDATE_BLOCK));
    switch (update_faces(st.maxCellFaces)) {
        case 4: k_update<4><<<grid, UPDATE_BLOCK, 0, s>>>(m, st, dt, Sp, Su); break;
        case 6: k_update<6><<<grid, UPDATE_BLOCK, 0, s>>>(m, st, dt, Sp, Su); break;
        case 8: k_update<8><<<grid, UPDATE_BLOCK, 0, s>>>(m, st, dt, Sp, Su); break;
        default: k_update<0><<<grid, UPDATE_BLOCK, 0, s>>>(m, st, dt, Sp, Su); break;
    }
}

__device__ inline double face_value(const DevMesh& m, const double* f, int facei) {
    if (facei < m.nIF) return f[facei];
    if (m.fEmpty[facei]) return 0.0;
    return f[facei];
}

__device__ inline bool face_downwind_of(const DevMesh& m, const double* phi, int facei, int c) {
    const double ph = face_value(m, phi, facei);
    return (m.own[facei] == c) ? (ph >= 0) : (ph < 0);
}


// limitFlux (advectionTemplates.C:118-215), one pass = the kernels below.
// The pass runs only while cells lie outside [-aTol, 1+aTol] (bound count of
// the update kernel, kept current by the apply kernel):
//   list              ascending candidates from the bitmap (one launch),
//                     then the clearing of what the previous pass left in
//                     the other half of the double buffers
//                     (dVfCorrectionValues = 0 for its active cells)
//   prep              every candidate's first redistribution loop, which
//                     decides whether it writes corrections and which faces
//                     open correctedFaces; the record of each writing cell
//   flow              all remaining boundFlux work in dependency order in
//                     one launch, and the correctedFaces order keys and
//                     touched cells (the default), or
//   scatter/round/tail  the same as Kahn rounds: grid-wide kernels, then the
//                     remaining rounds inside one workgroup
//   apply             corrections applied per cell in correctedFaces order,
//                     dVf of the corrected faces, bitmap and bound count,
//                     pass statistics
// bcount: [0] candidates, [1] corrected faces (prep), [2] touched cells,
// [3+r] frontier size of round r (r <= st.gridRounds; [3] = active cells of
// the dataflow pass), [BND_CTRL-1] rounds, then the dataflow ticket and
// completion count.
constexpr int BND_CTRL = 3 + MAX_GRID_ROUNDS + 2;
// dataflow pass (one launch in place of the rounds and the tail): chunk
// ticket and completion count, cleared with the per-pass counters
constexpr int BND_TICKET = BND_CTRL;
constexpr int BND_PROGRESS = BND_CTRL + 1;
static_assert(BND_PROGRESS < 16, "bcount holds 16 counters");
// correction of a face whose writer has not completed yet in the dataflow
// pass: a NaN payload that no arithmetic produces (stored and compared as bits)
constexpr unsigned long long CORR_PENDING = 0x7FF4D00DF00DD00Dull;

// Record of candidate c for the dependent part of the pass, with iteration 0
// of boundFlux's redistribution loop (advectionTemplates.C:260-330) already
// done: that iteration reads only the cell's own downwind faces, whose
// corrections start at zero because a face is written solely by the cell it
// is downwind of.  Its selected faces are the first loop that opens
// correctedFaces.  Returns the number of those faces; 0 means the cell writes
// no corrections this pass and gets no record.  Every face the cell will
// write is marked CORR_PENDING.
template <int MF>
__device__ inline int bound_record(const DevMesh& m, const DevState& st, int c, double dt, const double* Sp,
                                   const double* Su, BndRec<MF>& R) {
    const double aTol = 100.0 * SVP_SMALL;
    // the cell's own values in one batch (the face labels wait on cOff only)
    const double ac = st.alpha[c];
    const double Vi = m.V[c];
    const double aOld = st.alphaOld[c];
    const double suC = Su ? Su[c] : 0.0, spC = Sp ? Sp[c] : 0.0;
    const int b = m.cOff[c], nf = m.cOff[c + 1] - b;
    const double alphaOvershoot = pos0(ac - 1.0) * (ac - 1.0) + neg0(ac) * ac;
    if (fabs(alphaOvershoot) < aTol) return 0;
    const double fluidToPassOn = alphaOvershoot * Vi;
    int fid[MF], oth[MF];
    double phf[MF], dvf[MF], smax[MF];
    unsigned int flg[(MF + 7) / 8];
    unsigned int first = 0u;
    int n = 0;
    double dVftot = 0.0;
#pragma unroll
    for (int w = 0; w < (MF + 7) / 8; ++w) flg[w] = 0u;
    // all face loads first, unconditional (slots past nf repeat the last
    // face), so they are in flight together instead of one face at a time
    int fo[MF], fn[MF];
    double fph[MF], fdv[MF];
    unsigned char fem[MF];
#pragma unroll
    for (int k = 0; k < MF; ++k) fid[k] = m.cFaces[b + min(k, nf - 1)];
#pragma unroll
    for (int k = 0; k < MF; ++k) {
        const int f = fid[k];
        fo[k] = m.own[f];
        fn[k] = m.nbr[f];
        fph[k] = st.phi[f];
        fdv[k] = st.dVf[f];
        fem[k] = m.fEmpty[empty_slot(m, f)];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < MF; ++k) {
        oth[k] = -1;
        phf[k] = 0.0;
        dvf[k] = 0.0;
        smax[k] = 0.0;
        if (k < nf) {
            const int f = fid[k];
            const bool own = (fo[k] == c);
            if (f < m.nIF) oth[k] = own ? fn[k] : fo[k];
            // face_value: empty-patch faces carry nothing
            const bool live = (f < m.nIF) || !fem[k];
            const double ph = live ? fph[k] : 0.0;
            const double dv = live ? fdv[k] : 0.0;
            phf[k] = ph;
            dvf[k] = dv;
            const bool down = own ? (ph >= 0) : (ph < 0);
            const bool writable = live;
            const bool lowerWriter = oth[k] >= 0 && !down && oth[k] < c;
            flg[k >> 3] |= ((own ? 1u : 0u) | (down ? 2u : 0u) | (writable ? 4u : 0u) | (lowerWriter ? 8u : 0u))
                           << (4 * (k & 7));
            if (down) {
                const double dVff = dv + 0.0;
                const double maxExtra = fabs(pos0(fluidToPassOn) * ph * dt - dVff);
                if (maxExtra / Vi > aTol) {
                    first |= 1u << k;
                    n++;
                    smax[k] = maxExtra;
                    dVftot += fabs(ph * dt);
                }
            }
        }
    }
    if (n == 0) return 0;
    auto bits = [&](int k) { return flg[k >> 3] >> (4 * (k & 7)); };
    int nFaces = 0;
#pragma unroll
    for (int k = 0; k < MF; ++k) {
        R.cor0[k] = 0.0;
        if (!((first >> k) & 1u)) continue;
        double fl = fabs(fluidToPassOn) * fabs(phf[k] * dt) / dVftot;
        nFaces += (int)pos0(smax[k] - fl);
        fl = fmin(fl, smax[k]);
        double dVff = 0.0;
        dVff += sgn(phf[k]) * sgn(fluidToPassOn) * fl;
        if (bits(k) & 4u) R.cor0[k] = dVff;
    }
    double nfd = 0.0;
#pragma unroll
    for (int k = 0; k < MF; ++k)
        if (k < nf) nfd = (bits(k) & 1u) ? nfd + dvf[k] : nfd - dvf[k];
    const double rDeltaT = 1.0 / dt;
    double num = aOld * rDeltaT;
    if (Su) num = num + suC;
    R.num0 = num - nfd / Vi * rDeltaT;
    R.den = Sp ? (rDeltaT - spC) : rDeltaT;
#pragma unroll
    for (int k = 0; k < MF; ++k) {
        R.fid[k] = fid[k];
        R.oth[k] = oth[k];
        R.phf[k] = phf[k];
        R.dvf[k] = dvf[k];
        if (k < nf && (bits(k) & 6u) == 6u) st.corr[fid[k]] = __longlong_as_double((long long)CORR_PENDING);
    }
#pragma unroll
    for (int w = 0; w < (MF + 7) / 8; ++w) R.flg[w] = flg[w];
    R.nf = nf;
    R.first = first;
    R.nFaces0 = nFaces;
    R.Vi = Vi;
    return n;
}

// boundFlux of one cell (advectionTemplates.C:243-347) from iteration 0 of
// its record on: cin[k] is the correction of face slot k written by its
// lower-numbered upwind neighbour (0 where the flags mark no such writer; a
// higher-numbered writer has not run yet in the reference's ascending walk,
// so it contributes zero); out(k, f, v) stores the final correction of
// downwind face f in slot k.
template <int MF, class Out>
__device__ inline void bound_finish(const BndRec<MF>& R, const double (&cin)[MF], double dt, double rDeltaT,
                                    Out out) {
    const double aTol = 100.0 * SVP_SMALL;
    unsigned int flg[(MF + 7) / 8];
#pragma unroll
    for (int w = 0; w < (MF + 7) / 8; ++w) flg[w] = R.flg[w];
    const int nf = R.nf;
    const double Vi = R.Vi;
    auto bits = [&](int k) { return flg[k >> 3] >> (4 * (k & 7)); };
    auto isOwn = [&](int k) { return (bits(k) & 1u) != 0u; };
    auto down = [&](int k) { return (bits(k) & 2u) != 0u; };
    auto writable = [&](int k) { return (bits(k) & 4u) != 0u; };
    double phf[MF], cor[MF];
#pragma unroll
    for (int k = 0; k < MF; ++k) {
        phf[k] = R.phf[k];
        cor[k] = down(k) ? R.cor0[k] : cin[k];
    }
    double alphaOvershoot = 0.0, fluidToPassOn = 0.0;
    auto settle = [&]() {
        double nfc = 0.0;
#pragma unroll
        for (int k = 0; k < MF; ++k)
            if (k < nf) nfc = isOwn(k) ? nfc + cor[k] : nfc - cor[k];
        const double alpha1New = (R.num0 - nfc / Vi * rDeltaT) / R.den;
        alphaOvershoot = pos0(alpha1New - 1.0) * (alpha1New - 1.0) + neg0(alpha1New) * alpha1New;
        fluidToPassOn = alphaOvershoot * Vi;
    };
    settle();
    int nFaces = R.nFaces0;
    for (int iter = 1; iter < 10; iter++) {
        if (fabs(alphaOvershoot) < aTol || nFaces == 0) break;
        bool sel[MF];
        double smax[MF];
        double dVftot = 0.0;
        nFaces = 0;
#pragma unroll
        for (int k = 0; k < MF; ++k) {
            sel[k] = false;
            smax[k] = 0.0;
            if (k < nf && down(k)) {
                const double dVff = R.dvf[k] + cor[k];
                const double maxExtra = fabs(pos0(fluidToPassOn) * phf[k] * dt - dVff);
                if (maxExtra / Vi > aTol) {
                    sel[k] = true;
                    smax[k] = maxExtra;
                    dVftot += fabs(phf[k] * dt);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < MF; ++k) {
            if (!sel[k]) continue;
            double fl = fabs(fluidToPassOn) * fabs(phf[k] * dt) / dVftot;
            nFaces += (int)pos0(smax[k] - fl);
            fl = fmin(fl, smax[k]);
            double dVff = cor[k];
            dVff += sgn(phf[k]) * sgn(fluidToPassOn) * fl;
            if (writable(k)) cor[k] = dVff;
        }
        settle();
    }
#pragma unroll
    for (int k = 0; k < MF; ++k)
        if (k < nf && down(k) && writable(k)) out(k, R.fid[k], cor[k]);
}

// Ascending candidates of the pass from the bitmap in one launch: every
// block counts its st.bndListWords words, publishes the count and looks back over the
// blocks before it for their prefix (a decoupled look-back), then writes its
// candidates.  Blocks take tickets in the order they start, so a block only
// waits for blocks that are already running and each of those waits only on
// its own predecessors.  The published words carry the launch's epoch, which
// makes any word of an earlier launch read as not yet published.  With a
// zero bound count the pass is off (no candidates).  Ticket 0 also clears
// the per-pass counters.
// candidates of the previous pass, whose leftovers the next list clears
constexpr int BND_PREV_COUNT = 16;
// any cell out of bounds: sum of the spread counters (each slot is one cache
// line, so the update kernel's blocks do not queue on a single address)
__device__ inline bool oob_any(const DevState& st) {
    int n = 0;
#pragma unroll 8
    for (int k = 0; k < OOB_SLOTS; ++k) n += st.oobCount[OOB_STRIDE * k];
    return n != 0;
}

__device__ __attribute__((always_inline)) inline void k_bnd_list_body(const DevMesh& m, const DevState& st, int nWords,
                                                                      int nBlocks, const BndPrev& prev) {
    __shared__ int lds[16];
    __shared__ int sTicket, sBase, sTotal;
    __shared__ unsigned sEpoch;
    if (threadIdx.x == 0) sTicket = take_ticket(st.bndTicket, nBlocks, sEpoch);
    __syncthreads();
    const int b = sTicket;
    const unsigned epoch = sEpoch;
    // what the previous pass (of this step or the last one) left: the
    // corrections and correctedFaces keys on the faces of its active cells
    // and their candidate index, back to 0 / -1 before this pass's records.
    // Done by the waves that wait while wave 0 looks back (thread t of nt
    // clearing threads per block), off the look-back chain.
    auto clearPrevious = [&](int t, int nt) {
        const int nPrev = st.bcount[BND_PREV_COUNT];
        for (int i = b * nt + t; i < nPrev; i += nBlocks * nt) {
            if (!prev.active[i]) continue;
            const int c = prev.cand[i];
            const int fb = m.cOff[c], nf = m.cOff[c + 1] - fb;
            if (nf <= 8) {
                // the face labels loaded together, then the stores
                int fs[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) fs[k] = m.cFaces[fb + min(k, nf - 1)];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (k < nf) {
                        prev.corr[fs[k]] = 0.0;
                        prev.corrPos[fs[k]] = -1;
                    }
            } else {
                for (int k = fb; k < fb + nf; ++k) {
                    const int f = m.cFaces[k];
                    prev.corr[f] = 0.0;
                    prev.corrPos[f] = -1;
                }
            }
            prev.actIdx[c] = -1;
        }
    };
    // the counters the later kernels of the pass count into (the corrected
    // faces of [1] are counted here, so the previous apply cleared them)
    if (b == 0 && threadIdx.x >= 2 && threadIdx.x <= BND_PROGRESS) st.bcount[threadIdx.x] = 0;
    if (!oob_any(st)) {
        if (b == 0 && threadIdx.x == 0) st.bcount[0] = 0;
        clearPrevious(threadIdx.x, blockDim.x);
        return;
    }
    const int wBeg = b * st.bndListWords, wEnd = min(nWords, wBeg + st.bndListWords);
    int cnt = 0;
    for (int w = wBeg + threadIdx.x; w < wEnd; w += 256) cnt += __popcll(st.candBits[w]);
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x < 64) {
        const unsigned long long total = (unsigned long long)(lds[0] + lds[1] + lds[2] + lds[3]);
        const unsigned long long excl = look_back(st.bndLook, b, epoch, total, st.errFlag, ERR_BOUND_STALL);
        if (threadIdx.x == 0) {
            if (b == nBlocks - 1) st.bcount[0] = (int)(excl + total);
            sBase = (int)excl;
            sTotal = (int)total;
        }
    } else {
        clearPrevious(threadIdx.x - 64, blockDim.x - 64);
    }
    __syncthreads();
    int base = sBase;
    // the block's words in chunks of 256: a block scan of the popcounts per chunk
    for (int w0 = wBeg; w0 < wEnd; w0 += 256) {
        const int w = w0 + threadIdx.x;
        unsigned long long bits = (w < wEnd) ? st.candBits[w] : 0ull;
        const int c = __popcll(bits);
        // exclusive scan over the 256 threads: wave scans, then the 4 wave totals
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        int x = c;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) lds[wv] = x;
        __syncthreads();
        int pre = 0;
        for (int k = 0; k < wv; ++k) pre += lds[k];
        const int tot = lds[0] + lds[1] + lds[2] + lds[3];
        int off = base + pre + x - c;
        while (bits) {
            const int bit = __ffsll((long long)bits) - 1;
            bits &= bits - 1;
            st.cand[off++] = w * 64 + bit;
        }
        base += tot;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_bnd_list(DevMesh m, DevState st, int nWords, int nBlocks, BndPrev prev) {
    k_bnd_list_body(m, st, nWords, nBlocks, prev);
}

// every candidate's record (iteration 0 of boundFlux done) and, when it
// writes corrections, its candidate index and the CORR_PENDING marks of the
// faces it will write; one candidate per thread over the whole grid (the
// candidates cluster in a few list blocks, so the list kernel's blocks would
// build them unevenly)
template <int MF>
__device__ __attribute__((always_inline)) inline void k_bnd_prep_body(const DevMesh& m, const DevState& st, double dt,
                                                                      const double* Sp, const double* Su) {
    const int nCand = st.bcount[0];
    int nOpen = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nCand; i += gridDim.x * blockDim.x) {
        const int c = st.cand[i];
        const int n = bound_record<MF>(m, st, c, dt, Sp, Su, ((BndRec<MF>*)st.bndRec)[i]);
        st.active[i] = n > 0 ? 1 : 0;
        if (n > 0) st.actIdx[c] = i;
        nOpen += n;
    }
    // correctedFaces length of the pass (statistics)
    for (int o = 32; o > 0; o >>= 1) nOpen += __shfl_xor(nOpen, o);
    if ((threadIdx.x & 63) == 0 && nOpen) atomicAdd(&st.bcount[1], nOpen);
}

template <int MF>
__global__ void __launch_bounds__(256) k_bnd_prep(DevMesh m, DevState st, double dt, const double* Sp, const double* Su) {
    k_bnd_prep_body<MF>(m, st, dt, Sp, Su);
}


// correctedFaces keys and the cells they touch (both schedules), for candidate
// c whose record is written.  correctedFaces lists the first-loop faces of
// the candidates in ascending cell order, so the key c * MF + (rank among
// the cell's first-loop faces) orders any two of its faces as their positions
// in that list would; only that order is used.  The touched-cell claims are
// issued together, then one counter add reserves the list entries.
template <int MF>
__device__ inline void bound_open_faces(const DevState& st, const BndRec<MF>& R, int c) {
    const unsigned int first = R.first;
    int key = c * MF;
    // claims first, results read afterwards, so that the claims overlap
    int old[MF];
#pragma unroll
    for (int k = 0; k < MF; ++k) {
        old[k] = 1;
        if (!((first >> k) & 1u)) continue;
        st.corrPos[R.fid[k]] = key++;
        if (R.oth[k] >= 0) old[k] = atomicCAS(&st.touchedFlag[R.oth[k]], 0, 1);
    }
    const int oldSelf = first ? atomicCAS(&st.touchedFlag[c], 0, 1) : 1;
    unsigned long long won = 0ull;
#pragma unroll
    for (int k = 0; k < MF; ++k)
        if (old[k] == 0) won |= 1ull << k;
    const bool self = oldSelf == 0;
    const int nWon = __popcll(won) + (self ? 1 : 0);
    if (nWon == 0) return;
    int slot = atomicAdd(&st.bcount[2], nWon);
    if (self) st.touched[slot++] = c;
#pragma unroll
    for (int k = 0; k < MF; ++k)
        if ((won >> k) & 1ull) st.touched[slot++] = R.oth[k];
}

// Kahn schedule: correctedFaces, successors of every active candidate,
// predecessor counts and frontier 0
template <int MF>
__global__ void __launch_bounds__(256) k_bnd_scatter(DevMesh m, DevState st) {
    const int nCand = st.bcount[0];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nCand; i += gridDim.x * blockDim.x) {
        if (!st.active[i]) continue;
        BndRec<MF>& R = ((BndRec<MF>*)st.bndRec)[i];
        const int c = st.cand[i];
        bound_open_faces<MF>(st, R, c);
        int np = 0;
#pragma unroll
        for (int k = 0; k < MF; ++k) {
            const unsigned int bits = R.flg[k >> 3] >> (4 * (k & 7));
            int sc = -1;
            if (k < R.nf && R.fid[k] < m.nIF) {
                const int f = R.fid[k];
                const int o = (bits & 1u) ? m.nbr[f] : m.own[f];
                if ((bits & 8u) && st.actIdx[o] >= 0) np++;
                if ((bits & 2u) && o > c) sc = st.actIdx[o];
            }
            R.succ[k] = sc;
        }
        st.pend[i] = np;
        if (np == 0) st.queue[atomicAdd(&st.bcount[3], 1)] = i;
    }
}

// boundFlux of candidate i from its record
template <int MF, class Out>
__device__ inline void bound_candidate(const DevState& st, const BndRec<MF>& R, const double (&cin)[MF], int i,
                                       double dt, double rDeltaT, Out out) {
    (void)st;
    (void)i;
    bound_finish<MF>(R, cin, dt, rDeltaT, out);
}

// completes cell i of frontier r and releases its successors into frontier
// r+1; release(a) drops one pending predecessor of candidate a and returns
// true when it was the last, push(a) appends a to the next frontier
template <int MF, class Release, class Push>
__device__ inline void bnd_complete(const DevState& st, int i, double dt, double rDeltaT, Release release,
                                    Push push) {
    const BndRec<MF>& R = ((const BndRec<MF>*)st.bndRec)[i];
    // predecessors wrote these corrections in an earlier round: the previous
    // kernel of the grid rounds, or before the workgroup barrier of the tail
    double cin[MF];
#pragma unroll
    for (int k = 0; k < MF; ++k)
        cin[k] = ((R.flg[k >> 3] >> (4 * (k & 7))) & 8u) ? st.corr[R.fid[k]] : 0.0;
    bound_candidate<MF>(st, R, cin, i, dt, rDeltaT, [&](int, int f, double v) { st.corr[f] = v; });
#pragma unroll
    for (int k = 0; k < MF; ++k) {
        const int a = R.succ[k];
        if (a < 0) continue;
        if (release(a)) push(a);
    }
}

template <int MF>
__global__ void __launch_bounds__(256) k_bnd_round(DevMesh m, DevState st, int r, double dt) {
    const int n = st.bcount[3 + r];
    const int* q = st.queue + (long)(r & 1) * m.nC;
    int* qnext = st.queue + (long)((r + 1) & 1) * m.nC;
    int* nNext = &st.bcount[3 + r + 1];
    const double rDeltaT = 1.0 / dt;
    auto release = [&](int a) { return atomicSub(&st.pend[a], 1) == 1; };
    auto push = [&](int a) { qnext[atomicAdd(nNext, 1)] = a; };
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        bnd_complete<MF>(st, q[k], dt, rDeltaT, release, push);
}

// Remaining rounds in one workgroup of TAIL_THREADS (frontiers are small by now); every round
// completes at least one cell because edges point to higher cell numbers.
// The frontiers live in LDS up to TAIL_Q entries (further entries in the
// global queue at the same positions), and with at most TAIL_PEND_CAP
// candidates the predecessor counts move to LDS as well (one byte each, four
// per word), so a round touches global memory only for the cells' records
// and corrections.
constexpr int TAIL_PEND_WORDS = 8192;
constexpr int TAIL_PEND_CAP = 4 * TAIL_PEND_WORDS;
constexpr int TAIL_Q = 4096;

constexpr int TAIL_THREADS = 512;

template <int MF>
__global__ void __launch_bounds__(TAIL_THREADS) k_bnd_tail(DevMesh m, DevState st, double dt) {
    __shared__ int sQ, sNext, sRounds;
    __shared__ unsigned int sPend[TAIL_PEND_WORDS];
    __shared__ int sQueue[2][TAIL_Q];
    const int t = threadIdx.x;
    const double rDeltaT = 1.0 / dt;
    const int R = st.gridRounds;
    const int nCand = st.bcount[0];
    const bool ldsPend = nCand <= TAIL_PEND_CAP;
    const int q0 = st.bcount[3 + R];
    if (t == 0) {
        sQ = q0;
        sNext = 0;
        sRounds = R;
    }
    if (ldsPend)
        for (int w = t; w < (nCand + 3) / 4; w += TAIL_THREADS) {
            unsigned int word = 0u;
            for (int b = 0; b < 4; ++b)
                if (4 * w + b < nCand) word |= ((unsigned int)st.pend[4 * w + b] & 0xFFu) << (8 * b);
            sPend[w] = word;
        }
    int gcur = R & 1;  // global queue buffer holding the overflow of the current frontier
    for (int k = t; k < min(q0, TAIL_Q); k += TAIL_THREADS) sQueue[0][k] = st.queue[(long)gcur * m.nC + k];
    __syncthreads();
    auto releaseLds = [&](int a) {
        const int sh = 8 * (a & 3);
        const unsigned int old = atomicSub(&sPend[a >> 2], 1u << sh);
        return ((old >> sh) & 0xFFu) == 1u;
    };
    auto releaseGlobal = [&](int a) { return atomicSub(&st.pend[a], 1) == 1; };
    int cur = 0;
    while (true) {
        const int qn = sQ;
        if (qn == 0) break;
        const int* gq = st.queue + (long)gcur * m.nC;
        int* gqNext = st.queue + (long)(gcur ^ 1) * m.nC;
        int* lqNext = sQueue[cur ^ 1];
        auto push = [&](int a) {
            const int idx = atomicAdd(&sNext, 1);
            if (idx < TAIL_Q) lqNext[idx] = a;
            else gqNext[idx] = a;
        };
        for (int k = t; k < qn; k += TAIL_THREADS) {
            const int i = k < TAIL_Q ? sQueue[cur][k] : gq[k];
            if (ldsPend) bnd_complete<MF>(st, i, dt, rDeltaT, releaseLds, push);
            else bnd_complete<MF>(st, i, dt, rDeltaT, releaseGlobal, push);
        }
        __syncthreads();
        if (t == 0) {
            sQ = sNext;
            sNext = 0;
            sRounds++;
        }
        cur ^= 1;
        gcur ^= 1;
        __syncthreads();
    }
    if (t == 0) st.bcount[BND_CTRL - 1] = sRounds;
}

// The dependent part of a pass in one launch.  Workgroups take chunks of
// consecutive candidates from a ticket, so every lower-numbered
// candidate is held by a workgroup that is already running; a candidate runs
// as soon as none of the corrections it reads (faces written by
// lower-numbered active upwind neighbours) is CORR_PENDING.  Those
// corrections are the only data passed between candidates, and each is
// written once, by one cell: to LDS for the readers of its own chunk and to
// global memory with an agent-scope atomic store for the others, so the
// value is its own ready flag.  One extra wave per workgroup polls global
// memory (agent-scope atomic loads) for the corrections its chunk reads from
// earlier chunks and copies them into LDS, so the computing waves only ever
// wait on LDS.  The lowest unfinished candidate is always ready, so the
// waits end; a wave that sees no completion anywhere in the grid for about a
// second flags ERR_BOUND_STALL and the pass stops.
// chunk of candidates per workgroup: one per thread of the computing waves;
// 7 + 1 waves at two waves per SIMD for cells of up to 8 faces (held
// records fill the registers), smaller chunks for larger cells (LDS)
template <int MF>
constexpr int flow_chunk() { return MF <= 8 ? 448 : 128; }
template <int MF>
constexpr int flow_threads() { return flow_chunk<MF>() + 64; }  // computing waves + the polling wave
constexpr int FLOW_POLL_BATCH = 8;          // loads a polling lane keeps in flight
constexpr int FLOW_STALL_CHECKS = 1 << 14;   // idle checks, 64 sleeps each, without progress

__device__ inline unsigned long long corr_load(const double* p) {
    return __hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// stall check of a waiting wave: true once the grid-wide completion count
// has not moved over FLOW_STALL_CHECKS checks
struct FlowStall {
    int idle = 0, seen = -1, stalls = 0;
    __device__ bool wait(const int* progress, int lane) {
        __builtin_amdgcn_s_sleep(1);
        if (++idle < 64) return false;
        idle = 0;
        int now = 0;
        if (lane == 0) now = __hip_atomic_load(progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        now = __shfl(now, 0);
        stalls = (now == seen) ? stalls + 1 : 0;
        seen = now;
        return stalls >= FLOW_STALL_CHECKS;
    }
};

template <int MF>
__device__ __attribute__((always_inline)) inline void k_bnd_flow_body(const DevMesh& m, const DevState& st, double dt) {
    constexpr int CH = flow_chunk<MF>();
    // LDS slots [0, CH*MF): corrections written by the chunk's cells (slot
    // t*MF+k: cell t, face slot k); [CH*MF, 2*CH*MF): corrections read from
    // earlier chunks, copied in by the polling wave
    constexpr int NS = CH * MF;
    __shared__ unsigned long long sVal[2 * NS];
    __shared__ int sFid[NS];                 // faces the chunk's cells write
    __shared__ int sPoll[NS][2];             // polled corrections: face, LDS slot
    __shared__ int sBase, sNPoll;
    // records of up to 8 faces stay in registers from the chunk's start, so
    // a cell that turns ready reads nothing but its upwind corrections
    constexpr bool kHold = MF <= 8;
    const int nCand = st.bcount[0];
    const int t = threadIdx.x, lane = t & 63;
    const bool poller = t >= CH;
    static_assert(NS <= 64 * 64, "a polling lane tracks at most 64 entries");
    int* progress = &st.bcount[BND_PROGRESS];
    const BndRec<MF>* rec = (const BndRec<MF>*)st.bndRec;
    while (true) {
        __syncthreads();
        if (t == 0) {
            // after a stall no more chunks are taken
            const bool stalled = __hip_atomic_load(st.errFlag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
                                 ERR_BOUND_STALL;
            sBase = stalled ? nCand : atomicAdd(&st.bcount[BND_TICKET], CH);
            sNPoll = 0;
        }
        __syncthreads();
        const int base = sBase;
        if (base >= nCand) break;
        const int i = base + t;
        const bool active = !poller && i < nCand && st.active[i] != 0;
        bool pending = active;
        const unsigned long long act = __ballot(active);
        if (lane == 0 && act) atomicAdd(&st.bcount[3], __popcll(act));
        BndRec<MF> held;
        if (kHold && active) held = rec[i];
        const BndRec<MF>& R = kHold ? held : rec[active ? i : 0];
        unsigned int flg[(MF + 7) / 8];
#pragma unroll
        for (int w = 0; w < (MF + 7) / 8; ++w) flg[w] = active ? R.flg[w] : 0u;
        auto bits = [&](int k) { return flg[k >> 3] >> (4 * (k & 7)); };
        // candidate index of each upwind writer (-1: it writes nothing this
        // pass); the loads are unconditional so that they are all in flight
        // together
        int wr[MF];
#pragma unroll
        for (int k = 0; k < MF; ++k) wr[k] = st.actIdx[(bits(k) & 8u) ? R.oth[k] : 0];
#pragma unroll
        for (int k = 0; k < MF; ++k)
            if (!(bits(k) & 8u)) wr[k] = -1;
        if (!poller) {
#pragma unroll
            for (int k = 0; k < MF; ++k) {
                sFid[t * MF + k] = ((bits(k) & 6u) == 6u) ? R.fid[k] : -1;
                sVal[t * MF + k] = CORR_PENDING;
                sVal[NS + t * MF + k] = CORR_PENDING;
            }
        }
        __syncthreads();
        // where each read correction comes from: the writer's slot when it is
        // a cell of this chunk, else a polled slot.  Nothing in the waiting
        // loop reads global memory, so the loop never waits on the global
        // stores of earlier completions.
        int src[MF];
        unsigned int need = 0u;
        double cin[MF];
#pragma unroll
        for (int k = 0; k < MF; ++k) {
            src[k] = -1;
            cin[k] = 0.0;
            const int p = wr[k];
            if (p < 0) continue;
            const int f = R.fid[k];
            need |= 1u << k;
            if (p >= base) {
                const int j = p - base;
                for (int q = 0; q < MF; ++q)
                    if (sFid[j * MF + q] == f) src[k] = j * MF + q;
            } else {
                const int g = atomicAdd(&sNPoll, 1);
                sPoll[g][0] = f;
                sPoll[g][1] = NS + t * MF + k;
                src[k] = NS + t * MF + k;
            }
        }
        const double dtc = dt;
        const double rdc = 1.0 / dtc;
        __syncthreads();
        FlowStall stall;
        if (poller) {
            // copy corrections of earlier chunks into LDS as they complete:
            // entry lane + 64 b of the list is this lane's bit b, and the
            // loads of FLOW_POLL_BATCH entries are issued together
            const int nPoll = sNPoll;
            unsigned long long left = 0ull;
            for (int b = 0; lane + 64 * b < nPoll; ++b) left |= 1ull << b;
            while (__any(left != 0ull)) {
                bool got = false;
                for (int b0 = 0; (left >> b0) != 0ull; b0 += FLOW_POLL_BATCH) {
                    unsigned long long u[FLOW_POLL_BATCH];
#pragma unroll
                    for (int b = 0; b < FLOW_POLL_BATCH; ++b) {
                        const int e = lane + 64 * (b0 + b);
                        u[b] = ((left >> (b0 + b)) & 1ull) ? corr_load(&st.corr[sPoll[e][0]]) : CORR_PENDING;
                    }
#pragma unroll
                    for (int b = 0; b < FLOW_POLL_BATCH; ++b) {
                        if (u[b] != CORR_PENDING) {
                            const int e = lane + 64 * (b0 + b);
                            __hip_atomic_store(&sVal[sPoll[e][1]], u[b], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                            left &= ~(1ull << (b0 + b));
                            got = true;
                        }
                    }
                }
                if (__any(got)) {
                    stall.idle = 0;
                    continue;
                }
                if (stall.wait(progress, lane)) {
                    if (lane == 0) atomicOr(st.errFlag, ERR_BOUND_STALL);
                    break;
                }
            }
        }
        auto out = [&](int k, int f, double v) {
            const unsigned long long u = (unsigned long long)__double_as_longlong(v);
            __hip_atomic_store(&sVal[t * MF + k], u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_store((unsigned long long*)&st.corr[f], u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        };
        while (!poller && __any(pending)) {
            bool fin = false;
            if (pending) {
                // corrections not seen yet; each one is read until it arrives
#pragma unroll
                for (int k = 0; k < MF; ++k) {
                    if (!((need >> k) & 1u)) continue;
                    const unsigned long long u =
                        __hip_atomic_load(&sVal[src[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (u != CORR_PENDING) {
                        cin[k] = __longlong_as_double((long long)u);
                        need &= ~(1u << k);
                    }
                }
                if (need == 0u) {
                    bound_finish<MF>(R, cin, dtc, rdc, out);
                    pending = false;
                    fin = true;
                }
            }
            const unsigned long long done = __ballot(fin);
            if (done) {
                if (lane == 0) atomicAdd(progress, __popcll(done));
                stall.idle = 0;
                continue;
            }
            if (stall.wait(progress, lane)) {
                if (lane == 0) atomicOr(st.errFlag, ERR_BOUND_STALL);
                break;
            }
        }
        if (active) bound_open_faces<MF>(st, R, st.cand[i]);
    }
}

template <int MF>
__global__ void __launch_bounds__(flow_threads<MF>()) k_bnd_flow(DevMesh m, DevState st, double dt) {
    k_bnd_flow_body<MF>(m, st, dt);
}


// alpha of every touched cell in correctedFaces order (advectionTemplates.C:164-192)
// and the dVf of every corrected face (the first-loop faces of the active
// candidates)
template <int MF>
__device__ __attribute__((always_inline)) inline void k_bnd_apply_body(const DevMesh& m, const DevState& st, int pass,
                                                                       int epoch) {
    const double aTol = 100.0 * SVP_SMALL;
    const int nT = st.bcount[2], nCand = st.bcount[0];
    // pass statistics, and the candidate count for the next list's clearing
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st.bcount[BND_PREV_COUNT] = nCand;
        if (pass < MAX_BOUND_PASSES && nCand > 0) {
            int nAct = 0;
            for (int r = 0; r <= st.gridRounds; ++r) nAct += st.bcount[3 + r];
            st.boundStats[4 * pass + 0] = nCand;
            st.boundStats[4 * pass + 1] = nAct;
            st.boundStats[4 * pass + 2] = st.bcount[BND_CTRL - 1];
            st.boundStats[4 * pass + 3] = st.bcount[1];
            st.boundTotals[0] += (unsigned long long)nCand;
            st.boundTotals[1] += (unsigned long long)nAct;
            st.boundTotals[2] += (unsigned long long)st.bcount[BND_CTRL - 1];
            st.boundTotals[3] += (unsigned long long)st.bcount[1];
            st.boundTotals[4] += 1ull;
            *st.candCount = nCand;
        }
        // the corrected-face count of the next pass starts at zero (its list
        // kernel counts into it)
        st.bcount[1] = 0;
    }
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < nT + nCand; idx += gridDim.x * blockDim.x) {
        if (idx >= nT) {
            const int i = idx - nT;
            if (!st.active[i]) continue;
            const BndRec<MF>& R = ((const BndRec<MF>*)st.bndRec)[i];
            const unsigned int first = R.first;
            // every slot's loads unconditional (the record's slots past the
            // cell's faces repeat its last face), then the writes: the
            // faces of the first redistribution take dVf + correction
            // (face_value of both; empty-patch faces keep theirs)
            int fid[MF];
            double dv[MF], cr[MF];
            unsigned char em[MF];
#pragma unroll
            for (int k = 0; k < MF; ++k) fid[k] = R.fid[k];
#pragma unroll
            for (int k = 0; k < MF; ++k) {
                dv[k] = st.dVf[fid[k]];
                cr[k] = st.corr[fid[k]];
                em[k] = m.fEmpty[empty_slot(m, fid[k])];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < MF; ++k)
                if (((first >> k) & 1u) && (fid[k] < m.nIF || !em[k])) st.dVf[fid[k]] = dv[k] + cr[k];
            continue;
        }
        const int c = st.touched[idx];
        const int b = m.cOff[c], nf = m.cOff[c + 1] - b;
        // every load unconditional (slots past nf repeat the last face), in
        // three batches: labels; keys, corrections, owners, empty flags; alpha
        int fid[MF], kpos[MF], kon[MF];
        double kval[MF];
        unsigned char kem[MF];
#pragma unroll
        for (int k = 0; k < MF; ++k) fid[k] = m.cFaces[b + min(k, nf - 1)];
#pragma unroll
        for (int k = 0; k < MF; ++k) {
            kpos[k] = st.corrPos[fid[k]];
            kval[k] = st.corr[fid[k]];
            kon[k] = m.own[fid[k]];
            kem[k] = m.fEmpty[empty_slot(m, fid[k])];
        }
        const double before = st.alpha[c];
        const double Vc = m.V[c];
        __builtin_amdgcn_sched_barrier(0);
        bool kown[MF];
#pragma unroll
        for (int k = 0; k < MF; ++k) {
            if (k >= nf) kpos[k] = -1;
            // face_value: empty-patch faces carry nothing
            kval[k] = (kpos[k] >= 0 && (fid[k] < m.nIF || !kem[k])) ? kval[k] : 0.0;
            kown[k] = kpos[k] >= 0 && kon[k] == c;
        }
        double a = before;
        // corrections in ascending correctedFaces position (selection over
        // the unrolled slots: positions are distinct)
        int last = -1;
        for (int q = 0; q < MF; ++q) {
            int best = -1, bpos = 0x7fffffff;
#pragma unroll
            for (int k = 0; k < MF; ++k)
                if (kpos[k] > last && kpos[k] < bpos) {
                    bpos = kpos[k];
                    best = k;
                }
            if (best < 0) break;
            double v = 0.0;
            bool own = false;
#pragma unroll
            for (int k = 0; k < MF; ++k)
                if (k == best) {
                    v = kval[k];
                    own = kown[k];
                }
            a = own ? a - v / Vc : a + v / Vc;
            last = bpos;
        }
        st.alpha[c] = a;
        const int dOob = (int)loop_oob(a, aTol) - (int)loop_oob(before, aTol);
        if (dOob) atomicAdd(&st.oobCount[OOB_STRIDE * ((c >> 8) & (OOB_SLOTS - 1))], dOob);
        const bool candNew = cell_oob(a, aTol) && need_bounding(st, c, epoch);
        const unsigned long long bit = 1ull << (c & 63);
        const bool candOld = (st.candBits[c >> 6] & bit) != 0ull;
        if (candNew != candOld) {
            if (candNew) atomicOr(&st.candBits[c >> 6], bit);
            else atomicAnd(&st.candBits[c >> 6], ~bit);
        }
        st.touchedFlag[c] = 0;
    }
}

template <int MF>
__global__ void __launch_bounds__(256) k_bnd_apply(DevMesh m, DevState st, int pass) {
    k_bnd_apply_body<MF>(m, st, pass, st.epoch);
}

template <int MF>
void launch_bound_pass(const DevMesh& m, const DevState& st, int pass, double dt, const double* Sp,
                       const double* Su, const BndPrev& prev, hipStream_t s) {
    const int nWords = (m.nC + 63) / 64;
    // candidate-sized work: a modest grid-stride grid launches faster than
    // one thread per cell of the mesh
    const int grid = std::max(1, std::min(st.bndGrid, cdiv(m.nC, 256)));
    const int nbw = cdiv(nWords, st.bndListWords);
    k_bnd_list<<<nbw, 256, 0, s>>>(m, st, nWords, nbw, prev);
    k_bnd_prep<MF><<<grid, 256, 0, s>>>(m, st, dt, Sp, Su);
    if (st.bndFlow) {
        k_bnd_flow<MF><<<std::max(1, std::min(st.bndFlow, cdiv(m.nC, flow_chunk<MF>()))), flow_threads<MF>(), 0,
                         s>>>(m, st, dt);
    } else {
        k_bnd_scatter<MF><<<grid, 256, 0, s>>>(m, st);
        for (int r = 0; r < st.gridRounds; ++r) k_bnd_round<MF><<<grid, 256, 0, s>>>(m, st, r, dt);
        k_bnd_tail<MF><<<1, TAIL_THREADS, 0, s>>>(m, st, dt);
    }
    k_bnd_apply<MF><<<grid, 256, 0, s>>>(m, st, pass);
}

// The state of bounding pass number seq of a handle: its candidate list,
// active flags, corrections, correctedFaces keys and writer indices are the
// half (seq & 1) of their double buffers, so the pass's list kernel can clear
// what the previous pass left in the other half while it writes its own.
DevState bound_pass_view(const DevMesh& m, const DevState& st, int seq) {
    DevState v = st;
    const long h = seq & 1;
    v.cand = st.cand + h * m.nC;
    v.active = st.active + h * m.nC;
    v.actIdx = st.actIdx + h * m.nC;
    v.corr = st.corr + h * m.nF;
    v.corrPos = st.corrPos + h * m.nF;
    return v;
}

BndPrev bound_prev(const DevState& other) {
    return BndPrev{other.cand, other.active, other.corr, other.corrPos, other.actIdx};
}

int bound_faces(int maxCellFaces) {
    if (maxCellFaces <= 4) return 4;
    if (maxCellFaces <= 6) return 6;
    if (maxCellFaces <= NS_CELL_FACES) return NS_CELL_FACES;
    return MAX_CELL_FACES;
}

size_t bound_record_bytes(int maxCellFaces) {
    switch (bound_faces(maxCellFaces)) {
        case 4: return sizeof(BndRec<4>);
        case 6: return sizeof(BndRec<6>);
        case NS_CELL_FACES: return sizeof(BndRec<NS_CELL_FACES>);
        default: return sizeof(BndRec<MAX_CELL_FACES>);
    }
}

void launch_bound(const DevMesh& m, const DevState& st, const Params& p, double dt, const double* Sp,
                  const double* Su, int* passSeq, hipStream_t s) {
    const int nPasses = std::min(p.nAlphaBounds, MAX_BOUND_PASSES);
    for (int pass = 0; pass < nPasses; ++pass) {
        const int seq = (*passSeq)++;
        const DevState sp = bound_pass_view(m, st, seq);
        const BndPrev prev = bound_prev(bound_pass_view(m, st, seq + 1));
        switch (bound_faces(st.maxCellFaces)) {
            case 4: launch_bound_pass<4>(m, sp, pass, dt, Sp, Su, prev, s); break;
            case 6: launch_bound_pass<6>(m, sp, pass, dt, Sp, Su, prev, s); break;
            case NS_CELL_FACES: launch_bound_pass<NS_CELL_FACES>(m, sp, pass, dt, Sp, Su, prev, s); break;
            default: launch_bound_pass<MAX_CELL_FACES>(m, sp, pass, dt, Sp, Su, prev, s); break;
        }
    }
}

// applyBruteForceBounding (advection.C:291-308)
__device__ __attribute__((always_inline)) inline void k_snap_clip_body(int nC, double* alpha, double snapTol, int clip, const DevState& st) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    double a = alpha[c];
    if (snapTol > 0.0) a = a * pos0(a - snapTol) * neg0(a - (1.0 - snapTol)) + pos0(a - (1.0 - snapTol));
    if (clip) a = fmin(1.0, fmax(0.0, a));
    alpha[c] = a;
}

__global__ void k_snap_clip(int nC, double* alpha, double snapTol, int clip, DevState st) {
    k_snap_clip_body(nC, alpha, snapTol, clip, st);
}

__global__ void k_zero_int(int* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

void launch_advect_stage(int stage, const DevMesh& m, const DevState& st, const Params& p, double dt,
                         const double* Sp, const double* Su, int* passSeq, hipStream_t s) {
    switch (stage) {
        case STAGE_UPWIND:
            k_upwind<<<cdiv(m.nF, 256), 256, 0, s>>>(m, st, dt);
            break;
        case STAGE_IFLUX:
            if (!st.un0Fused) k_un0<<<512, 64, 0, s>>>(m, st);
            if (st.maxCellFaces > 4) {
                k_flux_list<<<1024, 256, 0, s>>>(m, st, st.maxCellFaces);
                if (m.triQuadFaces && st.fluxFan) k_flux_fan<<<4096, 64, 0, s>>>(m, st, dt, st.maxCellFaces);
                else if (m.triQuadFaces) k_interface_flux_faces<true, true><<<2048, 64, 0, s>>>(m, st, dt, st.maxCellFaces);
                else k_interface_flux_faces<false, true><<<2048, 64, 0, s>>>(m, st, dt, st.maxCellFaces);
            } else {
                if (m.triQuadFaces) k_interface_flux_faces<true, false><<<2048, 64, 0, s>>>(m, st, dt, st.maxCellFaces);
                else k_interface_flux_faces<false, false><<<2048, 64, 0, s>>>(m, st, dt, st.maxCellFaces);
            }
            break;
        case STAGE_UPDATE:
            launch_update(m, st, dt, Sp, Su, s);
            break;
        case STAGE_BOUND:
            launch_bound(m, st, p, dt, Sp, Su, passSeq, s);
            break;
        case STAGE_SNAPCLIP:
            if (p.snapTol > 0.0 || p.clip)
                k_snap_clip<<<cdiv(m.nC, 256), 256, 0, s>>>(m.nC, st.alpha, p.snapTol, p.clip, st);
            break;
        default:
            break;
    }
}

__global__ void k_alpha_phi(DevMesh m, DevState st, double dtStep) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nF) return;
    st.alphaPhi[f] = dtStep > 0.0 ? st.dVf[f] / dtStep : 0.0;
}

void launch_alpha_phi(const DevMesh& m, const DevState& st, double dt, hipStream_t s) {
    k_alpha_phi<<<cdiv(m.nF, 256), 256, 0, s>>>(m, st, dt);
}

// advection::getRhoPhi with constant densities (advection.H:329-343):
// rhoPhi = (rho1 - rho2) alphaPhi + rho2 phi per face.  mode 0 writes it,
// mode 1 starts the sub-cycle sum rhoPhiSum = 0 + w rhoPhi and mode 2 adds
// w rhoPhi to it (alphaEqnSubCycle.H:33-52, w = subDeltaT / totalDeltaT).
__global__ void k_rho_phi(int nF, const double* __restrict__ ap, const double* __restrict__ phi, double drho,
                          double rho2, double w, double* __restrict__ out, int mode) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    const double rp = drho * ap[f] + rho2 * phi[f];
    if (mode == 0) out[f] = rp;
    else out[f] = (mode == 1 ? 0.0 : out[f]) + w * rp;
}

void launch_rho_phi(int nF, const double* ap, const double* phi, double rho1, double rho2, double w,
                    double* out, int mode, hipStream_t s) {
    if (nF > 0) k_rho_phi<<<cdiv(nF, 256), 256, 0, s>>>(nF, ap, phi, rho1 - rho2, rho2, w, out, mode);
}

// ---------------------------------------------------------------------------
// Batched cases: independent handles advanced together, one launch per stage
// with blockIdx.y = case.  Every case runs the single-case kernel body on its
// own mesh and state (its own tickets, look-back words and counters), so its
// results are those of its handle stepping alone; a case whose step is <= 0
// returns at once.  Blocks past a case's own look-back grid return before
// taking a ticket, so each case's ticket counter still grows by its own
// block count per launch.
// ---------------------------------------------------------------------------
__device__ inline const DevState& batch_view(const BatchCase& c, const BatchStepInfo& si, int swapped, int pass) {
    return c.st[((si.aPar ^ swapped) << 1) | (si.hPar ^ (pass & 1))];
}

__global__ void __launch_bounds__(DETECT_LIST_THREADS) kb_detect_list(const BatchCase* __restrict__ bc,
                                                                      const BatchStepInfo* __restrict__ info,
                                                                      double tol) {
    const BatchCase& c = bc[blockIdx.y];
    const BatchStepInfo si = info[blockIdx.y];
    if (!(si.dt > 0.0) || (int)blockIdx.x >= c.detBlocks) return;
    k_detect_list_body(c.m, batch_view(c, si, 0, 0), tol, c.detBlocks);
}

__global__ void __launch_bounds__(64, 1) kb_ls_normal(const BatchCase* __restrict__ bc,
                                                      const BatchStepInfo* __restrict__ info) {
    const BatchCase& c = bc[blockIdx.y];
    const BatchStepInfo si = info[blockIdx.y];
    if (!(si.dt > 0.0)) return;
    k_ls_normal_body(c.m, batch_view(c, si, 0, 0), nullptr, nullptr);
}

template <int G>
__global__ void __launch_bounds__(64, SVP_FSD_WAVES) kb_fsd_group(const BatchCase* __restrict__ bc,
                                                                  const BatchStepInfo* __restrict__ info) {
    const BatchCase& c = bc[blockIdx.y];
    const BatchStepInfo si = info[blockIdx.y];
    if (!(si.dt > 0.0)) return;
    k_fsd_group_body<G>(c.m, batch_view(c, si, 0, 0), nullptr, nullptr);
}

__global__ void __launch_bounds__(256) kb_upwind(const BatchCase* __restrict__ bc,
                                                 const BatchStepInfo* __restrict__ info) {
    const BatchCase& c = bc[blockIdx.y];
    const BatchStepInfo si = info[blockIdx.y];
    if (!(si.dt > 0.0)) return;
    k_upwind_body(c.m, batch_view(c, si, 1, 0), si.dt);
}

__global__ void __launch_bounds__(256) kb_flux_list(const BatchCase* __restrict__ bc,
                                                    const BatchStepInfo* __restrict__ info) {
    const BatchCase& c = bc[blockIdx.y];
    const BatchStepInfo si = info[blockIdx.y];
    if (!(si.dt > 0.0)) return;
    const DevState& st = batch_view(c, si, 1, 0);
    k_flux_list_body(c.m, st, st.maxCellFaces, si.epoch);
}

template <bool TRIQUAD, bool LISTED>
__global__ void __launch_bounds__(64, SVP_FLUX_WAVES) kb_interface_flux_faces(const BatchCase* __restrict__ bc,
                                                                              const BatchStepInfo* __restrict__ info) {
    const BatchCase& c = bc[blockIdx.y];
    const BatchStepInfo si = info[blockIdx.y];
    if (!(si.dt > 0.0)) return;
    const DevState& st = batch_view(c, si, 1, 0);
    k_interface_flux_faces_body<TRIQUAD, LISTED>(c.m, st, si.dt, st.maxCellFaces, si.epoch);
}

template <int MF>
__global__ void __launch_bounds__(UPDATE_BLOCK) kb_update(const BatchCase* __restrict__ bc,
                                                          const BatchStepInfo* __restrict__ info) {
    const BatchCase& c = bc[blockIdx.y];
    const BatchStepInfo si = info[blockIdx.y];
    if (!(si.dt > 0.0)) return;
    k_update_body<MF>(c.m, batch_view(c, si, 1, 0), si.dt, nullptr, nullptr, si.epoch);
}

__global__ void __launch_bounds__(256) kb_bnd_list(const BatchCase* __restrict__ bc,
                                                   const BatchStepInfo* __restrict__ info, int pass) {
    const BatchCase& c = bc[blockIdx.y];
    const BatchStepInfo si = info[blockIdx.y];
    if (!(si.dt > 0.0) || (int)blockIdx.x >= c.listBlocks) return;
    const DevState& o = batch_view(c, si, 1, pass + 1);
    const BndPrev prev{o.cand, o.active, o.corr, o.corrPos, o.actIdx};
    k_bnd_list_body(c.m, batch_view(c, si, 1, pass), c.listWords, c.listBlocks, prev);
}

template <int MF>
__global__ void __launch_bounds__(256) kb_bnd_prep(const BatchCase* __restrict__ bc,
                                                   const BatchStepInfo* __restrict__ info, int pass) {
    const BatchCase& c = bc[blockIdx.y];
    const BatchStepInfo si = info[blockIdx.y];
    if (!(si.dt > 0.0)) return;
    k_bnd_prep_body<MF>(c.m, batch_view(c, si, 1, pass), si.dt, nullptr, nullptr);
}

template <int MF>
__global__ void __launch_bounds__(flow_threads<MF>()) kb_bnd_flow(const BatchCase* __restrict__ bc,
                                                                  const BatchStepInfo* __restrict__ info, int pass) {
    const BatchCase& c = bc[blockIdx.y];
    const BatchStepInfo si = info[blockIdx.y];
    if (!(si.dt > 0.0)) return;
    k_bnd_flow_body<MF>(c.m, batch_view(c, si, 1, pass), si.dt);
}

template <int MF>
__global__ void __launch_bounds__(256) kb_bnd_apply(const BatchCase* __restrict__ bc,
                                                    const BatchStepInfo* __restrict__ info, int pass) {
    const BatchCase& c = bc[blockIdx.y];
    const BatchStepInfo si = info[blockIdx.y];
    if (!(si.dt > 0.0)) return;
    k_bnd_apply_body<MF>(c.m, batch_view(c, si, 1, pass), pass, si.epoch);
}

__global__ void kb_snap_clip(const BatchCase* __restrict__ bc, const BatchStepInfo* __restrict__ info,
                             double snapTol, int clip) {
    const BatchCase& c = bc[blockIdx.y];
    const BatchStepInfo si = info[blockIdx.y];
    if (!(si.dt > 0.0)) return;
    const DevState& st = batch_view(c, si, 1, 0);
    k_snap_clip_body(c.m.nC, st.alpha, snapTol, clip, st);
}

// pass = the pass of the step: its candidate-list half is the case's hPar
// flipped per pass, and it indexes the pass statistics
template <int MF>
static void launch_batch_bound(const BatchCase* bc, const BatchStepInfo* info, int n, int maxListBlocks,
                               const BatchDims& d, int nPasses, hipStream_t s) {
    for (int pass = 0; pass < nPasses; ++pass) {
        kb_bnd_list<<<dim3(maxListBlocks, n), 256, 0, s>>>(bc, info, pass);
        kb_bnd_prep<MF><<<dim3(d.prep, n), 256, 0, s>>>(bc, info, pass);
        kb_bnd_flow<MF><<<dim3(d.flow, n), flow_threads<MF>(), 0, s>>>(bc, info, pass);
        kb_bnd_apply<MF><<<dim3(d.apply, n), 256, 0, s>>>(bc, info, pass);
    }
}

void launch_batch_stage(int stage, const BatchCase* bc, const BatchStepInfo* info, int n, int maxC, int maxF,
                        int maxDetBlocks, int maxListBlocks, const BatchDims& d, const Params& p, int G, int MF,
                        int triQuad, hipStream_t s) {
    switch (stage) {
        case STAGE_DETECT:
            kb_detect_list<<<dim3(maxDetBlocks, n), DETECT_LIST_THREADS, 0, s>>>(bc, info, p.mixedCellTol);
            break;
        case STAGE_NORMAL:
            kb_ls_normal<<<dim3(d.ls, n), 64, 0, s>>>(bc, info);
            break;
        case STAGE_FSD:
            if (G == 4) kb_fsd_group<4><<<dim3(d.fsd, n), 64, 0, s>>>(bc, info);
            else if (G == 6) kb_fsd_group<6><<<dim3(d.fsd, n), 64, 0, s>>>(bc, info);
            else kb_fsd_group<8><<<dim3(d.fsd, n), 64, 0, s>>>(bc, info);
            break;
        case STAGE_UPWIND:
            kb_upwind<<<dim3(cdiv(maxF, 256), n), 256, 0, s>>>(bc, info);
            break;
        case STAGE_IFLUX:
            if (G > 4) {
                kb_flux_list<<<dim3(d.fluxList, n), 256, 0, s>>>(bc, info);
                if (triQuad) kb_interface_flux_faces<true, true><<<dim3(d.fluxFaces, n), 64, 0, s>>>(bc, info);
                else kb_interface_flux_faces<false, true><<<dim3(d.fluxFaces, n), 64, 0, s>>>(bc, info);
            } else {
                if (triQuad) kb_interface_flux_faces<true, false><<<dim3(d.fluxFaces, n), 64, 0, s>>>(bc, info);
                else kb_interface_flux_faces<false, false><<<dim3(d.fluxFaces, n), 64, 0, s>>>(bc, info);
            }
            break;
        case STAGE_UPDATE:
            // MF: the bounding slots of the largest cell (4, 6, 8 or 24)
            if (MF == 4) kb_update<4><<<dim3(cdiv(maxC, UPDATE_BLOCK), n), UPDATE_BLOCK, 0, s>>>(bc, info);
            else if (MF == 6) kb_update<6><<<dim3(cdiv(maxC, UPDATE_BLOCK), n), UPDATE_BLOCK, 0, s>>>(bc, info);
            else if (MF == 8) kb_update<8><<<dim3(cdiv(maxC, UPDATE_BLOCK), n), UPDATE_BLOCK, 0, s>>>(bc, info);
            else kb_update<0><<<dim3(cdiv(maxC, UPDATE_BLOCK), n), UPDATE_BLOCK, 0, s>>>(bc, info);
            break;
        case STAGE_BOUND: {
            const int nPasses = std::min(p.nAlphaBounds, MAX_BOUND_PASSES);
            switch (MF) {
                case 4: launch_batch_bound<4>(bc, info, n, maxListBlocks, d, nPasses, s); break;
                case 6: launch_batch_bound<6>(bc, info, n, maxListBlocks, d, nPasses, s); break;
                case NS_CELL_FACES: launch_batch_bound<NS_CELL_FACES>(bc, info, n, maxListBlocks, d, nPasses, s); break;
                default: launch_batch_bound<MAX_CELL_FACES>(bc, info, n, maxListBlocks, d, nPasses, s); break;
            }
            break;
        }
        case STAGE_SNAPCLIP:
            if (p.snapTol > 0.0 || p.clip)
                kb_snap_clip<<<dim3(cdiv(maxC, 256), n), 256, 0, s>>>(bc, info, p.snapTol, p.clip);
            break;
        default:
            break;
    }
}

}  // namespace svp
