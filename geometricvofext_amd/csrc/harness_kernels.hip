// Harness kernels of the deformation benchmark (plicVofAdvectionFoam) and
// input checks.
#include "simplic_device.hpp"
#include "simplic_types.hpp"

namespace svp {

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// updateU.H:117-127: phi *= uFactor; U *= uFactor
__device__ inline void scale_flow_at(int i, int nC, int nF, int nB, const double* U0, const double* phi0,
                                     const double* U0b, double factor, double* U, double* phi, double* Ub) {
    if (i < nF) phi[i] = phi0[i] * factor;
    if (i < 3 * nC) U[i] = U0[i] * factor;
    if (i < 3 * nB) Ub[i] = U0b ? U0b[i] * factor : 0.0;
}

__global__ void k_scale_flow(int nC, int nF, int nB, const double* U0, const double* phi0,
                             const double* U0b, double factor, double* U, double* phi, double* Ub) {
    scale_flow_at(blockIdx.x * blockDim.x + threadIdx.x, nC, nF, nB, U0, phi0, U0b, factor, U, phi, Ub);
}

// batch: every case's base flow scaled by its own factor (blockIdx.y = case)
__global__ void kb_scale_flow(const BatchCase* __restrict__ bc, const double* __restrict__ factors) {
    const BatchCase& c = bc[blockIdx.y];
    const DevState& st = c.st[0];
    scale_flow_at(blockIdx.x * blockDim.x + threadIdx.x, c.m.nC, c.m.nF, c.nB, c.U0, c.phi0, c.U0b,
                  factors[blockIdx.y], st.U, st.phi, st.Ub);
}

void launch_batch_scale_flow(const BatchCase* bc, const BatchStepInfo* info, const double* factors, int nCases,
                             long maxN, hipStream_t s) {
    (void)info;
    kb_scale_flow<<<dim3(cdiv(maxN, 256), nCases), 256, 0, s>>>(bc, factors);
}

void launch_scale_flow(int nC, int nF, int nB, const double* U0, const double* phi0, const double* U0b,
                       double factor, double* U, double* phi, double* Ub, hipStream_t s) {
    const long n = std::max<long>(std::max<long>(nF, 3L * nC), 3L * nB);
    k_scale_flow<<<cdiv(n, 256), 256, 0, s>>>(nC, nF, nB, U0, phi0, U0b, factor, U, phi, Ub);
}

// fvc::surfaceSum(mag(phi)) in ascending face order, then max of sum/V and of
// (pos0(alpha-0.01)*pos0(0.99-alpha)*sum)/V.  All values are >= 0, so the max
// is an integer max over the IEEE bit patterns.  Grid-stride blocks reduce in
// LDS and issue one atomic pair per block.
__device__ inline void courant_cell(const DevMesh& m, const double* phi, const double* alpha, int c, double& r,
                                    double& ra) {
    double s = 0.0;
    for (int k = m.cOff[c]; k < m.cOff[c + 1]; ++k) {
        const int e = m.cFacesAsc[k];
        const int f = e >= 0 ? e : ~e;
        if (f >= m.nIF && m.fEmpty[f]) continue;
        s += fabs(phi[f]);
    }
    const double a = alpha[c];
    r = s / m.V[c];
    ra = (pos0(a - 0.01) * pos0(0.99 - a) * s) / m.V[c];
}

__device__ inline unsigned long long umax64(unsigned long long a, unsigned long long b) { return a > b ? a : b; }

__device__ __attribute__((always_inline)) inline void k_courant_body(const DevMesh& m, const double* phi,
                                                                    const double* alpha, unsigned long long* out) {
    __shared__ unsigned long long lr[4], lra[4];
    unsigned long long br = 0ull, bra = 0ull;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < m.nC; c += gridDim.x * 256) {
        double r, ra;
        courant_cell(m, phi, alpha, c, r, ra);
        br = umax64(br, (unsigned long long)__double_as_longlong(r));
        bra = umax64(bra, (unsigned long long)__double_as_longlong(ra));
    }
    for (int o = 32; o > 0; o >>= 1) {
        br = umax64(br, __shfl_xor(br, o));
        bra = umax64(bra, __shfl_xor(bra, o));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lr[w] = br;
        lra[w] = bra;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicMax(out, umax64(umax64(lr[0], lr[1]), umax64(lr[2], lr[3])));
        atomicMax(out + 1, umax64(umax64(lra[0], lra[1]), umax64(lra[2], lra[3])));
    }
}

__global__ void __launch_bounds__(256) k_courant(DevMesh m, const double* phi, const double* alpha,
                                                 unsigned long long* out) {
    k_courant_body(m, phi, alpha, out);
}

// batch: the Courant sums of case blockIdx.y into out[2 y], out[2 y + 1],
// from its current alpha (view aPar)
__global__ void __launch_bounds__(256) kb_courant(const BatchCase* __restrict__ bc,
                                                  const BatchStepInfo* __restrict__ info, unsigned long long* out) {
    const BatchCase& c = bc[blockIdx.y];
    const DevState& st = c.st[info[blockIdx.y].aPar << 1];
    k_courant_body(c.m, st.phi, st.alpha, out + 2 * blockIdx.y);
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

void launch_courant(const DevMesh& m, const double* phi, const double* alpha, unsigned long long* out,
                    hipStream_t s) {
    k_zero_u64<<<1, 64, 0, s>>>(out, 2);
    k_courant<<<std::min(2048, cdiv(m.nC, 256)), 256, 0, s>>>(m, phi, alpha, out);
}


// ---------------------------------------------------------------------------
// Initial alpha of a union of spheres / ellipsoids by adaptive tetrahedral
// subdivision: each cell is split into tets (cell centre, face centre, face
// edge); a tet whose ball misses every shape is empty, a tet with all vertices
// inside one (convex) shape is full, any other tet is refined 1:8 (edge
// midpoints) down to maxDepth, where the linear interpolant of the shape
// function is integrated exactly.  Shape k: g_k(x) = |R_k (x - c_k) / a_k|^2 - 1,
// union f = min_k g_k.  The miss test runs in each shape's own frame
// q = R_k (x - c_k) / a_k, where the shape is the unit ball and a tet maps to
// the tet of its mapped vertices: a tet misses the shape when the ball about
// the mapped centroid through the farthest mapped vertex misses the unit
// ball.
// ---------------------------------------------------------------------------
struct Shape {
    double c[3], a[3], R[9];   // centre, semi-axes, rows of the rotation
};

// x in the shape's frame, where the shape is the unit ball
__device__ inline V3 shape_q(const Shape& s, V3 x) {
    const double d0 = x.x - s.c[0], d1 = x.y - s.c[1], d2 = x.z - s.c[2];
    return v3((s.R[0] * d0 + s.R[1] * d1 + s.R[2] * d2) / s.a[0], (s.R[3] * d0 + s.R[4] * d1 + s.R[5] * d2) / s.a[1],
              (s.R[6] * d0 + s.R[7] * d1 + s.R[8] * d2) / s.a[2]);
}

__device__ inline double shape_gq(V3 q) { return q.x * q.x + q.y * q.y + q.z * q.z - 1.0; }

__device__ inline double shape_g(const Shape& s, V3 x) { return shape_gq(shape_q(s, x)); }

// volume fraction of {f < 0} for f linear on a tet with vertex values v[4]:
// CDF of a linear function over a simplex, sum_i (-v_i)_+^3 / prod_j (v_j - v_i).
// Summed over all four vertices the terms give 1 (the third divided
// difference of (-x)^3), so with three negative values the fraction is one
// minus the positive vertex's term.
__device__ inline double linear_tet_fraction(double v0, double v1, double v2, double v3) {
    double v[4] = {v0, v1, v2, v3};
    int nneg = 0;
    for (int i = 0; i < 4; ++i) nneg += v[i] < 0.0;
    if (nneg == 0) return 0.0;
    if (nneg == 4) return 1.0;
    // separate nearly equal values so the divided differences stay finite
    double scale = 0.0;
    for (int i = 0; i < 4; ++i) scale = fmax(scale, fabs(v[i]));
    const double eps = 1e-9 * scale;
    for (int i = 1; i < 4; ++i)
        for (int j = 0; j < i; ++j)
            if (fabs(v[i] - v[j]) < eps) v[i] += (v[i] >= 0.0 ? 1.0 : -1.0) * eps * (1 + i);
    const bool useNeg = nneg <= 2;
    double sum = 0.0;
    for (int i = 0; i < 4; ++i) {
        const bool neg = v[i] < 0.0;
        if (neg != useNeg) continue;
        double den = 1.0;
        for (int j = 0; j < 4; ++j)
            if (j != i) den *= (v[j] - v[i]);
        const double t = -v[i];
        sum += (t * t * t) / den;
    }
    const double frac = useNeg ? sum : 1.0 - sum;
    return fmin(1.0, fmax(0.0, frac));
}

__device__ inline double tet_vol(V3 a, V3 b, V3 c, V3 d) {
    return fabs(dot(b - a, cross(c - a, d - a))) / 6.0;
}

__global__ void k_init_alpha(DevMesh m, const Shape* shapes, int nShapes, int maxDepth, double* alpha) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const V3 cc = ld3(m.C, c);
    // quick reject/accept of the whole cell by its bounding ball
    double R = 0.0;
    for (int k = m.cOff[c]; k < m.cOff[c + 1]; ++k) {
        const int f = m.cFaces[k];
        for (int q = m.fOff[f]; q < m.fOff[f + 1]; ++q) R = fmax(R, mag(ld3(m.pts, m.fPts[q]) - cc));
    }
    bool touch = false;
    for (int s = 0; s < nShapes; ++s) {
        // the cell's ball maps into the ball of radius R / (smallest semi-axis)
        const Shape& S = shapes[s];
        const double small = fmin(S.a[0], fmin(S.a[1], S.a[2]));
        if (mag(shape_q(S, cc)) - R / small <= 1.0) touch = true;
    }
    if (!touch) {
        alpha[c] = 0.0;
        return;
    }
    struct Tet { V3 p[4]; int depth; };
    Tet stack[64];
    double volIn = 0.0, volAll = 0.0;
    for (int k = m.cOff[c]; k < m.cOff[c + 1]; ++k) {
        const int f = m.cFaces[k];
        const int fs = m.fOff[f], n = m.fOff[f + 1] - fs;
        const V3 fc = ld3(m.Cf, f);
        for (int q = 0; q < n; ++q) {
            int sp = 0;
            stack[sp++] = Tet{{cc, fc, ld3(m.pts, m.fPts[fs + q]), ld3(m.pts, m.fPts[fs + (q + 1) % n])}, 0};
            while (sp > 0) {
                const Tet t = stack[--sp];
                const double vol = tet_vol(t.p[0], t.p[1], t.p[2], t.p[3]);
                volAll += vol;
                bool full = false, maybe = false;
                for (int s = 0; s < nShapes && !full; ++s) {
                    const Shape& S = shapes[s];
                    V3 q[4];
                    bool allIn = true;
                    for (int i = 0; i < 4; ++i) {
                        q[i] = shape_q(S, t.p[i]);
                        allIn &= shape_gq(q[i]) < 0.0;
                    }
                    if (allIn) full = true;
                    const V3 gq = 0.25 * (q[0] + q[1] + q[2] + q[3]);
                    double rad = 0.0;
                    for (int i = 0; i < 4; ++i) rad = fmax(rad, mag(q[i] - gq));
                    if (mag(gq) - rad <= 1.0) maybe = true;
                }
                if (full) { volIn += vol; continue; }
                if (!maybe) continue;
                if (t.depth >= maxDepth || sp + 8 > 64) {
                    double fv[4];
                    for (int i = 0; i < 4; ++i) {
                        double mn = 1e300;
                        for (int s = 0; s < nShapes; ++s) mn = fmin(mn, shape_g(shapes[s], t.p[i]));
                        fv[i] = mn;
                    }
                    volIn += vol * linear_tet_fraction(fv[0], fv[1], fv[2], fv[3]);
                    continue;
                }
                // not counted twice: remove this tet's volume, children add theirs
                volAll -= vol;
                const V3 a = t.p[0], b = t.p[1], cq = t.p[2], d = t.p[3];
                const V3 ab = 0.5 * (a + b), ac = 0.5 * (a + cq), ad = 0.5 * (a + d);
                const V3 bc = 0.5 * (b + cq), bd = 0.5 * (b + d), cd = 0.5 * (cq + d);
                const int nd = t.depth + 1;
                stack[sp++] = Tet{{a, ab, ac, ad}, nd};
                stack[sp++] = Tet{{ab, b, bc, bd}, nd};
                stack[sp++] = Tet{{ac, bc, cq, cd}, nd};
                stack[sp++] = Tet{{ad, bd, cd, d}, nd};
                stack[sp++] = Tet{{ab, ac, ad, bd}, nd};
                stack[sp++] = Tet{{ab, ac, bc, bd}, nd};
                stack[sp++] = Tet{{ac, ad, bd, cd}, nd};
                stack[sp++] = Tet{{ac, bc, bd, cd}, nd};
            }
        }
    }
    alpha[c] = volAll > 0.0 ? fmin(1.0, fmax(0.0, volIn / volAll)) : 0.0;
}

void launch_init_alpha(const DevMesh& m, const void* shapes, int nShapes, int maxDepth, double* alpha,
                       hipStream_t s) {
    k_init_alpha<<<cdiv(m.nC, 128), 128, 0, s>>>(m, (const Shape*)shapes, nShapes, maxDepth, alpha);
}

__global__ void k_count_nonfinite(const double* x, long n, int* count) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool bad = (i < n) && !isfinite(x[i]);
    const unsigned long long b = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(count, __popcll(b));
}

void launch_count_nonfinite(const double* x, long n, int* count, hipStream_t s) {
    if (n > 0) k_count_nonfinite<<<cdiv(n, 256), 256, 0, s>>>(x, n, count);
}

void launch_batch_courant(const BatchCase* bc, const BatchStepInfo* info, int nCases, int maxC,
                          unsigned long long* out, hipStream_t s) {
    k_zero_u64<<<cdiv(2L * nCases, 256), 256, 0, s>>>(out, 2 * nCases);
    kb_courant<<<dim3(std::min(2048, cdiv(maxC, 256)), nCases), 256, 0, s>>>(bc, info, out);
}

}  // namespace svp
