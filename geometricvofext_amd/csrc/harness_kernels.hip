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

__global__ void k_zero_one_int(int* p) {
    if (threadIdx.x == 0) *p = 0;
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

// ---------------------------------------------------------------------------
// Initial alpha from a closed triangulated surface: the volume of each cell
// inside the surface, exactly (the setVofField utility's CGAL boolean,
// setAlphaField.H:41-62, without CGAL).  With O any point, the indicator of
// the enclosed region is the signed sum of the cones tet(O, t) over the
// surface triangles t (signs from the triangles' orientation), so
//     vol(cell n inside) = sum_t sign_t * vol(tet(O, t) n cell).
// The cell is split into the tets (cell centre, face centre, face edge) of
// its faces; O is the cell centre, and each cone-tet intersection is the cone
// tet clipped by the cell tet's four planes (a tet cut by a plane leaves a
// tet or a prism, itself three tets).  Cells whose centre lies farther from
// every triangle than their farthest point take 1 or 0 from the winding
// number of their centre (solid angles, van Oosterom & Strackee).
// ---------------------------------------------------------------------------
struct Tet4 {
    V3 p[4];
};

__device__ inline double tet_abs_vol(const Tet4& t) { return tet_vol(t.p[0], t.p[1], t.p[2], t.p[3]); }

// point-triangle distance (closest point on the triangle; Ericson, Real-Time
// Collision Detection, 5.1.5)
__device__ inline double point_tri_dist(V3 p, V3 a, V3 b, V3 c) {
    const V3 ab = b - a, ac = c - a, ap = p - a;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return mag(ap);
    const V3 bp = p - b;
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return mag(bp);
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return mag(ap - (d1 / (d1 - d3)) * ab);
    const V3 cp = p - c;
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return mag(cp);
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return mag(ap - (d2 / (d2 - d6)) * ac);
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0)
        return mag(bp - ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b));
    const double den = 1.0 / (va + vb + vc);
    return mag(ap - (vb * den) * ab - (vc * den) * ac);
}

// solid angle of triangle (a, b, c) seen from p (van Oosterom & Strackee 1983)
__device__ inline double tri_solid_angle(V3 p, V3 a, V3 b, V3 c) {
    const V3 ra = a - p, rb = b - p, rc = c - p;
    const double la = mag(ra), lb = mag(rb), lc = mag(rc);
    const double num = dot(ra, cross(rb, rc));
    const double den = la * lb * lc + dot(ra, rb) * lc + dot(ra, rc) * lb + dot(rb, rc) * la;
    return 2.0 * atan2(num, den);
}

// volume of the cone tet clipped by the four planes of the cell tet T
__device__ inline double clip_tet_volume(const Tet4& cone, const Tet4& T) {
    // planes of T: the side of each face that holds the opposite vertex
    V3 pn[4];
    double pd[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const V3 a = T.p[(k + 1) & 3], b = T.p[(k + 2) & 3], c = T.p[(k + 3) & 3];
        V3 n = cross(b - a, c - a);
        if (dot(n, T.p[k] - a) < 0.0) n = (-1.0) * n;
        pn[k] = n;
        pd[k] = dot(n, a);
    }
    // depth-first clipping, plane by plane
    Tet4 stack[16];
    int level[16];
    int sp = 0;
    stack[sp] = cone;
    level[sp++] = 0;
    double vol = 0.0;
    while (sp > 0) {
        --sp;
        const Tet4 t = stack[sp];
        const int lv = level[sp];
        if (lv == 4) {
            vol += tet_abs_vol(t);
            continue;
        }
        double sd[4];
        int in[4], out[4], ni = 0, no = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            sd[i] = dot(pn[lv], t.p[i]) - pd[lv];
            if (sd[i] >= 0.0) in[ni++] = i;
            else out[no++] = i;
        }
        if (ni == 0) continue;
        if (sp + 3 > 16) continue;  // cannot happen: at most 3 pieces per level, 4 levels deep
        auto cut = [&](int i, int o) {
            const double w = sd[i] / (sd[i] - sd[o]);
            return t.p[i] + w * (t.p[o] - t.p[i]);
        };
        auto push = [&](V3 a, V3 b, V3 c, V3 d) {
            stack[sp] = Tet4{{a, b, c, d}};
            level[sp++] = lv + 1;
        };
        if (ni == 4) {
            stack[sp] = t;
            level[sp++] = lv + 1;
        } else if (ni == 1) {
            const int a = in[0];
            push(t.p[a], cut(a, out[0]), cut(a, out[1]), cut(a, out[2]));
        } else {
            // prism with bottom (p0, p1, p2) and top (q0, q1, q2), pi - qi edges:
            // tets (p0 p1 p2 q0), (p1 p2 q0 q1), (p2 q0 q1 q2)
            V3 P[3], Q[3];
            if (ni == 3) {
                const int d = out[0];
                for (int k = 0; k < 3; ++k) {
                    P[k] = t.p[in[k]];
                    Q[k] = cut(in[k], d);
                }
            } else {
                const int a = in[0], b = in[1];
                P[0] = t.p[a];
                P[1] = cut(a, out[0]);
                P[2] = cut(a, out[1]);
                Q[0] = t.p[b];
                Q[1] = cut(b, out[0]);
                Q[2] = cut(b, out[1]);
            }
            push(P[0], P[1], P[2], Q[0]);
            push(P[1], P[2], Q[0], Q[1]);
            push(P[2], Q[0], Q[1], Q[2]);
        }
    }
    return vol;
}

// per cell: 0 / 1 from the winding number for cells away from the surface,
// else the cell joins the list of cut cells
__global__ void __launch_bounds__(256) k_surface_classify(DevMesh m, const double* tris, int nT, double* frac,
                                                          int* cutList, int* nCut) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const V3 cc = ld3(m.C, c);
    double R = 0.0;
    for (int k = m.cOff[c]; k < m.cOff[c + 1]; ++k) {
        const int f = m.cFaces[k];
        for (int q = m.fOff[f]; q < m.fOff[f + 1]; ++q) R = fmax(R, mag(ld3(m.pts, m.fPts[q]) - cc));
    }
    double dmin = 1e300, omega = 0.0;
    for (int t = 0; t < nT; ++t) {
        const V3 a = ld3(tris, 3 * t), b = ld3(tris, 3 * t + 1), cq = ld3(tris, 3 * t + 2);
        dmin = fmin(dmin, point_tri_dist(cc, a, b, cq));
        omega += tri_solid_angle(cc, a, b, cq);
    }
    if (dmin > R * (1.0 + 1e-9)) {
        frac[c] = fabs(omega) > 2.0 * M_PI ? 1.0 : 0.0;   // winding number 1 or 0
    } else {
        frac[c] = -1.0;
        cutList[atomicAdd(nCut, 1)] = c;
    }
}

// one workgroup per cut cell: its (cell tet, triangle) pairs over the
// threads, each thread's sum in a fixed order, then a fixed-order tree
constexpr int SURF_THREADS = 256;

__global__ void __launch_bounds__(SURF_THREADS) k_surface_cut(DevMesh m, const double* tris, int nT,
                                                              const int* cutList, const int* nCut, double* frac) {
    __shared__ double sv[SURF_THREADS], sw[SURF_THREADS];
    const int n = *nCut;
    for (int w = blockIdx.x; w < n; w += gridDim.x) {
        const int c = cutList[w];
        const V3 cc = ld3(m.C, c);
        const int f0 = m.cOff[c], f1 = m.cOff[c + 1];
        // the cell's tets, enumerated as (face slot, edge)
        int nTet = 0;
        for (int k = f0; k < f1; ++k) {
            const int f = m.cFaces[k];
            nTet += m.fOff[f + 1] - m.fOff[f];
        }
        double acc = 0.0, vcell = 0.0;
        for (long j = threadIdx.x; j < (long)nTet * nT; j += SURF_THREADS) {
            const int it = (int)(j / nT), t = (int)(j % nT);
            // cell tet number it
            int k = f0, e = it;
            for (; k < f1; ++k) {
                const int f = m.cFaces[k];
                const int nf = m.fOff[f + 1] - m.fOff[f];
                if (e < nf) break;
                e -= nf;
            }
            const int f = m.cFaces[k];
            const int s0 = m.fOff[f], nf = m.fOff[f + 1] - s0;
            const Tet4 T{{cc, ld3(m.Cf, f), ld3(m.pts, m.fPts[s0 + e]), ld3(m.pts, m.fPts[s0 + (e + 1) % nf])}};
            if (t == 0) vcell += tet_abs_vol(T);
            const V3 a = ld3(tris, 3 * t), b = ld3(tris, 3 * t + 1), cq = ld3(tris, 3 * t + 2);
            const double sgn6 = dot(a - cc, cross(b - cc, cq - cc));
            if (sgn6 == 0.0) continue;
            const double v = clip_tet_volume(Tet4{{cc, a, b, cq}}, T);
            acc += sgn6 > 0.0 ? v : -v;
        }
        sv[threadIdx.x] = acc;
        sw[threadIdx.x] = vcell;
        __syncthreads();
        for (int o = SURF_THREADS / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) {
                sv[threadIdx.x] += sv[threadIdx.x + o];
                sw[threadIdx.x] += sw[threadIdx.x + o];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) frac[c] = sw[0] > 0.0 ? sv[0] / sw[0] : 0.0;
        __syncthreads();
    }
}

// alpha = fraction (mode 0), alpha + fraction (1) or alpha - fraction (2),
// then clipped to [0, 1]
__global__ void __launch_bounds__(256) k_surface_apply(int nC, const double* frac, int mode, double* alpha) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    const double v = mode == 0 ? frac[c] : (mode == 1 ? alpha[c] + frac[c] : alpha[c] - frac[c]);
    alpha[c] = fmin(1.0, fmax(0.0, v));
}

void launch_init_alpha_surface(const DevMesh& m, const double* tris, int nT, int mode, double* frac, int* cutList,
                               int* nCut, double* alpha, hipStream_t s) {
    k_zero_one_int<<<1, 64, 0, s>>>(nCut);
    k_surface_classify<<<cdiv(m.nC, 256), 256, 0, s>>>(m, tris, nT, frac, cutList, nCut);
    k_surface_cut<<<2048, SURF_THREADS, 0, s>>>(m, tris, nT, cutList, nCut, frac);
    k_surface_apply<<<cdiv(m.nC, 256), 256, 0, s>>>(m.nC, frac, mode, alpha);
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
