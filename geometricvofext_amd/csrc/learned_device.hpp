// Device math of the learned interface-normal estimator: BF16 rounding, the
// liquid fraction of an axis-aligned cube cut by a plane, the symmetry
// canonicalisation of a 3x3x3 volume-fraction stencil, GELU, and the layout
// of the 27 -> 128 -> 128 -> 3 network.
//
// The estimator has no counterpart in the reference: it is a new entry of the
// orientationMethod selector (reconstruction.C:60-68) that feeds the unchanged
// FP64 findSignedDistance (cutCell.C:611-799).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svp {
namespace nm {

// ---------------------------------------------------------------------------
// network layout
// ---------------------------------------------------------------------------
constexpr int NIN = 27;    // stencil values
constexpr int KIN = 32;    // padded input width (one K block of the MFMA)
constexpr int NH = 128;    // hidden width
constexpr int NOUT = 3;
constexpr int NOUTP = 16;  // padded output rows (one MFMA tile)

// packed parameters (the public order): W1[128][27], b1[128], W2[128][128],
// b2[128], W3[3][128], b3[3]
constexpr int P_W1 = 0;
constexpr int P_B1 = P_W1 + NH * NIN;
constexpr int P_W2 = P_B1 + NH;
constexpr int P_B2 = P_W2 + NH * NH;
constexpr int P_W3 = P_B2 + NH;
constexpr int P_B3 = P_W3 + NOUT * NH;
constexpr int N_PARAMS = P_B3 + NOUT;  // 20483

// padded FP32 layout used on the device (gradients, Adam moments):
// W1[128][32], b1[128], W2[128][128], b2[128], W3[16][128], b3[16]
constexpr int Q_W1 = 0;
constexpr int Q_B1 = Q_W1 + NH * KIN;
constexpr int Q_W2 = Q_B1 + NH;
constexpr int Q_B2 = Q_W2 + NH * NH;
constexpr int Q_W3 = Q_B2 + NH;
constexpr int Q_B3 = Q_W3 + NOUTP * NH;
constexpr int N_PADDED = Q_B3 + NOUTP;  // 22800

// BF16 copies of the weights read by the MFMA kernels (element offsets):
// W1b[128][32], W2b[128][128], W2Tb[128][128] (transpose), W3b[16][128]
constexpr int B_W1 = 0;
constexpr int B_W2 = B_W1 + NH * KIN;
constexpr int B_W2T = B_W2 + NH * NH;
constexpr int B_W3 = B_W2T + NH * NH;
constexpr int N_BF16 = B_W3 + NOUTP * NH;

// ---------------------------------------------------------------------------
// BF16 (round to nearest even; inputs are finite)
// ---------------------------------------------------------------------------
__host__ __device__ inline uint16_t f2bf(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }
__host__ __device__ inline float rbf(float f) { return bf2f(f2bf(f)); }

// GELU with the exact erf form and its derivative
__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// ---------------------------------------------------------------------------
// Liquid fraction of the unit cube {0 <= xi <= 1} below the plane
// m . xi = a, with m >= 0 and m0 + m1 + m2 = 1 (Scardovelli & Zaleski,
// J. Comput. Phys. 164 (2000), with the inclusion-exclusion terms regrouped
// so that no case divides a cancelling difference by a small m).
// ---------------------------------------------------------------------------
__host__ __device__ inline double cube_fraction_sorted(double m1, double m2, double m3, double a) {
    // m1 <= m2 <= m3, sum 1, 0 <= a <= 1/2
    const double m12 = m1 + m2;
    if (a <= 0.0) return 0.0;
    if (a < m1) return a * a * a / (6.0 * m1 * m2 * m3);
    const double base = 3.0 * a * a - 3.0 * a * m1 + m1 * m1;
    if (a < m2) return base / (6.0 * m2 * m3);
    if (a < m12 && a < m3) {
        const double b = a - m2;
        return (base - b * b * b / m1) / (6.0 * m2 * m3);
    }
    if (m3 < m12) {  // m3 <= a < m12
        const double b = a - m2, c = a - m3;
        return (base - (b * b * b + c * c * c) / m1) / (6.0 * m2 * m3);
    }
    return (2.0 * a - m12) / (2.0 * m3);  // m12 <= a <= m3
}

// d cube_fraction_sorted / da: the area of the cut (same pieces and
// arguments as cube_fraction_sorted)
__host__ __device__ inline double cube_area_sorted(double m1, double m2, double m3, double a) {
    const double m12 = m1 + m2;
    if (a <= 0.0) return 0.0;
    if (a < m1) return a * a / (2.0 * m1 * m2 * m3);
    const double base = 6.0 * a - 3.0 * m1;
    if (a < m2) return base / (6.0 * m2 * m3);
    if (a < m12 && a < m3) {
        const double b = a - m2;
        return (base - 3.0 * b * b / m1) / (6.0 * m2 * m3);
    }
    if (m3 < m12) {
        const double b = a - m2, c = a - m3;
        return (base - 3.0 * (b * b + c * c) / m1) / (6.0 * m2 * m3);
    }
    return 1.0 / m3;
}

// Inverse of cube_fraction_sorted: the level a in [0, 1/2] at which the unit
// cube holds the volume V in [0, 1/2] below m . xi = a (m1 <= m2 <= m3, sum
// 1).  The first two pieces and the last linear one in closed form
// (Scardovelli & Zaleski, J. Comput. Phys. 164 (2000): a = (6 m1 m2 m3 V)^(1/3)
// and a = m1/2 + (2 m2 m3 V - m1^2/12)^(1/2)); the cubic pieces by Newton's
// method inside the piece's bracket, bisecting when a step leaves it.
__host__ __device__ inline double cube_level_sorted(double m1, double m2, double m3, double V) {
    if (V <= 0.0) return 0.0;
    const double m12 = m1 + m2;
    const double V1 = cube_fraction_sorted(m1, m2, m3, m1);
    if (V < V1) return cbrt(6.0 * m1 * m2 * m3 * V);
    const double V2 = cube_fraction_sorted(m1, m2, m3, m2);
    if (V < V2) return 0.5 * m1 + sqrt(fmax(2.0 * m2 * m3 * V - m1 * m1 / 12.0, 0.0));
    double lo, hi, Vlo, Vhi;
    if (m12 <= m3) {
        // cubic on [m2, m12), linear from m12 on
        const double V3 = cube_fraction_sorted(m1, m2, m3, m12);
        if (V >= V3) return m3 * V + 0.5 * m12;
        lo = m2; hi = m12; Vlo = V2; Vhi = V3;
    } else {
        // cubic on [m2, m3), another from m3 to the centre (m12 > 1/2 here)
        const double V3 = cube_fraction_sorted(m1, m2, m3, m3);
        if (V < V3) {
            lo = m2; hi = m3; Vlo = V2; Vhi = V3;
        } else {
            lo = m3; hi = 0.5; Vlo = V3; Vhi = cube_fraction_sorted(m1, m2, m3, 0.5);
        }
    }
    double a = (Vhi > Vlo) ? lo + (hi - lo) * ((V - Vlo) / (Vhi - Vlo)) : lo;
    for (int it = 0; it < 12; ++it) {
        const double f = cube_fraction_sorted(m1, m2, m3, a) - V;
        if (f == 0.0) break;
        if (f > 0.0) hi = a;
        else lo = a;
        const double fp = cube_area_sorted(m1, m2, m3, a);
        double an = fp > 0.0 ? a - f / fp : 0.5 * (lo + hi);
        if (fabs(an - a) <= 2e-16) {
            a = an;
            break;
        }
        if (!(an > lo && an < hi)) an = 0.5 * (lo + hi);
        a = an;
    }
    return a;
}

__host__ __device__ inline double cube_fraction(double m0, double m1, double m2, double a) {
    const double s = m0 + m1 + m2;
    if (!(s > 0.0)) return a > 0.0 ? 1.0 : 0.0;
    m0 /= s;
    m1 /= s;
    m2 /= s;
    a /= s;
    if (a <= 0.0) return 0.0;
    if (a >= 1.0) return 1.0;
    // sort ascending
    double t;
    if (m0 > m1) { t = m0; m0 = m1; m1 = t; }
    if (m1 > m2) { t = m1; m1 = m2; m2 = t; }
    if (m0 > m1) { t = m0; m0 = m1; m1 = t; }
    if (a <= 0.5) return cube_fraction_sorted(m0, m1, m2, a);
    return 1.0 - cube_fraction_sorted(m0, m1, m2, 1.0 - a);
}

// Liquid fraction of the cube [lo, lo + h]^3 on the side n . x + D < 0 (the
// side the interface normal points away from).
__host__ __device__ inline double plane_cube_fraction(double nx, double ny, double nz, double D, double lox,
                                                      double loy, double loz, double h) {
    // xi = (x - lo)/h; reflect axes with n < 0 so that every coefficient is >= 0
    double rhs = -D - (nx * lox + ny * loy + nz * loz);
    double m0 = nx * h, m1 = ny * h, m2 = nz * h;
    if (m0 < 0.0) { rhs -= m0; m0 = -m0; }
    if (m1 < 0.0) { rhs -= m1; m1 = -m1; }
    if (m2 < 0.0) { rhs -= m2; m2 = -m2; }
    return cube_fraction(m0, m1, m2, rhs);
}

// plane_cube_fraction for one normal and cube edge over many plane constants
// and cube corners: the reflected, normalised and sorted coefficients are
// computed once, and every evaluation runs the same operations on the same
// values as plane_cube_fraction, so the results are identical.
struct CubePlane {
    double nx, ny, nz, h;
    double neg[3];      // reflected coefficients -m_i (0 where m_i >= 0), subtracted in axis order
    double s;           // m0 + m1 + m2 after reflection
    double sm[3];       // normalised coefficients, ascending
    __host__ __device__ CubePlane(double nx_, double ny_, double nz_, double h_) : nx(nx_), ny(ny_), nz(nz_), h(h_) {
        double m0 = nx * h, m1 = ny * h, m2 = nz * h;
        neg[0] = m0 < 0.0 ? m0 : 0.0;
        neg[1] = m1 < 0.0 ? m1 : 0.0;
        neg[2] = m2 < 0.0 ? m2 : 0.0;
        if (m0 < 0.0) m0 = -m0;
        if (m1 < 0.0) m1 = -m1;
        if (m2 < 0.0) m2 = -m2;
        s = m0 + m1 + m2;
        if (s > 0.0) {
            m0 /= s;
            m1 /= s;
            m2 /= s;
        }
        double t;
        if (m0 > m1) { t = m0; m0 = m1; m1 = t; }
        if (m1 > m2) { t = m1; m1 = m2; m2 = t; }
        if (m0 > m1) { t = m0; m0 = m1; m1 = t; }
        sm[0] = m0;
        sm[1] = m1;
        sm[2] = m2;
    }
    // the plane constant D at which [lo, lo + h]^3 holds the liquid fraction
    // alpha on the side n . x + D < 0 (the inverse of fraction)
    __host__ __device__ double offset(double alpha, double lox, double loy, double loz) const {
        double a;
        if (!(s > 0.0)) a = 0.0;
        else if (alpha <= 0.0) a = 0.0;
        else if (alpha >= 1.0) a = 1.0;
        else if (alpha <= 0.5) a = cube_level_sorted(sm[0], sm[1], sm[2], alpha);
        else a = 1.0 - cube_level_sorted(sm[0], sm[1], sm[2], 1.0 - alpha);
        // fraction: rhs = -D - n . lo - (the reflected coefficients), a = rhs / s
        double rhs = a * s;
        if (neg[0] < 0.0) rhs += neg[0];
        if (neg[1] < 0.0) rhs += neg[1];
        if (neg[2] < 0.0) rhs += neg[2];
        return -rhs - (nx * lox + ny * loy + nz * loz);
    }
    // liquid fraction of [lo, lo + h]^3 on the side n . x + D < 0
    __host__ __device__ double fraction(double D, double lox, double loy, double loz) const {
        double rhs = -D - (nx * lox + ny * loy + nz * loz);
        if (neg[0] < 0.0) rhs -= neg[0];
        if (neg[1] < 0.0) rhs -= neg[1];
        if (neg[2] < 0.0) rhs -= neg[2];
        if (!(s > 0.0)) return rhs > 0.0 ? 1.0 : 0.0;
        const double a = rhs / s;
        if (a <= 0.0) return 0.0;
        if (a >= 1.0) return 1.0;
        if (a <= 0.5) return cube_fraction_sorted(sm[0], sm[1], sm[2], a);
        return 1.0 - cube_fraction_sorted(sm[0], sm[1], sm[2], 1.0 - a);
    }
};

// ---------------------------------------------------------------------------
// Stencil symmetry.  Lattice index of offset (dx, dy, dz) in {-1,0,1}^3:
// L = 9 (dx+1) + 3 (dy+1) + (dz+1).  The canonical frame reflects every axis
// along which the Youngs normal estimate (-grad alpha with weights 1-2-1) is
// negative and orders the axes by decreasing estimate magnitude (stable), so
// the network only sees one of the 48 symmetric copies of a configuration.
// ---------------------------------------------------------------------------
struct Canon {
    int perm[3];   // canonical axis c <- mesh axis perm[c]
    double sgn[3]; // sign applied to mesh axis perm[c]
};

__host__ __device__ inline int lat_index(int dx, int dy, int dz) { return 9 * (dx + 1) + 3 * (dy + 1) + (dz + 1); }

template <typename T>
__host__ __device__ inline Canon canon_of(const T* a) {
    double g[3] = {0.0, 0.0, 0.0};
    for (int u = -1; u <= 1; ++u)
        for (int v = -1; v <= 1; ++v) {
            const double w = (u == 0 ? 2.0 : 1.0) * (v == 0 ? 2.0 : 1.0);
            g[0] += w * ((double)a[lat_index(1, u, v)] - (double)a[lat_index(-1, u, v)]);
            g[1] += w * ((double)a[lat_index(u, 1, v)] - (double)a[lat_index(u, -1, v)]);
            g[2] += w * ((double)a[lat_index(u, v, 1)] - (double)a[lat_index(u, v, -1)]);
        }
    Canon c;
    double mag[3];
    for (int k = 0; k < 3; ++k) {
        const double nk = -g[k];
        c.sgn[k] = nk < 0.0 ? -1.0 : 1.0;
        mag[k] = nk < 0.0 ? -nk : nk;
    }
    // stable descending order of the magnitudes
    int p0 = 0, p1 = 1, p2 = 2, t;
    if (mag[p1] > mag[p0]) { t = p0; p0 = p1; p1 = t; }
    if (mag[p2] > mag[p1]) { t = p1; p1 = p2; p2 = t; }
    if (mag[p1] > mag[p0]) { t = p0; p0 = p1; p1 = t; }
    const double s0 = c.sgn[0], s1 = c.sgn[1], s2 = c.sgn[2];
    const double sg[3] = {s0, s1, s2};
    c.perm[0] = p0;
    c.perm[1] = p1;
    c.perm[2] = p2;
    c.sgn[0] = sg[p0];
    c.sgn[1] = sg[p1];
    c.sgn[2] = sg[p2];
    return c;
}

// mesh-frame lattice index of canonical lattice index Lc
__host__ __device__ inline int canon_source(const Canon& c, int Lc) {
    const int e[3] = {Lc / 9 - 1, (Lc / 3) % 3 - 1, Lc % 3 - 1};
    int s[3];
    for (int k = 0; k < 3; ++k) s[c.perm[k]] = c.sgn[k] < 0.0 ? -e[k] : e[k];
    return lat_index(s[0], s[1], s[2]);
}

// vector mesh frame -> canonical frame and back
__host__ __device__ inline void to_canon(const Canon& c, const double* v, double* vc) {
    for (int k = 0; k < 3; ++k) vc[k] = c.sgn[k] * v[c.perm[k]];
}
__host__ __device__ inline void from_canon(const Canon& c, const double* vc, double* v) {
    for (int k = 0; k < 3; ++k) v[c.perm[k]] = c.sgn[k] * vc[k];
}

// network input of stencil value alpha
__host__ __device__ inline uint16_t input_of(double alpha) { return f2bf((float)(2.0 * alpha - 1.0)); }

}  // namespace nm
}  // namespace svp
