// Learned interface normal: training-set generation from exact cuts, the
// 27-128-128-3 GELU network as BF16 MFMA kernels with FP32 accumulation
// (forward, backward, Adam), inference on the interface cells of a uniform
// hex mesh, and the volume-residual test that falls back to the
// least-squares normal.  No counterpart in the reference; the plane
// placement that consumes these normals is findSignedDistance
// (cutCell.C:611-799), unchanged.
#include <hip/hip_runtime.h>
#include <cstdlib>
#include <stdint.h>

#include <algorithm>

#include "learned_device.hpp"
#include "learned_kernels.hpp"
#include "simplic_types.hpp"

namespace svp {
namespace nm {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f4;

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// Counter-based random numbers (splitmix64)
// ---------------------------------------------------------------------------
__device__ inline uint64_t smix(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ inline double urand(uint64_t& s) { return (double)(smix(s) >> 11) * (1.0 / 9007199254740992.0); }

__device__ inline void unit_vector(uint64_t& s, double* v) {
    const double z = 2.0 * urand(s) - 1.0;
    const double ph = 6.283185307179586 * urand(s);
    const double r = sqrt(fmax(0.0, 1.0 - z * z));
    v[0] = r * cos(ph);
    v[1] = r * sin(ph);
    v[2] = z;
}

// Shape of training sample i.  The centre cell of the stencil is the unit
// cube around the origin; the surface passes through a random point p of it.
//   kind 0: plane through p with random normal
//   kind 1: sphere through p, radius R log-uniform in [1.5, 50] cells
//   kind 2: ellipsoid through p, random orientation, semi-axes 1.5 R [1, 1.5]
//           (principal curvature radii >= R)
// target normal: outward normal at the surface point nearest the cell centre
__device__ inline void sample_shape(uint64_t seed, long i, NmShape& S) {
    uint64_t s = seed * 0xD1342543DE82EF95ull + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
    smix(s);
    S.kind = (int)(i % 3);
    double p[3], n[3];
    for (int k = 0; k < 3; ++k) p[k] = urand(s) - 0.5;
    unit_vector(s, n);
    const double R = exp(log(1.5) + urand(s) * (log(50.0) - log(1.5)));
    for (int k = 0; k < 9; ++k) S.Q[k] = (k % 4 == 0) ? 1.0 : 0.0;
    if (S.kind == 0) {
        for (int k = 0; k < 3; ++k) {
            S.c[k] = p[k];
            S.a[k] = 0.0;
            S.nrm[k] = n[k];
        }
        return;
    }
    if (S.kind == 1) {
        double cc = 0.0;
        for (int k = 0; k < 3; ++k) {
            S.c[k] = p[k] - R * n[k];
            S.a[k] = R;
            cc += S.c[k] * S.c[k];
        }
        const double m = sqrt(cc);
        for (int k = 0; k < 3; ++k) S.nrm[k] = -S.c[k] / m;
        return;
    }
    // ellipsoid: rotation from a uniform random unit quaternion (Shoemake)
    for (int k = 0; k < 3; ++k) S.a[k] = 1.5 * R * (1.0 + 0.5 * urand(s));
    const double u1 = urand(s), u2 = urand(s), u3 = urand(s);
    const double qx = sqrt(1.0 - u1) * sin(6.283185307179586 * u2), qy = sqrt(1.0 - u1) * cos(6.283185307179586 * u2);
    const double qz = sqrt(u1) * sin(6.283185307179586 * u3), qw = sqrt(u1) * cos(6.283185307179586 * u3);
    // columns of Q are the ellipsoid axes in the world frame: x = c + Q e
    S.Q[0] = 1 - 2 * (qy * qy + qz * qz); S.Q[1] = 2 * (qx * qy - qz * qw);     S.Q[2] = 2 * (qx * qz + qy * qw);
    S.Q[3] = 2 * (qx * qy + qz * qw);     S.Q[4] = 1 - 2 * (qx * qx + qz * qz); S.Q[5] = 2 * (qy * qz - qx * qw);
    S.Q[6] = 2 * (qx * qz - qy * qw);     S.Q[7] = 2 * (qy * qz + qx * qw);     S.Q[8] = 1 - 2 * (qx * qx + qy * qy);
    // random surface point e (ellipsoid frame) placed at p
    double d[3], e[3], sq = 0.0;
    unit_vector(s, d);
    for (int k = 0; k < 3; ++k) sq += (d[k] / S.a[k]) * (d[k] / S.a[k]);
    for (int k = 0; k < 3; ++k) e[k] = d[k] / sqrt(sq);
    for (int r = 0; r < 3; ++r) S.c[r] = p[r] - (S.Q[3 * r] * e[0] + S.Q[3 * r + 1] * e[1] + S.Q[3 * r + 2] * e[2]);
    // nearest point to the origin: y = Q^T (0 - c); x_k = a_k^2 y_k / (a_k^2 + t)
    // with sum (a_k y_k / (a_k^2 + t))^2 = 1, t in (-a_min^2, a_max |y|]
    double y[3], ymag = 0.0, amin = S.a[0], amax = S.a[0];
    for (int k = 0; k < 3; ++k) {
        y[k] = -(S.Q[k] * S.c[0] + S.Q[3 + k] * S.c[1] + S.Q[6 + k] * S.c[2]);
        ymag += y[k] * y[k];
        amin = fmin(amin, S.a[k]);
        amax = fmax(amax, S.a[k]);
    }
    double lo = -amin * amin, hi = amax * sqrt(ymag);
    for (int it = 0; it < 200; ++it) {
        const double t = 0.5 * (lo + hi);
        if (t <= lo || t >= hi) break;
        double g = -1.0;
        for (int k = 0; k < 3; ++k) {
            const double q = S.a[k] * y[k] / (S.a[k] * S.a[k] + t);
            g += q * q;
        }
        if (g > 0.0) lo = t;
        else hi = t;
    }
    const double t = 0.5 * (lo + hi);
    double ne[3], nw[3], nm2 = 0.0;
    for (int k = 0; k < 3; ++k) ne[k] = y[k] / (S.a[k] * S.a[k] + t);
    for (int r = 0; r < 3; ++r) {
        nw[r] = S.Q[3 * r] * ne[0] + S.Q[3 * r + 1] * ne[1] + S.Q[3 * r + 2] * ne[2];
        nm2 += nw[r] * nw[r];
    }
    for (int r = 0; r < 3; ++r) S.nrm[r] = nw[r] / sqrt(nm2);
}

// implicit function of a curved shape (negative inside) and its gradient
__device__ inline double shape_g(const NmShape& S, const double* x, double* grad) {
    double r[3] = {x[0] - S.c[0], x[1] - S.c[1], x[2] - S.c[2]};
    if (S.kind == 1) {
        const double m = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        if (grad) for (int k = 0; k < 3; ++k) grad[k] = m > 0.0 ? r[k] / m : 0.0;
        return m - S.a[0];
    }
    double e[3], m2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        e[k] = (S.Q[k] * r[0] + S.Q[3 + k] * r[1] + S.Q[6 + k] * r[2]) / S.a[k];
        m2 += e[k] * e[k];
    }
    const double m = sqrt(m2);
    if (grad) {
        double ge[3];
        for (int k = 0; k < 3; ++k) ge[k] = m > 0.0 ? e[k] / (m * S.a[k]) : 0.0;
        for (int rr = 0; rr < 3; ++rr) grad[rr] = S.Q[3 * rr] * ge[0] + S.Q[3 * rr + 1] * ge[1] + S.Q[3 * rr + 2] * ge[2];
    }
    return m - 1.0;
}

// liquid fraction of the unit cube centred at o.  Planes are exact; curved
// shapes subdivide the cube into octants down to maxDepth levels, classify a
// box as full or empty when the Lipschitz bound of the implicit function
// allows, and cut the remaining leaves with the shape's tangent-plane
// linearisation at the leaf centre.
__device__ inline double shape_cell_fraction(const NmShape& S, const double* o, int maxDepth) {
    if (S.kind == 0) {
        const double D = -(S.nrm[0] * S.c[0] + S.nrm[1] * S.c[1] + S.nrm[2] * S.c[2]);
        return plane_cube_fraction(S.nrm[0], S.nrm[1], S.nrm[2], D, o[0] - 0.5, o[1] - 0.5, o[2] - 0.5, 1.0);
    }
    // Lipschitz constant of shape_g: 1 for the sphere distance, 1/a_min for the ellipsoid
    const double lip = (S.kind == 1) ? 1.0 : 1.0 / fmin(S.a[0], fmin(S.a[1], S.a[2]));
    struct Box { double x, y, z; int depth; };
    Box stack[8 * 8];
    int sp = 0;
    stack[sp++] = {o[0], o[1], o[2], 0};
    double vol = 0.0;
    while (sp > 0) {
        const Box b = stack[--sp];
        const double half = 0.5 * ldexp(1.0, -b.depth);
        const double xc[3] = {b.x, b.y, b.z};
        double grad[3];
        const double g = shape_g(S, xc, grad);
        const double reach = 1.7320508075688772 * half * lip;
        const double bv = 8.0 * half * half * half;
        if (g <= -reach) { vol += bv; continue; }
        if (g >= reach) continue;
        if (b.depth >= maxDepth) {
            const double D = g - (grad[0] * xc[0] + grad[1] * xc[1] + grad[2] * xc[2]);
            vol += bv * plane_cube_fraction(grad[0], grad[1], grad[2], D, xc[0] - half, xc[1] - half, xc[2] - half,
                                            2.0 * half);
            continue;
        }
        const double q = 0.5 * half;
        for (int k = 0; k < 8; ++k)
            stack[sp++] = {b.x + ((k & 1) ? q : -q), b.y + ((k & 2) ? q : -q), b.z + ((k & 4) ? q : -q), b.depth + 1};
    }
    return vol;
}

// thread per (sample, stencil cell)
__global__ void __launch_bounds__(256) k_gen_alpha(uint64_t seed, long first, long n, int depth, double* alpha27,
                                                   double* shape19) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= n * NIN) return;
    const long i = t / NIN;
    const int L = (int)(t % NIN);
    NmShape S;
    sample_shape(seed, first + i, S);
    const double o[3] = {(double)(L / 9 - 1), (double)((L / 3) % 3 - 1), (double)(L % 3 - 1)};
    alpha27[i * NIN + L] = shape_cell_fraction(S, o, depth);
    if (L == 13 && shape19) {
        double* r = shape19 + i * 19;
        r[0] = (double)S.kind;
        for (int k = 0; k < 3; ++k) {
            r[1 + k] = S.c[k];
            r[4 + k] = S.a[k];
            r[16 + k] = S.nrm[k];
        }
        for (int k = 0; k < 9; ++k) r[7 + k] = S.Q[k];
    }
}

// canonical network input (BF16 of 2 alpha - 1, zero padded to 32) and target
__global__ void __launch_bounds__(256) k_gen_pack(long n, const double* alpha27, const double* shape19, uint16_t* X,
                                                  float* Y) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double* a = alpha27 + i * NIN;
    const Canon cn = canon_of(a);
    for (int Lc = 0; Lc < KIN; ++Lc) X[i * KIN + Lc] = Lc < NIN ? input_of(a[canon_source(cn, Lc)]) : (uint16_t)0;
    double vc[3];
    to_canon(cn, shape19 + i * 19 + 16, vc);
    for (int k = 0; k < 3; ++k) Y[i * 4 + k] = (float)vc[k];
    Y[i * 4 + 3] = 0.0f;
}

// ---------------------------------------------------------------------------
// Network on MFMA.  One wave evaluates 16 samples in the transposed form
// Z^T = W H^T (rows = units, columns = samples) with
// v_mfma_f32_16x16x32_bf16: operand lanes l hold row/column l%16 and eight K
// values of group q = l/16; the result lane l holds column l%16, rows
// 4q..4q+3.  The result of one layer is therefore already the K operand of
// the next when the K order of a 32-wide block is taken as
// {4q..4q+3, 16+4q..16+4q+3} for group q, which the weight loads follow.
// ---------------------------------------------------------------------------
__device__ inline bf16x8 ld_bf16x8(const uint16_t* p) { return *(const bf16x8*)p; }

__device__ inline bf16x8 ld_perm(const uint16_t* row, int ks, int q) {
    // K block ks of a weight row in the result-chaining order
    typedef __attribute__((ext_vector_type(4))) __bf16 bf16x4;
    const bf16x4 lo = *(const bf16x4*)(row + 32 * ks + 4 * q);
    const bf16x4 hi = *(const bf16x4*)(row + 32 * ks + 16 + 4 * q);
    bf16x8 r;
    r[0] = lo[0]; r[1] = lo[1]; r[2] = lo[2]; r[3] = lo[3];
    r[4] = hi[0]; r[5] = hi[1]; r[6] = hi[2]; r[7] = hi[3];
    return r;
}

__device__ inline bf16x8 pack_pair(const uint16_t (&a)[4], const uint16_t (&b)[4]) {
    uint16_t t[8] = {a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
    return __builtin_bit_cast(bf16x8, t);
}

struct Fwd {
    float z1[8][4], z2[8][4];      // pre-activations (rows 16t+4q+j)
    uint16_t h1[8][4], h2[8][4];   // BF16 activations
    float o[4];                    // output rows 4q+j (rows 0..2 meaningful in group 0)
};

// xop: this lane's eight inputs (sample l%16, inputs 8q..8q+7)
__device__ inline void forward16(const bf16x8 xop, const uint16_t* wb, const float* pq, int r, int q, Fwd& F) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld_bf16x8(wb + B_W1 + (16 * t + r) * KIN + 8 * q), xop, acc, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            F.z1[t][j] = acc[j] + pq[Q_B1 + 16 * t + 4 * q + j];
            F.h1[t][j] = f2bf(gelu(F.z1[t][j]));
        }
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        const uint16_t* row = wb + B_W2 + (16 * t + r) * NH;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld_perm(row, ks, q), pack_pair(F.h1[2 * ks], F.h1[2 * ks + 1]),
                                                          acc, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            F.z2[t][j] = acc[j] + pq[Q_B2 + 16 * t + 4 * q + j];
            F.h2[t][j] = f2bf(gelu(F.z2[t][j]));
        }
    }
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    const uint16_t* row = wb + B_W3 + r * NH;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld_perm(row, ks, q), pack_pair(F.h2[2 * ks], F.h2[2 * ks + 1]),
                                                      acc, 0, 0, 0);
#pragma unroll
    for (int j = 0; j < 4; ++j) F.o[j] = acc[j] + pq[Q_B3 + 4 * q + j];
}

// forward16 with the weights of the first two layers in wb (layout of the
// BF16 copy up to W2T), W3 in w3 and the three bias vectors apart
__device__ inline void forward16_split(const bf16x8 xop, const uint16_t* wb, const uint16_t* w3, const float* b1,
                                       const float* b2, const float* b3, int r, int q, Fwd& F) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld_bf16x8(wb + B_W1 + (16 * t + r) * KIN + 8 * q), xop, acc, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) F.h1[t][j] = f2bf(gelu(acc[j] + b1[16 * t + 4 * q + j]));
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        const uint16_t* row = wb + B_W2 + (16 * t + r) * NH;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld_perm(row, ks, q), pack_pair(F.h1[2 * ks], F.h1[2 * ks + 1]),
                                                          acc, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) F.h2[t][j] = f2bf(gelu(acc[j] + b2[16 * t + 4 * q + j]));
    }
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    const uint16_t* row = w3 + r * NH;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld_perm(row, ks, q), pack_pair(F.h2[2 * ks], F.h2[2 * ks + 1]),
                                                      acc, 0, 0, 0);
#pragma unroll
    for (int j = 0; j < 4; ++j) F.o[j] = acc[j] + b3[4 * q + j];
}

// sample at position k of the epoch order (affine permutation mod n)
__device__ inline long sample_at(long k, long n, long pa, long pb) {
    return (long)(((unsigned __int128)pa * (unsigned __int128)k + (unsigned __int128)pb) % (unsigned __int128)n);
}

// forward of n samples (inference check): 64 samples per workgroup
__global__ void __launch_bounds__(256) k_forward(const uint16_t* X, long n, const uint16_t* wb, const float* pq,
                                                 float* out) {
    const int w = threadIdx.x / 64, l = threadIdx.x % 64, r = l % 16, q = l / 16;
    const long s = (long)blockIdx.x * 64 + 16 * w + r;
    bf16x8 xop;
    if (s < n) xop = ld_bf16x8(X + s * KIN + 8 * q);
    else {
        uint16_t z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        xop = __builtin_bit_cast(bf16x8, z);
    }
    Fwd F;
    forward16(xop, wb, pq, r, q, F);
    if (q == 0 && s < n)
        for (int j = 0; j < 3; ++j) out[s * 3 + j] = F.o[j];
}

// One training micro-batch of 64 samples per workgroup (16 per wave):
// forward, mean-squared-error loss against the canonical target, backward,
// and this workgroup's share of every gradient in FP32 (partial[wg][...],
// padded layout).  Weight gradients with K = samples run on MFMA from LDS
// tiles [unit][sample]; bias gradients are FP32 sums.
constexpr int TB = 64;

__global__ void __launch_bounds__(256) k_train(const uint16_t* X, const float* Y, long n, long pa, long pb, long base,
                                               const uint16_t* wb, const float* pq, float invB, float* partial,
                                               float* lossPart) {
    __shared__ __attribute__((aligned(16))) uint16_t sX[KIN][TB];
    __shared__ __attribute__((aligned(16))) uint16_t sA[NH][TB];   // H1, then dZ1
    __shared__ __attribute__((aligned(16))) uint16_t sB[NH][TB];   // H2, then dZ2
    __shared__ float sD3[4][TB];
    __shared__ float red2[4][NH];  // per-wave bias sums of dZ2
    __shared__ float red1[4][NH];  // and of dZ1
    __shared__ float lred[4];
    const int tid = threadIdx.x, w = tid / 64, l = tid % 64, r = l % 16, q = l / 16;
    const int sl = 16 * w + r;  // sample slot in the workgroup
    const long gi = sample_at(base + (long)blockIdx.x * TB + sl, n, pa, pb);
    float* P = partial + (long)blockIdx.x * N_PADDED;

    const bf16x8 xop = ld_bf16x8(X + gi * KIN + 8 * q);
    {
        const uint16_t* xs = X + gi * KIN + 8 * q;
#pragma unroll
        for (int j = 0; j < 8; ++j) sX[8 * q + j][sl] = xs[j];
    }
    Fwd F;
    forward16(xop, wb, pq, r, q, F);
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            sA[16 * t + 4 * q + j][sl] = F.h1[t][j];
            sB[16 * t + 4 * q + j][sl] = F.h2[t][j];
        }
    // loss and output gradient (group 0 holds the outputs of sample r)
    float d3[3] = {0.f, 0.f, 0.f};
    float lossS = 0.f;
    if (q == 0) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float diff = F.o[j] - Y[gi * 4 + j];
            lossS += diff * diff;
            d3[j] = 2.0f * diff * invB;
            sD3[j][sl] = d3[j];
        }
        sD3[3][sl] = 0.f;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) d3[j] = __shfl(d3[j], r);
    for (int o = 1; o < 64; o <<= 1) lossS += __shfl_xor(lossS, o);
    if (l == 0) lred[w] = lossS;
    __syncthreads();

    // layer 3 gradients (VALU): dW3[o][h] = sum_s dZ3[o][s] H2[h][s]
    for (int e = tid; e < NOUTP * NH; e += 256) {
        const int o = e / NH, h = e % NH;
        float acc = 0.f;
        if (o < NOUT)
            for (int s = 0; s < TB; ++s) acc += sD3[o][s] * bf2f(sB[h][s]);
        P[Q_W3 + e] = acc;
    }
    if (tid < NOUTP) {
        float acc = 0.f;
        if (tid < NOUT)
            for (int s = 0; s < TB; ++s) acc += sD3[tid][s];
        P[Q_B3 + tid] = acc;
    }
    if (tid == 0) lossPart[blockIdx.x] = lred[0] + lred[1] + lred[2] + lred[3];
    __syncthreads();

    // dZ2 = (W3^T dZ3) * gelu'(Z2)
    uint16_t dz2b[8][4];
    float bsum[8][4];
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int h = 16 * t + 4 * q + j;
            const float dh = bf2f(wb[B_W3 + h]) * d3[0] + bf2f(wb[B_W3 + NH + h]) * d3[1] +
                             bf2f(wb[B_W3 + 2 * NH + h]) * d3[2];
            const float dz = dh * gelu_grad(F.z2[t][j]);
            dz2b[t][j] = f2bf(dz);
            bsum[t][j] = dz;
            sB[h][sl] = dz2b[t][j];
        }
    // db2: sum over the 16 samples of the wave, then over the waves
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float v = bsum[t][j];
            v += __shfl_xor(v, 1);
            v += __shfl_xor(v, 2);
            v += __shfl_xor(v, 4);
            v += __shfl_xor(v, 8);
            if (r == 0) red2[w][16 * t + 4 * q + j] = v;
        }
    // dH1^T = W2^T dZ2^T, dZ1 = dH1 * gelu'(Z1)
    uint16_t dz1b[8][4];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        const uint16_t* row = wb + B_W2T + (16 * t + r) * NH;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld_perm(row, ks, q), pack_pair(dz2b[2 * ks], dz2b[2 * ks + 1]),
                                                          acc, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float dz = acc[j] * gelu_grad(F.z1[t][j]);
            dz1b[t][j] = f2bf(dz);
            bsum[t][j] = dz;
        }
    }
    __syncthreads();
    if (tid < NH) P[Q_B2 + tid] = red2[0][tid] + red2[1][tid] + red2[2][tid] + red2[3][tid];

    // dW2[m][k] = sum_s dZ2[m][s] H1[k][s]: wave w owns m tiles 2w, 2w+1
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int tm = 2 * w + a;
#pragma unroll
        for (int tn = 0; tn < 8; ++tn) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld_bf16x8(&sB[16 * tm + r][32 * ks + 8 * q]),
                                                              ld_bf16x8(&sA[16 * tn + r][32 * ks + 8 * q]), acc, 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; ++j) P[Q_W2 + (16 * tm + 4 * q + j) * NH + 16 * tn + r] = acc[j];
        }
    }
    // db1: same reduction
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float v = bsum[t][j];
            v += __shfl_xor(v, 1);
            v += __shfl_xor(v, 2);
            v += __shfl_xor(v, 4);
            v += __shfl_xor(v, 8);
            if (r == 0) red1[w][16 * t + 4 * q + j] = v;
        }
    __syncthreads();  // every wave is done with H1 in sA
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) sA[16 * t + 4 * q + j][sl] = dz1b[t][j];
    if (tid < NH) P[Q_B1 + tid] = red1[0][tid] + red1[1][tid] + red1[2][tid] + red1[3][tid];
    __syncthreads();
    // dW1[m][i] = sum_s dZ1[m][s] X[s][i]
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int tm = 2 * w + a;
#pragma unroll
        for (int tn = 0; tn < 2; ++tn) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld_bf16x8(&sA[16 * tm + r][32 * ks + 8 * q]),
                                                              ld_bf16x8(&sX[16 * tn + r][32 * ks + 8 * q]), acc, 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; ++j) P[Q_W1 + (16 * tm + 4 * q + j) * KIN + 16 * tn + r] = acc[j];
        }
    }
}

// gradient = sum of the workgroup partials in workgroup order
__global__ void __launch_bounds__(256) k_reduce(const float* partial, int nWG, float* grad) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= N_PADDED) return;
    float g = 0.f;
    for (int w = 0; w < nWG; ++w) g += partial[(long)w * N_PADDED + p];
    grad[p] = g;
}

// BF16 copies of one padded parameter
__device__ inline void write_bf16(int p, float v, uint16_t* wb) {
    if (p < Q_B1) wb[B_W1 + p] = f2bf(v);
    else if (p >= Q_W2 && p < Q_B2) {
        const int e = p - Q_W2, m = e / NH, k = e % NH;
        wb[B_W2 + e] = f2bf(v);
        wb[B_W2T + k * NH + m] = f2bf(v);
    } else if (p >= Q_W3 && p < Q_B3) wb[B_W3 + (p - Q_W3)] = f2bf(v);
}

// Adam on the FP32 master parameters, gradient = sum of the partials
__global__ void __launch_bounds__(256) k_adam(const float* partial, int nWG, float* pq, float* m1, float* m2,
                                              uint16_t* wb, float lr, float b1, float b2, float eps, float c1,
                                              float c2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= N_PADDED) return;
    float g = 0.f;
    for (int w = 0; w < nWG; ++w) g += partial[(long)w * N_PADDED + p];
    const float a = b1 * m1[p] + (1.f - b1) * g;
    const float v = b2 * m2[p] + (1.f - b2) * g * g;
    m1[p] = a;
    m2[p] = v;
    const float th = pq[p] - lr * (a / c1) / (sqrtf(v / c2) + eps);
    pq[p] = th;
    write_bf16(p, th, wb);
}

__global__ void __launch_bounds__(256) k_sync_bf16(const float* pq, uint16_t* wb) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < N_PADDED) write_bf16(p, pq[p], wb);
}

// ---------------------------------------------------------------------------
// Inference on interface cells.  One wave per 16 cells: the four lane
// groups gather the 27 stencil values, lane group 0 canonicalises, the wave
// runs the network, and group 0 maps the output back to the mesh frame.
// Cells without a full uniform 3x3x3 lattice stencil (lat[27c] < 0) and
// degenerate outputs are listed for the least-squares normal.
// ---------------------------------------------------------------------------
constexpr int INFER_WAVES = 4;

// Inference, one wave per 16 interface cells (lane r + 16 q: cell r, part q).
// The four lanes of a cell gather its 27 stencil values, derive the same
// canonical frame and each build eight network inputs; the wave runs the
// network; group 0 holds the outputs, which every lane of the cell maps back
// to the mesh frame.  With tol < 1 the residual test follows: the plane of
// the network normal that leaves the cell's own alpha in its cube (the
// inverse of the cube fraction, CubePlane::offset), and the residual is the
// mean |fraction - alpha| over the 26 neighbour cubes
// lo + e h of the lattice (lane q takes the neighbours L = q, q+4, ...).
// Cells without a full uniform 3x3x3 stencil (lat[27c] < 0), degenerate
// outputs and cells above tol are listed for the least-squares normal.
__global__ void __launch_bounds__(64 * INFER_WAVES, 2) k_infer(const int* lat, LatStride ls, int nC, const double* alpha,
                                                             const int* mixed, const int* nMixedP, const double* C,
                                                             const double* V, const uint16_t* wbG, const float* pqG,
                                                             double tol, double* ifN, unsigned char* learned,
                                                             double* resid, int* lsList, int* nLs, int* nFb) {
    // weights and biases of the network, staged once per workgroup
    __shared__ __attribute__((aligned(16))) uint16_t wb[B_W2T];           // W1b, W2b (W2T is not needed)
    __shared__ __attribute__((aligned(16))) uint16_t w3[NOUTP * NH];
    __shared__ float bias[NH + NH + NOUTP];
    __shared__ double raw[INFER_WAVES][16][NIN];
    {
        // all of a thread's weight loads in flight together, then the LDS stores
        static_assert(B_W2T * 2 / 16 % (64 * INFER_WAVES) == 0 && NOUTP * NH * 2 / 16 == 64 * INFER_WAVES,
                      "staging covers the weights exactly");
        constexpr int PER = B_W2T * 2 / 16 / (64 * INFER_WAVES);
        const uint4* src = (const uint4*)wbG;
        uint4* dst = (uint4*)wb;
        uint4 tmp[PER + 1];
#pragma unroll
        for (int j = 0; j < PER; ++j) tmp[j] = src[threadIdx.x + 64 * INFER_WAVES * j];
        tmp[PER] = ((const uint4*)(wbG + B_W3))[threadIdx.x];
#pragma unroll
        for (int j = 0; j < PER; ++j) dst[threadIdx.x + 64 * INFER_WAVES * j] = tmp[j];
        ((uint4*)w3)[threadIdx.x] = tmp[PER];
        for (int k = threadIdx.x; k < NH; k += blockDim.x) {
            bias[k] = pqG[Q_B1 + k];
            bias[NH + k] = pqG[Q_B2 + k];
        }
        for (int k = threadIdx.x; k < NOUTP; k += blockDim.x) bias[2 * NH + k] = pqG[Q_B3 + k];
    }
    __syncthreads();
    const int wv = threadIdx.x / 64, l = threadIdx.x % 64, r = l % 16, q = l / 16;
    const int nMixed = *nMixedP;
    double (*rw)[NIN] = raw[wv];
    for (int i0 = (blockIdx.x * INFER_WAVES + wv) * 16; i0 < nMixed; i0 += gridDim.x * INFER_WAVES * 16) {
        const int i = i0 + r;
        const bool live = i < nMixed;
        const int c = live ? mixed[i] : 0;
        // this lane's stencil entries L = q + 4j: all labels, then all
        // alphas, in flight together; the cell's centre and volume with them
        constexpr int NJ = (NIN + 3) / 4;
        int sl[NJ];
        double av[NJ];
        bool valid;
        if (ls.affine) {
            // every full stencil is c + off[L]: the alpha loads go out with
            // the load of the validity entry (indices clamped into the mesh;
            // values of cells without a full stencil are not used)
            const int e0 = lat[(long)c * NIN];
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const int L = q + 4 * j;
                sl[j] = min(max(c + ls.off[L < NIN ? L : 0], 0), nC - 1);
                av[j] = alpha[sl[j]];
            }
            valid = live && e0 >= 0;
#pragma unroll
            for (int j = 0; j < NJ; ++j) av[j] = valid ? av[j] : 0.0;
        } else {
#pragma unroll
            for (int j = 0; j < NJ; ++j) sl[j] = (q + 4 * j < NIN) ? lat[(long)c * NIN + q + 4 * j] : -1;
            valid = live && __shfl(sl[0], r) >= 0;  // lattice entry 0 is held by lane r (q = 0)
#pragma unroll
            for (int j = 0; j < NJ; ++j) av[j] = (valid && sl[j] >= 0) ? alpha[sl[j]] : 0.0;
        }
        const double cV = V[c];
        const double cCx = C[3L * c], cCy = C[3L * c + 1], cCz = C[3L * c + 2];
#pragma unroll
        for (int j = 0; j < NJ; ++j)
            if (q + 4 * j < NIN) rw[r][q + 4 * j] = av[j];
        __builtin_amdgcn_wave_barrier();
        // every lane of the cell derives the same frame and builds its own
        // eight inputs (K slots 8q..8q+7 of the first layer)
        const Canon cn = canon_of(&rw[r][0]);
        uint16_t xin[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int Lc = 8 * q + j;
            xin[j] = (valid && Lc < NIN) ? input_of(rw[r][canon_source(cn, Lc)]) : (uint16_t)0;
        }
        __builtin_amdgcn_wave_barrier();
        Fwd F;
        forward16_split(__builtin_bit_cast(bf16x8, xin), wb, w3, bias, bias + NH, bias + 2 * NH, r, q, F);
        // outputs of cell r sit in lane r (group 0): every lane of the cell takes them
        const double vc[3] = {(double)__shfl(F.o[0], r), (double)__shfl(F.o[1], r), (double)__shfl(F.o[2], r)};
        double v[3];
        from_canon(cn, vc, v);
        const double mg = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        bool ok = valid && mg > 1e-6;
        const double nx = ok ? v[0] / mg : 0.0, ny = ok ? v[1] / mg : 0.0, nz = ok ? v[2] / mg : 0.0;
        double res = -1.0;
        if (tol < 1.0) {
            // cell cube (lattice cell: an axis-aligned cube of edge h at C)
            const double h = live ? cbrt(cV) : 1.0;
            const double lox = live ? cCx - 0.5 * h : 0.0, loy = live ? cCy - 0.5 * h : 0.0,
                         loz = live ? cCz - 0.5 * h : 0.0;
            const double a = rw[r][13];
            const CubePlane cp(nx, ny, nz, h);
            const double d = cp.offset(a, lox, loy, loz);
            double part = 0.0;
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const int L = q + 4 * j;
                if (L >= NIN || L == 13) continue;
                const double ex = (double)(L / 9 - 1), ey = (double)((L / 3) % 3 - 1), ez = (double)(L % 3 - 1);
                const double f = cp.fraction(d, lox + ex * h, loy + ey * h, loz + ez * h);
                part += fabs(f - rw[r][L]);
            }
            part += __shfl_xor(part, 16);
            part += __shfl_xor(part, 32);
            res = part * (1.0 / 26.0);
        }
        if (q == 0 && live) {
            if (ok) {
                ifN[3L * i] = nx;
                ifN[3L * i + 1] = ny;
                ifN[3L * i + 2] = nz;
            }
            if (resid) resid[i] = ok ? res : -1.0;
            const bool fb = ok && tol < 1.0 && res > tol;
            learned[i] = (ok && !fb) ? 1 : 0;
            if (!ok || fb) lsList[atomicAdd(nLs, 1)] = i;
            if (fb) atomicAdd(nFb, 1);
        }
        __builtin_amdgcn_wave_barrier();  // raw is rewritten by the next cells
    }
}


// cube fractions for the closed-form check: in[k*8] = n, d, lo, h
__global__ void __launch_bounds__(256) k_cube_fractions(const double* in, long k, double* out) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= k) return;
    const double* a = in + 8 * i;
    out[i] = plane_cube_fraction(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]);
}

// ---------------------------------------------------------------------------
// Uniform-lattice stencil map.  A cell is a lattice cube when it has 6 faces
// and 8 points, each point at C +- h/2 per axis (h = V^(1/3), tolerance
// 1e-9 h).  A cube whose cell-point-cell stencil is 27 cubes of its volume at
// the offsets {-h, 0, h}^3 gets lat[27c + L] = stencil cell at lattice index
// L; every other cell gets -1.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_cube_flags(DevMesh m, unsigned char* cube) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= m.nC) return;
    const double h = cbrt(m.V[c]), tol = 1e-9 * h;
    bool ok = (m.cOff[c + 1] - m.cOff[c] == 6) && (m.cpOff[c + 1] - m.cpOff[c] == 8);
    for (int k = m.cpOff[c]; ok && k < m.cpOff[c + 1]; ++k) {
        const int p = m.cPts[k];
        for (int a = 0; a < 3; ++a) ok &= fabs(fabs(m.pts[3L * p + a] - m.C[3L * c + a]) - 0.5 * h) <= tol;
    }
    cube[c] = ok ? 1 : 0;
}

__global__ void __launch_bounds__(256) k_lattice(DevMesh m, const unsigned char* cube, int* lat) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= m.nC) return;
    int* L = lat + (long)c * NIN;
    for (int k = 0; k < NIN; ++k) L[k] = -1;
    bool ok = cube[c] && (m.stOff[c + 1] - m.stOff[c] == NIN);
    const double h = cbrt(m.V[c]);
    for (int k = m.stOff[c]; ok && k < m.stOff[c + 1]; ++k) {
        const int j = m.st[k];
        int r[3];
        for (int a = 0; a < 3; ++a) {
            const double o = (m.C[3L * j + a] - m.C[3L * c + a]) / h;
            r[a] = (int)rint(o);
            ok &= fabs(o - r[a]) < 1e-6 && r[a] >= -1 && r[a] <= 1;
        }
        ok &= cube[j] && fabs(m.V[j] - m.V[c]) <= 1e-9 * m.V[c];
        if (!ok) break;
        const int li = lat_index(r[0], r[1], r[2]);
        ok &= L[li] < 0;
        L[li] = j;
    }
    if (!ok)
        for (int k = 0; k < NIN; ++k) L[k] = -1;
}

// Is the lattice map affine (lat[27c + L] = c + off[L] for every cell with a
// full stencil, one off for all)?  probe: [0] lowest cell with a full
// stencil, [1] affine flag, [2 + L] off[L] of that cell.
__global__ void k_lat_probe_init(int* probe) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        probe[0] = 0x7fffffff;
        probe[1] = 1;
    }
}

__global__ void __launch_bounds__(256) k_lat_probe_first(const int* lat, int nC, int* probe) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c < nC && lat[(long)c * NIN] >= 0) atomicMin(&probe[0], c);
}

__global__ void k_lat_probe_off(const int* lat, int* probe) {
    const int L = threadIdx.x;
    const int c0 = probe[0];
    if (L < NIN) probe[2 + L] = c0 == 0x7fffffff ? 0 : lat[(long)c0 * NIN + L] - c0;
    if (L == 0 && c0 == 0x7fffffff) probe[1] = 0;
}

__global__ void __launch_bounds__(256) k_lat_probe_check(const int* lat, int nC, int* probe) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= nC || lat[(long)c * NIN] < 0) return;
    bool ok = true;
    for (int L = 0; L < NIN; ++L) ok &= lat[(long)c * NIN + L] == c + probe[2 + L];
    if (!ok) atomicAnd(&probe[1], 0);
}

__global__ void k_zero_ints(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_cube_fractions(const double* in, long k, double* out, hipStream_t s) {
    k_cube_fractions<<<cdiv(k, 256), 256, 0, s>>>(in, k, out);
}

void launch_lattice(const DevMesh& m, unsigned char* cubeScratch, int* lat, hipStream_t s) {
    k_cube_flags<<<cdiv(m.nC, 256), 256, 0, s>>>(m, cubeScratch);
    k_lattice<<<cdiv(m.nC, 256), 256, 0, s>>>(m, cubeScratch, lat);
}

void launch_lattice_probe(const int* lat, int nC, int* probe, hipStream_t s) {
    k_lat_probe_init<<<1, 64, 0, s>>>(probe);
    k_lat_probe_first<<<cdiv(nC, 256), 256, 0, s>>>(lat, nC, probe);
    k_lat_probe_off<<<1, 64, 0, s>>>(lat, probe);
    k_lat_probe_check<<<cdiv(nC, 256), 256, 0, s>>>(lat, nC, probe);
}

void launch_zero_ints(int* p, int n, hipStream_t s) { k_zero_ints<<<cdiv(n, 256), 256, 0, s>>>(p, n); }
void launch_gen(uint64_t seed, long first, long n, int depth, double* alpha27, double* shape19, uint16_t* X, float* Y,
                hipStream_t s) {
    if (n <= 0) return;
    k_gen_alpha<<<cdiv(n * NIN, 256), 256, 0, s>>>(seed, first, n, depth, alpha27, shape19);
    k_gen_pack<<<cdiv(n, 256), 256, 0, s>>>(n, alpha27, shape19, X, Y);
}

void launch_forward(const uint16_t* X, long n, const uint16_t* wb, const float* pq, float* out, hipStream_t s) {
    if (n <= 0) return;
    k_forward<<<cdiv(n, 64), 256, 0, s>>>(X, n, wb, pq, out);
}

void launch_train_grad(const uint16_t* X, const float* Y, long n, long pa, long pb, long base, int nWG,
                       const uint16_t* wb, const float* pq, float invB, float* partial, float* lossPart,
                       hipStream_t s) {
    k_train<<<nWG, 256, 0, s>>>(X, Y, n, pa, pb, base, wb, pq, invB, partial, lossPart);
}

void launch_reduce(const float* partial, int nWG, float* grad, hipStream_t s) {
    k_reduce<<<cdiv(N_PADDED, 256), 256, 0, s>>>(partial, nWG, grad);
}

void launch_adam(const float* partial, int nWG, float* pq, float* m1, float* m2, uint16_t* wb, float lr, float b1,
                 float b2, float eps, float c1, float c2, hipStream_t s) {
    k_adam<<<cdiv(N_PADDED, 256), 256, 0, s>>>(partial, nWG, pq, m1, m2, wb, lr, b1, b2, eps, c1, c2);
}

void launch_sync_bf16(const float* pq, uint16_t* wb, hipStream_t s) {
    k_sync_bf16<<<cdiv(N_PADDED, 256), 256, 0, s>>>(pq, wb);
}

void launch_infer(const int* lat, const LatStride& ls, const double* alpha, const int* mixed, const int* nMixed,
                  const double* C, const double* V, const uint16_t* wb, const float* pq, double tol, double* ifN,
                  unsigned char* learned, double* resid, int* lsList, int* nLs, int* nFb, int nC, hipStream_t s) {
    // workgroups (each stages the 45 KB of weights once): knob SVP_INFER_GRID
    static const int maxGrid = [] {
        const char* e = std::getenv("SVP_INFER_GRID");
        const int g = e ? std::atoi(e) : 0;
        return (g >= 1 && g <= 4096) ? g : 512;
    }();
    k_infer<<<std::max(1, std::min(maxGrid, cdiv(nC, 16 * INFER_WAVES))), 64 * INFER_WAVES, 0, s>>>(
        lat, ls, nC, alpha, mixed, nMixed, C, V, wb, pq, tol, ifN, learned, resid, lsList, nLs, nFb);
}


}  // namespace nm
}  // namespace svp
