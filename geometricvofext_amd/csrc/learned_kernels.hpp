// Launchers of the learned-normal kernels (learned_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svp {
struct DevMesh;
namespace nm {

// one training shape; record layout of the 19 doubles kept per sample:
// kind, c[3], a[3], Q[9] (row-major, columns = shape axes), nrm[3]
struct NmShape {
    int kind;        // 0 plane, 1 sphere, 2 ellipsoid
    double c[3];     // plane point / centre
    double a[3];     // semi-axes (sphere: radius)
    double Q[9];
    double nrm[3];   // target normal (outward)
};

void launch_gen(uint64_t seed, long first, long n, int depth, double* alpha27, double* shape19, uint16_t* X, float* Y,
                hipStream_t s);
void launch_forward(const uint16_t* X, long n, const uint16_t* wb, const float* pq, float* out, hipStream_t s);
void launch_train_grad(const uint16_t* X, const float* Y, long n, long pa, long pb, long base, int nWG,
                       const uint16_t* wb, const float* pq, float invB, float* partial, float* lossPart,
                       hipStream_t s);
void launch_reduce(const float* partial, int nWG, float* grad, hipStream_t s);
void launch_adam(const float* partial, int nWG, float* pq, float* m1, float* m2, uint16_t* wb, float lr, float b1,
                 float b2, float eps, float c1, float c2, hipStream_t s);
void launch_sync_bf16(const float* pq, uint16_t* wb, hipStream_t s);
// lattice map of the form lat[27c + L] = c + off[L] on every cell with a full
// stencil (a lattice numbered in axis order): affine = 1 and the offsets
struct LatStride {
    int affine;
    int off[27];
};
void launch_infer(const int* lat, const LatStride& ls, const double* alpha, const int* mixed, const int* nMixed,
                  const double* C, const double* V, const uint16_t* wb, const float* pq, double tol, double* ifN,
                  unsigned char* learned, double* resid, int* lsList, int* nLs, int* nFb, int nC, hipStream_t s);

void launch_cube_fractions(const double* in, long k, double* out, hipStream_t s);
void launch_lattice(const DevMesh& m, unsigned char* cubeScratch, int* lat, hipStream_t s);
void launch_zero_ints(int* p, int n, hipStream_t s);
// probe[2 + 27]: see LatStride (probe[1] = affine, probe[2 + L] = off[L])
void launch_lattice_probe(const int* lat, int nC, int* probe, hipStream_t s);

}  // namespace nm
}  // namespace svp
