// Harness kernels of the deformation benchmark (plicVofAdvectionFoam) and
// input checks.
#include "simplic_device.hpp"
#include "simplic_types.hpp"

namespace svp {

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// updateU.H:117-127: phi *= uFactor; U *= uFactor
__global__ void k_scale_flow(int nC, int nF, int nB, const double* U0, const double* phi0,
                             const double* U0b, double factor, double* U, double* phi, double* Ub) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nF) phi[i] = phi0[i] * factor;
    if (i < 3 * nC) U[i] = U0[i] * factor;
    if (i < 3 * nB) Ub[i] = U0b ? U0b[i] * factor : 0.0;
}

void launch_scale_flow(int nC, int nF, int nB, const double* U0, const double* phi0, const double* U0b,
                       double factor, double* U, double* phi, double* Ub, hipStream_t s) {
    const long n = std::max<long>(std::max<long>(nF, 3L * nC), 3L * nB);
    k_scale_flow<<<cdiv(n, 256), 256, 0, s>>>(nC, nF, nB, U0, phi0, U0b, factor, U, phi, Ub);
}

// fvc::surfaceSum(mag(phi)) in ascending face order, then max of sum/V and of
// (pos0(alpha-0.01)*pos0(0.99-alpha)*sum)/V.  All values are >= 0, so the max
// is an integer max over the IEEE bit patterns.
__global__ void k_courant(DevMesh m, const double* phi, const double* alpha, unsigned long long* out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double r = 0.0, ra = 0.0;
    if (c < m.nC) {
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
    // wave max then one atomic per wave
    unsigned long long br = __double_as_longlong(r), bra = __double_as_longlong(ra);
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(br, o), y = __shfl_xor(bra, o);
        br = x > br ? x : br;
        bra = y > bra ? y : bra;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(out, br);
        atomicMax(out + 1, bra);
    }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    const int i = threadIdx.x;
    if (i < n) p[i] = 0ull;
}

void launch_courant(const DevMesh& m, const double* phi, const double* alpha, unsigned long long* out,
                    hipStream_t s) {
    k_zero_u64<<<1, 64, 0, s>>>(out, 2);
    k_courant<<<cdiv(m.nC, 256), 256, 0, s>>>(m, phi, alpha, out);
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

}  // namespace svp
