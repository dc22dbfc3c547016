// C-ABI of the learned-normal model (include/simplic_learned.h): parameter
// storage, training-set generation, training, and the forward/gradient
// entry points the tests check against the CPU network.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "api_common.hpp"
#include "learned_device.hpp"
#include "learned_kernels.hpp"
#include "learned_model.hpp"
#include "simplic_learned.h"

using namespace svp::api;
using namespace svp::nm;

namespace {

uint64_t host_mix(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double host_urand(uint64_t& s) { return (double)(host_mix(s) >> 11) * (1.0 / 9007199254740992.0); }

svp_normal_model& M(svp_normal_model* m) {
    if (!m) throw SvpError(SVP_E_ARG, "null model");
    HIPCHECK(hipSetDevice(m->device));
    return *m;
}

void upload_packed(svp_normal_model& m, const float* packed) {
    std::vector<float> q;
    packed_to_padded(packed, q);
    HIPCHECK(hipMemcpy(m.pq.p, q.data(), sizeof(float) * N_PADDED, hipMemcpyHostToDevice));
    HIPCHECK(hipMemsetAsync(m.m1.p, 0, sizeof(float) * N_PADDED, m.stream));
    HIPCHECK(hipMemsetAsync(m.m2.p, 0, sizeof(float) * N_PADDED, m.stream));
    m.adamT = 0;
    launch_sync_bf16((const float*)m.pq.p, (uint16_t*)m.wb.p, m.stream);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(m.stream));
}

// device copy of a host or device array
template <class T>
const T* stage_in(svp_normal_model& m, DBuf& tmp, const T* src, long n, int where) {
    if (where == SVP_DEVICE) return src;
    tmp.alloc<T>(n);
    HIPCHECK(hipMemcpy(tmp.p, src, sizeof(T) * n, hipMemcpyHostToDevice));
    return (const T*)tmp.p;
}

}  // namespace

extern "C" {

int svp_nm_create(int device, uint64_t seed, svp_normal_model** out) {
    return guarded([&] {
        if (!out) throw SvpError(SVP_E_ARG, "svp_nm_create: null argument");
        HIPCHECK(hipSetDevice(device));
        auto m = std::make_unique<svp_normal_model>();
        m->device = device;
        HIPCHECK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
        m->pq.alloc<float>(N_PADDED);
        m->m1.alloc<float>(N_PADDED);
        m->m2.alloc<float>(N_PADDED);
        m->wb.alloc<uint16_t>(N_BF16);
        m->grad.alloc<float>(N_PADDED);
        std::vector<float> p(N_PARAMS);
        uint64_t s = seed ^ 0x5DEECE66Dull;
        auto fill = [&](int off, int n, int fanIn) {
            const double b = 1.0 / std::sqrt((double)fanIn);
            for (int i = 0; i < n; ++i) p[off + i] = (float)((2.0 * host_urand(s) - 1.0) * b);
        };
        fill(P_W1, NH * NIN, NIN);
        fill(P_B1, NH, NIN);
        fill(P_W2, NH * NH, NH);
        fill(P_B2, NH, NH);
        fill(P_W3, NOUT * NH, NH);
        fill(P_B3, NOUT, NH);
        upload_packed(*m, p.data());
        *out = m.release();
    });
}

int svp_nm_destroy(svp_normal_model* m) {
    return guarded([&] {
        if (!m) return;
        (void)hipSetDevice(m->device);
        if (m->stream) {
            (void)hipStreamSynchronize(m->stream);
            (void)hipStreamDestroy(m->stream);
        }
        delete m;
    });
}

int svp_nm_set_params(svp_normal_model* m, const float* params, int32_t n) {
    return guarded([&] {
        auto& s = M(m);
        if (!params || n != N_PARAMS) throw SvpError(SVP_E_ARG, "svp_nm_set_params: need SVP_NM_N_PARAMS values");
        for (int i = 0; i < n; ++i)
            if (!std::isfinite(params[i])) throw SvpError(SVP_E_NAN, "svp_nm_set_params: non-finite parameter");
        upload_packed(s, params);
    });
}

int svp_nm_get_params(svp_normal_model* m, float* params, int32_t n) {
    return guarded([&] {
        auto& s = M(m);
        if (!params || n != N_PARAMS) throw SvpError(SVP_E_ARG, "svp_nm_get_params: need SVP_NM_N_PARAMS values");
        std::vector<float> q(N_PADDED);
        HIPCHECK(hipStreamSynchronize(s.stream));
        HIPCHECK(hipMemcpy(q.data(), s.pq.p, sizeof(float) * N_PADDED, hipMemcpyDeviceToHost));
        padded_to_packed(q.data(), params);
    });
}

int svp_nm_generate(svp_normal_model* m, int64_t n, uint64_t seed, int32_t depth) {
    return guarded([&] {
        auto& s = M(m);
        if (n <= 0 || n > (int64_t)1 << 28) throw SvpError(SVP_E_ARG, "svp_nm_generate: n out of range");
        if (depth < 1 || depth > 7) throw SvpError(SVP_E_ARG, "svp_nm_generate: depth must be 1..7");
        HIPCHECK(hipStreamSynchronize(s.stream));
        s.alpha27.alloc<double>((size_t)n * NIN);
        s.shape19.alloc<double>((size_t)n * 19);
        s.X.alloc<uint16_t>((size_t)n * KIN);
        s.Y.alloc<float>((size_t)n * 4);
        s.nData = n;
        // chunks keep one launch within a few seconds
        const long chunk = 1L << 20;
        for (long f = 0; f < n; f += chunk) {
            const long c = std::min<long>(chunk, n - f);
            launch_gen(seed, f, c, depth, (double*)s.alpha27.p + f * NIN, (double*)s.shape19.p + f * 19,
                       (uint16_t*)s.X.p + f * KIN, (float*)s.Y.p + f * 4, s.stream);
            HIPCHECK(hipGetLastError());
        }
        HIPCHECK(hipStreamSynchronize(s.stream));
    });
}

int svp_nm_dataset_get(svp_normal_model* m, int64_t first, int64_t count, double* alpha27, double* shape,
                       uint16_t* x, float* y) {
    return guarded([&] {
        auto& s = M(m);
        if (first < 0 || count < 0 || first + count > s.nData)
            throw SvpError(SVP_E_ARG, "svp_nm_dataset_get: range outside the training set");
        HIPCHECK(hipStreamSynchronize(s.stream));
        if (alpha27)
            HIPCHECK(hipMemcpy(alpha27, (double*)s.alpha27.p + first * NIN, sizeof(double) * count * NIN,
                               hipMemcpyDeviceToHost));
        if (shape)
            HIPCHECK(hipMemcpy(shape, (double*)s.shape19.p + first * 19, sizeof(double) * count * 19,
                               hipMemcpyDeviceToHost));
        if (x) HIPCHECK(hipMemcpy(x, (uint16_t*)s.X.p + first * KIN, 2 * count * KIN, hipMemcpyDeviceToHost));
        if (y) HIPCHECK(hipMemcpy(y, (float*)s.Y.p + first * 4, sizeof(float) * count * 4, hipMemcpyDeviceToHost));
    });
}

int svp_nm_train(svp_normal_model* m, int32_t epochs, int32_t batch, float lr0, float lr1, uint64_t seed,
                 float* loss_per_epoch, double* seconds) {
    return guarded([&] {
        auto& s = M(m);
        if (epochs < 0 || batch < 64 || batch % 64 != 0 || batch > (1 << 20))
            throw SvpError(SVP_E_ARG, "svp_nm_train: batch must be a positive multiple of 64");
        if (s.nData < batch) throw SvpError(SVP_E_ARG, "svp_nm_train: training set smaller than one batch");
        if (!(lr0 > 0.f) || !(lr1 >= 0.f)) throw SvpError(SVP_E_ARG, "svp_nm_train: bad learning rate");
        const int nWG = batch / 64;
        const long steps = s.nData / batch;
        if (s.partialWG < nWG) {
            s.partial.alloc<float>((size_t)nWG * N_PADDED);
            s.partialWG = nWG;
        }
        if (s.lossSlots < steps * nWG) {
            s.lossPart.alloc<float>((size_t)steps * nWG);
            s.lossSlots = (int)(steps * nWG);
        }
        hipEvent_t e0, e1;
        HIPCHECK(hipEventCreate(&e0));
        HIPCHECK(hipEventCreate(&e1));
        HIPCHECK(hipEventRecord(e0, s.stream));
        const double b1 = 0.9, b2 = 0.999;
        const long total = (long)epochs * steps;
        uint64_t rs = seed ^ 0xA0761D6478BD642Full;
        std::vector<float> lp((size_t)steps * nWG);
        for (int ep = 0; ep < epochs; ++ep) {
            long pa;
            do {
                pa = 1 + (long)(host_mix(rs) % (uint64_t)std::max<long>(1, s.nData - 1));
            } while (std::gcd(pa, s.nData) != 1);
            const long pb = (long)(host_mix(rs) % (uint64_t)s.nData);
            for (long st = 0; st < steps; ++st) {
                const long t = (long)ep * steps + st;
                const double lr = lr1 + 0.5 * (lr0 - lr1) * (1.0 + std::cos(M_PI * (double)t / (double)total));
                launch_train_grad((const uint16_t*)s.X.p, (const float*)s.Y.p, s.nData, pa, pb, st * batch, nWG,
                                  (const uint16_t*)s.wb.p, (const float*)s.pq.p, 1.0f / (float)batch,
                                  (float*)s.partial.p, (float*)s.lossPart.p + st * nWG, s.stream);
                s.adamT++;
                const float c1 = (float)(1.0 - std::pow(b1, (double)s.adamT));
                const float c2 = (float)(1.0 - std::pow(b2, (double)s.adamT));
                launch_adam((const float*)s.partial.p, nWG, (float*)s.pq.p, (float*)s.m1.p, (float*)s.m2.p,
                            (uint16_t*)s.wb.p, (float)lr, (float)b1, (float)b2, 1e-8f, c1, c2, s.stream);
            }
            HIPCHECK(hipGetLastError());
            if (loss_per_epoch) {
                HIPCHECK(hipMemcpyAsync(lp.data(), s.lossPart.p, sizeof(float) * lp.size(), hipMemcpyDeviceToHost,
                                        s.stream));
                HIPCHECK(hipStreamSynchronize(s.stream));
                double acc = 0.0;
                for (float v : lp) acc += v;
                loss_per_epoch[ep] = (float)(acc / ((double)steps * batch));
            }
        }
        HIPCHECK(hipEventRecord(e1, s.stream));
        HIPCHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (seconds) *seconds = 1e-3 * ms;
    });
}

int svp_nm_forward(svp_normal_model* m, const uint16_t* x, int64_t n, float* out, int where) {
    return guarded([&] {
        auto& s = M(m);
        if (!x || !out || n < 0) throw SvpError(SVP_E_ARG, "svp_nm_forward: bad argument");
        if (n == 0) return;
        DBuf tx, to;
        const uint16_t* dx = stage_in(s, tx, x, n * KIN, where);
        float* dout = where == SVP_DEVICE ? out : to.alloc<float>(n * 3);
        launch_forward(dx, n, (const uint16_t*)s.wb.p, (const float*)s.pq.p, dout, s.stream);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(s.stream));
        if (where != SVP_DEVICE) HIPCHECK(hipMemcpy(out, dout, sizeof(float) * n * 3, hipMemcpyDeviceToHost));
    });
}

int svp_nm_gradient(svp_normal_model* m, const uint16_t* x, const float* y, int64_t n, float* grad, double* loss,
                    int where) {
    return guarded([&] {
        auto& s = M(m);
        if (!x || !y || !grad || n <= 0 || n % 64 != 0 || n > (1 << 20))
            throw SvpError(SVP_E_ARG, "svp_nm_gradient: n must be a positive multiple of 64");
        const int nWG = (int)(n / 64);
        DBuf tx, ty, part, lpart;
        const uint16_t* dx = stage_in(s, tx, x, n * KIN, where);
        const float* dy = stage_in(s, ty, y, n * 4, where);
        part.alloc<float>((size_t)nWG * N_PADDED);
        lpart.alloc<float>(nWG);
        launch_train_grad(dx, dy, n, 1, 0, 0, nWG, (const uint16_t*)s.wb.p, (const float*)s.pq.p, 1.0f / (float)n,
                          (float*)part.p, (float*)lpart.p, s.stream);
        launch_reduce((const float*)part.p, nWG, (float*)s.grad.p, s.stream);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(s.stream));
        std::vector<float> q(N_PADDED), lp(nWG);
        HIPCHECK(hipMemcpy(q.data(), s.grad.p, sizeof(float) * N_PADDED, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(lp.data(), lpart.p, sizeof(float) * nWG, hipMemcpyDeviceToHost));
        padded_to_packed(q.data(), grad);
        if (loss) {
            double acc = 0.0;
            for (float v : lp) acc += v;
            *loss = acc / (double)n;
        }
    });
}

int svp_nm_cube_fractions(const double* in, int64_t k, double* out, int device) {
    return guarded([&] {
        if (!in || !out || k < 0) throw SvpError(SVP_E_ARG, "svp_nm_cube_fractions: bad argument");
        if (k == 0) return;
        HIPCHECK(hipSetDevice(device));
        DBuf di, dout;
        di.alloc<double>(k * 8);
        dout.alloc<double>(k);
        HIPCHECK(hipMemcpy(di.p, in, sizeof(double) * k * 8, hipMemcpyHostToDevice));
        launch_cube_fractions((const double*)di.p, k, (double*)dout.p, nullptr);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipDeviceSynchronize());
        HIPCHECK(hipMemcpy(out, dout.p, sizeof(double) * k, hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
