// The learned-normal model object behind the svp_nm_* C-ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "api_common.hpp"
#include "learned_device.hpp"

struct svp_normal_model {
    int device = 0;
    hipStream_t stream = nullptr;
    svp::api::DBuf pq;        // FP32 master parameters, padded layout
    svp::api::DBuf m1, m2;    // Adam moments
    svp::api::DBuf wb;        // BF16 weight copies read by the kernels
    svp::api::DBuf partial;   // per-workgroup gradient shares
    svp::api::DBuf lossPart;
    svp::api::DBuf grad;
    int partialWG = 0;
    int lossSlots = 0;
    // training set
    svp::api::DBuf alpha27, shape19, X, Y;
    long nData = 0;
    long adamT = 0;
};

namespace svp {
namespace nm {

// packed (public) <-> padded (device) parameter order
inline void packed_to_padded(const float* p, std::vector<float>& q) {
    q.assign(N_PADDED, 0.f);
    for (int m = 0; m < NH; ++m)
        for (int i = 0; i < NIN; ++i) q[Q_W1 + m * KIN + i] = p[P_W1 + m * NIN + i];
    for (int m = 0; m < NH; ++m) q[Q_B1 + m] = p[P_B1 + m];
    for (int e = 0; e < NH * NH; ++e) q[Q_W2 + e] = p[P_W2 + e];
    for (int m = 0; m < NH; ++m) q[Q_B2 + m] = p[P_B2 + m];
    for (int e = 0; e < NOUT * NH; ++e) q[Q_W3 + e] = p[P_W3 + e];
    for (int o = 0; o < NOUT; ++o) q[Q_B3 + o] = p[P_B3 + o];
}
inline void padded_to_packed(const float* q, float* p) {
    for (int m = 0; m < NH; ++m)
        for (int i = 0; i < NIN; ++i) p[P_W1 + m * NIN + i] = q[Q_W1 + m * KIN + i];
    for (int m = 0; m < NH; ++m) p[P_B1 + m] = q[Q_B1 + m];
    for (int e = 0; e < NH * NH; ++e) p[P_W2 + e] = q[Q_W2 + e];
    for (int m = 0; m < NH; ++m) p[P_B2 + m] = q[Q_B2 + m];
    for (int e = 0; e < NOUT * NH; ++e) p[P_W3 + e] = q[Q_W3 + e];
    for (int o = 0; o < NOUT; ++o) p[P_B3 + o] = q[Q_B3 + o];
}

}  // namespace nm
}  // namespace svp
