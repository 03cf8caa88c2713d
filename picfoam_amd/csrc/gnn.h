// Learned Poisson initial guess: device buffers and launch wrappers (k_gnn.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace picmi {

struct Gnn {
    int n = 0, K = 0, L = 0, np = 0;
    double dbar = 1.0;               // mean diagonal of the Laplacian
    hipStream_t stream = nullptr;
    const int32_t* nb = nullptr;     // cell graph (Laplacian couplings), [n*K]
    // parameters: W_enc[3][64], b_enc[64], per layer W[128][64], b[64], w_dec[64], b_dec
    float *theta = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    uint16_t* wbf = nullptr;         // per layer: Wf[128][64], WT[64][128] (BF16)
    // activations (training keeps every layer; inference ping-pongs two)
    float* x = nullptr;              // [n][3] features
    const float* cur_x = nullptr;    // features of the last forward pass
    uint16_t *h = nullptr, *mm = nullptr, *z = nullptr, *dz = nullptr;
    float *out = nullptr, *g = nullptr, *dm = nullptr, *part = nullptr;
    double *q = nullptr, *q2 = nullptr, *red = nullptr, *scal = nullptr;
    int loss_kind = 0;               // 0: A-norm, 1: residual norm, 2: both
    std::vector<void*> bufs;
    // snapshots: features, targets, normalisation
    int snap_cap = 0, snap_count = 0, snap_next = 0;
    float* snap_x = nullptr;         // [cap][n][3]
    float* snap_t = nullptr;         // [cap][n]
    std::vector<double> snap_norm;   // d^T A d
    std::vector<double> snap_rnorm;  // |A d|^2
    int64_t adam_t = 0;

    static int n_params(int L);
    const int32_t* nb_ptr() const { return nb; }
    hipError_t alloc(int n, int K, int L, hipStream_t s);
    void release();
    void refresh_weights();
    // features from r = b - A phi_prev, b and phi_prev (FP64 device vectors)
    hipError_t features(const double* r, const double* b, const double* phi);
    // target d = (phi* - phi_prev) dbar / s_r (call after features)
    hipError_t target(const double* phi_star, const double* phi_prev, float* t);
    hipError_t forward(const float* xin, bool train);
    // guess = phi_prev + out * s_r / dbar (after features + forward)
    hipError_t guess(const double* phi_prev, double* out_phi);
    // A-norm loss against t, gradient into grad; loss_out = e^T A e / norm
    hipError_t loss_backward(const float* t, double norm, const int32_t* anb, const double* aa, const double* ad,
                             int aK, double* loss_out);
    hipError_t loss_backward2(const float* t, double norm_a, double norm_r, const int32_t* anb, const double* aa,
                              const double* ad, int aK, double* loss_out);
    hipError_t adam(float lr, int64_t t);
};

}  // namespace picmi

namespace picmi {
// t^T A t on the device (synchronises)
double gnn_rnorm(Gnn& G, const float* t, const int32_t* anb, const double* aa, const double* ad, int aK);
double gnn_anorm(Gnn& G, const float* t, const int32_t* anb, const double* aa, const double* ad, int aK);
}  // namespace picmi
