// Per-node MLP of the learned warm start: sizes, parameter layout, device views and
// launchers (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"

namespace mm {

constexpr int kMlpIn = 32;            // input features per node
constexpr int kMlpW = 128;            // hidden width
constexpr int kMlpP1 = kMlpIn + 8;    // padded LDS / bf16 row lengths (conflict-free 8-byte reads)
constexpr int kMlpPW = kMlpW + 8;

// Flat FP32 parameter vector: W_l row-major [in][out], then b_l.
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpIn * kMlpW;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpW;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpW * kMlpW;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpW;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpW * kMlpW;
constexpr int kMlpOffW4 = kMlpOffB3 + kMlpW;
constexpr int kMlpOffB4 = kMlpOffW4 + kMlpW;
constexpr int kMlpParams = kMlpOffB4 + 1;   // 37377

// What the kernels read: bf16 copies (forward: W^T rows; backward: W rows) and the
// FP32 biases / last layer inside the flat parameter vector.
struct MlpDev {
    const uint16_t *w1t = nullptr;              // [128][kMlpP1]
    const uint16_t *w2t = nullptr, *w3t = nullptr;   // [128][kMlpPW]
    const uint16_t *w2n = nullptr, *w3n = nullptr;   // [128][kMlpPW]
    const float *b1 = nullptr, *b2 = nullptr, *b3 = nullptr, *w4 = nullptr, *b4 = nullptr;
};

// Structured grid as the feature kernel sees it.
struct FeatGrid {
    int nx = 0, ny = 0, nz = 0;
    const double *x = nullptr, *y = nullptr, *z = nullptr;   // node coordinates per axis
    double inv_half = 0;        // 1 / half box edge
    double inv_V = 0;           // 1 / magnet volume
    double inv_js = 0;          // 1 / largest Js
    int m0x = 0, m0y = 0, m0z = 0, mn = 0;
    const uint16_t *vmat = nullptr;
};

void launch_mlp_forward(const Launch &L, const MlpDev &w, int64_t n, const uint16_t *X, float *y);
void launch_mlp_forward_train(const Launch &L, const MlpDev &w, int64_t n, const uint16_t *X, float *y,
                              uint16_t *h1, uint16_t *h2, uint16_t *h3, int64_t ldh);
void launch_mlp_loss(const Launch &L, int64_t n, const float *y, const float *t, float *dy, double *loss);
void launch_mlp_backward(const Launch &L, const MlpDev &w, int64_t n, const float *dy, const uint16_t *h1,
                         const uint16_t *h2, const uint16_t *h3, uint16_t *dz1, uint16_t *dz2, uint16_t *dz3,
                         int64_t ldh);
void launch_mlp_param_grads(const Launch &L, int64_t n, int64_t ldh, const uint16_t *X, const float *dy,
                            const uint16_t *h1, const uint16_t *h2, const uint16_t *h3, const uint16_t *dz1,
                            const uint16_t *dz2, const uint16_t *dz3, float *grad);
void launch_mlp_adam(const Launch &L, float *p, const float *g, float *m1, float *m2, float lr, int step);
void launch_mlp_publish(const Launch &L, const float *p, uint16_t *w1t, uint16_t *w2t, uint16_t *w3t, uint16_t *w2n,
                        uint16_t *w3n);
void launch_mlp_features(const Launch &L, const FeatGrid &F, const double *m, const double *mbar, uint16_t *X);
void launch_mlp_direction(const Launch &L, const FeatGrid &F, const float *pred, const float *prev, const double *u,
                          double s, double *d);
// learned preconditioner: w = r / diag(K); residual features of w (ss = w.w on the device);
// z = c s y with s = sqrt(ss / N) and c the network's output scale, zero on the faces
void launch_lp_scaled(const Launch &L, int64_t n, const double *r, const double *invD, double *w);
void launch_lp_features(const Launch &L, const FeatGrid &F, const PoissonGrid<double> &P, const double *w,
                        const double *ss, uint16_t *X);
void launch_lp_apply(const Launch &L, const FeatGrid &F, const float *y, const double *ss, double c, double *z);
void launch_mlp_galerkin(const Launch &L, int64_t n, const double *num, const double *den, const double *d,
                         double *u, double *gamma_out);

}  // namespace mm
