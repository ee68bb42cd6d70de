// Surrogate network interface shared by the step (api.cpp) and the network
// kernels (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

struct bw_mlp;

namespace bw {

constexpr int MLP_IN = 144, MLP_INP = 160, MLP_H = 256, MLP_OUT = 12, MLP_OUTP = 32, MLP_T = 6;
constexpr int MLP_L = 4;

// activations of one forward/backward pass over up to `cap` rows
struct MlpBuf {
    int cap = 0;
    unsigned short *Xb = nullptr, *Hb[3] = {}, *dYb = nullptr, *dZb = nullptr;
    float *Z[3] = {}, *H[3] = {}, *Y = nullptr, *dY = nullptr, *dH = nullptr, *dZ = nullptr, *loss = nullptr;
    float *part = nullptr;  // partial sums of the backward reductions (split-K products, column sums)
};
int mlp_buf_alloc(MlpBuf &b, int rows);  // hipMalloc; 0 on success
void mlp_buf_free(MlpBuf &b);

// features x (n x 144 raw units, device) -> y_out (n x 12, device): impulse
// mean (6) and log-variance (6) in physical units; launches only, on `st`
void mlp_forward_device(bw_mlp *m, MlpBuf &b, const float *x, int n, float *y_out, hipStream_t st);

}  // namespace bw
