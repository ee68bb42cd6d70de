// Surrogate closure network of libbrownian: a multilayer perceptron
//   144 -> 256 -> 256 -> 256 -> 12, SiLU hidden activations,
// run and trained on the GPU with its own kernels.  Matrix products use the
// gfx950 matrix cores (V_MFMA_F32_32X32X16_BF16: BF16 operands, FP32
// accumulation, operands staged through LDS); weights are kept in FP32 (Adam
// state too) with BF16 copies for the products.  The output is the per-interval colloid impulse mean (6)
// and log-variance (6) in standardised units; the loss is the Gaussian
// negative log-likelihood.
//
// Padded shapes (multiples of 32): input 144 -> 160 (zero columns), output
// 12 -> 32 (zero weight columns).  Batch rows are padded to 32 with zeros.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <cmath>
#include <type_traits>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/brownian.h"
#include "mlp.h"

namespace bw {

typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// C (M x N) = op(A) (M x K) . op(B) (K x N) on the gfx950 matrix cores
// (V_MFMA_F32_32X32X16_BF16: BF16 operands, FP32 accumulation).  One
// workgroup of four waves per 64 x 64 block of C, each wave a 32 x 32
// quadrant; K in steps of 32 staged through LDS by 128-bit global loads (a
// transposed operand is transposed on its way into LDS), A as As[m][k] and B
// as Bs[n][k], so that a lane's 8 consecutive k are one 128-bit LDS read.
// M, N and K are multiples of 32; rows or columns of a block past M or N are
// zero-filled and not stored.
//   TA = 0: A is M x K row-major (lda = row stride); TA = 1: A is stored K x M.
//   TB = 0: B is K x N row-major;                    TB = 1: B is stored N x K.
// EPI 0: C = acc.  EPI 1: Z = acc + bias, C = silu(Z), Cb = bf16(C).
// EPI 2: C = acc + bias.
// Operand layout of one 32x32x16 product: lane l supplies A[l%32][k] and
// B[k][l%32] for 8 values of k selected by l/32 (the same selection for A and
// B, so that the sum over k does not depend on it); accumulator element i of
// lane l is C[8(i/4) + 4(l/32) + i%4][l%32].
template <int TA, int TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(const unsigned short *__restrict__ A,
                                              const unsigned short *__restrict__ B, int M, int N, int K, int lda,
                                              int ldb, float *__restrict__ C, int ldc, const float *__restrict__ bias,
                                              float *__restrict__ Z, unsigned short *__restrict__ Cb) {
    constexpr int BM = 64, BN = 64, BK = 32, LD = BK + 8;  // LDS rows of 40 BF16 (80 bytes, 16-byte aligned)
    __shared__ __align__(16) unsigned short As[BM * LD], Bs[BN * LD];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int nbn = (N + BN - 1) / BN;
    const int m0 = (blockIdx.x / nbn) * BM, n0 = (blockIdx.x % nbn) * BN;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;  // the wave's quadrant
    const int r = lane & 31, kh = (lane >> 5) * 8;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += BK) {
        // 64 x 32 BF16 of each operand: 256 chunks of 8, one per thread
        u32x4 va = {0u, 0u, 0u, 0u}, vb = {0u, 0u, 0u, 0u};
        const int ri = tid >> 2, rk = (tid & 3) * 8;  // row-major chunk: row, first k
        const int ck = tid >> 3, cm = (tid & 7) * 8;  // transposed chunk: k, first row
        if (TA == 0) {
            if (m0 + ri < M) va = *(const u32x4 *)(A + (size_t)(m0 + ri) * lda + k0 + rk);
        } else {
            if (m0 + cm < M) va = *(const u32x4 *)(A + (size_t)(k0 + ck) * lda + m0 + cm);
        }
        if (TB == 1) {
            if (n0 + ri < N) vb = *(const u32x4 *)(B + (size_t)(n0 + ri) * ldb + k0 + rk);
        } else {
            if (n0 + cm < N) vb = *(const u32x4 *)(B + (size_t)(k0 + ck) * ldb + n0 + cm);
        }
        if (TA == 0) {
            *(u32x4 *)(As + ri * LD + rk) = va;
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++)
                As[(cm + j) * LD + ck] = (unsigned short)(va[j >> 1] >> (16 * (j & 1)));
        }
        if (TB == 1) {
            *(u32x4 *)(Bs + ri * LD + rk) = vb;
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++)
                Bs[(cm + j) * LD + ck] = (unsigned short)(vb[j >> 1] >> (16 * (j & 1)));
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            const u32x4 a = *(const u32x4 *)(As + (wm + r) * LD + kk + kh);
            const u32x4 b = *(const u32x4 *)(Bs + (wn + r) * LD + kk + kh);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                          acc, 0, 0, 0);
        }
        __syncthreads();
    }
    if (m0 + wm >= M || n0 + wn >= N) return;
    const int kq = (lane >> 5) * 4;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int m = m0 + wm + (i >> 2) * 8 + kq + (i & 3), n = n0 + wn + r;
        const size_t o = (size_t)m * ldc + n;
        const float v = acc[i];
        if (EPI == 0) {
            C[o] = v;
        } else if (EPI == 1) {
            const float z = v + bias[n];
            const float h = z / (1.0f + expf(-z));
            Z[o] = z;
            C[o] = h;
            Cb[o] = f2bf(h);
        } else {
            C[o] = v + bias[n];
        }
    }
}

__global__ void k_to_bf16(const float *__restrict__ x, unsigned short *__restrict__ y, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// standardised, padded BF16 input rows: xb[i][f] = bf16((x[idx[i]][f] - mean[f]) * istd[f])
__global__ void k_mlp_input(const float *__restrict__ x, const int *__restrict__ idx, int n, int n_pad,
                            const float *__restrict__ mean, const float *__restrict__ istd,
                            unsigned short *__restrict__ xb) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n_pad * MLP_INP) return;
    const int i = (int)(e / MLP_INP), f = (int)(e - (long long)i * MLP_INP);
    float v = 0.0f;
    if (i < n && f < MLP_IN) {
        const long long row = idx ? idx[i] : i;
        v = (x[row * MLP_IN + f] - mean[f]) * istd[f];
    }
    xb[e] = f2bf(v);
}

// Gaussian NLL on standardised targets: loss = 0.5 sum_k (s_k + (t_k - mu_k)^2 e^{-s_k})
// with mu = y[0:6], s = y[6:12] (clamped to +-20); gradients scaled by 1/n.
__global__ void k_mlp_nll(const float *__restrict__ y, const float *__restrict__ t, const int *__restrict__ idx,
                          int n, int n_pad, const float *__restrict__ tmean, const float *__restrict__ tistd,
                          float *__restrict__ dy, unsigned short *__restrict__ dyb, float *__restrict__ loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    float l = 0.0f;
    for (int c = 0; c < MLP_OUTP; c++) dy[(size_t)i * MLP_OUTP + c] = 0.0f;
    if (i < n) {
        const long long row = idx ? idx[i] : i;
        const float inv_n = 1.0f / (float)n;
        for (int k = 0; k < MLP_T; k++) {
            const float tk = (t[row * MLP_T + k] - tmean[k]) * tistd[k];
            const float mu = y[(size_t)i * MLP_OUTP + k];
            const float s = fminf(fmaxf(y[(size_t)i * MLP_OUTP + MLP_T + k], -20.0f), 20.0f);
            const float e = expf(-s), d = tk - mu;
            l += 0.5f * (s + d * d * e);
            dy[(size_t)i * MLP_OUTP + k] = -d * e * inv_n;
            dy[(size_t)i * MLP_OUTP + MLP_T + k] = 0.5f * (1.0f - d * d * e) * inv_n;
        }
    }
    for (int c = 0; c < MLP_OUTP; c++) dyb[(size_t)i * MLP_OUTP + c] = f2bf(dy[(size_t)i * MLP_OUTP + c]);
    loss[i] = l;
}

// dZ = dH * silu'(Z), silu'(z) = s (1 + z (1 - s)), s = sigmoid(z); FP32 and BF16 copies
__global__ void k_silu_back(const float *__restrict__ dh, const float *__restrict__ z, float *__restrict__ dz,
                            unsigned short *__restrict__ dzb, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float zz = z[i], s = 1.0f / (1.0f + expf(-zz));
    const float g = dh[i] * (s * (1.0f + zz * (1.0f - s)));
    dz[i] = g;
    dzb[i] = f2bf(g);
}

// column sums of a (rows x cols) matrix: bias gradients
// Column sums of g (rows x cols) in two fixed-order levels: workgroup (x, y)
// sums rows [y COLSUM_ROWS, (y+1) COLSUM_ROWS) of 64 columns, four row
// phases per column combined in order, into part[y][col]; k_sum_parts then
// adds the row blocks in order.
constexpr int COLSUM_ROWS = 128;
constexpr int WGRAD_SPLITS = 8;  // chunks of the batch in a weight-gradient product
__global__ __launch_bounds__(256) void k_colsum_part(const float *__restrict__ g, int rows, int cols,
                                                     float *__restrict__ part) {
    __shared__ float s[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63), ph = threadIdx.x >> 6;
    const int r0 = blockIdx.y * COLSUM_ROWS, r1 = min(rows, r0 + COLSUM_ROWS);
    float a = 0.0f;
    if (c < cols)
        for (int r = r0 + ph; r < r1; r += 4) a += g[(size_t)r * cols + c];
    s[ph][threadIdx.x & 63] = a;
    __syncthreads();
    if (threadIdx.x < 64 && c < cols) part[(size_t)blockIdx.y * cols + c] = ((s[0][threadIdx.x] + s[1][threadIdx.x]) +
                                                                           s[2][threadIdx.x]) + s[3][threadIdx.x];
}
// out[i] = sum over p < np of part[p * n + i], in order of p
__global__ void k_sum_parts(const float *__restrict__ part, int np, long long n, float *__restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a = 0.0f;
    for (int p = 0; p < np; p++) a += part[(size_t)p * n + i];
    out[i] = a;
}

// Adam (Kingma & Ba, 2015) on FP32 master weights; refreshes the BF16 copy.
// mask: parameters outside the logical (unpadded) shape stay zero.
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, unsigned short *__restrict__ wb, long long n, float lr, float b1,
                       float b2, float c1, float c2, float eps) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

// Activations of bw_mlp_forward_noalloc for up to NOALLOC_ROWS rows, in a
// file-scope device array instead of allocations: input (BF16), three hidden
// layers (pre-activation and activation in FP32, activation in BF16) and the
// padded output.  Each GPU holds its own copy until the process ends.
constexpr int NOALLOC_ROWS = 1024;
constexpr size_t NOALLOC_ROW_BYTES = (size_t)MLP_INP * 2 + (size_t)MLP_OUTP * 4 + 3 * ((size_t)MLP_H * 10);
static __device__ __align__(256) unsigned char scratch[NOALLOC_ROWS * NOALLOC_ROW_BYTES];

// rows [0, n) of the padded output (stride MLP_OUTP) -> n x MLP_OUT
__global__ void k_mlp_rows(const float *__restrict__ y, int n, float *__restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * MLP_OUT) return;
    const int i = e / MLP_OUT, k = e - i * MLP_OUT;
    out[e] = y[(size_t)i * MLP_OUTP + k];
}

}  // namespace bw

using namespace bw;

// ------------------------------------------------------------------- host
struct bw_mlp {
    std::string err;
    hipStream_t st = nullptr;
    int K[MLP_L] = {MLP_INP, MLP_H, MLP_H, MLP_H};
    int N[MLP_L] = {MLP_H, MLP_H, MLP_H, MLP_OUTP};
    int Kl[MLP_L] = {MLP_IN, MLP_H, MLP_H, MLP_H};   // logical (unpadded)
    int Nl[MLP_L] = {MLP_H, MLP_H, MLP_H, MLP_OUT};
    float *W[MLP_L] = {}, *b[MLP_L] = {}, *gW[MLP_L] = {}, *gb[MLP_L] = {};
    float *mW[MLP_L] = {}, *vW[MLP_L] = {}, *mb[MLP_L] = {}, *vb[MLP_L] = {};
    unsigned short *Wb[MLP_L] = {};
    float *xmean = nullptr, *xistd = nullptr, *tmean = nullptr, *tistd = nullptr;
    long long adam_t = 0;
    MlpBuf tb;  // training / host-call activations
    unsigned char *scratch = nullptr;  // device address of the scratch array on the network's GPU
    // data set on the device
    float *dX = nullptr, *dT = nullptr;
    int *didx = nullptr;
    long long dn = 0;
    std::vector<void *> allocs;
};

namespace bw {
int mlp_buf_alloc(MlpBuf &b, int rows) {
    int cap = (rows + 31) / 32 * 32;
    if (cap <= b.cap) return 0;
    mlp_buf_free(b);
    auto al = [](auto **p, size_t n) {
        void *q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(**p)) != hipSuccess) return 1;
        hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(**p));
        *p = (std::remove_reference_t<decltype(*p)>)q;
        return 0;
    };
    int bad = al(&b.Xb, (size_t)cap * MLP_INP) | al(&b.Y, (size_t)cap * MLP_OUTP) | al(&b.dY, (size_t)cap * MLP_OUTP) |
              al(&b.dYb, (size_t)cap * MLP_OUTP) | al(&b.dH, (size_t)cap * MLP_H) | al(&b.dZ, (size_t)cap * MLP_H) |
              al(&b.dZb, (size_t)cap * MLP_H) | al(&b.loss, (size_t)cap) |
              al(&b.part, std::max((size_t)WGRAD_SPLITS * MLP_H * MLP_H, ((size_t)cap / COLSUM_ROWS + 1) * MLP_H));
    for (int l = 0; l < 3; l++)
        bad |= al(&b.Z[l], (size_t)cap * MLP_H) | al(&b.H[l], (size_t)cap * MLP_H) | al(&b.Hb[l], (size_t)cap * MLP_H);
    hipDeviceSynchronize();
    b.cap = bad ? 0 : cap;
    return bad;
}
void mlp_buf_free(MlpBuf &b) {
    void *ps[] = {b.Xb, b.Y, b.dY, b.dYb, b.dH, b.dZ, b.dZb, b.loss, b.Z[0], b.Z[1], b.Z[2], b.H[0], b.H[1], b.H[2],
                  b.Hb[0], b.Hb[1], b.Hb[2], b.part};
    if (b.cap) hipDeviceSynchronize();
    for (void *q : ps)
        if (q) hipFree(q);
    b = MlpBuf();
}
}  // namespace bw

namespace {

int mfail(bw_mlp *m, int code, const std::string &s) {
    if (m) m->err = s;
    return code;
}
#define MCHK(m, x)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return mfail(m, BW_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int malloc0(bw_mlp *m, T **p, size_t n) {
    void *q = nullptr;
    MCHK(m, hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    MCHK(m, hipMemsetAsync(q, 0, std::max<size_t>(n, 1) * sizeof(T), m->st));
    MCHK(m, hipStreamSynchronize(m->st));
    m->allocs.push_back(q);
    *p = (T *)q;
    return BW_OK;
}

// The scratch array's device address on the current GPU, looked up once per
// device (the array's host-side name is not a device address).
int scratch_address(bw_mlp *m, unsigned char **out) {
    static std::mutex mu;
    static unsigned char *addr[256] = {};
    int dev = 0;
    MCHK(m, hipGetDevice(&dev));
    if (dev < 0 || dev >= 256) return mfail(m, BW_ERR_HIP, "device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        MCHK(m, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        addr[dev] = (unsigned char *)p;
    }
    *out = addr[dev];
    return BW_OK;
}

// MlpBuf of n_pad rows carved from the scratch array (forward buffers only)
MlpBuf scratch_buf(unsigned char *base, int n_pad) {
    MlpBuf b;
    b.cap = n_pad;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        unsigned char *p = base + o;
        o += (bytes + 255) / 256 * 256;
        return p;
    };
    b.Xb = (unsigned short *)take((size_t)n_pad * MLP_INP * 2);
    b.Y = (float *)take((size_t)n_pad * MLP_OUTP * 4);
    for (int l = 0; l < 3; l++) {
        b.Z[l] = (float *)take((size_t)n_pad * MLP_H * 4);
        b.H[l] = (float *)take((size_t)n_pad * MLP_H * 4);
        b.Hb[l] = (unsigned short *)take((size_t)n_pad * MLP_H * 2);
    }
    return b;
}

unsigned short host_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

int refresh_bf16(bw_mlp *m) {
    for (int l = 0; l < MLP_L; l++) {
        long long n = (long long)m->K[l] * m->N[l];
        hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->st, m->W[l], m->Wb[l], n);
    }
    MCHK(m, hipStreamSynchronize(m->st));
    return BW_OK;
}

int ensure_batch(bw_mlp *m, int rows) {
    if (rows <= m->tb.cap) return BW_OK;
    MCHK(m, hipStreamSynchronize(m->st));
    if (mlp_buf_alloc(m->tb, rows)) return mfail(m, BW_ERR_HIP, "hipMalloc (network activations)");
    return BW_OK;
}

template <int TA, int TB, int EPI>
void gemm(bw_mlp *m, const unsigned short *A, const unsigned short *B, int M, int N, int K, int lda, int ldb,
          float *C, int ldc, const float *bias, float *Z, unsigned short *Cb, hipStream_t st) {
    const int blocks = ((M + 63) / 64) * ((N + 63) / 64);
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), dim3(blocks), dim3(256), 0, st, A, B, M, N, K, lda, ldb, C, ldc, bias,
                       Z, Cb);
}

// Weight gradient C (M x N, ldc = N) = A^T B with A stored K x M and B K x N,
// K (the batch) long: split into up to WGRAD_SPLITS chunks of whole
// multiples of 32 rows, one product each into `part`, added in chunk order.
void gemm_wgrad(const unsigned short *A, const unsigned short *B, int M, int N, int K, float *C, float *part,
                hipStream_t st) {
    const int splits = std::max(1, std::min(WGRAD_SPLITS, K / 512));
    const int kc = (K / 32 + splits - 1) / splits * 32;
    const int blocks = ((M + 63) / 64) * ((N + 63) / 64);
    int np = 0;
    for (int k0 = 0; k0 < K; k0 += kc, np++)
        hipLaunchKernelGGL((k_gemm<1, 0, 0>), dim3(blocks), dim3(256), 0, st, A + (size_t)k0 * M, B + (size_t)k0 * N,
                           M, N, std::min(kc, K - k0), M, N, part + (size_t)np * M * N, N, (const float *)nullptr,
                           (float *)nullptr, (unsigned short *)nullptr);
    const long long n = (long long)M * N;
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, part, np, n, C);
}
void colsum(const float *g, int rows, int cols, float *out, float *part, hipStream_t st) {
    const int nb = (rows + COLSUM_ROWS - 1) / COLSUM_ROWS;
    hipLaunchKernelGGL(k_colsum_part, dim3((cols + 63) / 64, nb), dim3(256), 0, st, g, rows, cols, part);
    hipLaunchKernelGGL(k_sum_parts, dim3((cols + 255) / 256), dim3(256), 0, st, part, nb, (long long)cols, out);
}

// forward over rows [0, n) of device features x (n x 144, raw units; idx
// selects rows, may be null); output in m->Y (n_pad x 32)
void forward(bw_mlp *m, MlpBuf &B, const float *x, const int *idx, int n, int n_pad, hipStream_t st) {
    long long e = (long long)n_pad * MLP_INP;
    hipLaunchKernelGGL(k_mlp_input, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, st, x, idx, n, n_pad, m->xmean,
                       m->xistd, B.Xb);
    const unsigned short *in = B.Xb;
    for (int l = 0; l < 3; l++) {
        gemm<0, 0, 1>(m, in, m->Wb[l], n_pad, m->N[l], m->K[l], m->K[l], m->N[l], B.H[l], m->N[l], m->b[l], B.Z[l],
                      B.Hb[l], st);
        in = B.Hb[l];
    }
    gemm<0, 0, 2>(m, in, m->Wb[3], n_pad, MLP_OUTP, MLP_H, MLP_H, MLP_OUTP, B.Y, MLP_OUTP, m->b[3], nullptr, nullptr,
                  st);
}

// backward from m->dY / m->dYb: parameter gradients into gW, gb
void backward(bw_mlp *m, MlpBuf &B, int n_pad, hipStream_t st) {
    const unsigned short *dzb = B.dYb;
    const float *dz = B.dY;
    for (int l = MLP_L - 1; l >= 0; l--) {
        const unsigned short *hin = l == 0 ? B.Xb : B.Hb[l - 1];
        // gW_l (K x N) = H_{l-1}^T (K x n_pad) . dZ_l (n_pad x N)
        gemm_wgrad(hin, dzb, m->K[l], m->N[l], n_pad, m->gW[l], B.part, st);
        colsum(dz, n_pad, m->N[l], m->gb[l], B.part, st);
        if (l == 0) break;
        // dH_{l-1} (n_pad x K) = dZ_l (n_pad x N) . W_l^T (N x K)
        gemm<0, 1, 0>(m, dzb, m->Wb[l], n_pad, m->K[l], m->N[l], m->N[l], m->N[l], B.dH, m->K[l], nullptr, nullptr,
                      nullptr, st);
        long long e = (long long)n_pad * m->K[l];
        hipLaunchKernelGGL(k_silu_back, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, st, B.dH, B.Z[l - 1], B.dZ,
                           B.dZb, e);
        dz = B.dZ;
        dzb = B.dZb;
    }
}

void nll(bw_mlp *m, MlpBuf &B, const float *t, const int *idx, int n, int n_pad, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_nll, dim3((n_pad + 255) / 256), dim3(256), 0, st, B.Y, t, idx, n, n_pad, m->tmean,
                       m->tistd, B.dY, B.dYb, B.loss);
}

// parameters in the logical layout: W1 (144x256), b1, W2, b2, W3, b3, W4 (256x12), b4
long long n_params() {
    return (long long)MLP_IN * MLP_H + MLP_H + 2LL * (MLP_H * MLP_H + MLP_H) + (long long)MLP_H * MLP_OUT + MLP_OUT;
}

int copy_params(bw_mlp *m, float *out, bool grads) {
    MCHK(m, hipStreamSynchronize(m->st));
    long long o = 0;
    for (int l = 0; l < MLP_L; l++) {
        std::vector<float> w((size_t)m->K[l] * m->N[l]), bb(m->N[l]);
        MCHK(m, hipMemcpy(w.data(), grads ? m->gW[l] : m->W[l], sizeof(float) * w.size(), hipMemcpyDeviceToHost));
        MCHK(m, hipMemcpy(bb.data(), grads ? m->gb[l] : m->b[l], sizeof(float) * bb.size(), hipMemcpyDeviceToHost));
        for (int k = 0; k < m->Kl[l]; k++)
            for (int j = 0; j < m->Nl[l]; j++) out[o++] = w[(size_t)k * m->N[l] + j];
        for (int j = 0; j < m->Nl[l]; j++) out[o++] = bb[j];
    }
    return BW_OK;
}

}  // namespace

extern "C" {

BW_API int bw_mlp_create(int seed, bw_mlp **out) {
    if (!out) return BW_ERR_ARG;
    bw_mlp *m = new bw_mlp();
    *out = m;
    MCHK(m, hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
    int rc;
    if ((rc = scratch_address(m, &m->scratch))) return rc;
    for (int l = 0; l < MLP_L; l++) {
        size_t n = (size_t)m->K[l] * m->N[l];
        if ((rc = malloc0(m, &m->W[l], n)) || (rc = malloc0(m, &m->gW[l], n)) || (rc = malloc0(m, &m->mW[l], n)) ||
            (rc = malloc0(m, &m->vW[l], n)) || (rc = malloc0(m, &m->Wb[l], n)) ||
            (rc = malloc0(m, &m->b[l], m->N[l])) || (rc = malloc0(m, &m->gb[l], m->N[l])) ||
            (rc = malloc0(m, &m->mb[l], m->N[l])) || (rc = malloc0(m, &m->vb[l], m->N[l])))
            return rc;
    }
    if ((rc = malloc0(m, &m->xmean, MLP_IN)) || (rc = malloc0(m, &m->xistd, MLP_IN)) ||
        (rc = malloc0(m, &m->tmean, MLP_T)) || (rc = malloc0(m, &m->tistd, MLP_T)))
        return rc;
    // He-uniform initialisation from a splitmix64 stream (logical entries only)
    uint64_t s = (uint64_t)(uint32_t)seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    auto next = [&]() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < MLP_L; l++) {
        std::vector<float> w((size_t)m->K[l] * m->N[l], 0.0f);
        double a = std::sqrt(6.0 / m->Kl[l]) * (l == MLP_L - 1 ? 0.1 : 1.0);
        for (int k = 0; k < m->Kl[l]; k++)
            for (int j = 0; j < m->Nl[l]; j++) w[(size_t)k * m->N[l] + j] = (float)((2.0 * next() - 1.0) * a);
        MCHK(m, hipMemcpy(m->W[l], w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
    }
    std::vector<float> one(MLP_IN, 1.0f);
    MCHK(m, hipMemcpy(m->xistd, one.data(), sizeof(float) * MLP_IN, hipMemcpyHostToDevice));
    MCHK(m, hipMemcpy(m->tistd, one.data(), sizeof(float) * MLP_T, hipMemcpyHostToDevice));
    return refresh_bf16(m);
}

BW_API int bw_mlp_destroy(bw_mlp *m) {
    if (!m) return BW_OK;
    if (m->st) hipStreamSynchronize(m->st);
    for (void *q : m->allocs) hipFree(q);
    mlp_buf_free(m->tb);
    if (m->dX) { hipFree(m->dX); hipFree(m->dT); hipFree(m->didx); }
    if (m->st) hipStreamDestroy(m->st);
    delete m;
    return BW_OK;
}

BW_API const char *bw_mlp_last_error(bw_mlp *m) { return m ? m->err.c_str() : "no network"; }
BW_API long long bw_mlp_n_params(void) { return n_params(); }

BW_API int bw_mlp_get_params(bw_mlp *m, float *out) {
    if (!m || !out) return BW_ERR_ARG;
    return copy_params(m, out, false);
}

BW_API int bw_mlp_set_params(bw_mlp *m, const float *in) {
    if (!m || !in) return BW_ERR_ARG;
    MCHK(m, hipStreamSynchronize(m->st));
    long long o = 0;
    for (int l = 0; l < MLP_L; l++) {
        std::vector<float> w((size_t)m->K[l] * m->N[l], 0.0f), bb(m->N[l], 0.0f);
        for (int k = 0; k < m->Kl[l]; k++)
            for (int j = 0; j < m->Nl[l]; j++) w[(size_t)k * m->N[l] + j] = in[o++];
        for (int j = 0; j < m->Nl[l]; j++) bb[j] = in[o++];
        MCHK(m, hipMemcpy(m->W[l], w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
        MCHK(m, hipMemcpy(m->b[l], bb.data(), sizeof(float) * bb.size(), hipMemcpyHostToDevice));
    }
    return refresh_bf16(m);
}

// standardisation: x' = (x - x_mean) / x_std (144 each), t' = (t - t_mean) / t_std (6 each)
BW_API int bw_mlp_set_normalisation(bw_mlp *m, const float *x_mean, const float *x_std, const float *t_mean,
                                    const float *t_std) {
    if (!m || !x_mean || !x_std || !t_mean || !t_std) return BW_ERR_ARG;
    std::vector<float> xi(MLP_IN), ti(MLP_T);
    for (int k = 0; k < MLP_IN; k++) xi[k] = x_std[k] > 0 ? 1.0f / x_std[k] : 0.0f;
    for (int k = 0; k < MLP_T; k++) ti[k] = t_std[k] > 0 ? 1.0f / t_std[k] : 0.0f;
    MCHK(m, hipStreamSynchronize(m->st));
    MCHK(m, hipMemcpy(m->xmean, x_mean, sizeof(float) * MLP_IN, hipMemcpyHostToDevice));
    MCHK(m, hipMemcpy(m->xistd, xi.data(), sizeof(float) * MLP_IN, hipMemcpyHostToDevice));
    MCHK(m, hipMemcpy(m->tmean, t_mean, sizeof(float) * MLP_T, hipMemcpyHostToDevice));
    MCHK(m, hipMemcpy(m->tistd, ti.data(), sizeof(float) * MLP_T, hipMemcpyHostToDevice));
    return BW_OK;
}

// y (n x 12): standardised mean (6) and log-variance (6) for host features x (n x 144)
BW_API int bw_mlp_forward(bw_mlp *m, const float *x, int n, float *y) {
    if (!m || !x || !y || n < 1) return BW_ERR_ARG;
    int n_pad = (n + 31) / 32 * 32, rc;
    if ((rc = ensure_batch(m, n_pad))) return rc;
    float *dx = nullptr;
    MCHK(m, hipMalloc(&dx, sizeof(float) * (size_t)n * MLP_IN));
    MCHK(m, hipMemcpy(dx, x, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice));
    forward(m, m->tb, dx, nullptr, n, n_pad, m->st);
    MCHK(m, hipStreamSynchronize(m->st));
    std::vector<float> yy((size_t)n_pad * MLP_OUTP);
    MCHK(m, hipMemcpy(yy.data(), m->tb.Y, sizeof(float) * yy.size(), hipMemcpyDeviceToHost));
    hipFree(dx);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < MLP_OUT; k++) y[(size_t)i * MLP_OUT + k] = yy[(size_t)i * MLP_OUTP + k];
    return BW_OK;
}

// bw_mlp_forward on device buffers with no allocation: the activations live
// in the scratch array (n <= NOALLOC_ROWS).  Every buffer it reads is written
// first in the same call, so nothing needs clearing; calls that overlap on one
// GPU share the array and race.
BW_API int bw_mlp_forward_noalloc(bw_mlp *m, const float *x, int n, float *y) {
    if (!m || !x || !y || n < 1) return BW_ERR_ARG;
    if (n > NOALLOC_ROWS)
        return mfail(m, BW_ERR_ARG, "bw_mlp_forward_noalloc: more rows than the scratch array holds (1024)");
    if (!m->scratch) return mfail(m, BW_ERR_STATE, "bw_mlp_forward_noalloc: no scratch address");
    const int n_pad = (n + 31) / 32 * 32;
    MlpBuf B = scratch_buf(m->scratch, n_pad);
    forward(m, B, x, nullptr, n, n_pad, m->st);
    hipLaunchKernelGGL(k_mlp_rows, dim3((n * MLP_OUT + 255) / 256), dim3(256), 0, m->st, B.Y, n, y);
    MCHK(m, hipGetLastError());
    MCHK(m, hipStreamSynchronize(m->st));
    return BW_OK;
}

// mean loss over n samples and its gradient in the parameter layout
BW_API int bw_mlp_loss_grad(bw_mlp *m, const float *x, const float *t, int n, double *loss, float *grad) {
    if (!m || !x || !t || n < 1) return BW_ERR_ARG;
    int n_pad = (n + 31) / 32 * 32, rc;
    if ((rc = ensure_batch(m, n_pad))) return rc;
    float *dx = nullptr, *dt = nullptr;
    MCHK(m, hipMalloc(&dx, sizeof(float) * (size_t)n * MLP_IN));
    MCHK(m, hipMalloc(&dt, sizeof(float) * (size_t)n * MLP_T));
    MCHK(m, hipMemcpy(dx, x, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice));
    MCHK(m, hipMemcpy(dt, t, sizeof(float) * (size_t)n * MLP_T, hipMemcpyHostToDevice));
    forward(m, m->tb, dx, nullptr, n, n_pad, m->st);
    nll(m, m->tb, dt, nullptr, n, n_pad, m->st);
    backward(m, m->tb, n_pad, m->st);
    MCHK(m, hipStreamSynchronize(m->st));
    if (loss) {
        std::vector<float> l(n_pad);
        MCHK(m, hipMemcpy(l.data(), m->tb.loss, sizeof(float) * n_pad, hipMemcpyDeviceToHost));
        double s = 0;
        for (int i = 0; i < n; i++) s += l[i];
        *loss = s / n;
    }
    hipFree(dx);
    hipFree(dt);
    if (grad) return copy_params(m, grad, true);
    return BW_OK;
}

// Adam training on n samples (x n x 144 raw features, t n x 6 raw targets):
// epochs over shuffled mini-batches; loss_out[e] = mean loss of epoch e
BW_API int bw_mlp_train(bw_mlp *m, const float *x, const float *t, int n, int epochs, int batch, float lr,
                        int seed, double *loss_out) {
    if (!m || !x || !t || n < 1 || epochs < 1 || batch < 32) return BW_ERR_ARG;
    batch = batch / 32 * 32;
    int rc;
    if ((rc = ensure_batch(m, batch))) return rc;
    MCHK(m, hipStreamSynchronize(m->st));
    if (m->dn < n) {
        if (m->dX) { hipFree(m->dX); hipFree(m->dT); hipFree(m->didx); }
        MCHK(m, hipMalloc(&m->dX, sizeof(float) * (size_t)n * MLP_IN));
        MCHK(m, hipMalloc(&m->dT, sizeof(float) * (size_t)n * MLP_T));
        MCHK(m, hipMalloc(&m->didx, sizeof(int) * (size_t)n));
        m->dn = n;
    }
    MCHK(m, hipMemcpy(m->dX, x, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice));
    MCHK(m, hipMemcpy(m->dT, t, sizeof(float) * (size_t)n * MLP_T, hipMemcpyHostToDevice));
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    uint64_t s = (uint64_t)(uint32_t)seed * 0x9E3779B97F4A7C15ull + 1;
    std::vector<float> l(batch);
    for (int e = 0; e < epochs; e++) {
        for (int i = n - 1; i > 0; i--) {  // Fisher-Yates with splitmix64
            s += 0x9E3779B97F4A7C15ull;
            uint64_t z = s;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            int j = (int)(z % (uint64_t)(i + 1));
            std::swap(perm[i], perm[j]);
        }
        MCHK(m, hipMemcpy(m->didx, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
        double tot = 0;
        for (int b0 = 0; b0 + batch <= n || b0 == 0; b0 += batch) {
            const int nb = std::min(batch, n - b0);
            const int n_pad = (nb + 31) / 32 * 32;
            forward(m, m->tb, m->dX, m->didx + b0, nb, n_pad, m->st);
            nll(m, m->tb, m->dT, m->didx + b0, nb, n_pad, m->st);
            backward(m, m->tb, n_pad, m->st);
            m->adam_t++;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.0f - (float)std::pow(b1, (double)m->adam_t), c2 = 1.0f - (float)std::pow(b2, (double)m->adam_t);
            for (int q = 0; q < MLP_L; q++) {
                long long nw = (long long)m->K[q] * m->N[q];
                hipLaunchKernelGGL(k_adam, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, m->st, m->W[q], m->gW[q],
                                   m->mW[q], m->vW[q], m->Wb[q], nw, lr, b1, b2, c1, c2, 1e-8f);
                hipLaunchKernelGGL(k_adam, dim3((m->N[q] + 255) / 256), dim3(256), 0, m->st, m->b[q], m->gb[q],
                                   m->mb[q], m->vb[q], (unsigned short *)nullptr, (long long)m->N[q], lr, b1, b2, c1,
                                   c2, 1e-8f);
            }
            MCHK(m, hipMemcpyAsync(l.data(), m->tb.loss, sizeof(float) * nb, hipMemcpyDeviceToHost, m->st));
            MCHK(m, hipStreamSynchronize(m->st));
            for (int i = 0; i < nb; i++) tot += l[i];
            if (nb < batch) break;
        }
        if (loss_out) loss_out[e] = tot / n;
    }
    return BW_OK;
}

}  // extern "C"

namespace bw {
// features (n x 144 raw, device) -> y_out (n x 12): mean (6) and log-variance
// (6) converted back to physical units (mean = t_mean + t_std mu',
// log-variance = s' + 2 log t_std)
__global__ void k_mlp_unscale(const float *__restrict__ y, int n, const float *__restrict__ tmean,
                              const float *__restrict__ tistd, float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < MLP_T; k++) {
        const float sd = tistd[k] > 0 ? 1.0f / tistd[k] : 0.0f;
        out[(size_t)i * MLP_OUT + k] = tmean[k] + sd * y[(size_t)i * MLP_OUTP + k];
        const float s = fminf(fmaxf(y[(size_t)i * MLP_OUTP + MLP_T + k], -20.0f), 20.0f);
        out[(size_t)i * MLP_OUT + MLP_T + k] = s + 2.0f * logf(fmaxf(sd, 1e-30f));
    }
}

void mlp_forward_device(bw_mlp *m, MlpBuf &B, const float *x, int n, float *y_out, hipStream_t st) {
    const int n_pad = (n + 31) / 32 * 32;
    forward(m, B, x, nullptr, n, n_pad, st);
    hipLaunchKernelGGL(k_mlp_unscale, dim3((n + 255) / 256), dim3(256), 0, st, B.Y, n, m->tmean, m->tistd, y_out);
}
}  // namespace bw
