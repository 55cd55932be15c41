// nn.hpp -- kernels of the warm-start network (nn.hip).
#pragma once

#include "st_common.hpp"

namespace st {
namespace nn {

__host__ __device__ inline unsigned short f2bf(float f) {  // round to nearest even
#ifdef __HIP_DEVICE_COMPILE__
    unsigned u = __float_as_uint(f);
#else
    unsigned u;
    __builtin_memcpy(&u, &f, 4);
#endif
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ inline float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

struct GemmArgs {
    const unsigned short *A;
    const unsigned short *B;
    float *C;
    unsigned short *C16;
    const float *bias;
    long long M;
    int N, K;
    long long lda, ldb, ldc;
    int ta, tb, relu, mode, ksplit;
};

void gemm(hipStream_t st, const GemmArgs &g);
void im2col(hipStream_t st, const float *X, unsigned short *col, int ns, int n, int C, int ksz);
void col2im(hipStream_t st, const float *dcol, float *dX, int ns, int n, int C, int ksz);
void pool(hipStream_t st, const float *X, float *Y, int ns, int n, int C);
void pool_bf16(hipStream_t st, const unsigned short *X, unsigned short *Y, int ns, int n, int C);
void pool_bwd(hipStream_t st, const float *dY, float *dX, int ns, int n, int C);
void upsample(hipStream_t st, const float *X, float *Y, int ns, int m, int C, int Ctot, int off);
void upsample_bwd(hipStream_t st, const float *dY, float *dX, int ns, int m, int C, int Ctot, int off);
void chancopy(hipStream_t st, const float *X, float *Y, long long rows, int C, int Ctot, int off,
              bool backward);
void relu_backward(hipStream_t st, float *dY, const float *Y, unsigned short *d16, long long n,
                   bool relu);
void bias_gradient(hipStream_t st, const float *dY, float *db, long long rows, int C);
void mse_loss(hipStream_t st, const float *p, const float *t, const float *m, float *grad,
              double *acc, long long n, float under_weight);
void adam_step(hipStream_t st, float *w, const float *g, float *m, float *v, unsigned short *w16,
               long long n, float lr, int step);
void bf16_copy(hipStream_t st, const float *x, unsigned short *y, long long n);

// implicit-GEMM convolution forward (conv3d_fwd) and its weight packing
int conv_cinp(int cin);
int conv_coutp(int cout);
long long conv_packed_size(int ksz, int cin, int cout);
void pack_conv_weights(hipStream_t st, const unsigned short *W16, unsigned short *Wp, int ksz,
                       int cin, int cout);
void conv_forward(hipStream_t st, const float *X, const unsigned short *Wp, const float *bias,
                  float *Y, int ns, int e, int cin, int cout, int ksz, bool relu);
// the same with the input the concatenation [X0 (C0 channels, nearest-
// upsampled from half resolution when up0) | X1 (C1 channels)]
void conv_forward2(hipStream_t st, const float *X0, int C0, bool up0, const float *X1, int C1,
                   const unsigned short *Wp, const float *bias, float *Y, int ns, int e, int cout,
                   int ksz, bool relu);
// the same with FP32 or BF16 (in_bf16, out_bf16) activations
void conv_forward_any(hipStream_t st, const void *X0, int C0, bool up0, const void *X1, int C1,
                      bool in_bf16, const unsigned short *Wp, const float *bias, void *Y,
                      bool out_bf16, int ns, int e, int cout, int ksz, bool relu);

}  // namespace nn
}  // namespace st
