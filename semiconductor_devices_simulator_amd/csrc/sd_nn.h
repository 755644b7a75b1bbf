// sd_nn.h - launchers of the warm-start network kernels (k_nn.hip) and the
// network's host API used by the engine (nn.cpp).
#pragma once
#include "sd_common.h"

void nn_conv_fwd(hipStream_t s, int G, const __bf16 *in, int K, const __bf16 *W, const float *bias, int N,
                 __bf16 *out, float *outf, int relu);
void nn_conv_wgrad(hipStream_t s, int G, const __bf16 *dy, int N, const __bf16 *in, int K, int chunk, int nchunk,
                   float *part, float *pb);
// the chunks' partials summed in order into the real [27][Nr][Kr] block g
void nn_wgrad_reduce(hipStream_t s, const float *part, int nchunk, int N, int K, int Nr, int Kr, float *g);
// bias gradient of the Nr real outputs into gb from the nchunk x N partials pb
// that nn_conv_wgrad wrote
void nn_bgrad(hipStream_t s, const float *pb, int nchunk, int N, int Nr, float *gb);
void nn_pack_w(hipStream_t s, const float *W32, int Nr, int Kr, int N, int K, int Kb, __bf16 *Wf, __bf16 *Wb);
void nn_adam(hipStream_t s, int n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2,
             float eps, int step);
void nn_pool(hipStream_t s, int Gc, int C, const __bf16 *in, __bf16 *out);
void nn_pool_bwd(hipStream_t s, int G, int C, const __bf16 *dy, __bf16 *dx, int add);
void nn_up_add(hipStream_t s, int G, int C, const __bf16 *in, const __bf16 *skip, __bf16 *out);
void nn_up_bwd(hipStream_t s, int Gc, int C, const __bf16 *dy, __bf16 *dx);
void nn_relu_bwd(hipStream_t s, int n, const __bf16 *y, __bf16 *g, const __bf16 *add);
void nn_mse(hipStream_t s, int V, const float *out, int N, const float *tgt, int C3, __bf16 *dy, float *part,
            int nblk);
void nn_mesh_to_grid(hipStream_t s, int G, int Nx, int Ny, int Nz, const double *psi, const double *phin,
                     const double *phip, const double *Nd, const double *Na, float dv, const double *h0,
                     const double *h1, const double *h2, int C, __bf16 *in);
void nn_target_to_grid(hipStream_t s, int G, int Nx, int Ny, int Nz, const double *a0, const double *a1,
                       const double *a2, const double *b0, const double *b1, const double *b2, float iscale,
                       float *tgt);
void nn_grid_to_mesh_add(hipStream_t s, int G, int Nx, int Ny, int Nz, const float *out, int N, float scale,
                         const unsigned char *kind, const int *con_mask, double *psi, double *phin, double *phip,
                         int chmask, float thr);
void nn_planar_to_cl(hipStream_t s, int V, int Cr, int C, const float *src, __bf16 *dst);

// ---- network host API (nn.cpp), used by the engine's warm-start hook
#define NN_IN 8                // input channels used
#define NN_CIN 16              // padded input channels
#define NN_COUT 32             // padded output channels (3 used)
bool nn_ready();               // a network exists
int nn_grid();                 // its grid size G
// input of the current sample (channel-last BF16, V x NN_CIN) and the
// inference output (FP32, V x NN_COUT); both on the device
__bf16 *nn_input_buf();
const float *nn_infer(hipStream_t s);          // forward on nn_input_buf()
float nn_target_scale();
// append one training sample: nn_input_buf() and the target built by fill()
int nn_capture(hipStream_t s, void (*fill)(hipStream_t, float *tgt, void *ctx), void *ctx);
