// k_nn.hip - kernels of the warm-start network: 3x3x3 convolutions as
// implicit GEMMs on the BF16 matrix cores (v_mfma_f32_32x32x16_bf16, FP32
// accumulation), their data and weight gradients, pooling, upsampling,
// resampling between the device mesh and the network grid, loss and Adam.
//
// Activations are channel-last: voxel v = x + G (y + G z) of a G^3 grid holds
// C contiguous BF16 channels, C a multiple of 16 (one MFMA K-step).  Weights
// of a convolution with K input and N output channels are stored per tap t
// (t = dx+1 + 3 (dy+1) + 9 (dz+1)) as W[t][n][k] (BF16), so that a lane reads
// its eight K values as one 16-byte load.
//
// MFMA 32x32x16 operand layout (64 lanes, lane = l): A row l%32, K elements
// 8 (l/32) .. 8 (l/32) + 7; B column l%32, the same K elements; result
// register r of lane l holds row 8 (r/4) + 4 (l/32) + r%4, column l%32.
#include "sd_nn.h"

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ int nn_off(int G, int t, int x, int y, int z, int &u) {
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    const int xx = x + dx, yy = y + dy, zz = z + dz;
    const bool ok = xx >= 0 && xx < G && yy >= 0 && yy < G && zz >= 0 && zz < G;
    u = ok ? xx + G * (yy + G * zz) : 0;
    return ok;
}

// ---------------------------------------------------------------- forward
// out[v][n] = bias[n] + sum_t sum_k W[t][n][k] in[v + off(t)][k], zero padded.
// One wavefront computes 64 voxels (two 32-voxel row tiles) x all N (NT =
// N/32 column tiles, 1 or 2); four per workgroup.  Per tap and 16-channel
// step a lane loads its two activation operands once and each weight operand
// once, and issues 2 NT MFMAs: every operand load feeds two or more MFMAs.
// relu: out = max(out, 0).  outf (optional): the same values in FP32.
template <int NT>
__global__ void __launch_bounds__(256) k_conv_fwd(int G, const __bf16 *__restrict__ in, int K,
                                                  const __bf16 *__restrict__ W, const float *__restrict__ bias, int N,
                                                  __bf16 *__restrict__ out, float *__restrict__ outf, int relu) {
    const int V = G * G * G;
    const int lane = threadIdx.x & 63, kh = lane >> 5, row = lane & 31;
    const int v0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
    if (v0 >= V) return;                       // no barriers below
    int x[2], y[2], z[2];
    bool vin[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int v = v0 + 32 * h + row;
        vin[h] = v < V;                        // rows past the grid read zeros
        x[h] = v % G; y[h] = (v / G) % G; z[h] = v / (G * G);
    }
    const v8bf zero8 = {};
    v16f acc[2][NT];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) acc[h][nt] = v16f{};
    for (int t = 0; t < 27; t++) {
        int u[2], ok[2];
#pragma unroll
        for (int h = 0; h < 2; h++) ok[h] = vin[h] && nn_off(G, t, x[h], y[h], z[h], u[h]);
        const __bf16 *ip0 = in + (size_t)u[0] * K + kh * 8, *ip1 = in + (size_t)u[1] * K + kh * 8;
        const __bf16 *wp = W + ((size_t)(t * N + row)) * K + kh * 8;
        for (int kc = 0; kc < K; kc += 16) {
            const v8bf a0 = ok[0] ? *(const v8bf *)(ip0 + kc) : zero8;
            const v8bf a1 = ok[1] ? *(const v8bf *)(ip1 + kc) : zero8;
#pragma unroll
            for (int nt = 0; nt < NT; nt++) {
                const v8bf b = *(const v8bf *)(wp + (size_t)nt * 32 * K + kc);
                acc[0][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b, acc[0][nt], 0, 0, 0);
                acc[1][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b, acc[1][nt], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
        const int n = nt * 32 + row;
        const float bn = bias ? bias[n] : 0.f;
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int i = 32 * h + 8 * (r >> 2) + 4 * kh + (r & 3);
                if (v0 + i >= V) continue;         // grids of fewer than 64 voxels
                float val = acc[h][nt][r] + bn;
                if (relu) val = fmaxf(val, 0.f);
                out[(size_t)(v0 + i) * N + n] = (__bf16)val;
                if (outf) outf[(size_t)(v0 + i) * N + n] = val;
            }
    }
}
void nn_conv_fwd(hipStream_t s, int G, const __bf16 *in, int K, const __bf16 *W, const float *bias, int N,
                 __bf16 *out, float *outf, int relu) {
    const int V = G * G * G, nb = ((V + 63) / 64 + 3) / 4;
    if (N == 64) k_conv_fwd<2><<<nb, 256, 0, s>>>(G, in, K, W, bias, N, out, outf, relu);
    else if (N == 32) k_conv_fwd<1><<<nb, 256, 0, s>>>(G, in, K, W, bias, N, out, outf, relu);
    else fprintf(stderr, "semisim-hip nn: convolution with %d outputs (32 or 64 supported)\n", N);
}

// ---------------------------------------------------------------- weight gradient
// part[c][t][n][k] = sum over the voxels of chunk c of dy[v][n] in[v + off(t)][k].
// As a GEMM the sum runs over voxels, but activations are channel-last, so a
// lane's MFMA operand (one channel, eight voxels) is strided in memory.  A
// workgroup (four wavefronts, one chunk and one tap t) therefore stages
// 128-voxel blocks of dy and of the shifted input with coalesced 16-byte
// loads and writes them to LDS transposed (channel-major, padded rows), where
// each lane reads its eight voxels as one 16-byte load.  Output tiles of 32 x
// 32 (N/32 x K/32 of them, at most four) are spread over the wavefronts; where
// fewer tiles than wavefronts, the wavefronts of a tile split the block's
// 16-voxel groups and sum their accumulators through LDS in a fixed order.
// K = 16 (the input layer): the rows past K are zero and are not stored.
// pb (optional): the centre tap's workgroups also sum dy over their chunk per
// channel from the staged rows, the bias gradient's partials pb[c][n].
#define WG_VB 128
#define WG_PITCH (WG_VB + 8)
__global__ void __launch_bounds__(256) k_conv_wgrad(int G, const __bf16 *__restrict__ dy, int N,
                                                    const __bf16 *__restrict__ in, int K, int chunk,
                                                    float *__restrict__ part, float *__restrict__ pb) {
    extern __shared__ __align__(16) unsigned char wg_sm[];
    __bf16 *s_dy = (__bf16 *)wg_sm;
    const int Kp = (K + 31) / 32 * 32;
    __bf16 *s_in = s_dy + (size_t)N * WG_PITCH;
    const int V = G * G * G;
    const int t = blockIdx.x % 27, c = blockIdx.x / 27;
    const int vb = c * chunk, ve = min(V, vb + chunk);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, kh = lane >> 5, row = lane & 31;
    const int nNt = N / 32, nKt = Kp / 32, ntiles = nNt * nKt;       // 1, 2 or 4
    const int per = 4 / ntiles;                                       // wavefronts per tile
    const int tile = wave % ntiles, phase = wave / ntiles;
    const int nt = tile / nKt, kt = tile % nKt;
    const v8bf zero8 = {};
    v16f acc = {};
    const bool bias = pb && t == 13 && (int)threadIdx.x < N;
    float bsum = 0.f;
    const int dvec = N / 8, kvec = Kp / 8;
    for (int v0 = vb; v0 < ve; v0 += WG_VB) {
        __syncthreads();                               // the previous block's operands are read
        for (int q = threadIdx.x; q < WG_VB * dvec; q += 256) {
            const int vl = q / dvec, cg = q - vl * dvec, v = v0 + vl;
            const v8bf val = v < ve ? *(const v8bf *)(dy + (size_t)v * N + cg * 8) : zero8;
#pragma unroll
            for (int i = 0; i < 8; i++) s_dy[(cg * 8 + i) * WG_PITCH + vl] = val[i];
        }
        for (int q = threadIdx.x; q < WG_VB * kvec; q += 256) {
            const int vl = q / kvec, cg = q - vl * kvec, v = v0 + vl;
            v8bf val = zero8;
            if (v < ve && cg * 8 < K) {
                const int x = v % G, y = (v / G) % G, z = v / (G * G);
                int u;
                if (nn_off(G, t, x, y, z, u)) val = *(const v8bf *)(in + (size_t)u * K + cg * 8);
            }
#pragma unroll
            for (int i = 0; i < 8; i++) s_in[(cg * 8 + i) * WG_PITCH + vl] = val[i];
        }
        __syncthreads();
        if (bias)
            for (int vl = 0; vl < WG_VB; vl++) bsum += (float)s_dy[threadIdx.x * WG_PITCH + vl];
        for (int g = phase; g < WG_VB / 16; g += per) {
            const v8bf a = *(const v8bf *)(s_dy + (nt * 32 + row) * WG_PITCH + g * 16 + kh * 8);
            const v8bf b = *(const v8bf *)(s_in + (kt * 32 + row) * WG_PITCH + g * 16 + kh * 8);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
    }
    if (bias) pb[(size_t)c * N + threadIdx.x] = bsum;
    // wavefronts that share a tile: sum in wavefront order through LDS
    if (per > 1) {
        __syncthreads();
        float *s_acc = (float *)wg_sm;                 // 4 x 16 x 64 floats
#pragma unroll
        for (int r = 0; r < 16; r++) s_acc[(wave * 16 + r) * 64 + lane] = acc[r];
        __syncthreads();
        if (phase != 0) return;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            float sum = 0.f;
            for (int p = 0; p < per; p++) sum += s_acc[((tile + p * ntiles) * 16 + r) * 64 + lane];
            acc[r] = sum;
        }
    }
    const int k = kt * 32 + row;
    const size_t base = ((size_t)c * 27 + t) * N;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int i = nt * 32 + 8 * (r >> 2) + 4 * kh + (r & 3);      // row n
        if (k < K) part[(base + i) * K + k] = acc[r];
    }
}
void nn_conv_wgrad(hipStream_t s, int G, const __bf16 *dy, int N, const __bf16 *in, int K, int chunk, int nchunk,
                   float *part, float *pb) {
    const int Kp = (K + 31) / 32 * 32;
    size_t lds = (size_t)(N + Kp) * WG_PITCH * sizeof(__bf16);
    if (lds < 4 * 16 * 64 * sizeof(float)) lds = 4 * 16 * 64 * sizeof(float);
    k_conv_wgrad<<<nchunk * 27, 256, lds, s>>>(G, dy, N, in, K, chunk, part, pb);
}
// g = sum_c part[c] (chunks in order) over the padded [27][N][K] layout,
// stored as the real [27][Nr][Kr] block of the gradient vector
__global__ void k_wgrad_reduce(const float *part, int nchunk, int N, int K, int Nr, int Kr, float *g) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int n = 27 * N * K;
    if (i >= n) return;
    const int k = i % K, nn = (i / K) % N, t = i / (K * N);
    if (nn >= Nr || k >= Kr) return;
    float s = 0.f;
    for (int c = 0; c < nchunk; c++) s += part[(size_t)c * n + i];
    g[((size_t)t * Nr + nn) * Kr + k] = s;
}
void nn_wgrad_reduce(hipStream_t s, const float *part, int nchunk, int N, int K, int Nr, int Kr, float *g) {
    const int n = 27 * N * K;
    k_wgrad_reduce<<<(n + 255) / 256, 256, 0, s>>>(part, nchunk, N, K, Nr, Kr, g);
}
// bias gradient gb[n] = sum over the chunks (in order) of the partials pb[c][n]
// the weight-gradient kernel's centre tap wrote, n < Nr
__global__ void k_bgrad_sum(const float *pb, int nchunk, int N, int Nr, float *gb) {
    const int n = threadIdx.x;
    if (n >= Nr) return;
    float s = 0.f;
    for (int c = 0; c < nchunk; c++) s += pb[(size_t)c * N + n];
    gb[n] = s;
}
void nn_bgrad(hipStream_t s, const float *pb, int nchunk, int N, int Nr, float *gb) {
    k_bgrad_sum<<<1, 64, 0, s>>>(pb, nchunk, N, Nr, gb);
}

// ---------------------------------------------------------------- weights
// BF16 forward weights from the FP32 master copy W32[t][n][k] (k < Kr real
// input channels, padded to K with zeros; n < Nr real outputs, padded to N),
// and the flipped, transposed copy for the data gradient:
// WB[26 - t][k][n] = W[t][n][k] (rows k padded to Kb, a multiple of 32).
__global__ void k_pack_w(const float *W32, int Nr, int Kr, int N, int K, int Kb, __bf16 *Wf, __bf16 *Wb) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 27 * N * Kb) return;
    // forward layout index
    if (i < 27 * N * K) {
        const int k = i % K, n = (i / K) % N, t = i / (K * N);
        const float w = (n < Nr && k < Kr) ? W32[((size_t)t * Nr + n) * Kr + k] : 0.f;
        Wf[i] = (__bf16)w;
    }
    if (Wb) {
        const int n = i % N, k = (i / N) % Kb, tb = i / (N * Kb);
        const int t = 26 - tb;
        const float w = (n < Nr && k < Kr) ? W32[((size_t)t * Nr + n) * Kr + k] : 0.f;
        Wb[i] = (__bf16)w;
    }
}
void nn_pack_w(hipStream_t s, const float *W32, int Nr, int Kr, int N, int K, int Kb, __bf16 *Wf, __bf16 *Wb) {
    const int n = 27 * N * (Kb > K ? Kb : K);
    k_pack_w<<<(n + 255) / 256, 256, 0, s>>>(W32, Nr, Kr, N, K, Kb, Wf, Wb);
}

// Adam on n FP32 parameters with gradient g (step count t >= 1)
__global__ void k_adam(int n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi; v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
void nn_adam(hipStream_t s, int n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2,
             float eps, int step) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    k_adam<<<(n + 255) / 256, 256, 0, s>>>(n, p, g, m, v, lr, b1, b2, eps, c1, c2);
}

// ---------------------------------------------------------------- elementwise
// 2x2x2 average pooling (G -> G/2), C channels
__global__ void k_pool(int Gc, int C, const __bf16 *in, __bf16 *out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int Vc = Gc * Gc * Gc;
    if (i >= Vc * C) return;
    const int c = i % C, vc = i / C, X = vc % Gc, Y = (vc / Gc) % Gc, Z = vc / (Gc * Gc), G = 2 * Gc;
    float s = 0.f;
    for (int d = 0; d < 8; d++) {
        const int x = 2 * X + (d & 1), y = 2 * Y + ((d >> 1) & 1), z = 2 * Z + (d >> 2);
        s += (float)in[((size_t)x + G * (y + (size_t)G * z)) * C + c];
    }
    out[i] = (__bf16)(0.125f * s);
}
void nn_pool(hipStream_t s, int Gc, int C, const __bf16 *in, __bf16 *out) {
    const int n = Gc * Gc * Gc * C;
    k_pool<<<(n + 255) / 256, 256, 0, s>>>(Gc, C, in, out);
}
// pooling backward: dx[child] = dy[parent] / 8  (mode 0: store, 1: add)
__global__ void k_pool_bwd(int G, int C, const __bf16 *dy, __bf16 *dx, int add) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int V = G * G * G;
    if (i >= V * C) return;
    const int c = i % C, v = i / C, x = v % G, y = (v / G) % G, z = v / (G * G), Gc = G / 2;
    const float g = 0.125f * (float)dy[((size_t)(x >> 1) + Gc * ((y >> 1) + (size_t)Gc * (z >> 1))) * C + c];
    dx[i] = (__bf16)(add ? (float)dx[i] + g : g);
}
void nn_pool_bwd(hipStream_t s, int G, int C, const __bf16 *dy, __bf16 *dx, int add) {
    const int n = G * G * G * C;
    k_pool_bwd<<<(n + 255) / 256, 256, 0, s>>>(G, C, dy, dx, add);
}
// nearest upsampling (G/2 -> G) plus skip: out = up(in) + skip
__global__ void k_up_add(int G, int C, const __bf16 *in, const __bf16 *skip, __bf16 *out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int V = G * G * G;
    if (i >= V * C) return;
    const int c = i % C, v = i / C, x = v % G, y = (v / G) % G, z = v / (G * G), Gc = G / 2;
    const float a = (float)in[((size_t)(x >> 1) + Gc * ((y >> 1) + (size_t)Gc * (z >> 1))) * C + c];
    out[i] = (__bf16)(a + (float)skip[i]);
}
void nn_up_add(hipStream_t s, int G, int C, const __bf16 *in, const __bf16 *skip, __bf16 *out) {
    const int n = G * G * G * C;
    k_up_add<<<(n + 255) / 256, 256, 0, s>>>(G, C, in, skip, out);
}
// upsampling backward: dx[parent] = sum of the 8 children's dy
__global__ void k_up_bwd(int Gc, int C, const __bf16 *dy, __bf16 *dx) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int Vc = Gc * Gc * Gc;
    if (i >= Vc * C) return;
    const int c = i % C, vc = i / C, X = vc % Gc, Y = (vc / Gc) % Gc, Z = vc / (Gc * Gc), G = 2 * Gc;
    float s = 0.f;
    for (int d = 0; d < 8; d++) {
        const int x = 2 * X + (d & 1), y = 2 * Y + ((d >> 1) & 1), z = 2 * Z + (d >> 2);
        s += (float)dy[((size_t)x + G * (y + (size_t)G * z)) * C + c];
    }
    dx[i] = (__bf16)s;
}
void nn_up_bwd(hipStream_t s, int Gc, int C, const __bf16 *dy, __bf16 *dx) {
    const int n = Gc * Gc * Gc * C;
    k_up_bwd<<<(n + 255) / 256, 256, 0, s>>>(Gc, C, dy, dx);
}
// g = g * (y > 0) (ReLU backward through the stored output y); add: g += a
__global__ void k_relu_bwd(int n, const __bf16 *y, __bf16 *g, const __bf16 *a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float gi = (float)g[i];
    if (a) gi += (float)a[i];
    g[i] = (__bf16)((float)y[i] > 0.f ? gi : 0.f);
}
void nn_relu_bwd(hipStream_t s, int n, const __bf16 *y, __bf16 *g, const __bf16 *add) {
    k_relu_bwd<<<(n + 255) / 256, 256, 0, s>>>(n, y, g, add);
}
// MSE over the first C3 of N channels: loss partials per workgroup and the
// output gradient dy = 2 (out - tgt) / (V C3) (BF16, padded channels 0)
__global__ void __launch_bounds__(256) k_mse(int V, const float *out, int N, const float *tgt, int C3, __bf16 *dy,
                                             float *part) {
    __shared__ float red[256];
    float s = 0.f;
    const float sc = 2.f / ((float)V * C3);
    for (int i = blockIdx.x * 256 + threadIdx.x; i < V * N; i += gridDim.x * 256) {
        const int c = i % N, v = i / N;
        float d = 0.f;
        if (c < C3) {
            d = out[i] - tgt[(size_t)v * C3 + c];
            s += d * d;
        }
        dy[i] = (__bf16)(sc * d);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
void nn_mse(hipStream_t s, int V, const float *out, int N, const float *tgt, int C3, __bf16 *dy, float *part,
            int nblk) {
    k_mse<<<nblk, 256, 0, s>>>(V, out, N, tgt, C3, dy, part);
}

// ---------------------------------------------------------------- resampling
// Device mesh <-> network grid in index space: voxel (a, b, c) of the G^3 grid
// sits at mesh index coordinates ((a + 0.5) Nx / G - 0.5, ...), read by
// trilinear interpolation of the node values; a mesh node reads the grid at
// (i + 0.5) G / Nx - 0.5 the same way (clamped at the borders).
__device__ __forceinline__ void nn_lin(float p, int n, int &i0, int &i1, float &w) {
    p = fminf(fmaxf(p, 0.f), (float)(n - 1));
    i0 = (int)floorf(p);
    if (i0 > n - 2) i0 = n > 1 ? n - 2 : 0;
    i1 = n > 1 ? i0 + 1 : 0;
    w = n > 1 ? p - (float)i0 : 0.f;
}
// network input channels of one sample (C channels, channel-last BF16):
// 0 psi, 1 phi_n, 2 phi_p (V), 3 signed doping log-magnitude, 4 bias step
// (units of 50 mV), 5-7 (psi, phi_n, phi_p) - h (the last converged state),
// in units of 50 mV (0 without h)
__global__ void k_mesh_to_grid(int G, int Nx, int Ny, int Nz, const double *psi, const double *phin,
                               const double *phip, const double *Nd, const double *Na, float dv, const double *h0,
                               const double *h1, const double *h2, int C, __bf16 *in) {
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= G * G * G) return;
    const int a = v % G, b = (v / G) % G, c = v / (G * G);
    int i0, i1, j0, j1, l0, l1;
    float wx, wy, wz;
    nn_lin((a + 0.5f) * Nx / G - 0.5f, Nx, i0, i1, wx);
    nn_lin((b + 0.5f) * Ny / G - 0.5f, Ny, j0, j1, wy);
    nn_lin((c + 0.5f) * Nz / G - 0.5f, Nz, l0, l1, wz);
    float f[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int d = 0; d < 8; d++) {
        const int i = (d & 1) ? i1 : i0, j = (d & 2) ? j1 : j0, l = (d & 4) ? l1 : l0;
        const float w = ((d & 1) ? wx : 1.f - wx) * ((d & 2) ? wy : 1.f - wy) * ((d & 4) ? wz : 1.f - wz);
        const size_t k = (size_t)i + (size_t)Nx * (j + (size_t)Ny * l);
        const double net = Nd[k] - Na[k];
        const double lg = log10(fabs(net) * 1e-6 + 1.0);       // cm^-3
        f[0] += w * (float)psi[k]; f[1] += w * (float)phin[k]; f[2] += w * (float)phip[k];
        f[3] += w * (float)(net >= 0. ? lg : -lg) / 20.f;
        if (h0) {
            f[4] += w * (float)((psi[k] - h0[k]) * 20.0);
            f[5] += w * (float)((phin[k] - h1[k]) * 20.0);
            f[6] += w * (float)((phip[k] - h2[k]) * 20.0);
        }
    }
    __bf16 *o = in + (size_t)v * C;
    for (int q = 0; q < C; q++) o[q] = (__bf16)0.f;
    o[0] = (__bf16)f[0]; o[1] = (__bf16)f[1]; o[2] = (__bf16)f[2]; o[3] = (__bf16)f[3]; o[4] = (__bf16)dv;
    o[5] = (__bf16)f[4]; o[6] = (__bf16)f[5]; o[7] = (__bf16)f[6];
}
void nn_mesh_to_grid(hipStream_t s, int G, int Nx, int Ny, int Nz, const double *psi, const double *phin,
                     const double *phip, const double *Nd, const double *Na, float dv, const double *h0,
                     const double *h1, const double *h2, int C, __bf16 *in) {
    k_mesh_to_grid<<<(G * G * G + 255) / 256, 256, 0, s>>>(G, Nx, Ny, Nz, psi, phin, phip, Nd, Na, dv, h0, h1, h2,
                                                            C, in);
}
// training target: (x1 - x0) / scale resampled to the grid, 3 FP32 channels
__global__ void k_target_to_grid(int G, int Nx, int Ny, int Nz, const double *a0, const double *a1, const double *a2,
                                 const double *b0, const double *b1, const double *b2, float iscale, float *tgt) {
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= G * G * G) return;
    const int a = v % G, b = (v / G) % G, c = v / (G * G);
    int i0, i1, j0, j1, l0, l1;
    float wx, wy, wz;
    nn_lin((a + 0.5f) * Nx / G - 0.5f, Nx, i0, i1, wx);
    nn_lin((b + 0.5f) * Ny / G - 0.5f, Ny, j0, j1, wy);
    nn_lin((c + 0.5f) * Nz / G - 0.5f, Nz, l0, l1, wz);
    float f[3] = {0.f, 0.f, 0.f};
    for (int d = 0; d < 8; d++) {
        const int i = (d & 1) ? i1 : i0, j = (d & 2) ? j1 : j0, l = (d & 4) ? l1 : l0;
        const float w = ((d & 1) ? wx : 1.f - wx) * ((d & 2) ? wy : 1.f - wy) * ((d & 4) ? wz : 1.f - wz);
        const size_t k = (size_t)i + (size_t)Nx * (j + (size_t)Ny * l);
        f[0] += w * (float)(b0[k] - a0[k]); f[1] += w * (float)(b1[k] - a1[k]); f[2] += w * (float)(b2[k] - a2[k]);
    }
    tgt[(size_t)v * 3 + 0] = f[0] * iscale; tgt[(size_t)v * 3 + 1] = f[1] * iscale; tgt[(size_t)v * 3 + 2] = f[2] * iscale;
}
void nn_target_to_grid(hipStream_t s, int G, int Nx, int Ny, int Nz, const double *a0, const double *a1,
                       const double *a2, const double *b0, const double *b1, const double *b2, float iscale,
                       float *tgt) {
    k_target_to_grid<<<(G * G * G + 255) / 256, 256, 0, s>>>(G, Nx, Ny, Nz, a0, a1, a2, b0, b1, b2, iscale, tgt);
}
// add the network's correction (grid output, N channel stride, first three
// channels, times scale) to the state on the mesh: psi on non-contact nodes,
// phi_n / phi_p on non-contact semiconductor nodes
__global__ void k_grid_to_mesh_add(int G, int Nx, int Ny, int Nz, const float *out, int N, float scale,
                                   const unsigned char *kind, const int *con_mask, double *psi, double *phin,
                                   double *phip, int chmask, float thr) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= Nx * Ny * Nz || con_mask[k]) return;
    const int i = k % Nx, j = (k / Nx) % Ny, l = k / (Nx * Ny);
    int a0, a1, b0, b1, c0, c1;
    float wx, wy, wz;
    nn_lin((i + 0.5f) * G / Nx - 0.5f, G, a0, a1, wx);
    nn_lin((j + 0.5f) * G / Ny - 0.5f, G, b0, b1, wy);
    nn_lin((l + 0.5f) * G / Nz - 0.5f, G, c0, c1, wz);
    float f[3] = {0.f, 0.f, 0.f};
    for (int d = 0; d < 8; d++) {
        const int a = (d & 1) ? a1 : a0, b = (d & 2) ? b1 : b0, c = (d & 4) ? c1 : c0;
        const float w = ((d & 1) ? wx : 1.f - wx) * ((d & 2) ? wy : 1.f - wy) * ((d & 4) ? wz : 1.f - wz);
        const size_t v = (size_t)a + (size_t)G * (b + (size_t)G * c);
        f[0] += w * out[v * N + 0]; f[1] += w * out[v * N + 1]; f[2] += w * out[v * N + 2];
    }
    // chmask: which of psi, phi_n, phi_p receive the correction; corrections
    // smaller than thr (V) are dropped
    for (int q = 0; q < 3; q++) f[q] = fabsf(f[q] * scale) >= thr ? f[q] * scale : 0.f;
    if (chmask & 1) psi[k] += (double)f[0];
    if (kind[k] == K_SEMI) {
        if (chmask & 2) phin[k] += (double)f[1];
        if (chmask & 4) phip[k] += (double)f[2];
    }
}
void nn_grid_to_mesh_add(hipStream_t s, int G, int Nx, int Ny, int Nz, const float *out, int N, float scale,
                         const unsigned char *kind, const int *con_mask, double *psi, double *phin, double *phip,
                         int chmask, float thr) {
    const int n = Nx * Ny * Nz;
    k_grid_to_mesh_add<<<(n + 255) / 256, 256, 0, s>>>(G, Nx, Ny, Nz, out, N, scale, kind, con_mask, psi, phin, phip,
                                                       chmask, thr);
}
// host-layout conversions for tests: in5 [C][V] float -> channel-last BF16
__global__ void k_planar_to_cl(int V, int Cr, int C, const float *src, __bf16 *dst) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= V * C) return;
    const int c = i % C, v = i / C;
    dst[i] = (__bf16)(c < Cr ? src[(size_t)c * V + v] : 0.f);
}
void nn_planar_to_cl(hipStream_t s, int V, int Cr, int C, const float *src, __bf16 *dst) {
    k_planar_to_cl<<<(V * C + 255) / 256, 256, 0, s>>>(V, Cr, C, src, dst);
}
