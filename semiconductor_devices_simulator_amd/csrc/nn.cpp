// nn.cpp - the warm-start network: a 3-D convolutional U-Net on a G^3 grid
// (four levels G, G/2, G/4, G/8 with c0..c3 channels), trained and run on the
// device with the kernels of k_nn.hip (BF16 matrix-core convolutions, FP32
// accumulation, FP32 master weights and Adam).  Input (NN_IN channels): psi,
// phi_n, phi_p of the plain continuation's predicted state, the signed doping
// log-magnitude, the bias step, and the predicted change of psi, phi_n, phi_p
// from the last converged state; output: the correction (delta psi, delta
// phi_n, delta phi_p) / scale that takes the predicted state toward the
// converged one.
//
//   e0 = relu(conv0(x))        G    c0      u2 = up(e3) + e2   G/4  c2
//   e1 = relu(conv1(pool e0))  G/2  c1      d2 = relu(conv4(u2))  c1
//   e2 = relu(conv2(pool e1))  G/4  c2      u1 = up(d2) + e1   G/2  c1
//   e3 = relu(conv3(pool e2))  G/8  c3=c2   d1 = relu(conv5(u1))  c0
//                                           u0 = up(d1) + e0   G    c0
//                                           d0 = relu(conv6(u0))  c0
//                                           y  = conv7(d0)        3
#include "sd_nn.h"
#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <random>
#include <vector>

#define NC(x)                                                                                   \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) {                                                                 \
            fprintf(stderr, "semisim-hip nn: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                                        \
            g_nn_err = 1;                                                                       \
        }                                                                                       \
    } while (0)

static int g_nn_err = 0;

namespace {
struct Layer {
    int G, Kr, K, Nr, N, relu;     // grid, input channels (real, padded), outputs (real, padded)
    size_t wofs, bofs;             // offsets of W[27][Nr][Kr] and b[Nr] in the parameter vector
    __bf16 *Wf = nullptr, *Wb = nullptr;
};
struct Net {
    int G = 0, c[4] = {0, 0, 0, 0};
    std::vector<Layer> L;
    size_t np = 0;
    float *P = nullptr, *Gr = nullptr, *M = nullptr, *Vv = nullptr;   // params, grads, Adam moments
    int step = 0;
    std::vector<void *> allocs;
    // activations (BF16 channel-last) and gradients
    __bf16 *x = nullptr, *e0, *p0, *e1, *p1, *e2, *p2, *e3, *u2, *d2, *u1, *d1, *u0, *d0, *y;
    float *yf = nullptr;
    __bf16 *gy, *gd0, *gu0, *gd1, *gu1, *gd2, *gu2, *ge3, *gp2, *ge2, *gp1, *ge1, *gp0, *ge0;
    float *part = nullptr, *lpart = nullptr;
    size_t part_n = 0;
    int lblk = 256;
    float *tgt = nullptr;          // target of a host-supplied sample (V x 3)
    float *planar = nullptr;       // host-layout staging (NN_IN V floats)
    // dataset on the device
    int ds_max = 0, ds_n = 0;
    __bf16 *ds_x = nullptr;
    float *ds_t = nullptr;
    float scale = 1e-3f;           // target scale (V)
    hipStream_t st = nullptr;
};
Net *g_net = nullptr;

template <class T> T *nalloc(Net &n, size_t count) {
    void *p = nullptr;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) { g_nn_err = 1; return nullptr; }
    n.allocs.push_back(p);
    return (T *)p;
}
void net_free(Net *n) {
    if (!n) return;
    if (n->st) hipStreamSynchronize(n->st);
    for (void *p : n->allocs) hipFree(p);
    if (n->ds_x) hipFree(n->ds_x);
    if (n->ds_t) hipFree(n->ds_t);
    if (n->st) hipStreamDestroy(n->st);
    delete n;
}
size_t vox(int G) { return (size_t)G * G * G; }

void pack_all(Net &n) {
    for (auto &l : n.L)
        nn_pack_w(n.st, n.P + l.wofs, l.Nr, l.Kr, l.N, l.K, l.K, l.Wf, l.Wb);
}

void forward(Net &n, const __bf16 *x) {
    hipStream_t s = n.st;
    const int G0 = n.G, G1 = G0 / 2, G2 = G0 / 4, G3 = G0 / 8;
    auto conv = [&](int i, const __bf16 *in, __bf16 *out, float *outf) {
        const Layer &l = n.L[i];
        nn_conv_fwd(s, l.G, in, l.K, l.Wf, n.P + l.bofs, l.N, out, outf, l.relu);
    };
    // biases are read from the FP32 parameter vector; padded outputs of the
    // last layer have no bias entry, so it gets a padded bias buffer
    conv(0, x, n.e0, nullptr);
    nn_pool(s, G1, n.c[0], n.e0, n.p0);
    conv(1, n.p0, n.e1, nullptr);
    nn_pool(s, G2, n.c[1], n.e1, n.p1);
    conv(2, n.p1, n.e2, nullptr);
    nn_pool(s, G3, n.c[2], n.e2, n.p2);
    conv(3, n.p2, n.e3, nullptr);
    nn_up_add(s, G2, n.c[2], n.e3, n.e2, n.u2);
    conv(4, n.u2, n.d2, nullptr);
    nn_up_add(s, G1, n.c[1], n.d2, n.e1, n.u1);
    conv(5, n.u1, n.d1, nullptr);
    nn_up_add(s, G0, n.c[0], n.d1, n.e0, n.u0);
    conv(6, n.u0, n.d0, nullptr);
    conv(7, n.d0, n.y, n.yf);
}

// data and weight gradients of layer i: dy at its output, in at its input;
// dx (optional) receives the input gradient
void conv_bwd(Net &n, int i, const __bf16 *dy, const __bf16 *in, __bf16 *dx) {
    hipStream_t s = n.st;
    const Layer &l = n.L[i];
    const size_t V = vox(l.G);
    int chunk = (int)(V / 64);
    if (chunk < 512) chunk = 512;
    chunk = (chunk + 15) / 16 * 16;
    const int nch = (int)((V + chunk - 1) / chunk);
    // the weight partials are laid out [chunk][t][N][K], the bias partials
    // [chunk][N] after them; their sums over the chunks go straight into the
    // real (Nr x Kr) block and the Nr biases of the gradient vector
    const int nw = 27 * l.N * l.K;
    float *pb = n.part + (size_t)nch * nw;
    nn_conv_wgrad(s, l.G, dy, l.N, in, l.K, chunk, nch, n.part, pb);
    nn_wgrad_reduce(s, n.part, nch, l.N, l.K, l.Nr, l.Kr, n.Gr + l.wofs);
    nn_bgrad(s, pb, nch, l.N, l.Nr, n.Gr + l.bofs);
    if (dx) nn_conv_fwd(s, l.G, dy, l.N, l.Wb, nullptr, l.K, dx, nullptr, 0);
}

// loss and all parameter gradients for input x and target tgt (V x 3)
float loss_grad(Net &n, const __bf16 *x, const float *tgt) {
    hipStream_t s = n.st;
    const int G0 = n.G, G1 = G0 / 2, G2 = G0 / 4, G3 = G0 / 8;
    forward(n, x);
    nn_mse(s, (int)vox(G0), n.yf, NN_COUT, tgt, 3, n.gy, n.lpart, n.lblk);
    conv_bwd(n, 7, n.gy, n.d0, n.gd0);
    nn_relu_bwd(s, (int)(vox(G0) * n.c[0]), n.d0, n.gd0, nullptr);
    conv_bwd(n, 6, n.gd0, n.u0, n.gu0);                         // d(u0)
    nn_up_bwd(s, G1, n.c[0], n.gu0, n.gd1);                      // d(d1)
    nn_relu_bwd(s, (int)(vox(G1) * n.c[0]), n.d1, n.gd1, nullptr);
    conv_bwd(n, 5, n.gd1, n.u1, n.gu1);
    nn_up_bwd(s, G2, n.c[1], n.gu1, n.gd2);
    nn_relu_bwd(s, (int)(vox(G2) * n.c[1]), n.d2, n.gd2, nullptr);
    conv_bwd(n, 4, n.gd2, n.u2, n.gu2);
    nn_up_bwd(s, G3, n.c[2], n.gu2, n.ge3);
    nn_relu_bwd(s, (int)(vox(G3) * n.c[3]), n.e3, n.ge3, nullptr);
    conv_bwd(n, 3, n.ge3, n.p2, n.gp2);
    nn_pool_bwd(s, G2, n.c[2], n.gp2, n.ge2, 0);
    nn_relu_bwd(s, (int)(vox(G2) * n.c[2]), n.e2, n.ge2, n.gu2);   // + skip of u2
    conv_bwd(n, 2, n.ge2, n.p1, n.gp1);
    nn_pool_bwd(s, G1, n.c[1], n.gp1, n.ge1, 0);
    nn_relu_bwd(s, (int)(vox(G1) * n.c[1]), n.e1, n.ge1, n.gu1);   // + skip of u1
    conv_bwd(n, 1, n.ge1, n.p0, n.gp0);
    nn_pool_bwd(s, G0, n.c[0], n.gp0, n.ge0, 0);
    nn_relu_bwd(s, (int)(vox(G0) * n.c[0]), n.e0, n.ge0, n.gu0);   // + skip of u0
    conv_bwd(n, 0, n.ge0, x, nullptr);
    std::vector<float> lp(n.lblk);
    NC(hipMemcpyAsync(lp.data(), n.lpart, n.lblk * sizeof(float), hipMemcpyDeviceToHost, s));
    NC(hipStreamSynchronize(s));
    double sum = 0.;
    for (float v : lp) sum += v;
    return (float)(sum / ((double)vox(G0) * 3.));
}
}  // namespace

// ------------------------------------------------------------------ engine side
bool nn_ready() { return g_net != nullptr && !g_nn_err; }
int nn_grid() { return g_net ? g_net->G : 0; }
__bf16 *nn_input_buf() { return g_net ? g_net->x : nullptr; }
float nn_target_scale() { return g_net ? g_net->scale : 1.f; }
const float *nn_infer(hipStream_t s) {
    Net &n = *g_net;
    NC(hipStreamSynchronize(s));         // the input was written on the engine's stream
    forward(n, n.x);
    NC(hipStreamSynchronize(n.st));
    return n.yf;
}
int nn_capture(hipStream_t s, void (*fill)(hipStream_t, float *, void *), void *ctx) {
    Net &n = *g_net;
    if (n.ds_n >= n.ds_max) return 0;
    const size_t V = vox(n.G);
    NC(hipMemcpyAsync(n.ds_x + (size_t)n.ds_n * V * NN_CIN, n.x, V * NN_CIN * sizeof(__bf16),
                      hipMemcpyDeviceToDevice, s));
    fill(s, n.ds_t + (size_t)n.ds_n * V * 3, ctx);
    n.ds_n++;
    return 1;
}

// ------------------------------------------------------------------ C ABI
extern "C" {

// Create the network (grid G, a power of two >= 8; channels c0..c3 multiples
// of 32 up to 64, c3 == c2) with He-initialised weights from seed.  1 = done.
int sdx_nn_init(int G, int c0, int c1, int c2, int c3, unsigned seed) {
    // channels up to 64: the weight-gradient kernel spreads at most four
    // 32 x 32 output tiles over its wavefronts
    if (G < 8 || (G & (G - 1)) || c0 % 32 || c1 % 32 || c2 % 32 || c3 != c2 || c0 > 64 || c1 > 64 || c2 > 64) {
        fprintf(stderr, "semisim-hip nn: unsupported shape G=%d channels %d %d %d %d\n", G, c0, c1, c2, c3);
        return 0;
    }
    net_free(g_net);
    g_net = nullptr;
    g_nn_err = 0;
    Net *np = new Net();
    Net &n = *np;
    n.G = G; n.c[0] = c0; n.c[1] = c1; n.c[2] = c2; n.c[3] = c3;
    if (hipStreamCreateWithFlags(&n.st, hipStreamNonBlocking) != hipSuccess) { delete np; return 0; }
    const int G0 = G, G1 = G / 2, G2 = G / 4, G3 = G / 8;
    // G, Kr, K, Nr, N, relu
    const int spec[8][6] = {{G0, NN_IN, NN_CIN, c0, c0, 1}, {G1, c0, c0, c1, c1, 1}, {G2, c1, c1, c2, c2, 1},
                            {G3, c2, c2, c3, c3, 1},    {G2, c2, c2, c1, c1, 1}, {G1, c1, c1, c0, c0, 1},
                            {G0, c0, c0, c0, c0, 1},    {G0, c0, c0, 3, NN_COUT, 0}};
    size_t ofs = 0;
    for (int i = 0; i < 8; i++) {
        Layer l;
        l.G = spec[i][0]; l.Kr = spec[i][1]; l.K = spec[i][2]; l.Nr = spec[i][3]; l.N = spec[i][4]; l.relu = spec[i][5];
        l.wofs = ofs; ofs += (size_t)27 * l.Nr * l.Kr;
        l.bofs = ofs; ofs += (size_t)l.Nr;
        n.L.push_back(l);
    }
    // the last layer's bias is read for N = 32 outputs: pad the vector
    n.np = ofs;
    const size_t nalloc_p = ofs + NN_COUT;
    n.P = nalloc<float>(n, nalloc_p); n.Gr = nalloc<float>(n, nalloc_p);
    n.M = nalloc<float>(n, nalloc_p); n.Vv = nalloc<float>(n, nalloc_p);
    for (int i = 0; i < 8; i++) {
        Layer &l = n.L[i];
        l.Wf = nalloc<__bf16>(n, (size_t)27 * l.N * l.K);
        l.Wb = i > 0 ? nalloc<__bf16>(n, (size_t)27 * l.N * l.K) : nullptr;
    }
    const size_t V0 = vox(G0), V1 = vox(G1), V2 = vox(G2), V3 = vox(G3);
    n.x = nalloc<__bf16>(n, V0 * NN_CIN);
    n.e0 = nalloc<__bf16>(n, V0 * c0); n.p0 = nalloc<__bf16>(n, V1 * c0);
    n.e1 = nalloc<__bf16>(n, V1 * c1); n.p1 = nalloc<__bf16>(n, V2 * c1);
    n.e2 = nalloc<__bf16>(n, V2 * c2); n.p2 = nalloc<__bf16>(n, V3 * c2);
    n.e3 = nalloc<__bf16>(n, V3 * c3);
    n.u2 = nalloc<__bf16>(n, V2 * c2); n.d2 = nalloc<__bf16>(n, V2 * c1);
    n.u1 = nalloc<__bf16>(n, V1 * c1); n.d1 = nalloc<__bf16>(n, V1 * c0);
    n.u0 = nalloc<__bf16>(n, V0 * c0); n.d0 = nalloc<__bf16>(n, V0 * c0);
    n.y = nalloc<__bf16>(n, V0 * NN_COUT); n.yf = nalloc<float>(n, V0 * NN_COUT);
    n.gy = nalloc<__bf16>(n, V0 * NN_COUT);
    n.gd0 = nalloc<__bf16>(n, V0 * c0); n.gu0 = nalloc<__bf16>(n, V0 * c0);
    n.gd1 = nalloc<__bf16>(n, V1 * c0); n.gu1 = nalloc<__bf16>(n, V1 * c1);
    n.gd2 = nalloc<__bf16>(n, V2 * c1); n.gu2 = nalloc<__bf16>(n, V2 * c2);
    n.ge3 = nalloc<__bf16>(n, V3 * c3); n.gp2 = nalloc<__bf16>(n, V3 * c2);
    n.ge2 = nalloc<__bf16>(n, V2 * c2); n.gp1 = nalloc<__bf16>(n, V2 * c1);
    n.ge1 = nalloc<__bf16>(n, V1 * c1); n.gp0 = nalloc<__bf16>(n, V1 * c0);
    n.ge0 = nalloc<__bf16>(n, V0 * c0);
    // weight-gradient partials: at most 64 chunks of the largest layer plus
    // one padded gradient
    size_t wmax = 0;
    for (auto &l : n.L) wmax = std::max(wmax, (size_t)27 * l.N * l.K);
    n.part_n = 65 * wmax;
    n.part = nalloc<float>(n, n.part_n);
    n.lpart = nalloc<float>(n, n.lblk);
    n.tgt = nalloc<float>(n, V0 * 3);
    n.planar = nalloc<float>(n, V0 * NN_IN);
    // He initialisation (the output layer scaled down)
    std::mt19937 rng(seed);
    std::vector<float> h(nalloc_p, 0.f);
    for (int i = 0; i < 8; i++) {
        const Layer &l = n.L[i];
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (27.f * l.Kr)) * (i == 7 ? 0.1f : 1.f));
        for (size_t q = 0; q < (size_t)27 * l.Nr * l.Kr; q++) h[l.wofs + q] = nd(rng);
    }
    NC(hipMemcpy(n.P, h.data(), nalloc_p * sizeof(float), hipMemcpyHostToDevice));
    NC(hipMemset(n.Gr, 0, nalloc_p * sizeof(float)));
    NC(hipMemset(n.M, 0, nalloc_p * sizeof(float)));
    NC(hipMemset(n.Vv, 0, nalloc_p * sizeof(float)));
    if (g_nn_err) { net_free(np); return 0; }
    g_net = np;
    pack_all(n);
    NC(hipStreamSynchronize(n.st));
    return g_nn_err ? 0 : 1;
}
void sdx_nn_free(void) { net_free(g_net); g_net = nullptr; }
int sdx_nn_nparams(void) { return g_net ? (int)g_net->np : 0; }
int sdx_nn_get_params(float *out, int n) {
    if (!g_net || n != (int)g_net->np) return 0;
    NC(hipMemcpy(out, g_net->P, n * sizeof(float), hipMemcpyDeviceToHost));
    return !g_nn_err;
}
int sdx_nn_set_params(const float *in, int n) {
    if (!g_net || n != (int)g_net->np) return 0;
    NC(hipMemcpy(g_net->P, in, n * sizeof(float), hipMemcpyHostToDevice));
    pack_all(*g_net);
    NC(hipStreamSynchronize(g_net->st));
    g_net->step = 0;
    NC(hipMemset(g_net->M, 0, g_net->np * sizeof(float)));
    NC(hipMemset(g_net->Vv, 0, g_net->np * sizeof(float)));
    return !g_nn_err;
}
void sdx_nn_set_scale(double s) { if (g_net && s > 0.) g_net->scale = (float)s; }
double sdx_nn_get_scale(void) { return g_net ? g_net->scale : 0.; }

// forward on a host sample: in = NN_IN planar channels of G^3 floats (voxel
// x + G (y + G z)); out3 = the 3 output channels, planar
int sdx_nn_forward_host(const float *in, float *out3) {
    if (!g_net) return 0;
    Net &n = *g_net;
    const size_t V = vox(n.G);
    NC(hipMemcpyAsync(n.planar, in, NN_IN * V * sizeof(float), hipMemcpyHostToDevice, n.st));
    nn_planar_to_cl(n.st, (int)V, NN_IN, NN_CIN, n.planar, n.x);
    forward(n, n.x);
    std::vector<float> y(V * NN_COUT);
    NC(hipMemcpyAsync(y.data(), n.yf, y.size() * sizeof(float), hipMemcpyDeviceToHost, n.st));
    NC(hipStreamSynchronize(n.st));
    for (int c = 0; c < 3; c++)
        for (size_t v = 0; v < V; v++) out3[c * V + v] = y[v * NN_COUT + c];
    return !g_nn_err;
}
// loss and parameter gradients (no update) for a host sample; tgt3 planar
double sdx_nn_grad_host(const float *in, const float *tgt3, float *grad) {
    if (!g_net) return -1.;
    Net &n = *g_net;
    const size_t V = vox(n.G);
    std::vector<float> t(V * 3);
    for (int c = 0; c < 3; c++)
        for (size_t v = 0; v < V; v++) t[v * 3 + c] = tgt3[c * V + v];
    NC(hipMemcpyAsync(n.planar, in, NN_IN * V * sizeof(float), hipMemcpyHostToDevice, n.st));
    NC(hipMemcpyAsync(n.tgt, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice, n.st));
    nn_planar_to_cl(n.st, (int)V, NN_IN, NN_CIN, n.planar, n.x);
    float loss = loss_grad(n, n.x, n.tgt);
    if (grad) NC(hipMemcpy(grad, n.Gr, n.np * sizeof(float), hipMemcpyDeviceToHost));
    return g_nn_err ? -1. : (double)loss;
}

// training samples kept on the device
int sdx_nn_dataset_alloc(int nmax) {
    if (!g_net || nmax < 1) return 0;
    Net &n = *g_net;
    const size_t V = vox(n.G);
    if (n.ds_x) { hipFree(n.ds_x); hipFree(n.ds_t); n.ds_x = nullptr; n.ds_t = nullptr; }
    if (hipMalloc(&n.ds_x, (size_t)nmax * V * NN_CIN * sizeof(__bf16)) != hipSuccess ||
        hipMalloc(&n.ds_t, (size_t)nmax * V * 3 * sizeof(float)) != hipSuccess) {
        fprintf(stderr, "semisim-hip nn: no memory for %d samples\n", nmax);
        n.ds_max = n.ds_n = 0;
        return 0;
    }
    n.ds_max = nmax; n.ds_n = 0;
    return 1;
}
int sdx_nn_dataset_size(void) { return g_net ? g_net->ds_n : 0; }
// read sample i back (planar, 5 input and 3 target channels) for inspection
int sdx_nn_dataset_get(int i, float *in, float *tgt3) {
    if (!g_net || i < 0 || i >= g_net->ds_n) return 0;
    Net &n = *g_net;
    const size_t V = vox(n.G);
    std::vector<__bf16> x(V * NN_CIN);
    std::vector<float> t(V * 3);
    NC(hipMemcpy(x.data(), n.ds_x + (size_t)i * V * NN_CIN, x.size() * sizeof(__bf16), hipMemcpyDeviceToHost));
    NC(hipMemcpy(t.data(), n.ds_t + (size_t)i * V * 3, t.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t v = 0; v < V; v++) {
        for (int c = 0; c < NN_IN; c++) if (in) in[c * V + v] = (float)x[v * NN_CIN + c];
        for (int c = 0; c < 3; c++) if (tgt3) tgt3[c * V + v] = t[v * 3 + c];
    }
    return !g_nn_err;
}
// append a host sample (planar: NN_IN input and 3 target channels) to the
// dataset, as the capture path stores it (input channel-last BF16, zero
// padded; target FP32); returns its index or -1
int sdx_nn_dataset_put(const float *in, const float *tgt3) {
    if (!g_net || !in || !tgt3 || g_net->ds_n >= g_net->ds_max) return -1;
    Net &n = *g_net;
    const size_t V = vox(n.G);
    std::vector<__bf16> x(V * NN_CIN, (__bf16)0.0f);
    std::vector<float> t(V * 3);
    for (size_t v = 0; v < V; v++) {
        for (int c = 0; c < NN_IN; c++) x[v * NN_CIN + c] = (__bf16)in[c * V + v];
        for (int c = 0; c < 3; c++) t[v * 3 + c] = tgt3[c * V + v];
    }
    NC(hipMemcpy(n.ds_x + (size_t)n.ds_n * V * NN_CIN, x.data(), x.size() * sizeof(__bf16), hipMemcpyHostToDevice));
    NC(hipMemcpy(n.ds_t + (size_t)n.ds_n * V * 3, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
    return g_nn_err ? -1 : n.ds_n++;
}
// Adam steps over samples drawn from the dataset (sample i < n_train only);
// per-step losses into loss (optional); returns the mean loss
double sdx_nn_train(int steps, int n_train, double lr, unsigned seed, float *loss) {
    if (!g_net || g_net->ds_n < 1) return -1.;
    Net &n = *g_net;
    const size_t V = vox(n.G);
    const int nt = n_train > 0 && n_train <= n.ds_n ? n_train : n.ds_n;
    std::mt19937 rng(seed);
    std::uniform_int_distribution<int> pick(0, nt - 1);
    double sum = 0.;
    for (int it = 0; it < steps; it++) {
        const int i = pick(rng);
        const float l = loss_grad(n, n.ds_x + (size_t)i * V * NN_CIN, n.ds_t + (size_t)i * V * 3);
        n.step++;
        nn_adam(n.st, (int)n.np, n.P, n.Gr, n.M, n.Vv, (float)lr, 0.9f, 0.999f, 1e-8f, n.step);
        pack_all(n);
        if (loss) loss[it] = l;
        sum += l;
        if (g_nn_err) return -1.;
    }
    NC(hipStreamSynchronize(n.st));
    return sum / (steps > 0 ? steps : 1);
}
// mean loss over dataset samples [i0, i1) without updates
double sdx_nn_eval(int i0, int i1) {
    if (!g_net) return -1.;
    Net &n = *g_net;
    const size_t V = vox(n.G);
    if (i1 > n.ds_n) i1 = n.ds_n;
    double sum = 0.;
    int cnt = 0;
    for (int i = i0; i < i1; i++, cnt++) sum += loss_grad(n, n.ds_x + (size_t)i * V * NN_CIN, n.ds_t + (size_t)i * V * 3);
    return cnt ? sum / cnt : -1.;
}

}  // extern "C"
