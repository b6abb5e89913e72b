// unet.hip -- learned preconditioner: a 4-level 3D U-Net (16/32/64/128 channels)
// with BF16 MFMA convolutions, trained on the GPU with the library's own forward
// and backward kernels, applied inside flexible CG with plain PCG as fallback.
//
// Preconditioner, on the Jacobi-scaled system A^ = D^-1/2 A D^-1/2 (unit diagonal):
//   z0 = mask * s * P U(R(r) / s, R(mask), R(c^x), R(c^y), R(c^z))
//   z_{i+1} = z_i + omega * mask * (r - A^ z_i),  i < kSmooth;   M(r) = z_kSmooth
// R: 2x2x2 average to the half-resolution grid, P: cell-centred trilinear
// prolongation back, s: the RMS of r over the unknowns (M is positively
// homogeneous), U: the U-Net; the weighted-Jacobi steps remove the high-frequency
// error that a half-resolution correction cannot represent.
// Training minimises the A^-norm error of M(b^) against the label x^* = A^-1 b^
// computed by the plain solver: L = (M(b^) - x^*)' A^ (M(b^) - x^*) / (x^*' b^).
//
// Network (m = half-resolution grid, each level halves it):
//   enc0 = relu(conv(feat 8 -> 16))                        m
//   enc1 = relu(conv(pool(enc0) 16 -> 32))                 m/2
//   enc2 = relu(conv(pool(enc1) 32 -> 64))                 m/4
//   bott = relu(conv(pool(enc2) 64 -> 128))                m/8
//   dec2 = relu(conv([up(bott) | enc2] 192 -> 64))         m/4
//   dec1 = relu(conv([up(dec2) | enc1] 96 -> 32))          m/2
//   dec0 = relu(conv([up(dec1) | enc0] 48 -> 16))          m
//   u    = conv(dec0 16 -> 16), channel 0                  m
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "ctx_internal.h"
#include "setup.h"
#include "unet_kernels.h"

namespace shk {

namespace {

struct LayerSpec {
    int cin, cout;
    bool relu;
    bool need_dgrad;
};
constexpr int kLayers = 8;
const LayerSpec kSpec[kLayers] = {{8, 16, true, false},  {16, 32, true, true},  {32, 64, true, true},
                                  {64, 128, true, true}, {192, 64, true, true}, {96, 32, true, true},
                                  {48, 16, true, true},  {16, 16, false, true}};
constexpr uint64_t kMagic = 0x324E45554B485353ull;  // "SSHKUEN2"
constexpr int kSmooth = 2;                          // weighted-Jacobi steps after the coarse correction

template <class T>
hipError_t dalloc(T *&p, size_t n) {
    if (p) {
        (void)hipFree(p);
        p = nullptr;
    }
    return hipMalloc((void **)&p, sizeof(T) * (n ? n : 1));
}
template <class T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Deterministic two-stage reductions: per-block partials, folded on the host in order.
constexpr int kRedBlocks = 1024;

__global__ __launch_bounds__(256) void k_dot2(int64_t n, const float *__restrict__ a, const float *__restrict__ b,
                                              const float *__restrict__ c, const float *__restrict__ d,
                                              const float *__restrict__ e, double *__restrict__ part) {
    __shared__ double lds[4];
    double s1 = 0.0, s2 = 0.0, m = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (a) s1 += (double)a[i] * (double)b[i];
        if (c) s2 += (double)c[i] * (double)d[i];
        if (e) m = fmax(m, fabs((double)e[i]));
    }
    const double r1 = block_sum<256>(s1, lds);
    const double r2 = block_sum<256>(s2, lds);
    const double r3 = block_max<256>(m, lds);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = r1;
        part[gridDim.x + blockIdx.x] = r2;
        part[2 * gridDim.x + blockIdx.x] = r3;
    }
}

__global__ void k_axpy_f32(int64_t n, float a, const float *__restrict__ x, float *__restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}
// y = x + b * y
__global__ void k_xpby_f32(int64_t n, const float *__restrict__ x, float b, float *__restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = x[i] + b * y[i];
}
// d = a - b
__global__ void k_sub_f32(int64_t n, const float *__restrict__ a, const float *__restrict__ b, float *__restrict__ d) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = a[i] - b[i];
}
// r = b - r
__global__ void k_rsub_f32(int64_t n, const float *__restrict__ b, float *__restrict__ r) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        r[i] = b[i] - r[i];
}
__global__ void k_scale_f32(int64_t n, float s, float *__restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] *= s;
}
// channel 0 of a 16-channel bf16 grid from an FP32 grid, other channels 0
__global__ void k_to_channel0(int64_t vox, const float *__restrict__ src, uint16_t *__restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < vox * 16;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i & 15);
        float v = c == 0 ? src[i >> 4] : 0.0f;
        const bf16 h = __float2bfloat16(v);
        dst[i] = *reinterpret_cast<const uint16_t *>(&h);
    }
}
// dense (padded) <-> brick layouts
template <class TS, class TD>
__global__ void k_brick_to_dense(int nx, int ny, int nz, int px, int py, int pz, const int *__restrict__ brick_slot,
                                 int nbx, int nby, const TS *__restrict__ src, TD *__restrict__ dst) {
    const int64_t n = (int64_t)px * py * pz;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(v % px), j = (int)((v / px) % py), k = (int)(v / ((int64_t)px * py));
        TD val = TD(0);
        if (i < nx && j < ny && k < nz) {
            const int s = brick_slot[(i / 8) + (int64_t)nbx * ((j / 8) + (int64_t)nby * (k / 8))];
            if (s >= 0) val = (TD)src[(int64_t)s * 512 + (i & 7) + 8 * (j & 7) + 64 * (k & 7)];
        }
        dst[v] = val;
    }
}
__global__ void k_dense_to_brick(int nx, int ny, int nz, int px, int py, const int4 *__restrict__ coord, int n_active,
                                 const float *__restrict__ src, float *__restrict__ dst) {
    const int64_t n = (int64_t)n_active * 512;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int s = (int)(e >> 9), l = (int)(e & 511);
        const int4 c = coord[s];
        const int i = c.x * 8 + (l & 7), j = c.y * 8 + ((l >> 3) & 7), k = c.z * 8 + (l >> 6);
        dst[e] = (i < nx && j < ny && k < nz) ? src[((int64_t)k * py + j) * px + i] : 0.0f;
    }
}
__global__ void k_label_mask(int nx, int ny, int nz, int px, int py, int pz, const uint8_t *__restrict__ label,
                             float *__restrict__ mask) {
    const int64_t n = (int64_t)px * py * pz;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(v % px), j = (int)((v / px) % py), k = (int)(v / ((int64_t)px * py));
        mask[v] = (i < nx && j < ny && k < nz && label[((int64_t)k * ny + j) * nx + i]) ? 1.0f : 0.0f;
    }
}

// out = sig * a (inv = 0) or a / sig (inv = 1, 0 where sig == 0)
__global__ void k_mul_sig(int64_t n, const double *__restrict__ a, const double *__restrict__ sig,
                          double *__restrict__ out, int inv) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = inv ? (sig[i] > 0.0 ? a[i] / sig[i] : 0.0) : sig[i] * a[i];
}

inline unsigned g1(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    return (unsigned)(g > 0 ? g : 1);
}

}  // namespace

// ===================================================================== network

// Step weight i of k for Chebyshev smoothing of the unit-diagonal operator on the
// eigenvalue interval [2 / ratio, 2] (2 bounds the spectrum: every row's couplings sum
// to at most 1): the reciprocals of the Chebyshev nodes of the interval, taken
// alternately from its two ends so that no intermediate product grows far.
double cheb_weight(int i, int k, double ratio) {
    const double hi = 2.0, lo = 2.0 / ratio;
    const int j = (i & 1) ? k - 1 - i / 2 : i / 2;  // 0, k-1, 1, k-2, ...
    const double node = 0.5 * (hi + lo) + 0.5 * (hi - lo) * std::cos(M_PI * (2 * j + 1) / (2.0 * k));
    return 1.0 / node;
}

struct UnetNet {
    hipStream_t stream = nullptr;
    int device = 0;
    // parameters (FP32 master), gradients, Adam moments: one contiguous block
    int64_t nparam = 0;
    int64_t woff[kLayers], boff[kLayers];
    float *param = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    int64_t step = 0;
    double omega = 0.8;  // smoothing weight the network was trained with (saved with it)
    // the smoothing of the dense application and of the training loss: smooth_steps steps,
    // each with omega, or with Chebyshev weights on [2 / smooth_cheb, 2] when smooth_cheb > 1
    int smooth_steps = kSmooth;
    double smooth_cheb = 0.0;
    bf16 *pk[kLayers] = {}, *pkt[kLayers] = {};
    bf16 *pkr[kLayers] = {};      // row-reuse packing of the 16/32-output layers (inference)
    uint32_t *head_pk = nullptr;  // the head's output-channel-0 weights as BF16 pairs (inference)
    const int32_t *stop = nullptr;  // inference kernels return at once while *stop is set
    bool packed_dirty = true;
    // activations for the current half-resolution grid
    Grid3 L[4] = {};
    bf16 *feat = nullptr, *enc0 = nullptr, *p1 = nullptr, *enc1 = nullptr, *p2 = nullptr, *enc2 = nullptr,
         *p3 = nullptr, *bott = nullptr, *cat2 = nullptr, *dec2 = nullptr, *cat1 = nullptr, *dec1 = nullptr,
         *cat0 = nullptr, *dec0 = nullptr;
    bf16 *dout = nullptr, *d16 = nullptr, *dcat0 = nullptr, *dd1 = nullptr, *dcat1 = nullptr, *dd2 = nullptr,
         *dcat2 = nullptr, *dbott = nullptr, *dp3 = nullptr, *de2 = nullptr, *dp2 = nullptr, *de1 = nullptr,
         *dp1 = nullptr, *de0 = nullptr, *tmp = nullptr;
    float *u = nullptr, *du = nullptr;
    // fine-level training work arrays
    int64_t fine_cap = 0;
    float *fz = nullptr, *fe = nullptr, *fae = nullptr, *ftmp = nullptr;
    double *part = nullptr, *h_part = nullptr;
    std::string err;
};

namespace {

int fail(UnetNet *n, int code, const char *msg) {
    n->err = msg;
    return code;
}

#define UHIP(net, expr)                                                                       \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) {                                                               \
            (net)->err = std::string(#expr) + ": " + hipGetErrorString(_e);                   \
            return _e == hipErrorOutOfMemory ? SHK_OUT_OF_MEMORY : SHK_HIP_ERROR;             \
        }                                                                                     \
    } while (0)

int repack(UnetNet *n) {
    if (!n->packed_dirty) return SHK_OK;
    for (int l = 0; l < kLayers; ++l) {
        launch_conv_pack(n->stream, n->param + n->woff[l], kSpec[l].cin, kSpec[l].cout, false, n->pk[l]);
        if (kSpec[l].need_dgrad)
            launch_conv_pack(n->stream, n->param + n->woff[l], kSpec[l].cin, kSpec[l].cout, true, n->pkt[l]);
        if (n->pkr[l]) launch_conv_pack_rows(n->stream, n->param + n->woff[l], kSpec[l].cin, kSpec[l].cout, n->pkr[l]);
    }
    launch_head_pack(n->stream, n->param + n->woff[kLayers - 1], kSpec[kLayers - 1].cout, n->head_pk);
    UHIP(n, hipGetLastError());
    n->packed_dirty = false;
    return SHK_OK;
}

int ensure_grid(UnetNet *n, Grid3 m) {
    if (m.nx == n->L[0].nx && m.ny == n->L[0].ny && m.nz == n->L[0].nz && n->feat) return SHK_OK;
    if (m.nx % 8 || m.ny % 8 || m.nz % 8 || m.nx <= 0) return fail(n, SHK_BAD_ARG, "U-Net grid must be a multiple of 8");
    n->L[0] = m;
    for (int i = 1; i < 4; ++i) n->L[i] = Grid3{n->L[i - 1].nx / 2, n->L[i - 1].ny / 2, n->L[i - 1].nz / 2};
    const int64_t v0 = n->L[0].vox(), v1 = n->L[1].vox(), v2 = n->L[2].vox(), v3 = n->L[3].vox();
    struct A {
        bf16 **p;
        int64_t n;
    } a[] = {{&n->feat, v0 * 8},   {&n->enc0, v0 * 16}, {&n->p1, v1 * 16},   {&n->enc1, v1 * 32},  {&n->p2, v2 * 32},
             {&n->enc2, v2 * 64},  {&n->p3, v3 * 64},   {&n->bott, v3 * 128}, {&n->cat2, v2 * 192}, {&n->dec2, v2 * 64},
             {&n->cat1, v1 * 96},  {&n->dec1, v1 * 32}, {&n->cat0, v0 * 48},  {&n->dec0, v0 * 16},  {&n->dout, v0 * 16},
             {&n->d16, v0 * 16},   {&n->dcat0, v0 * 48}, {&n->dd1, v1 * 32},  {&n->dcat1, v1 * 96}, {&n->dd2, v2 * 64},
             {&n->dcat2, v2 * 192}, {&n->dbott, v3 * 128}, {&n->dp3, v3 * 64}, {&n->de2, v2 * 64},  {&n->dp2, v2 * 32},
             {&n->de1, v1 * 32},   {&n->dp1, v1 * 16},  {&n->de0, v0 * 16},   {&n->tmp, v0 * 64}};
    for (auto &x : a) UHIP(n, dalloc(*x.p, (size_t)x.n));
    UHIP(n, dalloc(n->u, (size_t)v0));
    UHIP(n, dalloc(n->du, (size_t)v0));
    return SHK_OK;
}

int ensure_fine(UnetNet *n, int64_t nf) {
    if (nf <= n->fine_cap) return SHK_OK;
    UHIP(n, dalloc(n->fz, (size_t)nf));
    UHIP(n, dalloc(n->fe, (size_t)nf));
    UHIP(n, dalloc(n->fae, (size_t)nf));
    UHIP(n, dalloc(n->ftmp, (size_t)nf));
    n->fine_cap = nf;
    return SHK_OK;
}

// sum(a.b), sum(c.d), max|e| over n floats (any pair may be null)
int reduce3(UnetNet *n, int64_t len, const float *a, const float *b, const float *c, const float *d, const float *e,
            double out[3]) {
    int blocks = (int)std::min<int64_t>(kRedBlocks, (len + 255) / 256);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_dot2, dim3(blocks), dim3(256), 0, n->stream, len, a, b, c, d, e, n->part);
    UHIP(n, hipGetLastError());
    UHIP(n, hipMemcpyAsync(n->h_part, n->part, sizeof(double) * 3 * blocks, hipMemcpyDeviceToHost, n->stream));
    UHIP(n, hipStreamSynchronize(n->stream));
    out[0] = out[1] = out[2] = 0.0;
    for (int i = 0; i < blocks; ++i) {
        out[0] += n->h_part[i];
        out[1] += n->h_part[blocks + i];
        out[2] = std::max(out[2], n->h_part[2 * blocks + i]);
    }
    return SHK_OK;
}

void conv(UnetNet *n, int l, Grid3 g, const bf16 *x, bf16 *y, float *out_f32 = nullptr) {
    launch_conv_fwd(n->stream, g, x, kSpec[l].cin, n->pk[l], n->param + n->boff[l], kSpec[l].cout, kSpec[l].relu, y,
                    out_f32);
}
// input gradient of layer l: dx (cin channels) from dz (cout channels)
void dgrad(UnetNet *n, int l, Grid3 g, const bf16 *dz, bf16 *dx) {
    launch_conv_fwd(n->stream, g, dz, kSpec[l].cout, n->pkt[l], nullptr, kSpec[l].cin, false, dx, nullptr);
}
void wgrad(UnetNet *n, int l, Grid3 g, const bf16 *x, const bf16 *dz) {
    launch_conv_wgrad(n->stream, g, x, kSpec[l].cin, dz, kSpec[l].cout, n->grad + n->woff[l], n->grad + n->boff[l]);
}

// forward on the features already in n->feat; u = network output (FP32, level 0)
// train: materialise the decoder's concatenations (the backward pass reads them);
// otherwise each decoder convolution reads [upsample(coarse), skip] directly.
void forward(UnetNet *n, bool train) {
    const Grid3 *L = n->L;
    // inference runs the row-reuse kernel where a layer's grid fills the chip
    auto rows = [&](int l, Grid3 g) { return !train && n->pkr[l] && conv_rows_ok(g, kSpec[l].cin, kSpec[l].cout); };
    // and the direct-load kernel on the grids too small for the tiled kernel
    auto direct = [&](int l, Grid3 g) { return !train && conv_direct_ok(g, kSpec[l].cin, kSpec[l].cout); };
    auto enc = [&](int l, Grid3 g, const bf16 *x, bf16 *y) {
        if (rows(l, g))
            launch_conv_rows(n->stream, g, x, kSpec[l].cin, n->pkr[l], n->param + n->boff[l], kSpec[l].cout,
                             kSpec[l].relu, y, nullptr, 0, n->stop);
        else if (direct(l, g))
            launch_conv_direct(n->stream, g, x, kSpec[l].cin, n->pk[l], n->param + n->boff[l], kSpec[l].cout,
                               kSpec[l].relu, y, nullptr, 0, n->stop);
        else
            launch_conv_fwd(n->stream, g, x, kSpec[l].cin, n->pk[l], n->param + n->boff[l], kSpec[l].cout,
                            kSpec[l].relu, y, nullptr, n->stop);
    };
    enc(0, L[0], n->feat, n->enc0);
    launch_pool(n->stream, L[1], 16, n->enc0, n->p1, n->stop);
    enc(1, L[1], n->p1, n->enc1);
    launch_pool(n->stream, L[2], 32, n->enc1, n->p2, n->stop);
    enc(2, L[2], n->p2, n->enc2);
    launch_pool(n->stream, L[3], 64, n->enc2, n->p3, n->stop);
    enc(3, L[3], n->p3, n->bott);
    if (!train) {
        auto cat = [&](int l, Grid3 g, const bf16 *coarse, int c1, const bf16 *skip, bf16 *y) {
            if (rows(l, g))
                launch_conv_rows(n->stream, g, skip, kSpec[l].cin, n->pkr[l], n->param + n->boff[l], kSpec[l].cout,
                                 kSpec[l].relu, y, coarse, c1, n->stop);
            else if (direct(l, g))
                launch_conv_direct(n->stream, g, skip, kSpec[l].cin, n->pk[l], n->param + n->boff[l],
                                   kSpec[l].cout, kSpec[l].relu, y, coarse, c1, n->stop);
            else
                launch_conv_fwd_cat(n->stream, g, coarse, c1, skip, kSpec[l].cin, n->pk[l], n->param + n->boff[l],
                                    kSpec[l].cout, kSpec[l].relu, y, n->stop);
        };
        cat(4, L[2], n->bott, 128, n->enc2, n->dec2);
        cat(5, L[1], n->dec2, 64, n->enc1, n->dec1);
        cat(6, L[0], n->dec1, 32, n->enc0, n->dec0);
        // only output channel 0 of the head is used: the vector-ALU head kernel
        launch_conv_head(n->stream, L[0], n->dec0, n->head_pk, n->param + n->boff[7], n->u, n->stop);
        return;
    }
    launch_upsample(n->stream, L[3], 128, n->bott, 192, n->cat2);
    launch_copy_channels(n->stream, L[2].vox(), 64, n->enc2, 64, 0, n->cat2, 192, 128, false);
    conv(n, 4, L[2], n->cat2, n->dec2);
    launch_upsample(n->stream, L[2], 64, n->dec2, 96, n->cat1);
    launch_copy_channels(n->stream, L[1].vox(), 32, n->enc1, 32, 0, n->cat1, 96, 64, false);
    conv(n, 5, L[1], n->cat1, n->dec1);
    launch_upsample(n->stream, L[1], 32, n->dec1, 48, n->cat0);
    launch_copy_channels(n->stream, L[0].vox(), 16, n->enc0, 16, 0, n->cat0, 48, 32, false);
    conv(n, 6, L[0], n->cat0, n->dec0);
    conv(n, 7, L[0], n->dec0, nullptr, n->u);
}

// backward from n->du; accumulates parameter gradients into n->grad
void backward(UnetNet *n) {
    const Grid3 *L = n->L;
    hipStream_t st = n->stream;
    hipLaunchKernelGGL(k_to_channel0, dim3(g1(L[0].vox() * 16)), dim3(256), 0, st, L[0].vox(), n->du,
                       reinterpret_cast<uint16_t *>(n->dout));
    wgrad(n, 7, L[0], n->dec0, n->dout);
    dgrad(n, 7, L[0], n->dout, n->d16);
    launch_relu_bwd(st, L[0].vox() * 16, n->dec0, n->d16);
    wgrad(n, 6, L[0], n->cat0, n->d16);
    dgrad(n, 6, L[0], n->d16, n->dcat0);
    launch_upsample_bwd(st, L[1], 32, n->dcat0, 48, n->dd1);
    launch_copy_channels(st, L[0].vox(), 16, n->dcat0, 48, 32, n->de0, 16, 0, false);
    launch_relu_bwd(st, L[1].vox() * 32, n->dec1, n->dd1);
    wgrad(n, 5, L[1], n->cat1, n->dd1);
    dgrad(n, 5, L[1], n->dd1, n->dcat1);
    launch_upsample_bwd(st, L[2], 64, n->dcat1, 96, n->dd2);
    launch_copy_channels(st, L[1].vox(), 32, n->dcat1, 96, 64, n->de1, 32, 0, false);
    launch_relu_bwd(st, L[2].vox() * 64, n->dec2, n->dd2);
    wgrad(n, 4, L[2], n->cat2, n->dd2);
    dgrad(n, 4, L[2], n->dd2, n->dcat2);
    launch_upsample_bwd(st, L[3], 128, n->dcat2, 192, n->dbott);
    launch_copy_channels(st, L[2].vox(), 64, n->dcat2, 192, 128, n->de2, 64, 0, false);
    launch_relu_bwd(st, L[3].vox() * 128, n->bott, n->dbott);
    wgrad(n, 3, L[3], n->p3, n->dbott);
    dgrad(n, 3, L[3], n->dbott, n->dp3);
    launch_pool_bwd(st, L[3], 64, n->dp3, n->tmp);
    launch_copy_channels(st, L[2].vox(), 64, n->tmp, 64, 0, n->de2, 64, 0, true);
    launch_relu_bwd(st, L[2].vox() * 64, n->enc2, n->de2);
    wgrad(n, 2, L[2], n->p2, n->de2);
    dgrad(n, 2, L[2], n->de2, n->dp2);
    launch_pool_bwd(st, L[2], 32, n->dp2, n->tmp);
    launch_copy_channels(st, L[1].vox(), 32, n->tmp, 32, 0, n->de1, 32, 0, true);
    launch_relu_bwd(st, L[1].vox() * 32, n->enc1, n->de1);
    wgrad(n, 1, L[1], n->p1, n->de1);
    dgrad(n, 1, L[1], n->de1, n->dp1);
    launch_pool_bwd(st, L[1], 16, n->dp1, n->tmp);
    launch_copy_channels(st, L[0].vox(), 16, n->tmp, 16, 0, n->de0, 16, 0, true);
    launch_relu_bwd(st, L[0].vox() * 16, n->enc0, n->de0);
    wgrad(n, 0, L[0], n->feat, n->de0);
}

// z = M(r) on the fine grid (fine dims multiples of 16); rms = RMS of r over the unknowns
// weight of smoothing step i of the dense application / training loss
double step_weight(const UnetNet *n, int i, double omega) {
    return n->smooth_cheb > 1.0 ? cheb_weight(i, n->smooth_steps, n->smooth_cheb) : omega;
}

int precondition(UnetNet *n, Grid3 fine, const float *r, const float *mask, const float *cx, const float *cy,
                 const float *cz, double omega, double rms, float *z, bool train = false) {
    int rc;
    const Grid3 m{fine.nx / 2, fine.ny / 2, fine.nz / 2};
    if ((rc = ensure_grid(n, m))) return rc;
    if ((rc = ensure_fine(n, fine.vox()))) return rc;
    if ((rc = repack(n))) return rc;
    const double s = rms > 0 ? rms : 1.0;
    launch_features(n->stream, m, r, 1.0 / s, mask, cx, cy, cz, n->feat);
    forward(n, train);
    launch_prolong(n->stream, fine, n->u, s, mask, z);
    for (int i = 0; i < n->smooth_steps; ++i) {
        launch_dense_stencil(n->stream, fine, cx, cy, cz, z, n->ftmp);
        launch_smooth(n->stream, fine.vox(), step_weight(n, i, omega), r, mask, n->ftmp, z);
    }
    UHIP(n, hipGetLastError());
    return SHK_OK;
}

}  // namespace

void unet_net_destroy(UnetNet *n) {
    if (!n) return;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    dfree(n->param);
    dfree(n->grad);
    dfree(n->m1);
    dfree(n->m2);
    for (int l = 0; l < kLayers; ++l) {
        dfree(n->pk[l]);
        dfree(n->pkt[l]);
        dfree(n->pkr[l]);
    }
    dfree(n->head_pk);
    bf16 **acts[] = {&n->feat, &n->enc0, &n->p1,   &n->enc1, &n->p2,  &n->enc2,  &n->p3,   &n->bott,
                     &n->cat2, &n->dec2, &n->cat1, &n->dec1, &n->cat0, &n->dec0, &n->dout, &n->d16,
                     &n->dcat0, &n->dd1, &n->dcat1, &n->dd2, &n->dcat2, &n->dbott, &n->dp3, &n->de2,
                     &n->dp2,  &n->de1,  &n->dp1,  &n->de0,  &n->tmp};
    for (auto p : acts) dfree(*p);
    dfree(n->u);
    dfree(n->du);
    dfree(n->fz);
    dfree(n->fe);
    dfree(n->fae);
    dfree(n->ftmp);
    dfree(n->part);
    if (n->h_part) (void)hipHostFree(n->h_part);
    delete n;
}

int unet_net_create(uint64_t seed, hipStream_t stream, UnetNet **out) {
    UnetNet *n = new UnetNet();
    n->stream = stream;
    int64_t off = 0;
    for (int l = 0; l < kLayers; ++l) {
        n->woff[l] = off;
        off += 27LL * kSpec[l].cin * kSpec[l].cout;
        n->boff[l] = off;
        off += kSpec[l].cout;
    }
    n->nparam = off;
    auto bail = [&](hipError_t e) {
        n->err = hipGetErrorString(e);
        unet_net_destroy(n);
        return e == hipErrorOutOfMemory ? SHK_OUT_OF_MEMORY : SHK_HIP_ERROR;
    };
    hipError_t e;
    for (float **p : {&n->param, &n->grad, &n->m1, &n->m2})
        if ((e = dalloc(*p, (size_t)off)) != hipSuccess) return bail(e);
    for (int l = 0; l < kLayers; ++l) {
        if ((e = dalloc(n->pk[l], (size_t)conv_packed_elems(kSpec[l].cin, kSpec[l].cout))) != hipSuccess) return bail(e);
        if (kSpec[l].need_dgrad &&
            (e = dalloc(n->pkt[l], (size_t)conv_packed_elems(kSpec[l].cout, kSpec[l].cin))) != hipSuccess)
            return bail(e);
        if (l < kLayers - 1 && (kSpec[l].cout == 16 || kSpec[l].cout == 32) &&
            (e = dalloc(n->pkr[l], (size_t)conv_rows_packed_elems(kSpec[l].cin, kSpec[l].cout))) != hipSuccess)
            return bail(e);
    }
    if ((e = dalloc(n->head_pk, 27 * 8)) != hipSuccess) return bail(e);
    if ((e = dalloc(n->part, 3 * kRedBlocks)) != hipSuccess) return bail(e);
    if ((e = hipHostMalloc((void **)&n->h_part, sizeof(double) * 3 * kRedBlocks, hipHostMallocDefault)) != hipSuccess)
        return bail(e);
    // He-normal initialisation from a seeded generator (host), biases 0
    std::vector<float> h(off, 0.0f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < kLayers; ++l) {
        const double sd = std::sqrt(2.0 / (27.0 * kSpec[l].cin)) * (l == kLayers - 1 ? 0.1 : 1.0);
        std::normal_distribution<double> nd(0.0, sd);
        for (int64_t i = 0; i < 27LL * kSpec[l].cin * kSpec[l].cout; ++i) h[n->woff[l] + i] = (float)nd(rng);
    }
    if ((e = hipMemcpy(n->param, h.data(), sizeof(float) * off, hipMemcpyHostToDevice)) != hipSuccess) return bail(e);
    for (float *p : {n->grad, n->m1, n->m2})
        if ((e = hipMemset(p, 0, sizeof(float) * off)) != hipSuccess) return bail(e);
    *out = n;
    return SHK_OK;
}

// One training sample: loss and gradients (accumulated into n->grad); with
// lr > 0 an Adam step follows.  Fine grid dims multiples of 16, device arrays.
int unet_net_train(UnetNet *n, Grid3 fine, const float *r, const float *mask, const float *cx, const float *cy,
                   const float *cz, const float *xstar, double omega, float lr, double *loss) {
    int rc;
    const int64_t nf = fine.vox();
    if ((rc = ensure_fine(n, nf))) return rc;
    n->omega = omega;
    double d3[3];
    // s = RMS of r over unknowns; den = x*' b
    if ((rc = reduce3(n, nf, r, r, xstar, r, mask, d3))) return rc;
    double count[3];
    if ((rc = reduce3(n, nf, mask, mask, nullptr, nullptr, nullptr, count))) return rc;
    const double rms = count[0] > 0 ? std::sqrt(d3[0] / count[0]) : 1.0;
    const double den = d3[1];
    if (!(den > 0.0)) return fail(n, SHK_BAD_ARG, "training sample has x*'b <= 0");
    if ((rc = precondition(n, fine, r, mask, cx, cy, cz, omega, rms, n->fz, true))) return rc;
    // e = z - x*, Ae, num = e'Ae
    hipLaunchKernelGGL(k_sub_f32, dim3(g1(nf)), dim3(256), 0, n->stream, nf, n->fz, xstar, n->fe);
    launch_dense_stencil(n->stream, fine, cx, cy, cz, n->fe, n->fae);
    if ((rc = reduce3(n, nf, n->fe, n->fae, nullptr, nullptr, nullptr, d3))) return rc;
    *loss = d3[0] / den;
    // dL/dz = 2 A e / den; through the smoothing steps, last first, g <- g - w_i A (mask g);
    // then through the prolongation (masked, times s)
    hipLaunchKernelGGL(k_scale_f32, dim3(g1(nf)), dim3(256), 0, n->stream, nf, (float)(2.0 / den), n->fae);
    for (int i = n->smooth_steps - 1; i >= 0; --i) {
        launch_mask_mul(n->stream, nf, mask, n->fae, n->fe);
        launch_dense_stencil(n->stream, fine, cx, cy, cz, n->fe, n->ftmp);
        hipLaunchKernelGGL(k_axpy_f32, dim3(g1(nf)), dim3(256), 0, n->stream, nf, (float)-step_weight(n, i, omega),
                           n->ftmp, n->fae);
    }
    launch_prolong_bwd(n->stream, n->L[0], n->fae, mask, rms, n->du);
    backward(n);
    if (lr > 0.0f) {
        n->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float bc1 = 1.0f - std::pow(b1, (float)n->step), bc2 = 1.0f - std::pow(b2, (float)n->step);
        launch_adam(n->stream, n->nparam, n->param, n->grad, n->m1, n->m2, lr, b1, b2, 1e-8f, bc1, bc2);
        n->packed_dirty = true;
    }
    UHIP(n, hipGetLastError());
    UHIP(n, hipStreamSynchronize(n->stream));
    return SHK_OK;
}

int unet_net_apply(UnetNet *n, Grid3 fine, const float *r, const float *mask, const float *cx, const float *cy,
                   const float *cz, double omega, float *z) {
    int rc;
    double d3[3], count[3];
    if ((rc = reduce3(n, fine.vox(), r, r, nullptr, nullptr, nullptr, d3))) return rc;
    if ((rc = reduce3(n, fine.vox(), mask, mask, nullptr, nullptr, nullptr, count))) return rc;
    const double rms = count[0] > 0 ? std::sqrt(d3[0] / count[0]) : 1.0;
    if ((rc = precondition(n, fine, r, mask, cx, cy, cz, omega, rms, z))) return rc;
    UHIP(n, hipStreamSynchronize(n->stream));
    return SHK_OK;
}

int unet_net_save(UnetNet *n, const char *path) {
    std::vector<float> h(n->nparam);
    UHIP(n, hipMemcpy(h.data(), n->param, sizeof(float) * n->nparam, hipMemcpyDeviceToHost));
    FILE *f = fopen(path, "wb");
    if (!f) return fail(n, SHK_BAD_ARG, "cannot open weights file for writing");
    const int32_t nl = kLayers;
    bool ok = fwrite(&kMagic, 8, 1, f) == 1 && fwrite(&nl, 4, 1, f) == 1;
    for (int l = 0; l < kLayers && ok; ++l) {
        const int32_t io[2] = {kSpec[l].cin, kSpec[l].cout};
        ok = fwrite(io, 4, 2, f) == 2;
    }
    ok = ok && fwrite(&n->nparam, 8, 1, f) == 1 && fwrite(&n->omega, 8, 1, f) == 1 &&
         fwrite(h.data(), 4, n->nparam, f) == (size_t)n->nparam;
    fclose(f);
    return ok ? SHK_OK : fail(n, SHK_BAD_ARG, "short write of weights file");
}

int unet_net_load(UnetNet *n, const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return fail(n, SHK_BAD_ARG, "cannot open weights file");
    uint64_t magic = 0;
    int32_t nl = 0;
    bool ok = fread(&magic, 8, 1, f) == 1 && magic == kMagic && fread(&nl, 4, 1, f) == 1 && nl == kLayers;
    for (int l = 0; l < kLayers && ok; ++l) {
        int32_t io[2];
        ok = fread(io, 4, 2, f) == 2 && io[0] == kSpec[l].cin && io[1] == kSpec[l].cout;
    }
    int64_t np = 0;
    double omega = 0.0;
    ok = ok && fread(&np, 8, 1, f) == 1 && np == n->nparam && fread(&omega, 8, 1, f) == 1 && omega >= 0.0 &&
         omega <= 2.0;
    std::vector<float> h(ok ? n->nparam : 0);
    ok = ok && fread(h.data(), 4, n->nparam, f) == (size_t)n->nparam;
    fclose(f);
    if (!ok) return fail(n, SHK_BAD_ARG, "weights file does not match this network");
    UHIP(n, hipMemcpy(n->param, h.data(), sizeof(float) * n->nparam, hipMemcpyHostToDevice));
    n->omega = omega;
    n->packed_dirty = true;
    return SHK_OK;
}

// ===================================================================== flexible CG in shk_project

struct UnetPrecond {
    UnetNet *net = nullptr;
    Grid3 pad{0, 0, 0};
    int64_t cap = 0;                    // brick elements the vectors below hold
    float *mask = nullptr, *z = nullptr;  // brick vectors: 1 on unknowns; z = M r
    float *p2 = nullptr;                  // the second search-direction buffer
    _Float16 *c16 = nullptr;              // FP16 copies of the couplings hx, hy, hz (3 x cap): the
                                          // solve's smoothing steps read them
    double *part1 = nullptr, *part2 = nullptr, *part3 = nullptr, *fst = nullptr;
    double *h_out = nullptr;            // pinned: part2 (2 x 1024) and fst (3)
    static constexpr int kRing = 2;     // inner-solve states in flight to the host
    FcgState *fs = nullptr;             // the inner solve's device-side state (two: before, after)
    FcgState *h_fs = nullptr;           // pinned, coherent: kRing copies of it, stored by the GPU
    hipEvent_t fev[kRing] = {};         // each copy's completion
};

void unet_destroy(UnetPrecond *u) {
    if (!u) return;
    unet_net_destroy(u->net);
    for (float **p : {&u->mask, &u->z, &u->p2}) dfree(*p);
    dfree(u->c16);
    for (double **p : {&u->part1, &u->part2, &u->part3, &u->fst}) dfree(*p);
    if (u->h_out) (void)hipHostFree(u->h_out);
    if (u->fs) (void)hipFree(u->fs);
    if (u->h_fs) (void)hipHostFree(u->h_fs);
    for (hipEvent_t e : u->fev)
        if (e) (void)hipEventDestroy(e);
    delete u;
}

namespace {

constexpr int kPartCap = 2 * kMaxBlocks;
static_assert(kPartCap >= kSmoothParts, "part3 holds the last smoothing step's partials");

__global__ void k_brick_mask(int nx, int ny, int nz, const int4 *__restrict__ coord, const uint8_t *__restrict__ label,
                             float *__restrict__ mask) {
    const int slot = blockIdx.x, l = threadIdx.x;
    const int4 c = coord[slot];
    const int i = c.x * 8 + (l & 7), j = c.y * 8 + ((l >> 3) & 7), k = c.z * 8 + (l >> 6);
    mask[(int64_t)slot * 512 + l] =
        (i < nx && j < ny && k < nz && label[((int64_t)k * ny + j) * nx + i]) ? 1.0f : 0.0f;
}

int ensure_precond(shk_ctx *c) {
    if (!c->unet) {
        if (c->prm.unet_weights.empty())
            return c->fail(SHK_BAD_ARG, "Preconditioner=unet needs UnetWeights=<path of a trained network>");
        UnetPrecond *u = new UnetPrecond();
        int rc = unet_net_create(0, c->stream, &u->net);
        if (rc) {
            delete u;
            return c->fail(rc, "cannot create the U-Net");
        }
        c->unet = u;
        if ((rc = unet_net_load(u->net, c->prm.unet_weights.c_str()))) {
            const std::string m = u->net->err;
            unet_destroy(u);
            c->unet = nullptr;
            return c->fail(rc, "UnetWeights: %s", m.c_str());
        }
        HIPC(c, dalloc(u->part1, (size_t)kPartCap));
        HIPC(c, dalloc(u->part2, (size_t)kPartCap));
        HIPC(c, dalloc(u->part3, (size_t)kPartCap));
        HIPC(c, dalloc(u->fst, 4));
        HIPC(c, hipHostMalloc((void **)&u->h_out, sizeof(double) * (kPartCap + 4), hipHostMallocDefault));
        HIPC(c, hipMalloc((void **)&u->fs, 2 * sizeof(FcgState)));
        HIPC(c, hipHostMalloc((void **)&u->h_fs, sizeof(FcgState) * UnetPrecond::kRing, hipHostMallocCoherent));
        for (hipEvent_t &e : u->fev) HIPC(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    UnetPrecond *u = c->unet;
    u->net->stream = c->stream;
    u->pad = Grid3{(c->dims.nx + 15) / 16 * 16, (c->dims.ny + 15) / 16 * 16, (c->dims.nz + 15) / 16 * 16};
    const int64_t n = (int64_t)c->n_active * kBrickCells;
    if (n > u->cap) {
        for (float **p : {&u->mask, &u->z, &u->p2}) HIPC(c, dalloc(*p, (size_t)n));
        HIPC(c, dalloc(u->c16, (size_t)(3 * n)));
        u->cap = n;
    }
    return SHK_OK;
}

BrickView brick_view(shk_ctx *c) {
    return BrickView{c->coord.p, c->brick_slot.p, c->nbr.p, c->n_active, c->nbx, c->nby,
                     c->dims.nx, c->dims.ny, c->dims.nz};
}

// z = M(r) on the projection's brick vectors: features of r (all channels when
// `full`, else the residual channel alone), the network on the half grid of the
// padded grid, prolongation into the bricks and UnetSmoothing damped-Jacobi steps
// (the training loss uses kSmooth; more steps at solve time cost one bandwidth-bound
// pass each and save flexible-CG iterations).  With
// q: the last step also writes the z.q partials (their count into *nz).  ctl (not
// with full): the inner solve's stop rules run first, and the input scale and the
// stop flag come from its state.  solver: r is the solve's own residual (zero off the
// unknowns), so the smoothing steps skip the mask, and they use the FP16 couplings
// (prepared at the solve's start).
int precondition_bricks(shk_ctx *c, const float *r, double rms, bool full, const float *q, int *nz,
                        const FcgCtl *ctl = nullptr, bool solver = false) {
    UnetPrecond *u = c->unet;
    UnetNet *n = u->net;
    int rc;
    const Grid3 P = u->pad;
    const Grid3 m{P.nx / 2, P.ny / 2, P.nz / 2};
    if ((rc = ensure_grid(n, m))) return c->fail(rc, "%s", n->err.c_str());
    if ((rc = repack(n))) return c->fail(rc, "%s", n->err.c_str());
    const BrickView bv = brick_view(c);
    const double s = rms > 0 ? rms : 1.0;
    const double *dscale = ctl ? &ctl->out->scale : nullptr;
    const int32_t *stop = ctl ? &ctl->out->stop : nullptr;
    launch_features_bricks(c->stream, m, bv, r, 1.0 / s, u->mask, c->hx32.p, c->hy32.p, c->hz32.p, !full, n->feat,
                           full ? nullptr : ctl);
    n->stop = stop;
    forward(n, false);
    n->stop = nullptr;
    // the steps alternate between z and t32: an odd count starts from t32, so that the
    // result lands in z
    const int steps = c->prm.unet_smooth;
    float *src = (steps & 1) ? c->t32.p : u->z, *dst = (steps & 1) ? u->z : c->t32.p;
    launch_prolong_bricks(c->stream, P, bv, n->u, s, u->mask, src, dscale, stop);
    for (int i = 0; i < steps; ++i) {
        const bool last = i + 1 == steps;
        const double w = c->prm.unet_cheb > 1.0 ? cheb_weight(i, steps, c->prm.unet_cheb) : n->omega;
        // inside the solver r and z are zero off the unknowns: the steps skip the mask and
        // read the FP16 copies of the couplings
        const int64_t nb = (int64_t)c->n_active * kBrickCells;
        const int blocks =
            solver ? launch_smooth_bricks(c->stream, bv, w, r, nullptr, u->c16, u->c16 + nb, u->c16 + 2 * nb, src,
                                          dst, last ? q : nullptr, last && q ? u->part3 : nullptr, stop)
                   : launch_smooth_bricks(c->stream, bv, w, r, u->mask, c->hx32.p, c->hy32.p, c->hz32.p, src, dst,
                                          last ? q : nullptr, last && q ? u->part3 : nullptr, stop);
        if (last && nz) *nz = blocks;
        std::swap(src, dst);
    }
    HIPC(c, hipGetLastError());
    return SHK_OK;
}

// sum and max of the 2 x blocks partials of the last launch_fcg_axpy, and fst
int read_norms(shk_ctx *c, int blocks, double &rr, double &rmax, double fst[3]) {
    UnetPrecond *u = c->unet;
    HIPC(c, hipMemcpyAsync(u->h_out, u->part2, sizeof(double) * 2 * blocks, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipMemcpyAsync(u->h_out + kPartCap, u->fst, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    rr = 0.0;
    rmax = 0.0;
    for (int i = 0; i < blocks; ++i) {
        rr += u->h_out[i];
        rmax = std::max(rmax, u->h_out[blocks + i]);
    }
    for (int i = 0; i < 3; ++i) fst[i] = u->h_out[kPartCap + i];
    return SHK_OK;
}

}  // namespace

// z = M(r) inside flexible CG; with ProfileKernels=1 the application is timed with
// events on the context's stream (read back at the next stream synchronisation).
int timed_precondition(shk_ctx *c, const float *r, double rms, bool full, const float *q, int *nz,
                       const FcgCtl *ctl = nullptr) {
    // called from the flexible-CG solve only: its vectors are zero off the unknowns
    if (c->profile && !c->uev[0]) {
        HIPC(c, hipEventCreate(&c->uev[0]));
        HIPC(c, hipEventCreate(&c->uev[1]));
    }
    if (c->profile) HIPC(c, hipEventRecord(c->uev[0], c->stream));
    const int rc = precondition_bricks(c, r, rms, full, q, nz, ctl, true);
    if (rc) return rc;
    if (c->profile) {
        HIPC(c, hipEventRecord(c->uev[1], c->stream));
        HIPC(c, hipEventSynchronize(c->uev[1]));
        float ms = 0.0f;
        HIPC(c, hipEventElapsedTime(&ms, c->uev[0], c->uev[1]));
        c->ucount += 1;
        c->ums += ms;
    }
    return SHK_OK;
}

// Learned flexible CG (FCG(1), Notay 2000) on the projection's brick vectors of the
// Jacobi-scaled system in FP32 with FP64 dots, inside the same FP64 refinement loop
// as the mixed solver.  The network runs on the half grid of the padded grid; every
// vector operation, the smoothing and the prolongation work on the bricks, and the
// scalars stay on the device: the stop rules and the network's input scale are
// applied by the first kernel of every network application (each of its blocks folds
// the update's partials; the state is double-buffered, read from one copy and written
// to the other), a stop makes every kernel of the following iterations return at
// once, and the host reads the state one iteration late, so the stream never waits
// for it.  Falls back to plain Jacobi-PCG inner solves when an
// inner FCG stalls (less than 10% residual reduction in 20 iterations) or hits its
// cap; the result records the fallback.
int unet_solve(shk_ctx *c, shk_result *res) {
    int rc;
    if ((rc = ensure_precond(c))) return rc;
    UnetPrecond *u = c->unet;
    const Dims d = c->dims;
    hipStream_t st = c->stream;
    const Grid &g = c->grid;
    hipLaunchKernelGGL(k_brick_mask, dim3(c->n_active), dim3(512), 0, st, d.nx, d.ny, d.nz, c->coord.p, c->label.p,
                       u->mask);
    {
        const int64_t nb = (int64_t)c->n_active * kBrickCells;
        launch_to_half(st, nb, c->hx32.p, u->c16);
        launch_to_half(st, nb, c->hy32.p, u->c16 + nb);
        launch_to_half(st, nb, c->hz32.p, u->c16 + 2 * nb);
    }
    HIPC(c, hipGetLastError());
    const double nunk = (double)c->unknowns;
    HIPC(c, hipMemsetAsync(c->x64.p, 0, sizeof(double) * g.n, st));
    double rr[4];
    if ((rc = true_residual(c, c->x64.p, c->r64.p, rr, c->sig64.p, c->rs32.p))) return rc;
    const double b_inf = rr[3];
    const double tol_abs = std::max(c->prm.residual, 1e-30) * b_inf;
    double r_inf = rr[0];
    int64_t total = 0;
    int outer = 0;
    int status = SHK_OK;
    bool fell_back = false;
    const int64_t fcg_cap = 400;
    const BrickOp<float> op{c->nbr.p, nullptr, c->hx32.p, c->hy32.p, c->hz32.p, c->n_active, c->kappa32.p};
    float *x = c->x32.p, *r = c->r32.p, *p = c->s32.p, *q = c->q32.p;
    while (r_inf > tol_abs && outer < std::max(1, c->prm.outer_max) + (fell_back ? 2 : 0)) {
        double inner_rel = std::max(c->prm.inner_residual, 0.5 * tol_abs / r_inf);
        if (inner_rel > 0.5) inner_rel = 0.5;
        ++outer;
        if (!fell_back) {
            // inner FCG on A^ d = r^ (bricks)
            HIPC(c, hipMemcpyAsync(r, c->rs32.p, sizeof(float) * g.n, hipMemcpyDeviceToDevice, st));
            HIPC(c, hipMemsetAsync(x, 0, sizeof(float) * g.n, st));
            double fst[3], r2, r0;
            launch_fcg_axpy(st, g, x, r, p, q, nullptr, u->part2, u->fst, true);
            if ((rc = read_norms(c, g.blocks, r2, r0, fst))) return rc;
            double rn = r0;
            if ((rc = timed_precondition(c, r, std::sqrt(r2 / nunk), true, nullptr, nullptr))) return rc;
            FcgState s0{};
            s0.scale = 1.0;
            s0.rn = rn;
            s0.target = inner_rel * r0;
            s0.best = rn;
            s0.cap = std::min<int64_t>(fcg_cap, c->prm.max_iterations - total);
            launch_fcg_state_init(st, u->fs, s0);
            int64_t it = 0;
            bool stalled = false;
            if (rn > s0.target && s0.cap > 0) {
                constexpr int R = UnetPrecond::kRing;
                int64_t enq = 0;
                // p = z, q = A p (the search directions alternate between two buffers: the
                // stencil reads the old direction's neighbours while the new one is written)
                float *P[2] = {p, u->p2};
                int cur = 0;
                launch_fcg_dir_spmv(st, g, op, u->z, u->z, P[0], q, FcgDir{u->fs, nullptr, 0, u->fst, r, u->part1});
                for (;;) {
                    // the state before this iteration's update, and after it
                    FcgState *fin = u->fs + (enq & 1), *fout = u->fs + ((enq + 1) & 1);
                    launch_fcg_axpy(st, g, x, r, P[cur], q, u->part1, u->part2, u->fst, false, fin);
                    const int slot = (int)(enq % R);
                    // the features kernel stores the state after this update in h_fs[slot] too
                    const FcgCtl ctl{u->part2, u->fst, g.blocks, nunk, fin, fout, u->h_fs + slot};
                    int nz = 0;
                    if ((rc = timed_precondition(c, r, 1.0, false, q, &nz, &ctl))) return rc;
                    // the next direction and its stencil product
                    launch_fcg_dir_spmv(st, g, op, u->z, P[cur], P[cur ^ 1], q,
                                        FcgDir{fout, u->part3, nz, u->fst, r, u->part1});
                    cur ^= 1;
                    HIPC(c, hipGetLastError());
                    HIPC(c, hipEventRecord(u->fev[slot], st));
                    ++enq;
                    // the state after the previous iteration (this one is still running)
                    if (enq >= R) {
                        const int prev = (int)((enq - R) % R);
                        HIPC(c, hipEventSynchronize(u->fev[prev]));
                        if (u->h_fs[prev].stop) break;
                    }
                    // every iteration either stops or counts: a device state that does
                    // neither within the cap is an error, not a reason to spin
                    if (enq > s0.cap + R) return c->fail(SHK_HIP_ERROR, "inner FCG state did not advance");
                }
                HIPC(c, hipStreamSynchronize(st));
                const FcgState &last = u->h_fs[(enq - 1) % R];
                it = last.it;
                stalled = last.stop == 2;
            }
            total += it;
            if (stalled || it >= fcg_cap) fell_back = true;
            // correction into x64 (x += sig * d)
            launch_add_correction(st, g.n, c->x64.p, x, c->sig64.p);
        } else {
            // plain Jacobi-PCG inner solve (the fallback), FP32 as in the mixed path
            const Reduction red = reduction(c);
            launch_cg_init<float, float>(st, g, c->rs32.p, nullptr, c->x32.p, c->r32.p, c->s32.p, red, inner_rel,
                                         c->prm.max_iterations - total);
            float *S[2] = {c->s32.p, c->t32.p};
            int m = 0;
            launch_dir_spmv<float, float>(st, g, op, c->r32.p, S[0], S[1], c->q32.p, red);
            ++m;
            rc = drive(c, [&]() {
                launch_axpy_norm<float>(st, g, nullptr, c->x32.p, c->r32.p, S[m & 1], c->q32.p, red, c->next_events());
                launch_dir_spmv<float, float>(st, g, op, c->r32.p, S[m & 1], S[(m + 1) & 1], c->q32.p, red, c->next_events());
                ++m;
            });
            if (rc) return rc;
            total += c->h_state->iter;
            launch_add_correction(st, g.n, c->x64.p, c->x32.p, c->sig64.p);
        }
        if ((rc = true_residual(c, c->x64.p, c->r64.p, rr, c->sig64.p, c->rs32.p))) return rc;
        r_inf = rr[0];
        if (!(r_inf == r_inf)) {
            status = SHK_NAN;
            break;
        }
        if (total >= c->prm.max_iterations) break;
    }
    if (status == SHK_OK && r_inf > tol_abs) status = SHK_NOT_CONVERGED;
    res->count = total;
    res->outer = outer;
    res->status = status;
    res->fallback = fell_back ? 1 : 0;
    res->reresid = b_inf > 0 ? r_inf / b_inf : 0.0;
    fill_true_residual(res, rr);
    return SHK_OK;
}

// ===================================================================== dense exports for training

int export_scaled(shk_ctx *c, float *rhat, float *mask, float *hx, float *hy, float *hz, float *xhat) {
    const Dims d = c->dims;
    const int64_t n = (int64_t)d.nx * d.ny * d.nz;
    hipStream_t st = c->stream;
    auto dense = [&](const double *src, float *dst) {
        hipLaunchKernelGGL((k_brick_to_dense<double, float>), dim3(g1(n)), dim3(256), 0, st, d.nx, d.ny, d.nz, d.nx,
                           d.ny, d.nz, c->brick_slot.p, c->nbx, c->nby, src, dst);
    };
    // temporaries in brick layout: b^ = sig*b, x^ = x/sig, couplings of the scaled system
    const int64_t ne = (int64_t)c->n_active * kBrickCells;
    double *tmp = nullptr, *th[3] = {nullptr, nullptr, nullptr};
    HIPC(c, dalloc(tmp, (size_t)ne));
    for (auto &t : th) HIPC(c, dalloc(t, (size_t)ne));
    launch_scale_coeff(st, c->dims.bx, c->n_active, c->nbr.p, c->sig64.p, c->cx64.p, c->cy64.p, c->cz64.p, th[0], th[1], th[2],
                       nullptr, nullptr, nullptr);
    if (hx) dense(th[0], hx);
    if (hy) dense(th[1], hy);
    if (hz) dense(th[2], hz);
    if (rhat) {
        hipLaunchKernelGGL(k_mul_sig, dim3(g1(ne)), dim3(256), 0, st, ne, c->b64.p, c->sig64.p, tmp, 0);
        dense(tmp, rhat);
    }
    if (xhat) {
        hipLaunchKernelGGL(k_mul_sig, dim3(g1(ne)), dim3(256), 0, st, ne, c->x64.p, c->sig64.p, tmp, 1);
        dense(tmp, xhat);
    }
    if (mask)
        hipLaunchKernelGGL(k_label_mask, dim3(g1(n)), dim3(256), 0, st, d.nx, d.ny, d.nz, d.nx, d.ny, d.nz,
                           c->label.p, mask);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(st));
    dfree(tmp);
    for (auto &t : th) dfree(t);
    return SHK_OK;
}

}  // namespace shk

// ===================================================================== C-ABI

struct shk_unet {
    shk::UnetNet *net = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;
    std::string err;
};

using namespace shk;

namespace {

int unet_fail(shk_unet *u, int code) {
    if (u && u->net) u->err = u->net->err;
    return code;
}

bool bad_fine(int nx, int ny, int nz) {
    return nx <= 0 || ny <= 0 || nz <= 0 || nx % 16 || ny % 16 || nz % 16;
}

}  // namespace

extern "C" {

int shk_unet_create(uint64_t seed, int device, shk_unet **out) {
    if (!out) return SHK_BAD_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SHK_NO_DEVICE;
    if (device < 0 && hipGetDevice(&device) != hipSuccess) device = 0;
    if (device >= ndev) return SHK_BAD_ARG;
    if (hipSetDevice(device) != hipSuccess) return SHK_HIP_ERROR;
    shk_unet *u = new shk_unet();
    u->device = device;
    if (hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking) != hipSuccess) {
        delete u;
        return SHK_HIP_ERROR;
    }
    const int rc = unet_net_create(seed, u->stream, &u->net);
    if (rc) {
        (void)hipStreamDestroy(u->stream);
        delete u;
        return rc;
    }
    *out = u;
    return SHK_OK;
}

void shk_unet_destroy(shk_unet *u) {
    if (!u) return;
    (void)hipSetDevice(u->device);
    unet_net_destroy(u->net);
    if (u->stream) (void)hipStreamDestroy(u->stream);
    delete u;
}

const char *shk_unet_last_error(const shk_unet *u) { return u ? u->err.c_str() : ""; }

int shk_unet_num_params(const shk_unet *u, int64_t *n) {
    if (!u || !n) return SHK_BAD_ARG;
    *n = u->net->nparam;
    return SHK_OK;
}

int shk_unet_get_params(shk_unet *u, float *host, int grads) {
    if (!u || !host) return SHK_BAD_ARG;
    (void)hipSetDevice(u->device);
    if (hipStreamSynchronize(u->stream) != hipSuccess) return SHK_HIP_ERROR;
    if (hipMemcpy(host, grads ? u->net->grad : u->net->param, sizeof(float) * u->net->nparam,
                  hipMemcpyDeviceToHost) != hipSuccess)
        return SHK_HIP_ERROR;
    return SHK_OK;
}

int shk_unet_set_params(shk_unet *u, const float *host) {
    if (!u || !host) return SHK_BAD_ARG;
    (void)hipSetDevice(u->device);
    if (hipMemcpy(u->net->param, host, sizeof(float) * u->net->nparam, hipMemcpyHostToDevice) != hipSuccess)
        return SHK_HIP_ERROR;
    if (hipMemset(u->net->grad, 0, sizeof(float) * u->net->nparam) != hipSuccess) return SHK_HIP_ERROR;
    u->net->packed_dirty = true;
    return SHK_OK;
}

int shk_unet_save(shk_unet *u, const char *path) {
    if (!u || !path) return SHK_BAD_ARG;
    (void)hipSetDevice(u->device);
    return unet_fail(u, unet_net_save(u->net, path));
}

int shk_unet_set_smoothing(shk_unet *u, int steps, double chebyshev) {
    if (!u || steps < 1 || steps > 64 || !(chebyshev == 0.0 || (chebyshev > 1.0 && chebyshev <= 1e6)))
        return SHK_BAD_ARG;
    u->net->smooth_steps = steps;
    u->net->smooth_cheb = chebyshev;
    return SHK_OK;
}

int shk_unet_load(shk_unet *u, const char *path) {
    if (!u || !path) return SHK_BAD_ARG;
    (void)hipSetDevice(u->device);
    return unet_fail(u, unet_net_load(u->net, path));
}

int shk_unet_train_step(shk_unet *u, int nx, int ny, int nz, const float *r, const float *mask, const float *cx,
                        const float *cy, const float *cz, const float *xstar, double omega, float lr, double *loss) {
    if (!u || !r || !mask || !cx || !cy || !cz || !xstar || !loss) return SHK_BAD_ARG;
    if (bad_fine(nx, ny, nz)) {
        u->err = "grid dims must be positive multiples of 16";
        return SHK_BAD_ARG;
    }
    (void)hipSetDevice(u->device);
    return unet_fail(u, unet_net_train(u->net, Grid3{nx, ny, nz}, r, mask, cx, cy, cz, xstar, omega, lr, loss));
}

int shk_unet_apply(shk_unet *u, int nx, int ny, int nz, const float *r, const float *mask, const float *cx,
                   const float *cy, const float *cz, double omega, float *z) {
    if (!u || !r || !mask || !cx || !cy || !cz || !z) return SHK_BAD_ARG;
    if (bad_fine(nx, ny, nz)) {
        u->err = "grid dims must be positive multiples of 16";
        return SHK_BAD_ARG;
    }
    (void)hipSetDevice(u->device);
    return unet_fail(u, unet_net_apply(u->net, Grid3{nx, ny, nz}, r, mask, cx, cy, cz, omega, z));
}

int shk_unet_residual(int nx, int ny, int nz, const float *cx, const float *cy, const float *cz, const float *x,
                      const float *b, float *r) {
    if (!cx || !cy || !cz || !x || !b || !r || nx <= 0 || ny <= 0 || nz <= 0) return SHK_BAD_ARG;
    const Grid3 g{nx, ny, nz};
    launch_dense_stencil(nullptr, g, cx, cy, cz, x, r);
    hipLaunchKernelGGL(k_rsub_f32, dim3(g1(g.vox())), dim3(256), 0, nullptr, g.vox(), b, r);
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipDeviceSynchronize();
    return (e1 == hipSuccess && e2 == hipSuccess) ? SHK_OK : SHK_HIP_ERROR;
}

int shk_unet_precondition(shk_ctx *c, const float *r, float *z) {
    if (!c || !r || !z) return SHK_BAD_ARG;
    if (!c->have_system || !c->unet || c->prm.precond != 2)
        return c->fail(SHK_BAD_ARG, "no learned preconditioner (project with Preconditioner=unet first)");
    if (c->prm.host_pointers) return c->fail(SHK_BAD_ARG, "shk_unet_precondition takes device pointers");
    HIPC(c, hipSetDevice(c->device));
    int rc;
    if ((rc = ensure_precond(c))) return rc;
    const Dims d = c->dims;
    const Grid &g = c->grid;
    hipStream_t st = c->stream;
    const int64_t nd = (int64_t)d.nx * d.ny * d.nz;
    hipLaunchKernelGGL(k_brick_mask, dim3(c->n_active), dim3(512), 0, st, d.nx, d.ny, d.nz, c->coord.p, c->label.p,
                       c->unet->mask);
    hipLaunchKernelGGL(k_dense_to_brick, dim3(g1(g.n)), dim3(256), 0, st, d.nx, d.ny, d.nz, d.nx, d.ny, c->coord.p,
                       c->n_active, r, c->r32.p);
    double rr, rmax, fst[3];
    launch_fcg_axpy(st, g, nullptr, c->r32.p, nullptr, nullptr, nullptr, c->unet->part2, c->unet->fst, true);
    if ((rc = read_norms(c, g.blocks, rr, rmax, fst))) return rc;
    const double rms = c->unknowns > 0 ? std::sqrt(rr / (double)c->unknowns) : 1.0;
    if ((rc = precondition_bricks(c, c->r32.p, rms, true, nullptr, nullptr))) return rc;
    hipLaunchKernelGGL((k_brick_to_dense<float, float>), dim3(g1(nd)), dim3(256), 0, st, d.nx, d.ny, d.nz, d.nx, d.ny,
                       d.nz, c->brick_slot.p, c->nbx, c->nby, c->unet->z, z);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(st));
    return SHK_OK;
}

int shk_export_scaled(shk_ctx *c, float *rhat, float *mask, float *hx, float *hy, float *hz, float *xhat) {
    if (!c) return SHK_BAD_ARG;
    if (!c->have_system) return c->fail(SHK_BAD_ARG, "no linear system (no projection with unknowns yet)");
    if (c->prm.host_pointers) return c->fail(SHK_BAD_ARG, "shk_export_scaled takes device pointers");
    if (c->dims.bx != 8) return c->fail(SHK_BAD_ARG, "shk_export_scaled needs a 3D context");
    HIPC(c, hipSetDevice(c->device));
    return export_scaled(c, rhat, mask, hx, hy, hz, xhat);
}

int shk_nn_conv3d(int nx, int ny, int nz, const void *x, int cin, const float *w, const float *bias, int cout,
                  int relu, void *y) {
    if (!x || !w || !y || nx <= 0 || ny <= 0 || nz <= 0 || cin <= 0 || cin % 8 || cout <= 0 || cout % 16 ||
        (cout / 16 != 1 && cout / 16 != 2 && cout / 16 != 3 && cout / 16 != 4 && cout / 16 != 6 && cout / 16 != 8 &&
         cout / 16 != 12))
        return SHK_BAD_ARG;
    bf16 *pk = nullptr;
    if (hipMalloc((void **)&pk, sizeof(bf16) * conv_packed_elems(cin, cout)) != hipSuccess) return SHK_OUT_OF_MEMORY;
    launch_conv_pack(nullptr, w, cin, cout, false, pk);
    launch_conv_fwd(nullptr, Grid3{nx, ny, nz}, static_cast<const bf16 *>(x), cin, pk, bias, cout, relu != 0,
                    static_cast<bf16 *>(y), nullptr);
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipDeviceSynchronize();
    (void)hipFree(pk);
    return (e1 == hipSuccess && e2 == hipSuccess) ? SHK_OK : SHK_HIP_ERROR;
}

int shk_nn_conv3d_cat(int nx, int ny, int nz, const void *coarse, int c1, const void *x, int cin, const float *w,
                      const float *bias, int cout, int relu, int kernel, void *y) {
    const int mt = cout / 16;
    if (!x || !w || !y || nx <= 0 || ny <= 0 || nz <= 0 || cin <= 0 || cin % 8 || cout <= 0 || cout % 16 ||
        kernel < 0 || kernel > 3 || (coarse && (c1 <= 0 || c1 % 8 || c1 >= cin || nx % 2 || ny % 2 || nz % 2)) ||
        (!coarse && c1 != 0))
        return SHK_BAD_ARG;
    const Grid3 g{nx, ny, nz};
    const bool rows = kernel == 2 || (kernel == 0 && conv_rows_ok(g, cin, cout));
    const bool direct = !rows && (kernel == 3 || (kernel == 0 && conv_direct_ok(g, cin, cout)));
    if (rows && cout != 16 && cout != 32) return SHK_BAD_ARG;
    if (!rows && !direct && mt != 1 && mt != 2 && mt != 3 && mt != 4 && mt != 6 && mt != 8 && mt != 12)
        return SHK_BAD_ARG;
    if (direct && g.vox() >= (1ll << 31)) return SHK_BAD_ARG;
    const int64_t elems = rows ? conv_rows_packed_elems(cin, cout) : conv_packed_elems(cin, cout);
    bf16 *pk = nullptr;
    if (hipMalloc((void **)&pk, sizeof(bf16) * elems) != hipSuccess) return SHK_OUT_OF_MEMORY;
    const bf16 *xc = static_cast<const bf16 *>(coarse);
    const bf16 *xx = static_cast<const bf16 *>(x);
    bf16 *yy = static_cast<bf16 *>(y);
    if (rows) {
        launch_conv_pack_rows(nullptr, w, cin, cout, pk);
        launch_conv_rows(nullptr, g, xx, cin, pk, bias, cout, relu != 0, yy, xc, c1);
    } else if (direct) {
        launch_conv_pack(nullptr, w, cin, cout, false, pk);
        launch_conv_direct(nullptr, g, xx, cin, pk, bias, cout, relu != 0, yy, xc, c1);
    } else {
        launch_conv_pack(nullptr, w, cin, cout, false, pk);
        if (xc)
            launch_conv_fwd_cat(nullptr, g, xc, c1, xx, cin, pk, bias, cout, relu != 0, yy);
        else
            launch_conv_fwd(nullptr, g, xx, cin, pk, bias, cout, relu != 0, yy, nullptr);
    }
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipDeviceSynchronize();
    (void)hipFree(pk);
    return (e1 == hipSuccess && e2 == hipSuccess) ? SHK_OK : SHK_HIP_ERROR;
}

int shk_nn_conv3d_backward(int nx, int ny, int nz, const void *x, int cin, const float *w, int cout, const void *dz,
                           void *dx, float *dw, float *db) {
    if (!x || !w || !dz || nx <= 0 || ny <= 0 || nz <= 0 || cin <= 0 || cin % 8 || cout <= 0 || cout % 16 ||
        cout > 256)
        return SHK_BAD_ARG;
    const Grid3 g{nx, ny, nz};
    if (dx) {
        if (cin % 16 || (cin / 16 != 1 && cin / 16 != 2 && cin / 16 != 3 && cin / 16 != 4 && cin / 16 != 6 &&
                         cin / 16 != 8 && cin / 16 != 12))
            return SHK_BAD_ARG;
        bf16 *pk = nullptr;
        if (hipMalloc((void **)&pk, sizeof(bf16) * conv_packed_elems(cout, cin)) != hipSuccess)
            return SHK_OUT_OF_MEMORY;
        launch_conv_pack(nullptr, w, cin, cout, true, pk);
        launch_conv_fwd(nullptr, g, static_cast<const bf16 *>(dz), cout, pk, nullptr, cin, false,
                        static_cast<bf16 *>(dx), nullptr);
        const hipError_t e = hipDeviceSynchronize();
        (void)hipFree(pk);
        if (e != hipSuccess) return SHK_HIP_ERROR;
    }
    if (dw) {
        launch_conv_wgrad(nullptr, g, static_cast<const bf16 *>(x), cin, static_cast<const bf16 *>(dz), cout, dw, db);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return SHK_HIP_ERROR;
    }
    return SHK_OK;
}

}  // extern "C"
