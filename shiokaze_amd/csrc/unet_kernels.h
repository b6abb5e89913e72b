// unet_kernels.h -- kernels of the learned preconditioner's 3D U-Net (gfx950).
//
// Activations are dense channels-last BF16 grids: element (voxel v, channel c) at
// v * C + c, voxel v = x + nx*(y + ny*z).  Convolutions are 3x3x3, stride 1, zero
// padding 1, computed as implicit GEMMs on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation over the 27 taps (tap = (dz+1)*9 + (dy+1)*3 + (dx+1)) and the input
// channels; Cin is a multiple of 8.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>

#include "shk_common.h"

namespace shk {

using bf16 = __hip_bfloat16;

struct Grid3 {
    int nx, ny, nz;
    __host__ __device__ int64_t vox() const { return (int64_t)nx * ny * nz; }
};

// The convolution stages its input through LDS CC channels at a time: CC is the
// largest divisor of cin that is a multiple of 8 and at most 32.  The reduction
// index runs stage by stage, k = (stage, tap, c), padded to whole chunks of 32.
inline int conv_cc(int cin) {
    for (int c = 32; c >= 8; c -= 8)
        if (cin % c == 0) return c;
    return 8;
}
// K chunks of 32 per stage, and in total
inline int conv_kpc(int cin) { return (27 * conv_cc(cin) + 31) / 32; }
inline int conv_kchunks(int cin) { return (cin / conv_cc(cin)) * conv_kpc(cin); }
// Packed weight elements (bf16) for cout outputs: cout/16 M-tiles x kchunks x 64 lanes x 8.
inline int64_t conv_packed_elems(int cin, int cout) { return (int64_t)(cout / 16) * conv_kchunks(cin) * 64 * 8; }

// Packs FP32 master weights W[tap][ci][co] into the MFMA A-fragment order:
// transposed = false: forward (A[m = co][k = (stage, tap, c)], ci = stage*CC + c);
// transposed = true : input gradient (A[m = ci][k = (stage, tap', c)] = W[26 - tap'][ci][co],
//                     co = stage*CC + c, CC = conv_cc(cout)).
void launch_conv_pack(hipStream_t st, const float *w, int cin, int cout, bool transposed, bf16 *packed);

// y[v][co] = act(bias[co] + sum_k w * x); act = ReLU when relu.  out_f32 (nullable)
// receives channel 0 in FP32 instead of bf16 output (the network's final layer).
void launch_conv_fwd(hipStream_t st, Grid3 g, const bf16 *x, int cin, const bf16 *packed, const float *bias, int cout,
                     bool relu, bf16 *y, float *out_f32, const int32_t *stop = nullptr);

// The same convolution of the concatenation [nearest upsampling of coarse (c1
// channels, grid g/2), skip (cin - c1 channels, grid g)], read from the two tensors
// directly; c1 and cin - c1 are multiples of 8.
void launch_conv_fwd_cat(hipStream_t st, Grid3 g, const bf16 *coarse, int c1, const bf16 *skip, int cin,
                         const bf16 *packed, const float *bias, int cout, bool relu, bf16 *y,
                         const int32_t *stop = nullptr);

// Direct-load forward convolution for grids too small for the tiled kernel to fill the
// chip (fewer than 1024 of its blocks): same packed weights as
// launch_conv_fwd, K split over the waves of a block and added in LDS, BF16 output
// (coarse != null: the input is [upsample(coarse) (c1 channels), x (cin - c1)]).
bool conv_direct_ok(Grid3 g, int cin, int cout);
void launch_conv_direct(hipStream_t st, Grid3 g, const bf16 *x, int cin, const bf16 *packed, const float *bias,
                        int cout, bool relu, bf16 *y, const bf16 *coarse = nullptr, int c1 = 0,
                        const int32_t *stop = nullptr);

// Row-reuse forward convolution for grids that fill the chip (inference).  The
// reduction runs by 32-channel stages of cin (the last one 8, 16 or 24 wide: G groups
// of 8 channels); within a stage, for each (dz, dy), the 3G entries (dx, group) in the
// order e = dx*G + group are cut into chunks of four (one per 16-lane quarter of the
// MFMA's K = 32, the last chunk padded with zero entries).  All entries of a chunk share
// (dz, dy), so one B fragment (a 16-voxel row) feeds the up-to-9 output rows it is a
// (dz, dy) neighbour of.
inline int conv_rows_chunks(int cin) {
    int q = 0;
    for (int ch0 = 0; ch0 < cin; ch0 += 32) {
        const int g = (cin - ch0) / 8 < 4 ? (cin - ch0) / 8 : 4;
        q += (3 * g + 3) / 4;
    }
    return q;
}
// Packed A fragments: [chunk][dz*3 + dy][16-output tile][lane][8] BF16.
inline int64_t conv_rows_packed_elems(int cin, int cout) {
    return (int64_t)conv_rows_chunks(cin) * 9 * (cout / 16) * 64 * 8;
}
void launch_conv_pack_rows(hipStream_t st, const float *w, int cin, int cout, bf16 *packed);
// Whether launch_conv_rows takes a layer: 1 or 2 output tiles of 16 channels, cin a
// multiple of 8 and at least 16, and at least 512 tiles of 16 x 4 x 8 voxels.
bool conv_rows_ok(Grid3 g, int cin, int cout);
// y = act(conv(x) + bias) as launch_conv_fwd (coarse != null: the input is
// [upsample(coarse) (c1 channels), x (cin - c1)] as in launch_conv_fwd_cat).
void launch_conv_rows(hipStream_t st, Grid3 g, const bf16 *x, int cin, const bf16 *packed, const float *bias,
                      int cout, bool relu, bf16 *y, const bf16 *coarse = nullptr, int c1 = 0,
                      const int32_t *stop = nullptr);

// y[v] = bias[0] + sum_{tap, ci} W[tap][ci][0] x[v + tap][ci] for a 16-channel BF16 input
// (the network's head, of which only output channel 0 is used); wpk from
// launch_head_pack (27 * 8 BF16 pairs)
void launch_conv_head(hipStream_t st, Grid3 g, const bf16 *x, const uint32_t *wpk, const float *bias, float *y,
                      const int32_t *stop = nullptr);
void launch_head_pack(hipStream_t st, const float *w, int cout, uint32_t *wpk);

// dw[tap][ci][co] += sum_v x[v + tap][ci] * dz[v][co]; db[co] += sum_v dz[v][co]
// (FP32 atomics, split over voxel chunks).
void launch_conv_wgrad(hipStream_t st, Grid3 g, const bf16 *x, int cin, const bf16 *dz, int cout, float *dw,
                       float *db);

// dz = dy * (y > 0) in place on dy (ReLU backward), n elements.
void launch_relu_bwd(hipStream_t st, int64_t n, const bf16 *y, bf16 *dy);
// coarse[V][c] = mean of the 8 fine children (fine grid is 2x in each axis).
void launch_pool(hipStream_t st, Grid3 coarse, int c, const bf16 *fine, bf16 *out, const int32_t *stop = nullptr);
// fine_grad[v][c] = coarse_grad[parent(v)][c] / 8 (pool backward).
void launch_pool_bwd(hipStream_t st, Grid3 coarse, int c, const bf16 *coarse_grad, bf16 *fine_grad);
// out[v][c_off + c] = coarse[parent(v)][c] for c < c_coarse (nearest upsampling into
// a concat buffer of cstride channels).
void launch_upsample(hipStream_t st, Grid3 cgrid, int c_coarse, const bf16 *coarse, int cstride, bf16 *out);
// coarse_grad[V][c] = sum over the 8 children of grad[child][c_off + c].
void launch_upsample_bwd(hipStream_t st, Grid3 coarse, int c_coarse, const bf16 *grad, int cstride, bf16 *coarse_grad);
// dst[v][d_off + c] = src[v][c], c < cs (copy a channel block between buffers);
// with add = true, dst += src.
void launch_copy_channels(hipStream_t st, int64_t vox, int cs, const bf16 *src, int sstride, int soff, bf16 *dst,
                          int dstride, int doff, bool add);
// out = a + b elementwise (bf16, n elements).
void launch_add_bf16(hipStream_t st, int64_t n, const bf16 *a, const bf16 *b, bf16 *out);

// Network input on the coarse grid (fine grid = 2x): 8 channels
//   [R(r)/scale, R(mask), R(cx), R(cy), R(cz), 0, 0, 0] with R = 2x2x2 average.
void launch_features(hipStream_t st, Grid3 coarse, const float *r, double inv_scale, const float *mask,
                     const float *cx, const float *cy, const float *cz, bf16 *feat);
// z = mask * scale * P u on the fine grid, P = cell-centred trilinear interpolation
// (weights 3/4, 1/4 per axis, clamped at the box edges).
void launch_prolong(hipStream_t st, Grid3 fine, const float *u, double scale, const float *mask, float *z);
// The same three steps on the projection's 8x8x8 brick vectors (brick coordinates
// coord[slot], slot table brick_slot over nbx x nby bricks, fine grid nx x ny x nz
// inside the padded grid of the network): every value is the one the dense kernels
// give on the padded grid, whose cells outside the bricks are zero.
struct BrickView {
    const int4 *coord;
    const int *slot;
    const int *nbr;
    int n_active, nbx, nby;
    int nx, ny, nz;
};
// only_r: write channel 0 (R(r)/scale) of the coarse voxels over active bricks and keep
// everything else (a full pass wrote the rest); with ctl (only_r), the inner solve's
// stop rules are applied first (fcg_decide) and the scale is the state's.  In the
// prolongation, dscale (nullable) is the scale read from device memory instead of the
// argument; stop (nullable, there and in the network's inference kernels): the kernel
// returns at once when *stop is set.
void launch_features_bricks(hipStream_t st, Grid3 coarse, BrickView b, const float *r, double inv_scale,
                            const float *mask, const float *cx, const float *cy, const float *cz, bool only_r,
                            bf16 *feat, const FcgCtl *ctl = nullptr);
void launch_prolong_bricks(hipStream_t st, Grid3 fine, BrickView b, const float *u, double scale, const float *mask,
                           float *z, const double *dscale = nullptr, const int32_t *stop = nullptr);
// zout = zin + omega * mask * (r - A zin), A the unit-diagonal operator with low-face
// couplings cx, cy, cz, evaluated in the order and rounding of the dense stencil and
// smoothing kernels.  mask == nullptr: 1 everywhere, for vectors that are zero off
// the unknowns with zin (there every coupling is zero too, so zout stays zero: the
// result is the masked one bit for bit).  With q and part: part[block] = sum zout * q (FP64), at most
// kSmoothParts blocks; returns the number of blocks.
constexpr int kSmoothParts = 2048;
int launch_smooth_bricks(hipStream_t st, BrickView b, double omega, const float *r, const float *mask,
                         const float *cx, const float *cy, const float *cz, const float *zin, float *zout,
                         const float *q, double *part, const int32_t *stop = nullptr);
// The same with FP16 copies of the couplings (a smoother's operator: it then differs
// from A by their rounding, 2^-11 relative).
int launch_smooth_bricks(hipStream_t st, BrickView b, double omega, const float *r, const float *mask,
                         const _Float16 *cx, const _Float16 *cy, const _Float16 *cz, const float *zin, float *zout,
                         const float *q, double *part, const int32_t *stop = nullptr);
// dst = (FP16) src, rounded to nearest, n elements.
void launch_to_half(hipStream_t st, int64_t n, const float *src, _Float16 *dst);

// grad_u = scale * P^T (mask * g).
void launch_prolong_bwd(hipStream_t st, Grid3 coarse, const float *g, const float *mask, double scale, float *grad_u);
// z += omega * mask * (r - t): one weighted-Jacobi step given t = A z.
void launch_smooth(hipStream_t st, int64_t n, double omega, const float *r, const float *mask, const float *t,
                   float *z);
// out = mask * g.
void launch_mask_mul(hipStream_t st, int64_t n, const float *mask, const float *g, float *out);

// Dense 7-point unit-diagonal stencil: y = x - sum c_face * x_nbr (couplings at
// the high cell of each face, zero on faces touching non-unknowns).
void launch_dense_stencil(hipStream_t st, Grid3 g, const float *cx, const float *cy, const float *cz, const float *x,
                          float *y);

// Adam step on n FP32 parameters with gradient g (zeroed afterwards).
void launch_adam(hipStream_t st, int64_t n, float *p, float *g, float *m, float *v, float lr, float b1, float b2,
                 float eps, float bc1, float bc2);

// float <-> bf16 conversion of a 1-channel grid into an 8-channel slot is handled by
// launch_features; this converts an FP32 array to bf16.
void launch_f32_to_bf16(hipStream_t st, int64_t n, const float *src, bf16 *dst);

}  // namespace shk
