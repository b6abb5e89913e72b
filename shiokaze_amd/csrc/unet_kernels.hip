// unet_kernels.hip -- BF16 MFMA 3D convolution (forward, input gradient, weight
// gradient) and the small kernels around it for the learned preconditioner.
//
// MFMA v_mfma_f32_16x16x32_bf16 (gfx950), one 64-lane wavefront per 16x16 tile:
//   A (16 x 32): lane l holds A[l % 16][8*(l/16) .. 8*(l/16)+7]
//   B (32 x 16): lane l holds B[8*(l/16) .. 8*(l/16)+7][l % 16]
//   D (16 x 16): lane l holds D[4*(l/16) + j][l % 16], j = 0..3
// A convolution is computed transposed, D[co][voxel] = sum_k W[co][k] X[k][voxel],
// so the B fragment of a lane is 8 consecutive channels of one (shifted) voxel
// (one 16-byte load from the channels-last grid) and the lane's results are 4
// consecutive output channels of one voxel (one 8-byte store).
#include "unet_kernels.h"

#include <type_traits>

#include "shk_common.h"

namespace shk {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

union Frag {
    uint4 u;
    bf16x8 v;
};

__device__ inline uint16_t bf16_bits(float f) {
    const bf16 h = __float2bfloat16(f);
    return *reinterpret_cast<const uint16_t *>(&h);
}
__device__ inline float bits_f32(uint16_t b) {
    return __uint_as_float((uint32_t)b << 16);
}

__device__ inline void tap_offset(int tap, int &dx, int &dy, int &dz) {
    dx = tap % 3 - 1;
    dy = (tap / 3) % 3 - 1;
    dz = tap / 9 - 1;
}

inline unsigned grid1(int64_t n, int nt) {
    int64_t g = (n + nt - 1) / nt;
    if (g > (1 << 22)) g = 1 << 22;
    return (unsigned)(g > 0 ? g : 1);
}

// ---------------------------------------------------------------- weight packing

__global__ void k_conv_pack(const float *__restrict__ w, int cin, int cout, int transposed, int cc, int kpc,
                            int kchunks, uint16_t *__restrict__ packed, int64_t total) {
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int j = (int)(idx & 7);
    const int lane = (int)((idx >> 3) & 63);
    const int64_t rest = idx >> 9;
    const int kcg = (int)(rest % kchunks);
    const int mt = (int)(rest / kchunks);
    const int stage = kcg / kpc, kc = kcg - stage * kpc;
    const int m = mt * 16 + (lane & 15);
    const int k = kc * 32 + 8 * (lane >> 4) + j;  // within the stage
    float val = 0.0f;
    if (k < 27 * cc) {
        const int tap = k / cc, c = stage * cc + (k - tap * cc);
        if (!transposed)
            val = w[((int64_t)tap * cin + c) * cout + m];             // A[co = m][ci = c]
        else
            val = w[((int64_t)(26 - tap) * cin + m) * cout + c];      // A[ci = m][co = c], flipped tap
    }
    packed[idx] = bf16_bits(val);
}

// ---------------------------------------------------------------- forward

// Block: 256 threads (4 waves) computing a 16 x 4 x 4 voxel tile (x, y, z) for MTB
// of the MT = cout/16 output tiles (blockIdx.z = z-tile * MT/MTB + output group; small
// grids split the outputs over blocks to fill the chip).  The 18 x 6 x 6 input halo
// of the tile is staged in LDS CC channels at a time (one 16-byte row per voxel and 8
// channels, the voxel stride padded to an odd number of 16-byte rows).  Wave w owns
// the tile's z-plane w: 4 rows of 16 voxels = 4 MFMA N-tiles, times MTB M-tiles of 16
// outputs, so each A fragment (weights, from L1/L2) feeds 4 MFMAs and each B fragment
// (inputs, from LDS) MTB.
constexpr int kTX = 16, kTY = 4, kTZ = 4;
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2;
constexpr int kHalo = kHX * kHY * kHZ;

// With xc: the input is the concatenation [upsample(xc), x] of the nearest upsampling
// of the half-size grid xc (its first c1 channels) and x (the other cin - c1), read
// straight from the two tensors while staging (no concatenated buffer).
template <int MTB, int CC>
__global__ __launch_bounds__(256) void k_conv_fwd(Grid3 g, const uint16_t *__restrict__ x, int cin, int mt_total,
                                                  const uint4 *__restrict__ packed, const float *__restrict__ bias,
                                                  int relu, uint16_t *__restrict__ y, float *__restrict__ out_f32,
                                                  const uint16_t *__restrict__ xc, int c1,
                                                  const int32_t *__restrict__ stop) {
    constexpr int G8 = CC / 8;        // 16-byte groups per voxel and stage
    // voxel stride in uint4, odd so that 16 lanes reading 16 consecutive voxels hit 16
    // disjoint 4-bank groups of the 64-bank LDS
    constexpr int VS = (G8 & 1) ? G8 : G8 + 1;
    constexpr int KPC = (27 * CC + 31) / 32;
    constexpr int PF = MTB <= 2 ? 4 : (MTB <= 4 ? 2 : 1);
    __shared__ uint4 lds[kHalo * VS];
    if (stop && *stop) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n = lane & 15, kq = lane >> 4;
    const int mgroups = mt_total / MTB;
    const int ztile = (int)blockIdx.z / mgroups, mg = (int)blockIdx.z - ztile * mgroups;
    const int x0 = blockIdx.x * kTX, y0 = blockIdx.y * kTY, z0 = ztile * kTZ;
    const int stages = cin / CC;
    const int kchunks = stages * KPC;
    f32x4 acc[MTB][4];
#pragma unroll
    for (int mt = 0; mt < MTB; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int stage = 0; stage < stages; ++stage) {
        __syncthreads();  // the previous stage's reads are done
        for (int idx = threadIdx.x; idx < kHalo * G8; idx += 256) {
            const int vox = idx / G8, part = idx - vox * G8;
            const int hx = vox % kHX, hy = (vox / kHX) % kHY, hz = vox / (kHX * kHY);
            const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny && gz >= 0 && gz < g.nz) {
                const int ch = stage * CC + part * 8;
                const int64_t vi = ((int64_t)gz * g.ny + gy) * g.nx + gx;
                const uint16_t *src;
                if (!xc) {
                    src = x + vi * cin + ch;
                } else if (ch < c1) {
                    const int64_t pv = ((int64_t)(gz >> 1) * (g.ny >> 1) + (gy >> 1)) * (g.nx >> 1) + (gx >> 1);
                    src = xc + pv * c1 + ch;
                } else {
                    src = x + vi * (cin - c1) + (ch - c1);
                }
                v = *reinterpret_cast<const uint4 *>(src);
            }
            lds[vox * VS + part] = v;
        }
        __syncthreads();
        // weight fragments run PF chunks ahead of the MFMAs that use them
        const uint4 *abase = packed + ((int64_t)(mg * MTB) * kchunks + (int64_t)stage * KPC) * 64 + lane;
        Frag ar[PF][MTB];
#pragma unroll
        for (int p = 0; p < PF; ++p)
#pragma unroll
            for (int mt = 0; mt < MTB; ++mt)
                if (p < KPC) ar[p][mt].u = abase[((int64_t)mt * kchunks + p) * 64];
#pragma unroll
        for (int kc = 0; kc < KPC; ++kc) {
            const int k0 = kc * 32 + kq * 8;
            Frag b[4];
            if (k0 < 27 * CC) {
                const int tap = k0 / CC, grp = (k0 - tap * CC) / 8;
                int dx, dy, dz;
                tap_offset(tap, dx, dy, dz);
                const int base = ((w + dz + 1) * kHY + (dy + 1)) * kHX + (n + dx + 1);
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) b[nt].u = lds[(base + nt * kHX) * VS + grp];
            } else {
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) b[nt].u = make_uint4(0u, 0u, 0u, 0u);
            }
            Frag a[MTB];
#pragma unroll
            for (int mt = 0; mt < MTB; ++mt) {
                a[mt] = ar[kc % PF][mt];
                if (kc + PF < KPC) ar[kc % PF][mt].u = abase[((int64_t)mt * kchunks + kc + PF) * 64];
            }
#pragma unroll
            for (int mt = 0; mt < MTB; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt].v, b[nt].v, acc[mt][nt], 0, 0, 0);
        }
    }
    // lane holds D[co = (mg*MTB + mt)*16 + 4*kq + j][voxel n of row nt]
    const int cout = mt_total * 16;
    const int gx = x0 + n, gz = z0 + w;
    if (gx >= g.nx || gz >= g.nz) return;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const int gy = y0 + nt;
        if (gy >= g.ny) continue;
        const int64_t vb = ((int64_t)gz * g.ny + gy) * g.nx + gx;
#pragma unroll
        for (int mt = 0; mt < MTB; ++mt) {
            const int co = (mg * MTB + mt) * 16 + 4 * kq;
            float o[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                o[j] = acc[mt][nt][j] + (bias ? bias[co + j] : 0.0f);
                if (relu) o[j] = o[j] > 0.0f ? o[j] : 0.0f;
            }
            if (out_f32) {
                if (co == 0) out_f32[vb] = o[0];
            } else {
                uint2 po;
                po.x = (uint32_t)bf16_bits(o[0]) | ((uint32_t)bf16_bits(o[1]) << 16);
                po.y = (uint32_t)bf16_bits(o[2]) | ((uint32_t)bf16_bits(o[3]) << 16);
                *reinterpret_cast<uint2 *>(y + vb * cout + co) = po;
            }
        }
    }
}

// ---------------------------------------------------------------- direct forward (small grids)

// For grids of 8^3 to 32^3, whose tiles cannot fill the chip: no LDS staging, and the
// K loop split over the waves of one block.  Block: S = blockDim.x / 64 waves, all on the
// same outputs, MW of the output tiles of 16 channels (blockIdx.y selects which) times
// NW N-tiles of 16 consecutive voxels in grid order (blockIdx.x, XCD-grouped so that
// neighbouring voxels, which share their inputs, run on one L2).  Wave w takes the
// weight chunks [w C / S, (w+1) C / S) of the C = stages x KPC chunks of the packed
// A fragments of launch_conv_pack (chunk (stage, kc), quarter kq: tap = k / CC and
// channel stage*CC + k % CC for k = 32 kc + 8 kq), and loads each B fragment (8
// channels of one shifted voxel per lane, one 16-byte load) straight from global
// memory, PF chunks ahead of its MFMAs.  The waves' FP32 partial sums meet in LDS and
// are added in wave order; bias, ReLU and the BF16 store follow.
template <int MW, int NW>
constexpr int direct_max_threads() {
    return 64 * (64 / (MW * NW) < 16 ? 64 / (MW * NW) : 16);
}

template <int MW, int NW, int CC>
__global__ __launch_bounds__((direct_max_threads<MW, NW>())) void k_conv_direct(
    Grid3 g, const uint16_t *__restrict__ x, int cin, int mt_total, const uint4 *__restrict__ packed,
    const float *__restrict__ bias, int relu, uint16_t *__restrict__ y, const uint16_t *__restrict__ xc, int c1,
    const int32_t *__restrict__ stop) {
    constexpr int KPC = (27 * CC + 31) / 32;
    constexpr int PF = MW + NW <= 2 ? 4 : (MW + NW <= 4 ? 3 : 2);
    constexpr int F = MW * NW;
    extern __shared__ f32x4 red[];  // [S][F][64]
    if (stop && *stop) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, S = blockDim.x >> 6;
    const int n = lane & 15, kq = lane >> 4;
    const int nb = xcd_slot((int)blockIdx.x, (int)gridDim.x), mg = blockIdx.y;
    const int nxy = g.nx * g.ny, vox = nxy * g.nz;
    const int kchunks = (cin / CC) * KPC;
    const int lo = (int)((int64_t)w * kchunks / S), hi = (int)((int64_t)(w + 1) * kchunks / S);
    // this lane's voxel in each N-tile
    int vx[NW], vy[NW], vz[NW], vv[NW];
#pragma unroll
    for (int t = 0; t < NW; ++t) {
        const int v = (nb * NW + t) * 16 + n;
        vv[t] = v < vox ? v : -1;
        const int vc = v < vox ? v : 0;
        vz[t] = vc / nxy;
        vy[t] = (vc - vz[t] * nxy) / g.nx;
        vx[t] = vc - vz[t] * nxy - vy[t] * g.nx;
    }
    const int hnx = g.nx >> 1, hny = g.ny >> 1;
    const int fstride = xc ? cin - c1 : cin;
    const uint4 *abase = packed + (int64_t)(mg * MW) * kchunks * 64 + lane;
    auto load = [&](int gc, Frag *a, Frag *b) {
#pragma unroll
        for (int m = 0; m < MW; ++m) a[m].u = abase[((int64_t)m * kchunks + gc) * 64];
        const int stage = gc / KPC, k = (gc - stage * KPC) * 32 + kq * 8;
        const int tap = k / CC, c = stage * CC + (k - tap * CC);
        const bool kin = k < 27 * CC;  // padding entries of a stage's last chunk: zero
        int dx, dy, dz;
        tap_offset(kin ? tap : 13, dx, dy, dz);
        const bool coarse = xc && c < c1;
#pragma unroll
        for (int t = 0; t < NW; ++t) {
            const int gx = vx[t] + dx, gy = vy[t] + dy, gz = vz[t] + dz;
            b[t].u = make_uint4(0u, 0u, 0u, 0u);
            if (kin && vv[t] >= 0 && (unsigned)gx < (unsigned)g.nx && (unsigned)gy < (unsigned)g.ny &&
                (unsigned)gz < (unsigned)g.nz) {
                const uint16_t *src =
                    coarse ? xc + (int64_t)(((gz >> 1) * hny + (gy >> 1)) * hnx + (gx >> 1)) * c1 + c
                           : x + (int64_t)(vv[t] + dz * nxy + dy * g.nx + dx) * fstride + (c - (xc ? c1 : 0));
                b[t].u = *reinterpret_cast<const uint4 *>(src);
            }
        }
    };
    f32x4 acc[MW][NW];
#pragma unroll
    for (int m = 0; m < MW; ++m)
#pragma unroll
        for (int t = 0; t < NW; ++t) acc[m][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    Frag fa[PF][MW], fb[PF][NW];
#pragma unroll
    for (int p = 0; p < PF; ++p)
        if (lo + p < hi) load(lo + p, fa[p], fb[p]);
    for (int base = lo; base < hi; base += PF) {
#pragma unroll
        for (int p = 0; p < PF; ++p) {
            const int gc = base + p;
            if (gc >= hi) break;  // wave-uniform
#pragma unroll
            for (int m = 0; m < MW; ++m)
#pragma unroll
                for (int t = 0; t < NW; ++t)
                    acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[p][m].v, fb[p][t].v, acc[m][t], 0, 0, 0);
            if (gc + PF < hi) load(gc + PF, fa[p], fb[p]);
        }
    }
#pragma unroll
    for (int m = 0; m < MW; ++m)
#pragma unroll
        for (int t = 0; t < NW; ++t) red[(w * F + m * NW + t) * 64 + lane] = acc[m][t];
    __syncthreads();
    // lane holds D[co = (mg*MW + m)*16 + 4*kq + j][voxel n of N-tile t]
    const int cout = mt_total * 16;
    for (int f = w; f < F; f += S) {
        const int m = f / NW, t = f - m * NW;
        f32x4 s = red[f * 64 + lane];
        for (int q = 1; q < S; ++q) s += red[(q * F + f) * 64 + lane];
        int v = -1;
#pragma unroll
        for (int u = 0; u < NW; ++u)
            if (u == t) v = vv[u];
        if (v < 0) continue;
        const int co = (mg * MW + m) * 16 + 4 * kq;
        float o[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            o[j] = s[j] + (bias ? bias[co + j] : 0.0f);
            if (relu) o[j] = o[j] > 0.0f ? o[j] : 0.0f;
        }
        uint2 po;
        po.x = (uint32_t)bf16_bits(o[0]) | ((uint32_t)bf16_bits(o[1]) << 16);
        po.y = (uint32_t)bf16_bits(o[2]) | ((uint32_t)bf16_bits(o[3]) << 16);
        *reinterpret_cast<uint2 *>(y + (int64_t)v * cout + co) = po;
    }
}

// ---------------------------------------------------------------- row-reuse forward

// i / g for the 1..4 channel groups of a stage, without an integer division
__device__ inline int div_groups(int i, int g) {
    return g == 4 ? i >> 2 : g == 2 ? i >> 1 : g == 1 ? i : (int)(((uint32_t)i * 0xAAABu) >> 17);
}

// Block: 256 threads (4 waves, 2 in y by 2 in z) computing a 16 x 2RY x 2RZ voxel tile
// for all MT output tiles of 16 channels; wave (wy, wz) owns RY x RZ rows of 16 voxels
// (one MFMA N-tile each).  A stage's input halo (18 x (2RY+2) x (2RZ+2) voxels, up to 4
// groups of 8 channels) sits in LDS group-major, lds[group * HP + voxel], so the 16
// lanes of a quarter read 16 consecutive voxels of one group (no bank conflicts).  For
// each chunk (four (dx, group) entries sharing (dz, dy)) the wave reads each of its
// (RY+2) x (RZ+2) input rows once and feeds it to the output rows it neighbours: 9 RY RZ
// MT MFMAs from (RY+2)(RZ+2) LDS reads.  A fragments (the chunk's 9 (dz, dy) taps x MT)
// come from global memory (L2: every block reads the same weights).  Tiles go to blocks
// XCD by XCD, so neighbouring tiles (which share halo rows) run on one L2.
//
// Staging: thread t always loads channel group t % 4 of the halo voxels t / 4 + 64 i, so
// its source tensor, channel offset and voxel list are fixed per stage and the address
// of each load is a few integer operations; all of a stage's loads are issued before
// any is written to LDS.
template <int MT, int RY, int RZ>
__global__ __launch_bounds__(256, 2) void k_conv_rows(Grid3 g, const uint16_t *__restrict__ x, int cin,
                                                      const uint4 *__restrict__ packed,
                                                      const float *__restrict__ bias, int relu,
                                                      uint16_t *__restrict__ y, const uint16_t *__restrict__ xc,
                                                      int c1, int tx, int ty, const int32_t *__restrict__ stop) {
    constexpr int TY = 2 * RY, TZ = 2 * RZ;
    constexpr int HX = 18, HY = TY + 2, HZ = TZ + 2;
    constexpr int HV = HX * HY * HZ;
    constexpr int HP = (HV + 11) / 16 * 16 + 4;  // >= HV, = 4 (mod 16)
    constexpr int NLD = (HV + 63) / 64;          // halo voxels per thread
    __shared__ uint4 lds[4 * HP];
    // element index of every halo voxel in x (fine) and in xc (its parent), -1 outside
    // the grid: computed once per block, read by every stage's loads
    __shared__ int off_f[HV], off_c[HV];
    if (stop && *stop) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n = lane & 15, kq = lane >> 4;
    const int wy = w & 1, wz = w >> 1;
    const int t = xcd_slot((int)blockIdx.x, (int)gridDim.x);
    const int x0 = (t % tx) * 16, y0 = ((t / tx) % ty) * TY, z0 = (t / (tx * ty)) * TZ;
    {
        const int hnx = g.nx >> 1, hny = g.ny >> 1;
        for (int vox = threadIdx.x; vox < HV; vox += 256) {
            const int hx = vox % HX, hy = (vox / HX) % HY, hz = vox / (HX * HY);
            const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
            const bool in = (unsigned)gx < (unsigned)g.nx && (unsigned)gy < (unsigned)g.ny && (unsigned)gz < (unsigned)g.nz;
            off_f[vox] = in ? (gz * g.ny + gy) * g.nx + gx : -1;
            off_c[vox] = in ? ((gz >> 1) * hny + (gy >> 1)) * hnx + (gx >> 1) : -1;
        }
    }
    f32x4 acc[RZ][RY][MT];
#pragma unroll
    for (int oz = 0; oz < RZ; ++oz)
#pragma unroll
        for (int oy = 0; oy < RY; ++oy)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) acc[oz][oy][mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    Frag a[9][MT];
    const int tgrp = threadIdx.x & 3, tv0 = threadIdx.x >> 2;
    uint4 st[NLD];
    int q = 0;
    for (int ch0 = 0; ch0 < cin; ch0 += 32) {
        const int G = (cin - ch0) / 8 < 4 ? (cin - ch0) / 8 : 4;
        // this thread's source for the stage: 8 channels of x, or of the coarse grid xc
        // read at the voxel's parent (the nearest upsampling)
        const int ch = ch0 + 8 * tgrp;
        const bool on = tgrp < G;
        const bool coarse = xc && ch < c1;
        const uint16_t *base = coarse ? xc + ch : (xc ? x + (ch - c1) : x + ch);
        const int64_t stride = coarse ? c1 : (xc ? cin - c1 : cin);
        if (ch0 == 0) __syncthreads();  // the offset tables are complete
        const int *offs = coarse ? off_c : off_f;
#pragma unroll
        for (int i = 0; i < NLD; ++i) {
            const int vox = tv0 + 64 * i;
            st[i] = make_uint4(0u, 0u, 0u, 0u);
            if (!on || vox >= HV) continue;
            const int o = offs[vox];
            if (o >= 0) st[i] = *reinterpret_cast<const uint4 *>(base + (int64_t)o * stride);
        }
        __syncthreads();  // the previous stage's reads are done
#pragma unroll
        for (int i = 0; i < NLD; ++i) {
            const int vox = tv0 + 64 * i;
            if (on && vox < HV) lds[tgrp * HP + vox] = st[i];
        }
        __syncthreads();
        const int nch = (3 * G + 3) / 4;
        for (int c = 0; c < nch; ++c, ++q) {
            const uint4 *cb = packed + (int64_t)q * 9 * MT * 64 + lane;
#pragma unroll
            for (int k = 0; k < 9; ++k)
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) a[k][mt].u = cb[(k * MT + mt) * 64];
            // this quarter's entry of the chunk: voxel offset dx, channel group grp
            const int e = 4 * c + kq;
            const bool valid = e < 3 * G;
            const int dx = valid ? div_groups(e, G) : 0;
            const int grp = valid ? e - dx * G : 0;
            const uint4 *lb = lds + grp * HP + ((wz * RZ) * HY + wy * RY) * HX + n + dx;
            auto rows = [&](auto masked) {
#pragma unroll
                for (int iz = 0; iz < RZ + 2; ++iz)
#pragma unroll
                    for (int iy = 0; iy < RY + 2; ++iy) {
                        Frag b;
                        b.u = lb[(iz * HY + iy) * HX];
                        if (decltype(masked)::value && !valid) b.u = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
                        for (int dz = 0; dz < 3; ++dz) {
                            const int oz = iz - dz;
                            if (oz < 0 || oz >= RZ) continue;
#pragma unroll
                            for (int dy = 0; dy < 3; ++dy) {
                                const int oy = iy - dy;
                                if (oy < 0 || oy >= RY) continue;
#pragma unroll
                                for (int mt = 0; mt < MT; ++mt)
                                    acc[oz][oy][mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                        a[dz * 3 + dy][mt].v, b.v, acc[oz][oy][mt], 0, 0, 0);
                            }
                        }
                    }
            };
            // only a stage's last chunk can hold padding entries (zero weights): zero
            // their inputs there, so that no Inf or NaN of a neighbour reaches a sum
            if (4 * c + 4 <= 3 * G)
                rows(std::false_type{});
            else
                rows(std::true_type{});
        }
    }
    // lane holds D[co = mt*16 + 4*kq + j][voxel n of row (oy, oz)]
    constexpr int cout = MT * 16;
    const int gx = x0 + n;
    if (gx >= g.nx) return;
#pragma unroll
    for (int oz = 0; oz < RZ; ++oz) {
        const int gz = z0 + wz * RZ + oz;
        if (gz >= g.nz) continue;
#pragma unroll
        for (int oy = 0; oy < RY; ++oy) {
            const int gy = y0 + wy * RY + oy;
            if (gy >= g.ny) continue;
            const int64_t vb = ((int64_t)gz * g.ny + gy) * g.nx + gx;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const int co = mt * 16 + 4 * kq;
                float o[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    o[j] = acc[oz][oy][mt][j] + (bias ? bias[co + j] : 0.0f);
                    if (relu) o[j] = o[j] > 0.0f ? o[j] : 0.0f;
                }
                uint2 po;
                po.x = (uint32_t)bf16_bits(o[0]) | ((uint32_t)bf16_bits(o[1]) << 16);
                po.y = (uint32_t)bf16_bits(o[2]) | ((uint32_t)bf16_bits(o[3]) << 16);
                *reinterpret_cast<uint2 *>(y + vb * cout + co) = po;
            }
        }
    }
}

// k_conv_rows for one output tile of 16 channels, persistent and software-pipelined:
// the grid holds as many blocks as fit on the chip at once, the tiles are cut into one
// contiguous range per XCD (blocks b, b + 8, ... share XCD b % 8 and take its tiles in
// turn), and within a chunk the B rows are read from LDS two rows ahead of the MFMAs
// that use them.  2 x WZ waves share a tile of 2 RY x WZ RZ rows, RY x RZ each: at
// WZ = 4 (2 x 2 rows a wave, 512 threads) two blocks per CU give 4 waves per SIMD
// within 128 VGPRs, so a chunk's A fragments are loaded at its start instead of a
// chunk ahead (measured 132.6 against 139.7 us for four waves of 2 x 4 rows with the
// prefetch, 48 -> 16 channels at 128^3).  Each output row sums its taps in
// k_conv_rows<1, RY, RZ>'s order.
template <int RY, int RZ, int WZ>
__global__ __launch_bounds__((128 * WZ), (WZ)) void k_conv_rows_pipe(Grid3 g, const uint16_t *__restrict__ x, int cin,
                                                           const uint4 *__restrict__ packed,
                                                           const float *__restrict__ bias, int relu,
                                                           uint16_t *__restrict__ y, const uint16_t *__restrict__ xc,
                                                           int c1, int tx, int ty, int ntiles,
                                                           const int32_t *__restrict__ stop) {
    constexpr int TY = 2 * RY, TZ = WZ * RZ;
    constexpr int HX = 18, HY = TY + 2, HZ = TZ + 2;
    constexpr int HV = HX * HY * HZ;
    constexpr int HP = (HV + 11) / 16 * 16 + 4;
    constexpr int VS = 32 * WZ;  // halo voxels staged per pass (4 threads each)
    constexpr int NLD = (HV + VS - 1) / VS;
    constexpr int NR = (RZ + 2) * (RY + 2);  // input rows of a wave per chunk
    __shared__ uint4 lds[4 * HP];
    if (stop && *stop) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n = lane & 15, kq = lane >> 4;
    const int wy = w & 1, wz = w >> 1;
    // this block's tiles: range [t_lo, t_hi) of its XCD, every nb-th from t_lo + idx
    const int xcd = (int)blockIdx.x % kXcds, idx = (int)blockIdx.x / kXcds;
    const int nb = ((int)gridDim.x - xcd + kXcds - 1) / kXcds;
    const int groups = (int)gridDim.x < kXcds ? (int)gridDim.x : kXcds;  // XCDs that hold blocks
    const int t_lo = (int)((int64_t)ntiles * xcd / groups), t_hi = (int)((int64_t)ntiles * (xcd + 1) / groups);
    const int stages = (cin + 31) / 32;
    const int my_tiles = t_lo + idx < t_hi ? (t_hi - t_lo - idx + nb - 1) / nb : 0;
    const int items = my_tiles * stages;
    const int tgrp = threadIdx.x & 3, tv0 = threadIdx.x >> 2;
    const int hnx = g.nx >> 1, hny = g.ny >> 1;
    f32x4 acc[RZ][RY];
#pragma unroll
    for (int oz = 0; oz < RZ; ++oz)
#pragma unroll
        for (int oy = 0; oy < RY; ++oy) acc[oz][oy] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    Frag a[9];
    int q = 0;  // chunk of the tile
    for (int i = 0; i < items; ++i) {
        const int t = t_lo + idx + (i / stages) * nb, s = i - (i / stages) * stages;
        const int x0 = (t % tx) * 16, y0 = ((t / tx) % ty) * TY, z0 = (t / (tx * ty)) * TZ;
        const int ch0 = 32 * s;
        const int G = (cin - ch0) / 8 < 4 ? (cin - ch0) / 8 : 4;
        if (s == 0) q = 0;
        {
            // this stage's halo: thread t loads channel group t % 4 of the halo voxels
            // t / 4 + 64 k, all loads issued before any is written to LDS
            const int ch = ch0 + 8 * tgrp;
            const bool on = tgrp < G;
            const bool coarse = xc && ch < c1;
            const uint16_t *base = coarse ? xc + ch : (xc ? x + (ch - c1) : x + ch);
            const int64_t stride = coarse ? c1 : (xc ? cin - c1 : cin);
            uint4 st[NLD];
#pragma unroll
            for (int k = 0; k < NLD; ++k) {
                const int vox = tv0 + VS * k;
                st[k] = make_uint4(0u, 0u, 0u, 0u);
                if (!on || vox >= HV) continue;
                const int hx = vox % HX, hy = (vox / HX) % HY, hz = vox / (HX * HY);
                const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
                if ((unsigned)gx < (unsigned)g.nx && (unsigned)gy < (unsigned)g.ny &&
                    (unsigned)gz < (unsigned)g.nz) {
                    const int o =
                        coarse ? ((gz >> 1) * hny + (gy >> 1)) * hnx + (gx >> 1) : (gz * g.ny + gy) * g.nx + gx;
                    st[k] = *reinterpret_cast<const uint4 *>(base + (int64_t)o * stride);
                }
            }
            __syncthreads();  // the previous item's reads are done
#pragma unroll
            for (int k = 0; k < NLD; ++k) {
                const int vox = tv0 + VS * k;
                if (on && vox < HV) lds[tgrp * HP + vox] = st[k];
            }
            __syncthreads();
        }
        const int nch = (3 * G + 3) / 4;
        for (int c = 0; c < nch; ++c, ++q) {
            // the next chunk's A fragments (chunk 0 again after the tile's last)
            const uint4 *cn = packed + (int64_t)q * 9 * 64 + lane;
#pragma unroll
            for (int k = 0; k < 9; ++k) a[k].u = cn[k * 64];
            const int e = 4 * c + kq;
            const bool valid = e < 3 * G;
            const int dx = valid ? div_groups(e, G) : 0;
            const int grp = valid ? e - dx * G : 0;
            const uint4 *lb = lds + grp * HP + ((wz * RZ) * HY + wy * RY) * HX + n + dx;
            auto rows = [&](auto masked) {
                Frag bq[3];
                bq[0].u = lb[0];
                bq[1].u = lb[(1 / (RY + 2) * HY + 1 % (RY + 2)) * HX];
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    const int iz = r / (RY + 2), iy = r % (RY + 2);
                    if (r + 2 < NR) {
                        const int r2 = r + 2;
                        bq[r2 % 3].u = lb[((r2 / (RY + 2)) * HY + r2 % (RY + 2)) * HX];
                    }
                    // keep the read above this row's MFMAs (the scheduler would sink it
                    // to its use, leaving the LDS latency uncovered)
                    __builtin_amdgcn_sched_barrier(0);
                    Frag b = bq[r % 3];
                    if (decltype(masked)::value && !valid) b.u = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
                    for (int dz = 0; dz < 3; ++dz) {
                        const int oz = iz - dz;
                        if (oz < 0 || oz >= RZ) continue;
#pragma unroll
                        for (int dy = 0; dy < 3; ++dy) {
                            const int oy = iy - dy;
                            if (oy < 0 || oy >= RY) continue;
                            acc[oz][oy] =
                                __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[dz * 3 + dy].v, b.v, acc[oz][oy], 0, 0, 0);
                        }
                    }
                }
            };
            if (4 * c + 4 <= 3 * G)
                rows(std::false_type{});
            else
                rows(std::true_type{});
        }
        if (s + 1 < stages) continue;
        // the tile is complete: lane holds D[co = 4*kq + j][voxel n of row (oy, oz)]
        const int gx = x0 + n;
#pragma unroll
        for (int oz = 0; oz < RZ; ++oz)
#pragma unroll
            for (int oy = 0; oy < RY; ++oy) {
                const int gy = y0 + wy * RY + oy, gz = z0 + wz * RZ + oz;
                if (gx < g.nx && gy < g.ny && gz < g.nz) {
                    const int co = 4 * kq;
                    float o[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        o[j] = acc[oz][oy][j] + (bias ? bias[co + j] : 0.0f);
                        if (relu) o[j] = o[j] > 0.0f ? o[j] : 0.0f;
                    }
                    uint2 po;
                    po.x = (uint32_t)bf16_bits(o[0]) | ((uint32_t)bf16_bits(o[1]) << 16);
                    po.y = (uint32_t)bf16_bits(o[2]) | ((uint32_t)bf16_bits(o[3]) << 16);
                    const int64_t vb = ((int64_t)gz * g.ny + gy) * g.nx + gx;
                    *reinterpret_cast<uint2 *>(y + vb * 16 + co) = po;
                }
                acc[oz][oy] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            }
    }
}

// A fragments of k_conv_rows: element idx = (((q * 9 + dz*3 + dy) * MT + mt) * 64 + lane) * 8 + j
// is W[tap = dz*9 + dy*3 + dx][ci][co] with co = mt*16 + lane % 16 and, for the chunk's
// entry e = 4c + lane/16 of its stage (ch0, G groups), dx = e / G and
// ci = ch0 + 8 (e % G) + j; 0 for the padding entries (e >= 3G).
__global__ void k_conv_pack_rows(const float *__restrict__ w, int cin, int cout, uint16_t *__restrict__ packed,
                                 int64_t total) {
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int mtn = cout / 16;
    const int j = (int)(idx & 7);
    const int lane = (int)((idx >> 3) & 63);
    int64_t rest = idx >> 9;
    const int mt = (int)(rest % mtn);
    rest /= mtn;
    const int t9 = (int)(rest % 9);
    int q = (int)(rest / 9);
    // the stage and chunk of global chunk q
    int ch0 = 0, G = 0, c = 0;
    for (ch0 = 0; ch0 < cin; ch0 += 32) {
        G = (cin - ch0) / 8 < 4 ? (cin - ch0) / 8 : 4;
        const int nch = (3 * G + 3) / 4;
        if (q < nch) {
            c = q;
            break;
        }
        q -= nch;
    }
    const int e = 4 * c + (lane >> 4);
    float val = 0.0f;
    if (ch0 < cin && e < 3 * G) {
        const int dx = e / G, grp = e - dx * G;
        const int tap = t9 * 3 + dx;  // (dz*3 + dy)*3 + dx
        const int ci = ch0 + 8 * grp + j;
        const int co = mt * 16 + (lane & 15);
        val = w[((int64_t)tap * cin + ci) * cout + co];
    }
    packed[idx] = bf16_bits(val);
}

// ---------------------------------------------------------------- head (16 -> 1 channel)

// The network's last convolution has one used output channel: y[v] = b + sum over the
// 27 taps and 16 input channels of W[tap][ci] x[v + tap][ci].  As an MFMA GEMM it would
// compute 16 output rows to keep one, so it runs on the vector ALUs with packed BF16
// channel-pair dot products (v_dot2c_f32_bf16, FP32 accumulation).  Block: 256 threads
// over a 16 x 8 x 8 voxel tile, each thread 4 consecutive x voxels; the 18 x 10 x 10
// input halo is staged in LDS as two arrays of 8 channels (one 16-byte row per voxel,
// rows of 19 voxels so that the 16 lanes of one LDS phase, 4 x-quads by 4 y rows, hit 16
// disjoint bank groups).  The weights, packed as BF16 pairs wpk[tap * 8 + pair], are
// wave-uniform and live in scalar registers.
constexpr int kHeadTX = 16, kHeadTY = 8, kHeadTZ = 8;
constexpr int kHeadHX = 19, kHeadHY = kHeadTY + 2, kHeadHZ = kHeadTZ + 2;
constexpr int kHeadRows = kHeadHX * kHeadHY * kHeadHZ;

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

__device__ inline float dot2(uint32_t a, uint32_t w, float acc) {
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, a), __builtin_bit_cast(bf16x2, w), acc, false);
}

__global__ __launch_bounds__(256) void k_conv_head(Grid3 g, const uint16_t *__restrict__ x,
                                                   const uint32_t *__restrict__ wpk, const float *__restrict__ bias,
                                                   float *__restrict__ y, const int32_t *__restrict__ stop) {
    __shared__ uint4 lds[2][kHeadRows];
    if (stop && *stop) return;
    const int t = threadIdx.x;
    const int x0 = blockIdx.x * kHeadTX, y0 = blockIdx.y * kHeadTY, z0 = blockIdx.z * kHeadTZ;
    // stage the halo (hx 0..17 used of the 19-voxel rows): thread t loads half t % 2 of
    // the halo voxels t / 2 + 128 i, all loads issued before any LDS write
    constexpr int kHV = 18 * kHeadHY * kHeadHZ, kNL = (kHV + 127) / 128;
    const int half = t & 1, v0 = t >> 1;
    uint4 st[kNL];
#pragma unroll
    for (int i = 0; i < kNL; ++i) {
        const int vox = v0 + 128 * i;
        const int hx = vox % 18, hy = (vox / 18) % kHeadHY, hz = vox / (18 * kHeadHY);
        const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
        st[i] = make_uint4(0u, 0u, 0u, 0u);
        if (vox < kHV && (unsigned)gx < (unsigned)g.nx && (unsigned)gy < (unsigned)g.ny && (unsigned)gz < (unsigned)g.nz)
            st[i] = *reinterpret_cast<const uint4 *>(x + (((int64_t)gz * g.ny + gy) * g.nx + gx) * 16 + half * 8);
    }
#pragma unroll
    for (int i = 0; i < kNL; ++i) {
        const int vox = v0 + 128 * i;
        const int hx = vox % 18, hy = (vox / 18) % kHeadHY, hz = vox / (18 * kHeadHY);
        if (vox < kHV) lds[half][(hz * kHeadHY + hy) * kHeadHX + hx] = st[i];
    }
    __syncthreads();
    const int xq = t & 3, ty = (t >> 2) & 7, tz = t >> 5;
    const float b0 = bias[0];
    float acc[4] = {b0, b0, b0, b0};
    // one (dz, dy) row of taps at a time keeps 12 staged rows live, not all 108
#pragma unroll 1
    for (int dzy = 0; dzy < 9; ++dzy) {
            const int dz = dzy / 3, dy = dzy - 3 * dz;
            const int row = ((tz + dz) * kHeadHY + (ty + dy)) * kHeadHX + 4 * xq;
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                uint4 in[6];
#pragma unroll
                for (int p = 0; p < 6; ++p) in[p] = lds[half][row + p];
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) {
                    const uint32_t *w = wpk + ((dz * 3 + dy) * 3 + dx) * 8 + half * 4;
                    const uint32_t w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const uint4 v = in[j + dx];
                        float a = acc[j];
                        a = dot2(v.x, w0, a);
                        a = dot2(v.y, w1, a);
                        a = dot2(v.z, w2, a);
                        a = dot2(v.w, w3, a);
                        acc[j] = a;
                    }
                }
            }
        }
    const int gy = y0 + ty, gz = z0 + tz;
    if (gy >= g.ny || gz >= g.nz) return;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int gx = x0 + 4 * xq + j;
        if (gx < g.nx) y[((int64_t)gz * g.ny + gy) * g.nx + gx] = acc[j];
    }
}

// wpk[tap * 8 + pair] = (bf16(W[tap][2 pair][0]), bf16(W[tap][2 pair + 1][0])) from the
// FP32 weights W[tap][ci][co] of a 16 -> cout convolution (output channel 0)
__global__ void k_head_pack(const float *__restrict__ w, int cout, uint32_t *__restrict__ wpk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 27 * 8) return;
    const int tap = i / 8, pair = i % 8;
    const float lo = w[((int64_t)tap * 16 + 2 * pair) * cout];
    const float hi = w[((int64_t)tap * 16 + 2 * pair + 1) * cout];
    wpk[i] = (uint32_t)bf16_bits(lo) | ((uint32_t)bf16_bits(hi) << 16);
}

// ---------------------------------------------------------------- weight gradient

// One wave per (voxel chunk, tap, 16x16 (co, ci) tile): D[co][ci] = sum over the
// chunk's voxels of dz[v][co] * x[v + tap][ci]; the tile is added with FP32 atomics.
constexpr int kWgradChunk = 512;  // voxels per wave (16 MFMA steps of 32)

__global__ __launch_bounds__(64) void k_conv_wgrad(Grid3 g, const uint16_t *__restrict__ x, int cin,
                                                   const uint16_t *__restrict__ dz, int cout, float *__restrict__ dw) {
    const int lane = threadIdx.x;
    const int n = lane & 15, kq = lane >> 4;
    const int64_t nvox = g.vox();
    const int64_t c0 = (int64_t)blockIdx.x * kWgradChunk;
    const int tap = blockIdx.y;
    const int ci_tiles = (cin + 15) / 16;
    const int co0 = (blockIdx.z / ci_tiles) * 16;
    const int ci0 = (blockIdx.z % ci_tiles) * 16;
    int dx, dy, dz_;
    tap_offset(tap, dx, dy, dz_);
    const int co = co0 + n;  // A row of this lane
    const int ci = ci0 + n;  // B column of this lane
    f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int step = 0; step < kWgradChunk; step += 32) {
        const int64_t vs = c0 + step + 8 * kq;
        Frag a, b;
        a.u = make_uint4(0u, 0u, 0u, 0u);
        b.u = make_uint4(0u, 0u, 0u, 0u);
        uint16_t av[8], bv[8];
        int64_t v = vs;
        int vx = (int)(v % g.nx), vy = (int)((v / g.nx) % g.ny), vz = (int)(v / ((int64_t)g.nx * g.ny));
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            av[j] = 0;
            bv[j] = 0;
            if (v < nvox) {
                av[j] = dz[v * cout + co];
                const int xx = vx + dx, yy = vy + dy, zz = vz + dz_;
                if (ci < cin && xx >= 0 && xx < g.nx && yy >= 0 && yy < g.ny && zz >= 0 && zz < g.nz)
                    bv[j] = x[(((int64_t)zz * g.ny + yy) * g.nx + xx) * cin + ci];
            }
            ++v;
            if (++vx == g.nx) {
                vx = 0;
                if (++vy == g.ny) {
                    vy = 0;
                    ++vz;
                }
            }
        }
        a.u.x = av[0] | ((uint32_t)av[1] << 16);
        a.u.y = av[2] | ((uint32_t)av[3] << 16);
        a.u.z = av[4] | ((uint32_t)av[5] << 16);
        a.u.w = av[6] | ((uint32_t)av[7] << 16);
        b.u.x = bv[0] | ((uint32_t)bv[1] << 16);
        b.u.y = bv[2] | ((uint32_t)bv[3] << 16);
        b.u.z = bv[4] | ((uint32_t)bv[5] << 16);
        b.u.w = bv[6] | ((uint32_t)bv[7] << 16);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc, 0, 0, 0);
    }
    // lane holds D[co0 + 4*kq + j][ci0 + n]
    if (ci < cin) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int cc = co0 + 4 * kq + j;
            atomicAdd(&dw[((int64_t)tap * cin + ci) * cout + cc], acc[j]);
        }
    }
}

__global__ __launch_bounds__(256) void k_bias_grad(int64_t nvox, const uint16_t *__restrict__ dz, int cout,
                                                   float *__restrict__ db) {
    // thread (block-local) t handles channel t % cout over a strided set of voxels
    const int c = threadIdx.x % cout;
    const int lanes_per_c = blockDim.x / cout;
    const int sub = threadIdx.x / cout;
    if (sub >= lanes_per_c) return;
    float acc = 0.0f;
    for (int64_t v = (int64_t)blockIdx.x * lanes_per_c + sub; v < nvox; v += (int64_t)gridDim.x * lanes_per_c)
        acc += bits_f32(dz[v * cout + c]);
    atomicAdd(&db[c], acc);
}

// ---------------------------------------------------------------- elementwise / resampling

__global__ void k_relu_bwd(int64_t n, const uint16_t *__restrict__ y, uint16_t *__restrict__ dy) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (!(bits_f32(y[i]) > 0.0f)) dy[i] = 0;
}

__device__ inline int64_t child(const Grid3 &c, int X, int Y, int Z, int a) {
    const int fx = 2 * c.nx, fy = 2 * c.ny;
    return ((int64_t)(2 * Z + (a >> 2)) * fy + (2 * Y + ((a >> 1) & 1))) * fx + (2 * X + (a & 1));
}

__global__ void k_pool(Grid3 c, int ch, const uint16_t *__restrict__ fine, uint16_t *__restrict__ out,
                       const int32_t *__restrict__ stop) {
    if (stop && *stop) return;
    const int64_t total = c.vox() * ch;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int cc = (int)(i % ch);
        const int64_t V = i / ch;
        const int X = (int)(V % c.nx), Y = (int)((V / c.nx) % c.ny), Z = (int)(V / ((int64_t)c.nx * c.ny));
        float s = 0.0f;
        for (int a = 0; a < 8; ++a) s += bits_f32(fine[child(c, X, Y, Z, a) * ch + cc]);
        out[i] = bf16_bits(0.125f * s);
    }
}

// k_pool for 8 channels per thread (16-byte loads and stores; the same per-channel sums)
__global__ void k_pool8(Grid3 c, int ch, const uint16_t *__restrict__ fine, uint16_t *__restrict__ out,
                        const int32_t *__restrict__ stop) {
    if (stop && *stop) return;
    const int groups = ch >> 3;
    const int64_t total = c.vox() * groups;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int cg = (int)(i % groups);
        const int64_t V = i / groups;
        const int X = (int)(V % c.nx), Y = (int)((V / c.nx) % c.ny), Z = (int)(V / ((int64_t)c.nx * c.ny));
        uint4 v[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) v[a] = *reinterpret_cast<const uint4 *>(fine + child(c, X, Y, Z, a) * ch + cg * 8);
        float s[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const uint32_t w[4] = {v[a].x, v[a].y, v[a].z, v[a].w};
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                s[2 * h] += bits_f32((uint16_t)(w[h] & 0xffffu));
                s[2 * h + 1] += bits_f32((uint16_t)(w[h] >> 16));
            }
        }
        uint4 o;
        o.x = (uint32_t)bf16_bits(0.125f * s[0]) | ((uint32_t)bf16_bits(0.125f * s[1]) << 16);
        o.y = (uint32_t)bf16_bits(0.125f * s[2]) | ((uint32_t)bf16_bits(0.125f * s[3]) << 16);
        o.z = (uint32_t)bf16_bits(0.125f * s[4]) | ((uint32_t)bf16_bits(0.125f * s[5]) << 16);
        o.w = (uint32_t)bf16_bits(0.125f * s[6]) | ((uint32_t)bf16_bits(0.125f * s[7]) << 16);
        *reinterpret_cast<uint4 *>(out + V * ch + cg * 8) = o;
    }
}

__global__ void k_pool_bwd(Grid3 c, int ch, const uint16_t *__restrict__ cg, uint16_t *__restrict__ fg) {
    const int64_t total = c.vox() * ch;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int cc = (int)(i % ch);
        const int64_t V = i / ch;
        const int X = (int)(V % c.nx), Y = (int)((V / c.nx) % c.ny), Z = (int)(V / ((int64_t)c.nx * c.ny));
        const uint16_t v = bf16_bits(0.125f * bits_f32(cg[i]));
        for (int a = 0; a < 8; ++a) fg[child(c, X, Y, Z, a) * ch + cc] = v;
    }
}

__global__ void k_upsample(Grid3 c, int cc_n, const uint16_t *__restrict__ coarse, int cstride,
                           uint16_t *__restrict__ out) {
    const int64_t total = c.vox() * cc_n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int cc = (int)(i % cc_n);
        const int64_t V = i / cc_n;
        const int X = (int)(V % c.nx), Y = (int)((V / c.nx) % c.ny), Z = (int)(V / ((int64_t)c.nx * c.ny));
        const uint16_t v = coarse[i];
        for (int a = 0; a < 8; ++a) out[child(c, X, Y, Z, a) * cstride + cc] = v;
    }
}

__global__ void k_upsample_bwd(Grid3 c, int cc_n, const uint16_t *__restrict__ grad, int cstride,
                               uint16_t *__restrict__ cgrad) {
    const int64_t total = c.vox() * cc_n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int cc = (int)(i % cc_n);
        const int64_t V = i / cc_n;
        const int X = (int)(V % c.nx), Y = (int)((V / c.nx) % c.ny), Z = (int)(V / ((int64_t)c.nx * c.ny));
        float s = 0.0f;
        for (int a = 0; a < 8; ++a) s += bits_f32(grad[child(c, X, Y, Z, a) * cstride + cc]);
        cgrad[i] = bf16_bits(s);
    }
}

__global__ void k_copy_channels(int64_t vox, int cs, const uint16_t *__restrict__ src, int sstride, int soff,
                                uint16_t *__restrict__ dst, int dstride, int doff, int add) {
    const int64_t total = vox * cs;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % cs);
        const int64_t v = i / cs;
        const uint16_t s = src[v * sstride + soff + c];
        uint16_t &d = dst[v * dstride + doff + c];
        d = add ? bf16_bits(bits_f32(d) + bits_f32(s)) : s;
    }
}

__global__ void k_add_bf16(int64_t n, const uint16_t *__restrict__ a, const uint16_t *__restrict__ b,
                           uint16_t *__restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = bf16_bits(bits_f32(a[i]) + bits_f32(b[i]));
}

__global__ void k_features(Grid3 c, const float *__restrict__ r, double inv_scale, const float *__restrict__ mask,
                           const float *__restrict__ cx, const float *__restrict__ cy, const float *__restrict__ cz,
                           uint16_t *__restrict__ feat) {
    const int64_t nv = c.vox();
    for (int64_t V = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; V < nv; V += (int64_t)gridDim.x * blockDim.x) {
        const int X = (int)(V % c.nx), Y = (int)((V / c.nx) % c.ny), Z = (int)(V / ((int64_t)c.nx * c.ny));
        float s[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        for (int a = 0; a < 8; ++a) {
            const int64_t f = child(c, X, Y, Z, a);
            s[0] += r[f];
            s[1] += mask[f];
            s[2] += cx[f];
            s[3] += cy[f];
            s[4] += cz[f];
        }
        uint4 o;
        o.x = (uint32_t)bf16_bits((float)(0.125 * s[0] * inv_scale)) | ((uint32_t)bf16_bits(0.125f * s[1]) << 16);
        o.y = (uint32_t)bf16_bits(0.125f * s[2]) | ((uint32_t)bf16_bits(0.125f * s[3]) << 16);
        o.z = (uint32_t)bf16_bits(0.125f * s[4]);
        o.w = 0u;
        *reinterpret_cast<uint4 *>(feat + V * 8) = o;
    }
}

// Cell-centred linear interpolation taps along one axis: fine index x has coarse taps
// x >> 1 (weight 3/4) and its neighbour on x's side (weight 1/4), clamped to [0, c).
__device__ inline void taps1(int x, int c, int &t0, int &t1) {
    t0 = x >> 1;
    t1 = t0 + ((x & 1) ? 1 : -1);
    t1 = t1 < 0 ? 0 : (t1 >= c ? c - 1 : t1);
}
// weight of coarse index X in the interpolation of fine index x
__device__ inline double tapw(int x, int X, int c) {
    int t0, t1;
    taps1(x, c, t0, t1);
    return (t0 == X ? 0.75 : 0.0) + (t1 == X ? 0.25 : 0.0);
}

__global__ void k_prolong(Grid3 f, const float *__restrict__ u, double scale, const float *__restrict__ mask,
                          float *__restrict__ z) {
    const int64_t nv = f.vox();
    const int cnx = f.nx / 2, cny = f.ny / 2, cnz = f.nz / 2;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(v % f.nx), y = (int)((v / f.nx) % f.ny), zc = (int)(v / ((int64_t)f.nx * f.ny));
        int X[2], Y[2], Z[2];
        taps1(x, cnx, X[0], X[1]);
        taps1(y, cny, Y[0], Y[1]);
        taps1(zc, cnz, Z[0], Z[1]);
        const double w[2] = {0.75, 0.25};
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int a = 0; a < 2; ++a)
                    acc += w[c] * w[b] * w[a] * (double)u[((int64_t)Z[c] * cny + Y[b]) * cnx + X[a]];
        z[v] = (float)((double)mask[v] * scale * acc);
    }
}

// grad_u[V] = scale * sum over the fine cells whose taps include V of weight * mask * g
__global__ void k_prolong_bwd(Grid3 c, const float *__restrict__ g, const float *__restrict__ mask, double scale,
                              float *__restrict__ gu) {
    const int64_t nv = c.vox();
    const int fx = 2 * c.nx, fy = 2 * c.ny, fz = 2 * c.nz;
    for (int64_t V = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; V < nv; V += (int64_t)gridDim.x * blockDim.x) {
        const int X = (int)(V % c.nx), Y = (int)((V / c.nx) % c.ny), Z = (int)(V / ((int64_t)c.nx * c.ny));
        double s = 0.0;
        for (int zz = 2 * Z - 1; zz <= 2 * Z + 2; ++zz) {
            if (zz < 0 || zz >= fz) continue;
            const double wz = tapw(zz, Z, c.nz);
            if (wz == 0.0) continue;
            for (int yy = 2 * Y - 1; yy <= 2 * Y + 2; ++yy) {
                if (yy < 0 || yy >= fy) continue;
                const double wy = tapw(yy, Y, c.ny);
                if (wy == 0.0) continue;
                for (int xx = 2 * X - 1; xx <= 2 * X + 2; ++xx) {
                    if (xx < 0 || xx >= fx) continue;
                    const double wx = tapw(xx, X, c.nx);
                    if (wx == 0.0) continue;
                    const int64_t f = ((int64_t)zz * fy + yy) * fx + xx;
                    s += wz * wy * wx * (double)mask[f] * (double)g[f];
                }
            }
        }
        gu[V] = (float)(scale * s);
    }
}

// z += omega * mask * (r - t)   (one weighted-Jacobi step with t = A z)
__global__ void k_smooth(int64_t n, double omega, const float *__restrict__ r, const float *__restrict__ mask,
                         const float *__restrict__ t, float *__restrict__ z) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        z[i] = (float)((double)z[i] + omega * (double)mask[i] * ((double)r[i] - (double)t[i]));
}

// out = mask * g
__global__ void k_mask_mul(int64_t n, const float *__restrict__ mask, const float *__restrict__ g,
                           float *__restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = mask[i] * g[i];
}

__global__ void k_dense_stencil(Grid3 g, const float *__restrict__ cx, const float *__restrict__ cy,
                                const float *__restrict__ cz, const float *__restrict__ x, float *__restrict__ y) {
    const int64_t nv = g.vox();
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(v % g.nx), j = (int)((v / g.nx) % g.ny), k = (int)(v / sz);
        float acc = x[v];
        if (i > 0) acc -= cx[v] * x[v - 1];
        if (i + 1 < g.nx) acc -= cx[v + 1] * x[v + 1];
        if (j > 0) acc -= cy[v] * x[v - sy];
        if (j + 1 < g.ny) acc -= cy[v + sy] * x[v + sy];
        if (k > 0) acc -= cz[v] * x[v - sz];
        if (k + 1 < g.nz) acc -= cz[v + sz] * x[v + sz];
        y[v] = acc;
    }
}

// ---------------------------------------------------------------- brick-layout preconditioner steps

// Brick element of fine cell (i, j, k), -1 outside the bricks or the fine grid.
__device__ inline int64_t brick_of(const BrickView &b, int i, int j, int k) {
    if (i >= b.nx || j >= b.ny || k >= b.nz) return -1;
    const int s = b.slot[(i >> 3) + (int64_t)b.nbx * ((j >> 3) + (int64_t)b.nby * (k >> 3))];
    if (s < 0) return -1;
    return (int64_t)s * 512 + (i & 7) + 8 * (j & 7) + 64 * (k & 7);
}

__global__ void k_features_bricks(Grid3 c, BrickView b, const float *__restrict__ r, double inv_scale,
                                  const double *__restrict__ dscale, const float *__restrict__ mask,
                                  const float *__restrict__ cx, const float *__restrict__ cy,
                                  const float *__restrict__ cz, int only_r, uint16_t *__restrict__ feat,
                                  const int32_t *__restrict__ stop) {
    if (stop && *stop) return;
    const int64_t nv = c.vox();
    if (dscale) inv_scale = 1.0 / *dscale;
    for (int64_t V = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; V < nv; V += (int64_t)gridDim.x * blockDim.x) {
        const int X = (int)(V % c.nx), Y = (int)((V / c.nx) % c.ny), Z = (int)(V / ((int64_t)c.nx * c.ny));
        // the 8 children lie in one brick (2x2x2 aligned in 8x8x8)
        const int64_t e0 = brick_of(b, 2 * X, 2 * Y, 2 * Z);
        float s[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (e0 >= 0) {
            // child a = (dx, dy, dz) bits in the dense kernel's order; +0.0f for a cell
            // outside the fine grid, as the dense padded grid holds there
            for (int a = 0; a < 8; ++a) {
                const int i = 2 * X + (a & 1), j = 2 * Y + ((a >> 1) & 1), k = 2 * Z + (a >> 2);
                const bool in = i < b.nx && j < b.ny && k < b.nz;
                const int64_t e = e0 + (a & 1) + 8 * ((a >> 1) & 1) + 64 * (a >> 2);
                s[0] += in ? r[e] : 0.0f;
                if (!only_r) {
                    s[1] += in ? mask[e] : 0.0f;
                    s[2] += in ? cx[e] : 0.0f;
                    s[3] += in ? cy[e] : 0.0f;
                    s[4] += in ? cz[e] : 0.0f;
                }
            }
        }
        const uint16_t f0 = bf16_bits((float)(0.125 * s[0] * inv_scale));
        if (only_r) {
            feat[V * 8] = f0;
        } else {
            uint4 o;
            o.x = (uint32_t)f0 | ((uint32_t)bf16_bits(0.125f * s[1]) << 16);
            o.y = (uint32_t)bf16_bits(0.125f * s[2]) | ((uint32_t)bf16_bits(0.125f * s[3]) << 16);
            o.z = (uint32_t)bf16_bits(0.125f * s[4]);
            o.w = 0u;
            *reinterpret_cast<uint4 *>(feat + V * 8) = o;
        }
    }
}

// Channel 0 of k_features_bricks for the coarse voxels whose children lie in an active
// brick: 8 bricks per 512-thread block, thread t -> coarse voxel t % 64 of brick t / 64.
// With ctl.in: every block first applies the inner solve's stop rules to the axpy's
// partials (one fixed fold order, so every block gets the same state; block 0 stores
// it in ctl.out) and takes the input scale from there.
__global__ __launch_bounds__(512) void k_features_bricks_r(Grid3 c, BrickView b, const float *__restrict__ r,
                                                           double inv_scale, FcgCtl ctl, uint16_t *__restrict__ feat) {
    __shared__ double lds[8];
    __shared__ double sc;
    __shared__ int stp;
    if (ctl.in) {
        const FcgState v0 = *ctl.in;
        if (v0.stop) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                *ctl.out = v0;
                if (ctl.host) *ctl.host = v0;
            }
            return;
        }
        const double r2 = reduce_partials_sum<512>(ctl.part2, ctl.nparts, lds);
        const double rn = reduce_partials_max<512>(ctl.part2 + ctl.nparts, ctl.nparts, lds);
        if (threadIdx.x == 0) {
            FcgState v = v0;
            fcg_decide(v, r2, rn, ctl.fst[2] != 0.0, ctl.nunk);
            if (blockIdx.x == 0) {
                *ctl.out = v;
                if (ctl.host) *ctl.host = v;
            }
            sc = v.scale;
            stp = v.stop;
        }
        __syncthreads();
        if (stp) return;
        inv_scale = 1.0 / sc;
    }
    const int slot = blockIdx.x * 8 + (threadIdx.x >> 6);
    if (slot >= b.n_active) return;
    const int v = threadIdx.x & 63, ci = v & 3, cj = (v >> 2) & 3, ck = v >> 4;
    const int4 bc = b.coord[slot];
    const int X = bc.x * 4 + ci, Y = bc.y * 4 + cj, Z = bc.z * 4 + ck;
    const int64_t e0 = (int64_t)slot * 512 + 2 * ci + 16 * cj + 128 * ck;
    float s0 = 0.0f;
    for (int a = 0; a < 8; ++a) {
        const int i = 2 * X + (a & 1), j = 2 * Y + ((a >> 1) & 1), k = 2 * Z + (a >> 2);
        const bool in = i < b.nx && j < b.ny && k < b.nz;
        s0 += in ? r[e0 + (a & 1) + 8 * ((a >> 1) & 1) + 64 * (a >> 2)] : 0.0f;
    }
    feat[(((int64_t)Z * c.ny + Y) * c.nx + X) * 8] = bf16_bits((float)(0.125 * s0 * inv_scale));
}

__global__ void k_prolong_bricks(Grid3 f, BrickView b, const float *__restrict__ u, double scale,
                                 const double *__restrict__ dscale, const float *__restrict__ mask,
                                 float *__restrict__ z, const int32_t *__restrict__ stop) {
    if (stop && *stop) return;
    const int slot = blockIdx.x, l = threadIdx.x;
    if (dscale) scale = *dscale;
    const int4 bc = b.coord[slot];
    const int x = bc.x * 8 + (l & 7), y = bc.y * 8 + ((l >> 3) & 7), zc = bc.z * 8 + (l >> 6);
    const int64_t e = (int64_t)slot * 512 + l;
    const int cnx = f.nx / 2, cny = f.ny / 2, cnz = f.nz / 2;
    int X[2], Y[2], Z[2];
    taps1(x, cnx, X[0], X[1]);
    taps1(y, cny, Y[0], Y[1]);
    taps1(zc, cnz, Z[0], Z[1]);
    const double w[2] = {0.75, 0.25};
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int a = 0; a < 2; ++a)
                acc += w[c] * w[bb] * w[a] * (double)u[((int64_t)Z[c] * cny + Y[bb]) * cnx + X[a]];
    z[e] = (float)((double)mask[e] * scale * acc);
}

// One thread per 4 consecutive x cells of a brick row (float4 accesses), 128 threads per
// brick, 4 bricks per 512-thread block.  Each cell's arithmetic is the dense stencil's:
// z - sum of coupling * neighbour in the order x-, x+, y-, y+, z-, z+, missing
// neighbours (no brick there) skipped, then the damped-Jacobi update in double.
// C: the couplings' type (float, or _Float16 copies inside the solve: the smoother's
// operator then differs from A by the couplings' FP16 rounding).
template <class C>
__global__ __launch_bounds__(512) void k_smooth_bricks(BrickView b, double omega, const float *__restrict__ r,
                                                       const float *__restrict__ mask, const C *__restrict__ cx,
                                                       const C *__restrict__ cy, const C *__restrict__ cz,
                                                       const float *__restrict__ zin, float *__restrict__ zout,
                                                       const float *__restrict__ q, double *__restrict__ part,
                                                       const int32_t *__restrict__ stop) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    typedef _Float16 h4 __attribute__((ext_vector_type(4)));
    __shared__ double lds[8];
    if (stop && *stop) return;
    const int l = threadIdx.x & 127, sub = threadIdx.x >> 7;
    const int h = l & 1, j = (l >> 1) & 7, k = l >> 4;
    const int off = 4 * h + 8 * j + 64 * k;  // first of the thread's 4 cells in its brick
    auto v4 = [](const float *p, int64_t e) { return *reinterpret_cast<const f4 *>(p + e); };
    auto c4 = [](const C *p, int64_t e) -> f4 {
        if constexpr (std::is_same<C, float>::value) {
            return *reinterpret_cast<const f4 *>(p + e);
        } else {
            const h4 v = *reinterpret_cast<const h4 *>(p + e);
            return f4{(float)v[0], (float)v[1], (float)v[2], (float)v[3]};
        }
    };
    double acc_zq = 0.0;
    for (int slot = blockIdx.x * 4 + sub; slot < b.n_active; slot += gridDim.x * 4) {
        const int *nb = b.nbr + 6 * slot;
        const int64_t e0 = (int64_t)slot * 512 + off;
        const f4 z = v4(zin, e0), cxv = c4(cx, e0), cyv = c4(cy, e0), czv = c4(cz, e0);
        // x: the neighbours outside the quad (the other quad, or the x neighbour brick)
        bool hxm = true, hxp = true;
        float zxm = 0.0f, zxp = 0.0f, cxp3 = 0.0f;
        if (h == 1) {
            zxm = zin[e0 - 1];
        } else if (nb[0] >= 0) {
            zxm = zin[(int64_t)nb[0] * 512 + 7 + 8 * j + 64 * k];
        } else {
            hxm = false;
        }
        if (h == 0) {
            zxp = zin[e0 + 4];
            cxp3 = (float)cx[e0 + 4];
        } else if (nb[1] >= 0) {
            const int64_t e = (int64_t)nb[1] * 512 + 8 * j + 64 * k;
            zxp = zin[e];
            cxp3 = (float)cx[e];
        } else {
            hxp = false;
        }
        // y and z: whole neighbouring rows
        bool hym = true, hyp = true, hzm = true, hzp = true;
        f4 zym = {0, 0, 0, 0}, zyp = {0, 0, 0, 0}, cyp = {0, 0, 0, 0};
        f4 zzm = {0, 0, 0, 0}, zzp = {0, 0, 0, 0}, czp = {0, 0, 0, 0};
        if (j > 0)
            zym = v4(zin, e0 - 8);
        else if (nb[2] >= 0)
            zym = v4(zin, (int64_t)nb[2] * 512 + off + 56);
        else
            hym = false;
        if (j < 7) {
            zyp = v4(zin, e0 + 8);
            cyp = c4(cy, e0 + 8);
        } else if (nb[3] >= 0) {
            const int64_t e = (int64_t)nb[3] * 512 + off - 56;
            zyp = v4(zin, e);
            cyp = c4(cy, e);
        } else {
            hyp = false;
        }
        if (k > 0)
            zzm = v4(zin, e0 - 64);
        else if (nb[4] >= 0)
            zzm = v4(zin, (int64_t)nb[4] * 512 + off + 448);
        else
            hzm = false;
        if (k < 7) {
            zzp = v4(zin, e0 + 64);
            czp = c4(cz, e0 + 64);
        } else if (nb[5] >= 0) {
            const int64_t e = (int64_t)nb[5] * 512 + off - 448;
            zzp = v4(zin, e);
            czp = c4(cz, e);
        } else {
            hzp = false;
        }
        const f4 rv = v4(r, e0), mv = mask ? v4(mask, e0) : f4{1.0f, 1.0f, 1.0f, 1.0f};
        f4 zn;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            // A z in the dense stencil's order: x-, x+, y-, y+, z-, z+
            float acc = z[u];
            if (u > 0) {
                acc -= cxv[u] * z[u - 1];
            } else if (hxm) {
                acc -= cxv[0] * zxm;
            }
            if (u < 3) {
                acc -= cxv[u + 1] * z[u + 1];
            } else if (hxp) {
                acc -= cxp3 * zxp;
            }
            if (hym) acc -= cyv[u] * zym[u];
            if (hyp) acc -= cyp[u] * zyp[u];
            if (hzm) acc -= czv[u] * zzm[u];
            if (hzp) acc -= czp[u] * zzp[u];
            zn[u] = (float)((double)z[u] + omega * (double)mv[u] * ((double)rv[u] - (double)acc));
        }
        *reinterpret_cast<f4 *>(zout + e0) = zn;
        if (part) {
            const f4 qv = v4(q, e0);
#pragma unroll
            for (int u = 0; u < 4; ++u) acc_zq += (double)zn[u] * (double)qv[u];
        }
    }
    if (part) {
        const double t = block_sum<512>(acc_zq, lds);
        if (threadIdx.x == 0) part[blockIdx.x] = t;
    }
}

__global__ void k_adam(int64_t n, float *__restrict__ p, float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
        g[i] = 0.0f;
    }
}

__global__ void k_f32_to_bf16(int64_t n, const float *__restrict__ s, uint16_t *__restrict__ d) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = bf16_bits(s[i]);
}

inline const uint16_t *U(const bf16 *p) { return reinterpret_cast<const uint16_t *>(p); }
inline uint16_t *U(bf16 *p) { return reinterpret_cast<uint16_t *>(p); }

}  // namespace

// ---------------------------------------------------------------- launchers

void launch_features_bricks(hipStream_t st, Grid3 coarse, BrickView b, const float *r, double inv_scale,
                            const float *mask, const float *cx, const float *cy, const float *cz, bool only_r,
                            bf16 *feat, const FcgCtl *ctl) {
    if (only_r) {
        // channel 0 of the coarse voxels over the active bricks (4^3 per brick); the
        // others hold the zero a full pass wrote there and keep it
        const FcgCtl none{nullptr, nullptr, 0, 1.0, nullptr, nullptr, nullptr};
        if (b.n_active > 0)
            hipLaunchKernelGGL(k_features_bricks_r, dim3((unsigned)((b.n_active + 7) / 8)), dim3(512), 0, st, coarse,
                               b, r, inv_scale, ctl ? *ctl : none, U(feat));
        return;
    }
    hipLaunchKernelGGL(k_features_bricks, dim3(grid1(coarse.vox(), 256)), dim3(256), 0, st, coarse, b, r, inv_scale,
                       nullptr, mask, cx, cy, cz, 0, U(feat), nullptr);
}

void launch_prolong_bricks(hipStream_t st, Grid3 fine, BrickView b, const float *u, double scale, const float *mask,
                           float *z, const double *dscale, const int32_t *stop) {
    if (b.n_active <= 0) return;
    hipLaunchKernelGGL(k_prolong_bricks, dim3(b.n_active), dim3(512), 0, st, fine, b, u, scale, dscale, mask, z, stop);
}

namespace {
template <class C>
int smooth_bricks(hipStream_t st, BrickView b, double omega, const float *r, const float *mask, const C *cx,
                  const C *cy, const C *cz, const float *zin, float *zout, const float *q, double *part,
                  const int32_t *stop) {
    // 4 bricks per block and pass; with part, at most kSmoothParts blocks (partials)
    const int want = (b.n_active + 3) / 4;
    const int blocks = want < 1 ? 1 : (part && want > kSmoothParts ? kSmoothParts : want);
    hipLaunchKernelGGL(k_smooth_bricks<C>, dim3(blocks), dim3(512), 0, st, b, omega, r, mask, cx, cy, cz, zin, zout,
                       q, part, stop);
    return blocks;
}
}  // namespace

int launch_smooth_bricks(hipStream_t st, BrickView b, double omega, const float *r, const float *mask,
                         const float *cx, const float *cy, const float *cz, const float *zin, float *zout,
                         const float *q, double *part, const int32_t *stop) {
    return smooth_bricks<float>(st, b, omega, r, mask, cx, cy, cz, zin, zout, q, part, stop);
}

int launch_smooth_bricks(hipStream_t st, BrickView b, double omega, const float *r, const float *mask,
                         const _Float16 *cx, const _Float16 *cy, const _Float16 *cz, const float *zin, float *zout,
                         const float *q, double *part, const int32_t *stop) {
    return smooth_bricks<_Float16>(st, b, omega, r, mask, cx, cy, cz, zin, zout, q, part, stop);
}

__global__ void k_to_half(int64_t n, const float *__restrict__ src, _Float16 *__restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (_Float16)src[i];
}

void launch_to_half(hipStream_t st, int64_t n, const float *src, _Float16 *dst) {
    hipLaunchKernelGGL(k_to_half, dim3(grid1(n, 256)), dim3(256), 0, st, n, src, dst);
}

void launch_conv_pack(hipStream_t st, const float *w, int cin, int cout, bool transposed, bf16 *packed) {
    const int gin = transposed ? cout : cin;  // GEMM reduction channels
    const int kchunks = conv_kchunks(gin);
    const int mtiles = transposed ? cin / 16 : cout / 16;
    const int64_t total = (int64_t)mtiles * kchunks * 64 * 8;
    hipLaunchKernelGGL(k_conv_pack, dim3(grid1(total, 256)), dim3(256), 0, st, w, cin, cout, transposed ? 1 : 0,
                       conv_cc(gin), conv_kpc(gin), kchunks, U(packed), total);
}

void launch_conv_head(hipStream_t st, Grid3 g, const bf16 *x, const uint32_t *wpk, const float *bias, float *y,
                      const int32_t *stop) {
    const dim3 grid((unsigned)((g.nx + kHeadTX - 1) / kHeadTX), (unsigned)((g.ny + kHeadTY - 1) / kHeadTY),
                    (unsigned)((g.nz + kHeadTZ - 1) / kHeadTZ));
    hipLaunchKernelGGL(k_conv_head, grid, dim3(256), 0, st, g, U(x), wpk, bias, y, stop);
}

void launch_head_pack(hipStream_t st, const float *w, int cout, uint32_t *wpk) {
    hipLaunchKernelGGL(k_head_pack, dim3(1), dim3(256), 0, st, w, cout, wpk);
}

namespace {
template <int MTB>
void conv_fwd_mt(hipStream_t st, dim3 grid, Grid3 g, const bf16 *x, int cin, int mt, const uint4 *pk,
                 const float *bias, bool relu, uint16_t *yy, float *out_f32, const uint16_t *xc, int c1,
                 const int32_t *stop) {
    switch (conv_cc(cin)) {
        case 8:
            hipLaunchKernelGGL((k_conv_fwd<MTB, 8>), grid, dim3(256), 0, st, g, U(x), cin, mt, pk, bias, relu, yy,
                               out_f32, xc, c1, stop);
            break;
        case 16:
            hipLaunchKernelGGL((k_conv_fwd<MTB, 16>), grid, dim3(256), 0, st, g, U(x), cin, mt, pk, bias, relu, yy,
                               out_f32, xc, c1, stop);
            break;
        case 24:
            hipLaunchKernelGGL((k_conv_fwd<MTB, 24>), grid, dim3(256), 0, st, g, U(x), cin, mt, pk, bias, relu, yy,
                               out_f32, xc, c1, stop);
            break;
        default:
            hipLaunchKernelGGL((k_conv_fwd<MTB, 32>), grid, dim3(256), 0, st, g, U(x), cin, mt, pk, bias, relu, yy,
                               out_f32, xc, c1, stop);
            break;
    }
}

struct ConvPlan {
    int bx, by, bz, mtb;
    bool small;  // fewer than 1024 blocks and at least 8 weight chunks: the direct kernel's grids
};

// Blocks of the convolution: output tiles per block (all of them, unless the grid is
// too small to fill the chip, then the largest divisor of mt that gives >= 1024
// blocks, or 1).  A layer that even then has fewer than 1024 blocks (and at least 8
// weight chunks to share out) is a small-grid layer: the direct-load kernel takes it.
ConvPlan conv_plan(Grid3 g, int cin, int cout) {
    ConvPlan p;
    const int mt = cout / 16;
    p.bx = (g.nx + kTX - 1) / kTX;
    p.by = (g.ny + kTY - 1) / kTY;
    p.bz = (g.nz + kTZ - 1) / kTZ;
    const int64_t spatial = (int64_t)p.bx * p.by * p.bz;
    p.mtb = mt;
    for (int d = mt; d >= 1; --d) {
        if (mt % d) continue;
        p.mtb = d;
        if (spatial * (mt / d) >= 1024) break;
    }
    const int64_t blocks = spatial * (mt / p.mtb);
    const int chunks = (cin / conv_cc(cin)) * conv_kpc(cin);
    p.small = blocks < 1024 && chunks >= 8;
    return p;
}

void conv_fwd_any(hipStream_t st, Grid3 g, const bf16 *x, int cin, const bf16 *packed, const float *bias, int cout,
                  bool relu, bf16 *y, float *out_f32, const uint16_t *xc, int c1, const int32_t *stop) {
    const int mt = cout / 16;
    const ConvPlan p = conv_plan(g, cin, cout);
    const dim3 grid((unsigned)p.bx, (unsigned)p.by, (unsigned)(p.bz * (mt / p.mtb)));
    const uint4 *pk = reinterpret_cast<const uint4 *>(packed);
    uint16_t *yy = y ? U(y) : nullptr;
    switch (p.mtb) {
        case 1: conv_fwd_mt<1>(st, grid, g, x, cin, mt, pk, bias, relu, yy, out_f32, xc, c1, stop); break;
        case 2: conv_fwd_mt<2>(st, grid, g, x, cin, mt, pk, bias, relu, yy, out_f32, xc, c1, stop); break;
        case 3: conv_fwd_mt<3>(st, grid, g, x, cin, mt, pk, bias, relu, yy, out_f32, xc, c1, stop); break;
        case 4: conv_fwd_mt<4>(st, grid, g, x, cin, mt, pk, bias, relu, yy, out_f32, xc, c1, stop); break;
        case 6: conv_fwd_mt<6>(st, grid, g, x, cin, mt, pk, bias, relu, yy, out_f32, xc, c1, stop); break;
        case 8: conv_fwd_mt<8>(st, grid, g, x, cin, mt, pk, bias, relu, yy, out_f32, xc, c1, stop); break;
        case 12: conv_fwd_mt<12>(st, grid, g, x, cin, mt, pk, bias, relu, yy, out_f32, xc, c1, stop); break;
        default: return;  // unsupported widths are rejected by the caller
    }
}
}  // namespace

namespace {
struct DirectPlan {
    int mw, nw, s;
};

// Output tile per block: the largest of 4x4, 4x2, 2x4, 2x2, 4x1, 2x1, 1x1 (output tiles x
// N-tiles) that still gives >= 256 blocks; waves per block: enough for about 4096 waves
// on the chip, at least 2 weight chunks per wave, at most 64 KB of partial sums in LDS
// (waves x tile <= 64) and 16 waves.
DirectPlan direct_plan(Grid3 g, int cin, int cout) {
    const int mt = cout / 16;
    const int64_t ntiles = (g.vox() + 15) / 16;
    static const int cand[7][2] = {{4, 4}, {4, 2}, {2, 4}, {2, 2}, {4, 1}, {2, 1}, {1, 1}};
    DirectPlan p{1, 1, 1};
    for (const auto &c : cand) {
        if (mt % c[0]) continue;
        p.mw = c[0];
        p.nw = c[1];
        if ((ntiles + c[1] - 1) / c[1] * (mt / c[0]) >= 256) break;
    }
    const int64_t blocks = (ntiles + p.nw - 1) / p.nw * (mt / p.mw);
    const int chunks = (cin / conv_cc(cin)) * conv_kpc(cin);
    const int cap = std::min(16, 64 / (p.mw * p.nw));
    while (p.s < cap && blocks * p.s < 4096 && 2 * p.s * 2 <= chunks) p.s *= 2;
    return p;
}

template <int MW, int NW>
void conv_direct_mw(hipStream_t st, const DirectPlan &p, Grid3 g, const uint16_t *x, int cin, int mt,
                    const uint4 *pk, const float *bias, int relu, uint16_t *y, const uint16_t *xc, int c1,
                    const int32_t *stop) {
    const dim3 grid((unsigned)((g.vox() + 16 * NW - 1) / (16 * NW)), (unsigned)(mt / MW));
    const dim3 block((unsigned)(64 * p.s));
    const size_t lds = sizeof(float) * 4 * 64 * MW * NW * p.s;
    switch (conv_cc(cin)) {
        case 8:
            hipLaunchKernelGGL((k_conv_direct<MW, NW, 8>), grid, block, lds, st, g, x, cin, mt, pk, bias, relu, y, xc,
                               c1, stop);
            break;
        case 16:
            hipLaunchKernelGGL((k_conv_direct<MW, NW, 16>), grid, block, lds, st, g, x, cin, mt, pk, bias, relu, y,
                               xc, c1, stop);
            break;
        case 24:
            hipLaunchKernelGGL((k_conv_direct<MW, NW, 24>), grid, block, lds, st, g, x, cin, mt, pk, bias, relu, y,
                               xc, c1, stop);
            break;
        default:
            hipLaunchKernelGGL((k_conv_direct<MW, NW, 32>), grid, block, lds, st, g, x, cin, mt, pk, bias, relu, y,
                               xc, c1, stop);
            break;
    }
}
}  // namespace

bool conv_direct_ok(Grid3 g, int cin, int cout) {
    if (cin % 8 || cout % 16 || cout <= 0 || g.vox() >= (1ll << 31)) return false;
    return conv_plan(g, cin, cout).small;  // the tiled kernel cannot fill the chip
}

void launch_conv_direct(hipStream_t st, Grid3 g, const bf16 *x, int cin, const bf16 *packed, const float *bias,
                        int cout, bool relu, bf16 *y, const bf16 *coarse, int c1, const int32_t *stop) {
    const DirectPlan p = direct_plan(g, cin, cout);
    const int mt = cout / 16, r = relu ? 1 : 0;
    const uint4 *pk = reinterpret_cast<const uint4 *>(packed);
    const uint16_t *xc = coarse ? U(coarse) : nullptr;
    const int key = p.mw * 8 + p.nw;
    switch (key) {
        case 4 * 8 + 4: conv_direct_mw<4, 4>(st, p, g, U(x), cin, mt, pk, bias, r, U(y), xc, c1, stop); break;
        case 4 * 8 + 2: conv_direct_mw<4, 2>(st, p, g, U(x), cin, mt, pk, bias, r, U(y), xc, c1, stop); break;
        case 2 * 8 + 4: conv_direct_mw<2, 4>(st, p, g, U(x), cin, mt, pk, bias, r, U(y), xc, c1, stop); break;
        case 2 * 8 + 2: conv_direct_mw<2, 2>(st, p, g, U(x), cin, mt, pk, bias, r, U(y), xc, c1, stop); break;
        case 4 * 8 + 1: conv_direct_mw<4, 1>(st, p, g, U(x), cin, mt, pk, bias, r, U(y), xc, c1, stop); break;
        case 2 * 8 + 1: conv_direct_mw<2, 1>(st, p, g, U(x), cin, mt, pk, bias, r, U(y), xc, c1, stop); break;
        default: conv_direct_mw<1, 1>(st, p, g, U(x), cin, mt, pk, bias, r, U(y), xc, c1, stop); break;
    }
}

void launch_conv_fwd(hipStream_t st, Grid3 g, const bf16 *x, int cin, const bf16 *packed, const float *bias, int cout,
                     bool relu, bf16 *y, float *out_f32, const int32_t *stop) {
    conv_fwd_any(st, g, x, cin, packed, bias, cout, relu, y, out_f32, nullptr, 0, stop);
}

void launch_conv_fwd_cat(hipStream_t st, Grid3 g, const bf16 *coarse, int c1, const bf16 *skip, int cin,
                         const bf16 *packed, const float *bias, int cout, bool relu, bf16 *y,
                         const int32_t *stop) {
    conv_fwd_any(st, g, skip, cin, packed, bias, cout, relu, y, nullptr, U(coarse), c1, stop);
}

namespace {
// compute units of the current device (cached per device)
int cu_count() {
    static int cus[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cus[dev]) {
        int n = 0;
        cus[dev] = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0 ? n : 256;
    }
    return cus[dev];
}

// rows per wave in y and z: tiles of 16 x 4 x 8 (measured 3% faster than 16 x 8 x 4 on
// every layer; 16 x 4 x 4 and 16 x 8 x 2 tiles re-read more halo and were slower)
constexpr int kRowsRY = 2, kRowsRZ = 4;
}  // namespace

void launch_conv_pack_rows(hipStream_t st, const float *w, int cin, int cout, bf16 *packed) {
    const int64_t total = conv_rows_packed_elems(cin, cout);
    hipLaunchKernelGGL(k_conv_pack_rows, dim3(grid1(total, 256)), dim3(256), 0, st, w, cin, cout, U(packed), total);
}

bool conv_rows_ok(Grid3 g, int cin, int cout) {
    // an 8-channel input fills a quarter of the staged groups and a chunk three quarters:
    // the tiled kernel is faster there (measured 37 us against 62 us for 8 -> 16 at 128^3)
    if (cin < 16 || cin % 8 || (cout != 16 && cout != 32)) return false;
    const int64_t tiles = (int64_t)((g.nx + 15) / 16) * ((g.ny + 2 * kRowsRY - 1) / (2 * kRowsRY)) *
                          ((g.nz + 2 * kRowsRZ - 1) / (2 * kRowsRZ));
    return tiles >= 512 && g.vox() < (1ll << 31);  // 32-bit voxel offsets
}

void launch_conv_rows(hipStream_t st, Grid3 g, const bf16 *x, int cin, const bf16 *packed, const float *bias,
                      int cout, bool relu, bf16 *y, const bf16 *coarse, int c1, const int32_t *stop) {
    const int tx = (g.nx + 15) / 16, ty = (g.ny + 2 * kRowsRY - 1) / (2 * kRowsRY);
    const int tz = (g.nz + 2 * kRowsRZ - 1) / (2 * kRowsRZ);
    const unsigned blocks = (unsigned)((int64_t)tx * ty * tz);
    const uint4 *pk = reinterpret_cast<const uint4 *>(packed);
    const uint16_t *xc = coarse ? U(coarse) : nullptr;
    if (cout == 16) {
        // persistent: two blocks of 8 waves per CU (the LDS holds two halos), the same
        // 16 x 2 kRowsRY x 2 kRowsRZ tiles as k_conv_rows
        const unsigned resident = 2u * (unsigned)cu_count();
        hipLaunchKernelGGL((k_conv_rows_pipe<kRowsRY, kRowsRZ / 2, 4>), dim3(blocks < resident ? blocks : resident),
                           dim3(512),
                           0, st, g, U(x), cin, pk, bias, relu ? 1 : 0, U(y), xc, c1, tx, ty, (int)blocks, stop);
    } else
        hipLaunchKernelGGL((k_conv_rows<2, kRowsRY, kRowsRZ>), dim3(blocks), dim3(256), 0, st, g, U(x), cin, pk, bias,
                           relu ? 1 : 0, U(y), xc, c1, tx, ty, stop);
}

void launch_conv_wgrad(hipStream_t st, Grid3 g, const bf16 *x, int cin, const bf16 *dz, int cout, float *dw,
                       float *db) {
    const int64_t chunks = (g.vox() + kWgradChunk - 1) / kWgradChunk;
    const dim3 grid((unsigned)chunks, 27, (unsigned)((cout / 16) * ((cin + 15) / 16)));
    hipLaunchKernelGGL(k_conv_wgrad, grid, dim3(64), 0, st, g, U(x), cin, U(dz), cout, dw);
    if (db) {
        const int64_t nb = (g.vox() + 255) / 256;
        hipLaunchKernelGGL(k_bias_grad, dim3((unsigned)(nb < 1024 ? nb : 1024)), dim3(256), 0, st, g.vox(), U(dz),
                           cout, db);
    }
}

void launch_relu_bwd(hipStream_t st, int64_t n, const bf16 *y, bf16 *dy) {
    hipLaunchKernelGGL(k_relu_bwd, dim3(grid1(n, 256)), dim3(256), 0, st, n, U(y), U(dy));
}

void launch_pool(hipStream_t st, Grid3 coarse, int c, const bf16 *fine, bf16 *out, const int32_t *stop) {
    if (c % 8 == 0)
        hipLaunchKernelGGL(k_pool8, dim3(grid1(coarse.vox() * (c / 8), 256)), dim3(256), 0, st, coarse, c, U(fine),
                           U(out), stop);
    else
        hipLaunchKernelGGL(k_pool, dim3(grid1(coarse.vox() * c, 256)), dim3(256), 0, st, coarse, c, U(fine), U(out),
                           stop);
}

void launch_pool_bwd(hipStream_t st, Grid3 coarse, int c, const bf16 *coarse_grad, bf16 *fine_grad) {
    hipLaunchKernelGGL(k_pool_bwd, dim3(grid1(coarse.vox() * c, 256)), dim3(256), 0, st, coarse, c, U(coarse_grad),
                       U(fine_grad));
}

void launch_upsample(hipStream_t st, Grid3 cgrid, int c_coarse, const bf16 *coarse, int cstride, bf16 *out) {
    hipLaunchKernelGGL(k_upsample, dim3(grid1(cgrid.vox() * c_coarse, 256)), dim3(256), 0, st, cgrid, c_coarse,
                       U(coarse), cstride, U(out));
}

void launch_upsample_bwd(hipStream_t st, Grid3 coarse, int c_coarse, const bf16 *grad, int cstride, bf16 *coarse_grad) {
    hipLaunchKernelGGL(k_upsample_bwd, dim3(grid1(coarse.vox() * c_coarse, 256)), dim3(256), 0, st, coarse, c_coarse,
                       U(grad), cstride, U(coarse_grad));
}

void launch_copy_channels(hipStream_t st, int64_t vox, int cs, const bf16 *src, int sstride, int soff, bf16 *dst,
                          int dstride, int doff, bool add) {
    hipLaunchKernelGGL(k_copy_channels, dim3(grid1(vox * cs, 256)), dim3(256), 0, st, vox, cs, U(src), sstride, soff,
                       U(dst), dstride, doff, add ? 1 : 0);
}

void launch_add_bf16(hipStream_t st, int64_t n, const bf16 *a, const bf16 *b, bf16 *out) {
    hipLaunchKernelGGL(k_add_bf16, dim3(grid1(n, 256)), dim3(256), 0, st, n, U(a), U(b), U(out));
}

void launch_features(hipStream_t st, Grid3 coarse, const float *r, double inv_scale, const float *mask,
                     const float *cx, const float *cy, const float *cz, bf16 *feat) {
    hipLaunchKernelGGL(k_features, dim3(grid1(coarse.vox(), 256)), dim3(256), 0, st, coarse, r, inv_scale, mask, cx,
                       cy, cz, U(feat));
}

void launch_prolong(hipStream_t st, Grid3 fine, const float *u, double scale, const float *mask, float *z) {
    hipLaunchKernelGGL(k_prolong, dim3(grid1(fine.vox(), 256)), dim3(256), 0, st, fine, u, scale, mask, z);
}

void launch_prolong_bwd(hipStream_t st, Grid3 coarse, const float *g, const float *mask, double scale,
                        float *grad_u) {
    hipLaunchKernelGGL(k_prolong_bwd, dim3(grid1(coarse.vox(), 256)), dim3(256), 0, st, coarse, g, mask, scale,
                       grad_u);
}

void launch_smooth(hipStream_t st, int64_t n, double omega, const float *r, const float *mask, const float *t,
                   float *z) {
    hipLaunchKernelGGL(k_smooth, dim3(grid1(n, 256)), dim3(256), 0, st, n, omega, r, mask, t, z);
}

void launch_mask_mul(hipStream_t st, int64_t n, const float *mask, const float *g, float *out) {
    hipLaunchKernelGGL(k_mask_mul, dim3(grid1(n, 256)), dim3(256), 0, st, n, mask, g, out);
}

void launch_dense_stencil(hipStream_t st, Grid3 g, const float *cx, const float *cy, const float *cz, const float *x,
                          float *y) {
    hipLaunchKernelGGL(k_dense_stencil, dim3(grid1(g.vox(), 256)), dim3(256), 0, st, g, cx, cy, cz, x, y);
}

void launch_adam(hipStream_t st, int64_t n, float *p, float *g, float *m, float *v, float lr, float b1, float b2,
                 float eps, float bc1, float bc2) {
    hipLaunchKernelGGL(k_adam, dim3(grid1(n, 256)), dim3(256), 0, st, n, p, g, m, v, lr, b1, b2, eps, bc1, bc2);
}

void launch_f32_to_bf16(hipStream_t st, int64_t n, const float *src, bf16 *dst) {
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid1(n, 256)), dim3(256), 0, st, n, src, U(dst));
}

}  // namespace shk
