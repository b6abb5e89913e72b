// mg.hip -- aggregation multigrid V-cycle kernels on the brick layout (see mg.h).
//
// Every kernel returns at once when the Krylov state it is given has stopped, so
// the V-cycles queued behind a converged iteration cost only their launches.
#include "mg.h"

namespace shk {

namespace {

constexpr int NT = 256;

inline unsigned blocks(int64_t n) { return (unsigned)((n + NT - 1) / NT > 0 ? (n + NT - 1) / NT : 1); }

__device__ inline bool stopped(const SolveState *s) { return s && s->done; }

// uniform coupling of the brick (level 0), 0 when none
__device__ inline float brick_kappa(const MgOp &op, int slot) { return op.kappa ? op.kappa[slot] : 0.0f; }

// diagonal of cell e: 1 in a uniform brick
__device__ inline float diag_of(const MgOp &op, int64_t e, float kap) { return kap != 0.0f ? 1.0f : op.d[e]; }

// (A v)_e for the 7-point brick operator of a level >= 1 (explicit diagonal and
// couplings, FP32), where v(s, m) is the value at local cell m of slot s; a missing
// neighbour brick adds nothing.  A 2D grid (Y == 1) has no y neighbours.
template <int BX, class V>
__device__ inline float apply_v(const MgOp &op, int slot, int l, V v) {
    using B = BrickShape<BX>;
    constexpr int X = B::X, Y = B::Y;
    const int64_t e = (int64_t)slot * kBrickCells + l;
    const int i = l % X, j = (l / X) % Y, k = l / kWave;
    const int *nb = op.nbr + 6 * slot;
    float acc = op.d[e] * v(slot, l);
    if (i > 0) {
        acc -= op.cx[e] * v(slot, l - 1);
    } else if (nb[0] >= 0) {
        acc -= op.cx[e] * v(nb[0], l + (X - 1));
    }
    if (i < X - 1) {
        acc -= op.cx[e + 1] * v(slot, l + 1);
    } else if (nb[1] >= 0) {
        acc -= op.cx[(int64_t)nb[1] * kBrickCells + l - (X - 1)] * v(nb[1], l - (X - 1));
    }
    if constexpr (Y > 1) {
        if (j > 0) {
            acc -= op.cy[e] * v(slot, l - X);
        } else if (nb[2] >= 0) {
            acc -= op.cy[e] * v(nb[2], l + X * (Y - 1));
        }
        if (j < Y - 1) {
            acc -= op.cy[e + X] * v(slot, l + X);
        } else if (nb[3] >= 0) {
            acc -= op.cy[(int64_t)nb[3] * kBrickCells + l - X * (Y - 1)] * v(nb[3], l - X * (Y - 1));
        }
    }
    if (k > 0) {
        acc -= op.cz[e] * v(slot, l - 64);
    } else if (nb[4] >= 0) {
        acc -= op.cz[e] * v(nb[4], l + 448);
    }
    if (k < 7) {
        acc -= op.cz[e + 64] * v(slot, l + 64);
    } else if (nb[5] >= 0) {
        acc -= op.cz[(int64_t)nb[5] * kBrickCells + l - 448] * v(nb[5], l - 448);
    }
    return acc;
}

// Aggregation: 2x2x2 children per coarse cell, 2x1x2 on a 2D grid (Y == 1).
template <int BX>
struct Agg {
    using B = BrickShape<BX>;
    static constexpr int X = B::X, Y = B::Y;
    static constexpr int AY = Y > 1 ? 2 : 1;       // aggregation factor along y
    static constexpr int NCH = 4 * AY;             // children per coarse cell
    // brick-local offset of child q (a along x, b along y, c along z)
    __device__ static int child(int q) {
        const int a = q & 1, b = AY > 1 ? (q >> 1) & 1 : 0, c = AY > 1 ? q >> 2 : q >> 1;
        return a + X * b + 64 * c;
    }
    __device__ static void bits(int q, int &a, int &b, int &c) {
        a = q & 1;
        b = AY > 1 ? (q >> 1) & 1 : 0;
        c = AY > 1 ? q >> 2 : q >> 1;
    }
};

// fine slot holding the children of coarse cell (slot_c, l) and the children's base
// local offset; -1 when the fine brick is outside the grid or inactive
template <int BX>
__device__ inline int children(const MgLevelDev &f, const MgLevelDev &c, int slot_c, int l, int &base) {
    using A = Agg<BX>;
    constexpr int X = A::X, Y = A::Y, HX = X / 2, HY = Y > 1 ? Y / 2 : 1;
    const int ci = l % X, cj = (l / X) % Y, ck = l / kWave;
    const int4 Bc = c.coord[slot_c];
    if (Bc.x * X + ci >= c.nx || Bc.y * Y + cj >= c.ny || Bc.z * 8 + ck >= c.nz) return -1;
    const int fbx = 2 * Bc.x + ci / HX, fby = A::AY * Bc.y + cj / HY, fbz = 2 * Bc.z + (ck >> 2);
    if (fbx >= f.nbx || fby >= f.nby || fbz >= f.nbz) return -1;
    base = 2 * (ci % HX) + X * (A::AY * (cj % HY)) + 64 * (2 * (ck & 3));
    return f.slot[fbx + f.nbx * (fby + f.nby * fbz)];
}

// cell index on the next level of the cell (brick coord Bf, local l) of a fine level
template <int BX>
__device__ inline int coarse_local(int4 Bf, int l) {
    using A = Agg<BX>;
    constexpr int X = A::X, Y = A::Y, HX = X / 2, HY = Y > 1 ? Y / 2 : 1;
    const int i = l % X, j = (l / X) % Y, k = l / kWave;
    const int cy = A::AY > 1 ? (Bf.y & 1) * HY + j / 2 : 0;
    return ((Bf.x & 1) * HX + i / 2) + X * cy + 64 * ((Bf.z & 1) * 4 + k / 2);
}

__global__ __launch_bounds__(NT) void k_mg_mask(int64_t n, const double *__restrict__ sig, float *__restrict__ d0) {
    const int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x;
    if (e < n) d0[e] = sig[e] > 0.0 ? 1.0f : 0.0f;
}

template <int BX>
__global__ __launch_bounds__(NT) void k_mg_coarsen(MgLevelDev f, MgLevelDev c) {
    using A = Agg<BX>;
    const int64_t E = blockIdx.x * (int64_t)NT + threadIdx.x;
    if (E >= (int64_t)c.n_active * kBrickCells) return;
    const int slot_c = (int)(E >> 9), l = (int)(E & 511);
    int base = 0;
    const int fs = children<BX>(f, c, slot_c, l, base);
    double d = 0.0, inner = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    if (fs >= 0) {
        const int64_t fb = (int64_t)fs * kBrickCells + base;
#pragma unroll
        for (int q = 0; q < A::NCH; ++q) {
            int a, b, cc;
            A::bits(q, a, b, cc);
            const int64_t e = fb + A::child(q);
            d += (double)f.d[e];
            if (a) inner += (double)f.cx[e]; else cx += (double)f.cx[e];
            if (A::AY > 1) {
                if (b) inner += (double)f.cy[e]; else cy += (double)f.cy[e];
            }
            if (cc) inner += (double)f.cz[e]; else cz += (double)f.cz[e];
        }
    }
    double dc = d - 2.0 * inner;
    if (!(dc > 1e-6 * d)) dc = 0.0;  // no unknown child, or an aggregate decoupled from every boundary
    c.d[E] = (float)dc;
    c.cx[E] = (float)cx;
    c.cy[E] = (float)cy;
    c.cz[E] = (float)cz;
    // k_mg_residual_restrict does not write the cells without a fine brick
    c.b[E] = 0.0f;
    c.x[E] = 0.0f;
}

// b_c = P0' (b_f - A_f x_f), x_c = w b_c / d_c (0 with w = 0) from a level >= 1, one
// thread per fine cell: each cell's residual goes to LDS and is never stored; the
// first child of each aggregate sums its children in the order q (the children of an
// aggregate lie in one 256-cell half brick, the z pairs (2m, 2m+1) never straddle
// the halves).  Coarse cells without a fine brick are not written: their b and x
// hold the zeros k_mg_coarsen writes, and nothing else writes there.
template <int BX>
__global__ __launch_bounds__(NT) void k_mg_residual_restrict(MgLevelDev f, MgLevelDev c, float omega,
                                                             const SolveState *state) {
    using A = Agg<BX>;
    constexpr int X = A::X, Y = A::Y;
    static_assert(kBrickCells % NT == 0 && NT % (2 * kWave) == 0, "a block holds whole z pairs of one brick");
    if (stopped(state)) return;
    __shared__ float ts[NT];
    const int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x;
    if (e >= (int64_t)f.n_active * kBrickCells) return;  // whole blocks: n_active * 512 is a multiple of NT
    const int slot = (int)(e >> 9), l = (int)(e & 511);
    // every load is issued before the barrier, none waits on the residual: the parent
    // cell's diagonal, then the residual's operands (all in bounds whatever d is)
    const int ps = f.parent[slot];  // >= 0: the parent of an active brick is active
    const int64_t E = (int64_t)(ps > 0 ? ps : 0) * kBrickCells + coarse_local<BX>(f.coord[slot], l);
    const float dc = c.d[E];
    const MgOp op = f.op(0.0f, 0.0f);
    const float *__restrict__ xf = f.x;
    const auto v = [xf](int sl, int m) { return xf[(int64_t)sl * kBrickCells + m]; };
    const float ax = apply_v<BX>(op, slot, l, v);
    ts[threadIdx.x] = f.d[e] > 0.0f ? f.b[e] - ax : 0.0f;
    __syncthreads();
    const int i = l % X, j = (l / X) % Y, k = l / kWave;
    if ((i & 1) || (A::AY > 1 && (j & 1)) || (k & 1) || ps < 0) return;
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < A::NCH; ++q) s += ts[threadIdx.x + A::child(q)];
    c.b[E] = dc > 0.0f ? s : 0.0f;
    c.x[E] = dc > 0.0f ? omega * s / dc : 0.0f;
}

// x_f += alpha P0 x_c on the unknowns of a level f >= 1, one thread per cell (the
// level above the coarsest; the others get theirs from k_mg_smooth)
template <int BX>
__global__ __launch_bounds__(NT) void k_mg_prolong(MgLevelDev f, MgLevelDev c, float alpha, const SolveState *state) {
    if (stopped(state)) return;
    const int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x;
    if (e >= (int64_t)f.n_active * kBrickCells) return;
    const int slot = (int)(e >> 9), l = (int)(e & 511);
    if (!(f.d[e] > 0.0f)) return;
    const int ps = f.parent[slot];
    if (ps < 0) return;  // cannot happen: the parent of an active brick is active
    f.x[e] = fmaf(alpha, c.x[(int64_t)ps * kBrickCells + coarse_local<BX>(f.coord[slot], l)], f.x[e]);
}

// c.t = y = x + w (b - A x) / d on a level >= 1, one thread per cell; with a finer
// level f >= 1 (prolong) the same thread also adds alpha y to the unknowns among its
// children there (f.x += alpha P0 y, the next coarse correction, which k_mg_prolong
// would make from the stored y)
template <int BX>
__global__ __launch_bounds__(NT) void k_mg_smooth(MgLevelDev c, MgLevelDev f, bool prolong, float omega, float alpha,
                                                  const SolveState *state) {
    using A = Agg<BX>;
    if (stopped(state)) return;
    const int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x;
    if (e >= (int64_t)c.n_active * kBrickCells) return;
    const int slot = (int)(e >> 9), l = (int)(e & 511);
    // the children's d and x are loaded with the smoothing's operands, not after them
    int base = 0;
    const int fs = prolong ? children<BX>(f, c, slot, l, base) : -1;
    const int64_t fb = (int64_t)(fs > 0 ? fs : 0) * kBrickCells + base;
    float fd[A::NCH], fx[A::NCH];
#pragma unroll
    for (int q = 0; q < A::NCH; ++q) {
        fd[q] = fs >= 0 ? f.d[fb + A::child(q)] : 0.0f;
        fx[q] = fs >= 0 ? f.x[fb + A::child(q)] : 0.0f;
    }
    const MgOp op = c.op(omega, alpha);
    const float *__restrict__ x = c.x;
    const auto v = [x](int sl, int m) { return x[(int64_t)sl * kBrickCells + m]; };
    const float dv = c.d[e], xe = x[e], be = c.b[e];
    const float ax = apply_v<BX>(op, slot, l, v);  // all loads in bounds whatever d is
    const float y = dv > 0.0f ? xe + omega * (be - ax) / dv : 0.0f;
    c.t[e] = y;
#pragma unroll
    for (int q = 0; q < A::NCH; ++q)
        if (fd[q] > 0.0f) f.x[fb + A::child(q)] = fmaf(alpha, y, fx[q]);  // one rounding, as k_mg_prolong
}

// ---- level 0, fused.  On the scaled system every level-0 vector is zero off the
// unknowns and the diagonal is 1 on them, so x0 = w r needs no division and is never
// stored: the pre-smoothing residual and the post-smoothing recompute it from r.

// value of x0' = w r + alpha P0 x1 at level-0 cell (slot, l) of brick B with parent ps
template <int BX>
__device__ inline float xprime(const MgOp &op, const float *__restrict__ r, const float *__restrict__ xc, int slot,
                               int l, int4 B, int ps) {
    return op.omega * r[(int64_t)slot * kBrickCells + l] +
           op.alpha * xc[(int64_t)ps * kBrickCells + coarse_local<BX>(B, l)];
}

// ---------------------------------------------------------------- level 0, one wave per brick
//
// The two level-0 passes with one wave per brick: lane = i + X j holds one (x, y)
// column (X x Y = 64: 8 x 8 for 3D bricks, 64 x 1 for the 64x1x8 bricks of 2D grids)
// and walks the brick's 8 z-planes.  In-brick x and y neighbours come by lane
// shuffles (taken by all 64 lanes, outside any lane-dependent branch), z neighbours
// from the lane's own registers, and the cells across the faces from one coalesced
// load per neighbour plane (the neighbour brick's slot, coordinates and parent are
// wave-uniform).  Each cell's arithmetic takes its terms in the face order of
// apply_v<BX> (-x, +x, -y, +y, -z, +z); a uniform brick sums its six neighbours
// first and scales the sum by its one coupling.

__device__ inline float lane_get(float v, int src) { return __shfl(v, src & 63, 64); }

// the cells across the faces of a brick as plane registers (0 where the neighbour
// brick does not exist): x planes hold (y = a, z = c) at lane a + Y c (the first 8 Y
// lanes), y planes (x = a, z = c) at lane a + X c, z planes (x, y) at the lane itself
template <int BX, class V>
__device__ inline void halo_planes(const int nb[6], int lane, V value, float h[6]) {
    using B = BrickShape<BX>;
    constexpr int X = B::X, Y = B::Y;
    const int xa = lane % Y, xc = lane / Y;
    const bool xl = lane < Y * 8;
    h[0] = nb[0] >= 0 && xl ? value(nb[0], (X - 1) + X * xa + 64 * xc) : 0.0f;
    h[1] = nb[1] >= 0 && xl ? value(nb[1], X * xa + 64 * xc) : 0.0f;
    h[2] = h[3] = 0.0f;
    if constexpr (Y > 1) {
        const int pa = lane % X, pc = lane / X;
        h[2] = nb[2] >= 0 ? value(nb[2], pa + X * (Y - 1) + 64 * pc) : 0.0f;
        h[3] = nb[3] >= 0 ? value(nb[3], pa + 64 * pc) : 0.0f;
    }
    h[4] = nb[4] >= 0 ? value(nb[4], lane + 448) : 0.0f;
    h[5] = nb[5] >= 0 ? value(nb[5], lane) : 0.0f;
}

// the six neighbour values of cell (i, j, k) in face order x-, x+, y-, y+, z-, z+
// (y values 0 without a y axis)
template <int BX>
__device__ inline void neighbours6(const float *v, const float h[6], int lane, int k, float n[6]) {
    using B = BrickShape<BX>;
    constexpr int X = B::X, Y = B::Y;
    const int i = lane % X, j = lane / X;
    const float x0 = v[k];
    float xm = lane_get(x0, lane - 1), xp = lane_get(x0, lane + 1);
    const float hxm = lane_get(h[0], j + Y * k), hxp = lane_get(h[1], j + Y * k);
    if (i == 0) xm = hxm;
    if (i == X - 1) xp = hxp;
    float ym = 0.0f, yp = 0.0f;
    if constexpr (Y > 1) {
        ym = lane_get(x0, lane - X);
        yp = lane_get(x0, lane + X);
        const float hym = lane_get(h[2], i + X * k), hyp = lane_get(h[3], i + X * k);
        if (j == 0) ym = hym;
        if (j == Y - 1) yp = hyp;
    }
    n[0] = xm;
    n[1] = xp;
    n[2] = ym;
    n[3] = yp;
    n[4] = k > 0 ? v[k - 1] : h[4];
    n[5] = k < 7 ? v[k + 1] : h[5];
}

// the couplings of cell e across its faces (low faces live at e, high faces at the
// neighbour cell) and whether each neighbour exists; a uniform brick has kap on all
template <int BX>
__device__ inline void couplings6(const MgOp &op, int64_t e, const int nb[6], int lane, int k, float kap, float c[6],
                                  bool has[6]) {
    using B = BrickShape<BX>;
    constexpr int X = B::X, Y = B::Y;
    const int i = lane % X, j = lane / X;
    has[0] = i > 0 || nb[0] >= 0;
    has[1] = i < X - 1 || nb[1] >= 0;
    has[2] = Y > 1 && (j > 0 || nb[2] >= 0);
    has[3] = Y > 1 && (j < Y - 1 || nb[3] >= 0);
    has[4] = k > 0 || nb[4] >= 0;
    has[5] = k < 7 || nb[5] >= 0;
    if (kap != 0.0f) {
#pragma unroll
        for (int f = 0; f < 6; ++f) c[f] = kap;
        return;
    }
    c[0] = op.cx[e];
    c[1] = i < X - 1 ? op.cx[e + 1]
                     : (nb[1] >= 0 ? op.cx[(int64_t)nb[1] * kBrickCells + lane - (X - 1) + 64 * k] : 0.0f);
    c[2] = c[3] = 0.0f;
    if constexpr (Y > 1) {
        c[2] = op.cy[e];
        c[3] = j < Y - 1 ? op.cy[e + X]
                         : (nb[3] >= 0 ? op.cy[(int64_t)nb[3] * kBrickCells + lane - X * (Y - 1) + 64 * k] : 0.0f);
    }
    c[4] = op.cz[e];
    c[5] = k < 7 ? op.cz[e + 64] : (nb[5] >= 0 ? op.cz[(int64_t)nb[5] * kBrickCells + lane] : 0.0f);
}

// t0 = r - A (w r) = r - w A r on the unknowns (0 elsewhere), by linearity, restricted
// at once to level 1 (b_1 = P0' t0, x_1 = w_c b_1 / d_1)
template <int BX>
__global__ __launch_bounds__(NT) void k_mg_pre0_w(MgOp op, const float *__restrict__ r, const int4 *__restrict__ coord,
                                                  const int *__restrict__ parent, MgLevelDev c, float omega_c,
                                                  const SolveState *state) {
    constexpr int X = BrickShape<BX>::X, Y = BrickShape<BX>::Y;
    if (stopped(state)) return;
    const int lane = threadIdx.x & 63;
    const int slot = blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    if (slot >= op.n_active) return;   // no block-wide synchronisation below
    int nb[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) nb[f] = op.nbr[6 * slot + f];
    const float kap = brick_kappa(op, slot);
    const int64_t base = (int64_t)slot * kBrickCells + lane;
    const int ps = parent[slot];  // >= 0: the parent of an active brick is active
    const int4 Bf = coord[slot];
    float dcm[4];  // level-1 diagonals of the cells this lane may restrict to, loaded early
#pragma unroll
    for (int m = 0; m < 4; ++m) dcm[m] = c.d[(int64_t)(ps > 0 ? ps : 0) * kBrickCells + coarse_local<BX>(Bf, lane + 128 * m)];
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = r[base + 64 * k];
    float h[6];
    halo_planes<BX>(nb, lane, [&](int ns, int l) { return r[(int64_t)ns * kBrickCells + l]; }, h);
    float t[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int64_t e = base + 64 * k;
        float n[6];
        neighbours6<BX>(v, h, lane, k, n);
        float ax;
        if (kap != 0.0f) {
            // the y terms are 0.0f without a y axis
            ax = v[k] - kap * (n[0] + n[1] + (Y > 1 ? n[2] : 0.0f) + (Y > 1 ? n[3] : 0.0f) + n[4] + n[5]);
        } else {
            float c[6];
            bool has[6];
            couplings6<BX>(op, e, nb, lane, k, kap, c, has);
            ax = op.d[e] * v[k];
#pragma unroll
            for (int f = 0; f < 6; ++f)
                if (has[f]) ax -= c[f] * n[f];
        }
        t[k] = diag_of(op, e, kap) > 0.0f ? v[k] - op.omega * ax : 0.0f;
    }
    // b_1 = P0' t and x_1 = w_c b_1 / d_1 on the coarse cells whose children are in
    // this brick: a cell of even (i, j, k) sums the children (i+a, j+b, k+c) in the
    // order a, b, c (that of Agg<BX>::child), taken from lanes +1, +X, +X+1 and its
    // own next register; t is not stored
    float sum[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        float q[8];
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) {
            const float tv = t[2 * m + cz];
            q[4 * cz] = tv;
            q[4 * cz + 1] = lane_get(tv, lane + 1);
            q[4 * cz + 2] = Y > 1 ? lane_get(tv, lane + X) : 0.0f;
            q[4 * cz + 3] = Y > 1 ? lane_get(tv, lane + X + 1) : 0.0f;
        }
        float sm = 0.0f;
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) {
            sm += q[4 * cz];
            sm += q[4 * cz + 1];
            if constexpr (Y > 1) {
                sm += q[4 * cz + 2];
                sm += q[4 * cz + 3];
            }
        }
        sum[m] = sm;
    }
    const int i = lane % X, j = lane / X;
    if ((i & 1) || (Y > 1 && (j & 1)) || ps < 0) return;  // ps < 0 cannot happen for an active brick
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int64_t E = (int64_t)ps * kBrickCells + coarse_local<BX>(Bf, lane + 64 * (2 * m));
        const float dc = dcm[m];
        c.b[E] = dc > 0.0f ? sum[m] : 0.0f;
        c.x[E] = dc > 0.0f ? omega_c * sum[m] / dc : 0.0f;
    }
}

// y = x0' + w (r - A x0') with x0' = w r + alpha P0 x1 computed on the fly (the
// prolongation fused into the post-smoothing); part[block] = sum r*y over the block's
// bricks
template <int BX>
__global__ __launch_bounds__(NT) void k_mg_post0_w(MgOp op, const int4 *__restrict__ coord,
                                                   const int *__restrict__ parent, const float *__restrict__ r,
                                                   const float *__restrict__ xc, float *__restrict__ y,
                                                   const SolveState *state, double *__restrict__ part) {
    if (stopped(state)) return;
    __shared__ double lds[NT / kWave];
    const int lane = threadIdx.x & 63;
    const int slot = blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    double acc = 0.0;
    if (slot < op.n_active) {
        int nb[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) nb[f] = op.nbr[6 * slot + f];
        const float kap = brick_kappa(op, slot);
        const int4 B = coord[slot];
        const int ps = parent[slot];
        const int64_t base = (int64_t)slot * kBrickCells + lane;
        float rv[8], xs[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            rv[k] = r[base + 64 * k];
            xs[k] = xprime<BX>(op, r, xc, slot, lane + 64 * k, B, ps);
        }
        float h[6];
        halo_planes<BX>(nb, lane,
                        [&](int ns, int l) { return xprime<BX>(op, r, xc, ns, l, coord[ns], parent[ns]); }, h);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int64_t e = base + 64 * k;
            float yv = 0.0f;
            // every lane takes part in the shuffles, whether its cell is an unknown or not
            float n[6];
            neighbours6<BX>(xs, h, lane, k, n);
            if (diag_of(op, e, kap) > 0.0f) {
                float c[6];
                bool has[6];
                couplings6<BX>(op, e, nb, lane, k, kap, c, has);
                float sum = 0.0f;
#pragma unroll
                for (int f = 0; f < 6; ++f)
                    if (has[f] && c[f] != 0.0f) sum += c[f] * n[f];
                yv = xs[k] + op.omega * (rv[k] - (xs[k] - sum));
            }
            y[e] = yv;
            acc += (double)rv[k] * (double)yv;
        }
    }
    const double s = block_sum<NT>(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[t] = sum of in[t*per .. (t+1)*per) for t < n_out, in a fixed order (the
// per-block partials of k_mg_post0_w folded to the Krylov grid's partial count)
__global__ __launch_bounds__(NT) void k_mg_fold(int n_in, const double *__restrict__ in, int n_out,
                                                double *__restrict__ out, const SolveState *state) {
    if (stopped(state)) return;
    const int t = blockIdx.x * NT + threadIdx.x;
    if (t >= n_out) return;
    const int per = (n_in + n_out - 1) / n_out;
    double s = 0.0;
    for (int q = t * per; q < (t + 1) * per && q < n_in; ++q) s += in[q];
    out[t] = s;
}

__global__ __launch_bounds__(kMgCoarseMax) void k_mg_coarse(int n, const int *__restrict__ idx,
                                                            const float *__restrict__ inv, const float *__restrict__ b,
                                                            float *__restrict__ x, int cells,
                                                            const SolveState *state) {
    if (stopped(state)) return;
    __shared__ float bs[kMgCoarseMax];
    const int t = threadIdx.x;
    if (t < n) bs[t] = b[idx[t]];
    __syncthreads();
    float acc = 0.0f;
    if (t < n)
        for (int j = 0; j < n; ++j) acc += inv[t * n + j] * bs[j];
    for (int e = t; e < cells; e += kMgCoarseMax) x[e] = 0.0f;
    __syncthreads();
    if (t < n) x[idx[t]] = acc;
}

}  // namespace

void launch_mg_mask(hipStream_t st, int64_t n, const double *sig, float *d0) {
    hipLaunchKernelGGL(k_mg_mask, dim3(blocks(n)), dim3(NT), 0, st, n, sig, d0);
}

#define SHK_MG_LAUNCH(kernel, bx, grid, ...)                                                   \
    do {                                                                                       \
        if ((bx) == 64)                                                                        \
            hipLaunchKernelGGL(kernel<64>, dim3(grid), dim3(NT), 0, __VA_ARGS__);              \
        else                                                                                   \
            hipLaunchKernelGGL(kernel<8>, dim3(grid), dim3(NT), 0, __VA_ARGS__);               \
    } while (0)

void launch_mg_coarsen(hipStream_t st, const MgLevelDev &f, const MgLevelDev &c) {
    SHK_MG_LAUNCH(k_mg_coarsen, c.bx, blocks((int64_t)c.n_active * kBrickCells), st, f, c);
}

void launch_mg_residual_restrict(hipStream_t st, const MgLevelDev &f, const MgLevelDev &c, float omega,
                                 const SolveState *state) {
    SHK_MG_LAUNCH(k_mg_residual_restrict, f.bx, blocks((int64_t)f.n_active * kBrickCells), st, f, c, omega, state);
}

void launch_mg_prolong(hipStream_t st, const MgLevelDev &f, const MgLevelDev &c, float alpha, const SolveState *state) {
    SHK_MG_LAUNCH(k_mg_prolong, f.bx, blocks((int64_t)f.n_active * kBrickCells), st, f, c, alpha, state);
}

void launch_mg_smooth(hipStream_t st, const MgLevelDev &c, const MgLevelDev *fine, float omega, float alpha,
                      const SolveState *state) {
    const MgLevelDev none{};
    SHK_MG_LAUNCH(k_mg_smooth, c.bx, blocks((int64_t)c.n_active * kBrickCells), st, c, fine ? *fine : none,
                  fine != nullptr, omega, alpha, state);
}

void launch_mg_pre0(hipStream_t st, const MgOp &op, const float *r, const int4 *coord, const int *parent,
                    const MgLevelDev &c1, float omega_c, const SolveState *state) {
    SHK_MG_LAUNCH(k_mg_pre0_w, op.bx, blocks((int64_t)op.n_active * 64), st, op, r, coord, parent, c1, omega_c, state);
}

void launch_mg_post0(hipStream_t st, const MgOp &op, const int4 *coord, const int *parent, const float *r,
                     const float *xc, float *y, const SolveState *state, const Grid &g, double *cell_part,
                     double *part) {
    const unsigned nb = blocks((int64_t)op.n_active * 64);   // one wave per brick
    SHK_MG_LAUNCH(k_mg_post0_w, op.bx, nb, st, op, coord, parent, r, xc, y, state, cell_part);
    hipLaunchKernelGGL(k_mg_fold, dim3((g.blocks + NT - 1) / NT), dim3(NT), 0, st, (int)nb, cell_part, g.blocks, part,
                       state);
}

void launch_mg_coarse(hipStream_t st, int n, const int *idx, const float *inv, const float *b, float *x, int cells,
                      const SolveState *state) {
    hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(kMgCoarseMax), 0, st, n, idx, inv, b, x, cells, state);
}

}  // namespace shk
