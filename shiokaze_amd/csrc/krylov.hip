// krylov.hip -- conjugate gradient kernels for gfx950.
//
// The iteration is PCGSolver::solve (local/include/pcgsolver/pcg_solver.h:246-295).
// Jacobi preconditioning is applied as a symmetric scaling of the system
// (D^-1/2 A D^-1/2 has a unit diagonal), which gives the Jacobi-PCG iterates
// while the kernels never read a diagonal; off-diagonal couplings are zero on
// every face that touches a non-unknown, so no masks are needed either.  In both
// the plain and the scaled form z = r.
//
// Two kernels per iteration, no host round trip and no grid-wide fences:
//   k_dir_spmv   every block folds the previous axpy's partials (r.r, max|r|) in
//                one fixed order: stop test and beta, bit-identical in every
//                block; then s' = r + beta s and q = A s' in one pass, partial s'.q
//   k_axpy_norm  every block folds the s'.q partials: alpha; x += alpha s',
//                r -= alpha q, partials r.r and max|r|
// s is double-buffered (the stencil reads neighbours' old s while s' is written).
// A tile is 512 elements: one 8x8x8 brick, or 512 CSR rows.  Block g of a grid of
// G <= kMaxBlocks processes tiles [g*per, (g+1)*per) after the XCD remap, so each
// XCD streams one contiguous range.
//
// The brick stencil runs one wavefront per brick with one z-column of 8 cells per
// lane (lane = i + 8 j): x and y neighbours come from lane shuffles, z neighbours
// from the lane's own registers.  Each lateral neighbour's facing plane (8 x 8
// cells) is read with one coalesced load per array across the wave and handed to
// the edge lanes by shuffles; the z planes map onto the lanes directly.
#include "krylov.h"

#include <hip/hip_ext.h>

#include "shk_projection.h"

namespace shk {

namespace {

constexpr int NB = 256;               // threads of the Krylov kernels (4 waves)
constexpr int NT = kBrickCells;       // threads of the residual kernel (one per brick cell)
constexpr int NF = 1024;              // threads of the one-workgroup finishing kernels

__device__ inline void chunk_range(int tiles, int per, int &t0, int &t1) {
    const int g = xcd_slot((int)blockIdx.x, (int)gridDim.x);
    t0 = g * per;
    t1 = t0 + per < tiles ? t0 + per : tiles;
}

// Deterministic fold of n partials, the result broadcast to every thread.
template <int NTH>
__device__ inline double fold_sum(const double *p, int n, double *lds, double *bcast) {
    const double v = reduce_partials_sum<NTH>(p, n, lds);
    if (threadIdx.x == 0) *bcast = v;
    __syncthreads();
    const double r = *bcast;
    __syncthreads();
    return r;
}
template <int NTH>
__device__ inline double fold_max(const double *p, int n, double *lds, double *bcast) {
    const double v = reduce_partials_max<NTH>(p, n, lds);
    if (threadIdx.x == 0) *bcast = v;
    __syncthreads();
    const double r = *bcast;
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------- init

template <class T, class TB, bool SCALED>
__global__ __launch_bounds__(NB) void k_cg_init(int64_t n, int tiles, int per, const TB *__restrict__ rhs,
                                                const double *__restrict__ sig, T *__restrict__ x, T *__restrict__ r,
                                                T *__restrict__ s, Reduction red) {
    __shared__ double lds[NB / kWave];
    int t0, t1;
    chunk_range(tiles, per, t0, t1);
    double rr = 0.0, rm = 0.0;
    const int64_t end = (int64_t)t1 * NT < n ? (int64_t)t1 * NT : n;
    for (int64_t e = (int64_t)t0 * NT + threadIdx.x; e < end; e += NB) {
        const double b = (double)rhs[e];
        const T rv = SCALED ? (T)(sig[e] * b) : (T)b;
        x[e] = T(0);
        r[e] = rv;
        s[e] = rv;
        rr += (double)rv * (double)rv;
        // with a scaled system the stop test is on the unscaled residual: max|b|
        rm = fmax(rm, SCALED ? (sig[e] > 0.0 ? fabs(b) : 0.0) : fabs((double)rv));
    }
    const double bs = block_sum<NB>(rr, lds);
    const double bm = block_max<NB>(rm, lds);
    if (threadIdx.x == 0) {
        red.pa[blockIdx.x] = bs;
        red.pb[blockIdx.x] = bm;
    }
}

__global__ __launch_bounds__(NF) void k_finish_init(int nblocks, Reduction red, double tol_rel, int64_t max_iter) {
    __shared__ double lds[NF / kWave];
    const double rho = reduce_partials_sum<NF>(red.pa, nblocks, lds);
    const double bmax = reduce_partials_max<NF>(red.pb, nblocks, lds);
    if (threadIdx.x == 0) {
        SolveState *st = red.st;
        const double tol = tol_rel < 1e-30 ? 1e-30 : tol_rel;  // pcg_solver.h:235
        st->rho_k1 = rho;
        st->rho_k2 = rho;
        st->alpha = 0.0;
        st->beta = 0.0;
        st->sigma = 0.0;
        st->rnorm = bmax;
        st->rnorm0 = bmax;
        st->tol_abs = tol * bmax;
        st->iter = 0;
        st->max_iter = max_iter;
        st->done = 0;
        st->status = SHK_OK;
        // pcg_solver.h:254-271: zero rhs, rho == 0, NaN
        if (bmax == 0.0) {
            st->done = 1;
        } else if (rho == 0.0) {
            st->done = 1;
            st->status = SHK_BREAKDOWN;
        } else if (rho != rho) {
            st->done = 1;
            st->status = SHK_NAN;
        } else if (max_iter <= 0) {
            st->done = 1;
            st->status = SHK_NOT_CONVERGED;
        }
    }
}

// ---------------------------------------------------------------- direction + stencil

// Folds the previous axpy's partials: the stop test of pcg_solver.h:281-284 and
// beta of :287-289.  Returns false when the solve has stopped.  With no
// iteration done yet beta = 0 (s = r, pcg_solver.h:272).
__device__ inline bool direction_step(const Reduction &red, double *lds, double *bc, double &beta) {
    SolveState *st = red.st;
    const int64_t it = st->iter;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (it == 0) {
        beta = 0.0;
        if (red.pz) {
            const double rz = fold_sum<NB>(red.pz, gridDim.x, lds, bc);
            if (lead) st->rho_k1 = rz;
        } else if (lead) {
            st->rho_k1 = st->rho_k2;
        }
        return true;
    }
    const double rho_new = fold_sum<NB>(red.pz ? red.pz : red.pb, gridDim.x, lds, bc);
    const double rn = fold_max<NB>(red.pb + gridDim.x, gridDim.x, lds, bc);
    const bool stop = rn <= st->tol_abs;
    const bool nan = rho_new != rho_new;
    const bool cap = it >= st->max_iter;
    beta = rho_new / st->rho_k2;
    if (lead) {
        st->rnorm = rn;
        if (stop) {
            st->done = 1;
            st->status = SHK_OK;
        } else if (nan) {
            st->done = 1;
            st->status = SHK_NAN;
        } else if (cap) {
            st->done = 1;
            st->status = SHK_NOT_CONVERGED;
        } else {
            st->beta = beta;
            st->rho_k1 = rho_new;
        }
    }
    return !(stop || nan || cap);
}

template <class T>
__device__ inline T lane_shfl(T v, int src) {
    return __shfl(v, src & (kWave - 1), kWave);
}

// Stores (and loads) of the vector streams with the non-temporal hint when NTS: past
// the 256 MB Infinity Cache the streams are touched once per kernel and streaming
// accesses measure faster; a working set that fits keeps the cached ones.
template <bool NTS, class V>
__device__ inline void st_stream(V v, V *p) {
    if constexpr (NTS)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}
template <bool NTS, class V>
__device__ inline V ld_stream(const V *p) {
    if constexpr (NTS)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// Brick order: with ilv == 0 workgroup g walks its own contiguous range of bricks (after
// the XCD remap, so neighbouring bricks share an L2); with ilv != 0 the workgroups walk
// the bricks together, wave w of workgroup g taking bricks 4g + w, 4g + w + 4G, ...
// (G workgroups), so that at any time the whole chip works on about 4G consecutive
// bricks (one z layer of bricks at 512^3) and a brick's y and z neighbours were read a
// layer before or are read now: their halo planes come from the Infinity Cache, not
// from HBM, when the vectors themselves do not fit in it.
//
// FCG: the same pass for flexible CG (r is the preconditioned residual z): beta =
// -(z.q) / delta from fd's z.q partials (none: beta = 0), and the partials s'.q and
// s'.res into fd.part; nothing once fd.fs->stop is set.
template <class T, class C, bool UNIT, int BX, bool NTS, bool FCG>
__global__ __launch_bounds__(NB) void k_dir_spmv(BrickOp<C> op, int per, int ilv, const T *__restrict__ r,
                                                 const T *__restrict__ sin, T *__restrict__ sout,
                                                 T *__restrict__ q, Reduction red, FcgDir fd) {
    using B = BrickShape<BX>;
    constexpr int X = B::X, Y = B::Y;
    __shared__ double lds[NB / kWave];
    __shared__ double bc;
    if constexpr (FCG) {
        if (fd.fs && fd.fs->stop) return;
    } else {
        if (red.st->done) return;
    }
    int t0, t1, step;
    if (ilv) {
        t0 = (int)blockIdx.x * (NB / kWave);
        t1 = op.n_active;
        step = (int)gridDim.x * (NB / kWave);
    } else {
        chunk_range(op.n_active, per, t0, t1);
        step = NB / kWave;
    }
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int i = lane % X, j = lane / X;
    // brick metadata (neighbour slots, uniform coupling) does not depend on beta: the
    // first brick's is loaded before the fold, each next one while the current brick runs
    int slot = t0 + wave;
    int mnb[6] = {-1, -1, -1, -1, -1, -1};
    // uniform bricks: one coupling (and, with a diagonal array, one diagonal value)
    const bool fast = op.kappa && (UNIT || op.kdiag);
    C mkap = C(0), mkd = C(0);
    if (slot < t1) {
#pragma unroll
        for (int f = 0; f < 6; ++f) mnb[f] = op.nbr[6 * slot + f];
        if (fast) {
            mkap = op.kappa[slot];
            if (!UNIT) mkd = op.kdiag[slot];
        }
    }
    double beta = 0.0;
    if constexpr (FCG) {
        if (fd.zq) beta = -fold_sum<NB>(fd.zq, fd.nz, lds, &bc) / fd.fst[0];
    } else {
        if (!direction_step(red, lds, &bc, beta)) return;
    }
    const T b = (T)beta;
    double acc = 0.0, acc2 = 0.0;
    for (; slot < t1; slot += step) {
        const int nxm = mnb[0], nxp = mnb[1], nym = mnb[2], nyp = mnb[3], nzm = mnb[4], nzp = mnb[5];
        const C kap = mkap, kdg = mkd;
        const int nslot = slot + step;
        if (nslot < t1) {
#pragma unroll
            for (int f = 0; f < 6; ++f) mnb[f] = op.nbr[6 * nslot + f];
            if (fast) {
                mkap = op.kappa[nslot];
                if (!UNIT) mkd = op.kdiag[nslot];
            }
        }
        const int64_t base = (int64_t)slot * NT + lane;
        // halo planes of the lateral neighbours, one coalesced load per plane.  x plane
        // (Y x 8 cells): lane L = a + Y c holds cell (y = a, z = c); y plane (X x 8, 3D
        // only): lane L = a + X c holds (x = a, z = c).  The edge lanes fetch them
        // below with a shuffle from lane j + Y k / i + X k.
        const int xa = lane % Y, xc = lane / Y;
        const bool xl = lane < Y * 8;  // lanes holding an x-plane cell (all in 3D, 8 in 2D)
        const int pa = lane % X, pc = lane / X;
        T hxl = T(0), hxh = T(0), hyl = T(0), hyh = T(0);
        if (nxm >= 0 && xl) {
            const int64_t e = (int64_t)nxm * NT + (X - 1) + X * xa + kWave * xc;
            hxl = r[e] + b * sin[e];
        }
        if (nxp >= 0 && xl) {
            const int64_t e = (int64_t)nxp * NT + X * xa + kWave * xc;
            hxh = r[e] + b * sin[e];
        }
        if constexpr (Y > 1) {
            if (nym >= 0) {
                const int64_t e = (int64_t)nym * NT + pa + X * (Y - 1) + kWave * pc;
                hyl = r[e] + b * sin[e];
            }
            if (nyp >= 0) {
                const int64_t e = (int64_t)nyp * NT + pa + kWave * pc;
                hyh = r[e] + b * sin[e];
            }
        }
        T szlo = T(0), szhi = T(0);
        if (nzm >= 0) {
            const int64_t e = (int64_t)nzm * NT + lane + 448;
            szlo = r[e] + b * sin[e];
        }
        if (nzp >= 0) {
            const int64_t e = (int64_t)nzp * NT + lane;
            szhi = r[e] + b * sin[e];
        }
        T sv[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int64_t e = base + 64 * k;
            sv[k] = r[e] + b * sin[e];
        }
        if (kap != C(0)) {
            // uniform brick: one coupling (and diagonal) value, no coefficient traffic
            const T kv = (T)kap;
            const T kd = (T)kdg;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const T sk = sv[k];
                T xm = lane_shfl(sk, lane - 1);
                T xp = lane_shfl(sk, lane + 1);
                const T hxm = lane_shfl(hxl, j + Y * k);
                const T hxp = lane_shfl(hxh, j + Y * k);
                if (i == 0) xm = hxm;
                if (i == X - 1) xp = hxp;
                T ym = T(0), yp = T(0);
                if constexpr (Y > 1) {
                    ym = lane_shfl(sk, lane - X);
                    yp = lane_shfl(sk, lane + X);
                    const T hym = lane_shfl(hyl, i + X * k);
                    const T hyp = lane_shfl(hyh, i + X * k);
                    if (j == 0) ym = hym;
                    if (j == Y - 1) yp = hyp;
                }
                const T zm = k > 0 ? sv[k - 1] : szlo;
                const T zp = k < 7 ? sv[k + 1] : szhi;
                // the general path's expression with the brick's constants: bit-identical
                T qv = UNIT ? sk : kd * sk;
                qv -= kv * xm + kv * xp + kv * ym + kv * yp + kv * zm + kv * zp;
                const int64_t e = base + 64 * k;
                st_stream<NTS>(sk, &sout[e]);
                st_stream<NTS>(qv, &q[e]);
                acc += (double)sk * (double)qv;
                if constexpr (FCG) acc2 += (double)sk * (double)ld_stream<NTS>(&fd.res[e]);
            }
            continue;
        }
        T cxv[8], cyv[8], czv[8], dv[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int64_t e = base + 64 * k;
            cxv[k] = (T)op.cx[e];
            cyv[k] = Y > 1 ? (T)op.cy[e] : T(0);
            czv[k] = (T)op.cz[e];
            if (!UNIT) dv[k] = (T)op.diag[e];
        }
        // couplings across the +x / +y / +z faces live in the neighbour bricks
        T cxh = T(0), cyh = T(0), czhi = T(0);
        if (nxp >= 0 && xl) cxh = (T)op.cx[(int64_t)nxp * NT + X * xa + kWave * xc];
        if constexpr (Y > 1) {
            if (nyp >= 0) cyh = (T)op.cy[(int64_t)nyp * NT + pa + kWave * pc];
        }
        if (nzp >= 0) czhi = (T)op.cz[(int64_t)nzp * NT + lane];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const T sk = sv[k];
            T xm = lane_shfl(sk, lane - 1);
            T xp = lane_shfl(sk, lane + 1);
            T cxp = lane_shfl(cxv[k], lane + 1);
            const T hxm = lane_shfl(hxl, j + Y * k);
            const T hxp = lane_shfl(hxh, j + Y * k);
            const T hcx = lane_shfl(cxh, j + Y * k);
            if (i == 0) xm = hxm;
            if (i == X - 1) {
                xp = hxp;
                cxp = hcx;
            }
            T ym = T(0), yp = T(0), cyp = T(0);
            if constexpr (Y > 1) {
                ym = lane_shfl(sk, lane - X);
                yp = lane_shfl(sk, lane + X);
                cyp = lane_shfl(cyv[k], lane + X);
                const T hym = lane_shfl(hyl, i + X * k);
                const T hyp = lane_shfl(hyh, i + X * k);
                const T hcy = lane_shfl(cyh, i + X * k);
                if (j == 0) ym = hym;
                if (j == Y - 1) {
                    yp = hyp;
                    cyp = hcy;
                }
            }
            const T zm = k > 0 ? sv[k - 1] : szlo;
            const T zp = k < 7 ? sv[k + 1] : szhi;
            const T czp = k < 7 ? czv[k + 1] : czhi;
            T qv = UNIT ? sk : dv[k] * sk;
            qv -= cxv[k] * xm + cxp * xp + cyv[k] * ym + cyp * yp + czv[k] * zm + czp * zp;
            const int64_t e = base + 64 * k;
            st_stream<NTS>(sk, &sout[e]);
            st_stream<NTS>(qv, &q[e]);
            acc += (double)sk * (double)qv;
            if constexpr (FCG) acc2 += (double)sk * (double)ld_stream<NTS>(&fd.res[e]);
        }
    }
    const double bs = block_sum<NB>(acc, lds);
    if constexpr (FCG) {
        const double b2 = block_sum<NB>(acc2, lds);
        if (threadIdx.x == 0) {
            fd.part[blockIdx.x] = bs;
            fd.part[gridDim.x + blockIdx.x] = b2;
        }
    } else {
        if (threadIdx.x == 0) red.pa[blockIdx.x] = bs;
    }
}

__global__ __launch_bounds__(NB) void k_dir_spmv_csr(CsrOp op, int per, const double *__restrict__ r,
                                                     const double *__restrict__ sin, double *__restrict__ sout,
                                                     double *__restrict__ q, Reduction red) {
    __shared__ double lds[NB / kWave];
    __shared__ double bc;
    if (red.st->done) return;
    double beta;
    if (!direction_step(red, lds, &bc, beta)) return;
    int t0, t1;
    chunk_range(blocks_for(op.rows), per, t0, t1);
    const int64_t end = (int64_t)t1 * NT < op.rows ? (int64_t)t1 * NT : op.rows;
    double acc = 0.0;
    for (int64_t e = (int64_t)t0 * NT + threadIdx.x; e < end; e += NB) {
        double a = 0.0;
        for (int64_t p = op.rowptr[e]; p < op.rowptr[e + 1]; ++p) {
            const int64_t c = op.cols[p];
            a += op.vals[p] * (r[c] + beta * sin[c]);
        }
        const double se = r[e] + beta * sin[e];
        sout[e] = se;
        q[e] = a;
        acc += se * a;
    }
    const double bs = block_sum<NB>(acc, lds);
    if (threadIdx.x == 0) red.pa[blockIdx.x] = bs;
}

// ---------------------------------------------------------------- axpy + norm

template <class T, bool SCALED, bool NTS>
__global__ __launch_bounds__(NB) void k_axpy_norm(int64_t n, int tiles, int per, const double *__restrict__ sig,
                                                  T *__restrict__ x, T *__restrict__ r, const T *__restrict__ s,
                                                  const T *__restrict__ q, Reduction red) {
    __shared__ double lds[NB / kWave];
    __shared__ double bc;
    SolveState *st = red.st;
    if (st->done) return;
    const double sigma = fold_sum<NB>(red.pa, gridDim.x, lds, &bc);
    const double rho = st->rho_k1;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (!(sigma > 0.0)) {
        if (lead) {
            st->sigma = sigma;
            st->done = 1;
            st->status = (sigma != sigma) ? SHK_NAN : SHK_BREAKDOWN;
        }
        return;
    }
    const double alpha = rho / sigma;
    if (lead) {
        st->sigma = sigma;
        st->alpha = alpha;
        st->rho_k2 = rho;
        st->iter = st->iter + 1;
    }
    const T a = (T)alpha;
    int t0, t1;
    chunk_range(tiles, per, t0, t1);
    const int64_t end = (int64_t)t1 * NT < n ? (int64_t)t1 * NT : n;
    double rr = 0.0, rm = 0.0;
    constexpr int V = 16 / (int)sizeof(T);  // elements per 16-byte access
    {
        // 16-byte vectors: V consecutive elements per thread (tile boundaries are
        // multiples of 4, the arrays 256-byte aligned; a remainder of fewer than V
        // elements at the end goes through the scalar loop); each element's arithmetic
        // is that of the scalar loop below
        typedef T TV __attribute__((ext_vector_type(V)));
        TV *__restrict__ xv4 = reinterpret_cast<TV *>(x);
        TV *__restrict__ rv4 = reinterpret_cast<TV *>(r);
        const TV *__restrict__ sv4 = reinterpret_cast<const TV *>(s);
        const TV *__restrict__ qv4 = reinterpret_cast<const TV *>(q);
        for (int64_t e = (int64_t)t0 * (NT / V) + threadIdx.x; e < end / V; e += NB) {
            TV xv = ld_stream<NTS>(&xv4[e]), rv = ld_stream<NTS>(&rv4[e]);
            const TV sv = ld_stream<NTS>(&sv4[e]), qv = ld_stream<NTS>(&qv4[e]);
#pragma unroll
            for (int u = 0; u < V; ++u) {
                xv[u] += a * sv[u];
                rv[u] = rv[u] - a * qv[u];
            }
            st_stream<NTS>(xv, &xv4[e]);
            st_stream<NTS>(rv, &rv4[e]);
#pragma unroll
            for (int u = 0; u < V; ++u) rr += (double)rv[u] * (double)rv[u];
            if constexpr (SCALED) {
                // max |r| / sig over the unknowns (sig > 0), sig read as double pairs
                typedef double D2 __attribute__((ext_vector_type(2)));
                const D2 *sg2 = reinterpret_cast<const D2 *>(sig) + e * (V / 2);
#pragma unroll
                for (int h = 0; h < V / 2; ++h) {
                    const D2 sg = ld_stream<NTS>(&sg2[h]);
#pragma unroll
                    for (int u = 0; u < 2; ++u)
                        if (sg[u] > 0.0) rm = fmax(rm, fabs((double)rv[2 * h + u]) / sg[u]);
                }
            } else if constexpr (V == 4) {
                rm = fmax(rm, fmax(fmax(fabs((double)rv[0]), fabs((double)rv[1])),
                                   fmax(fabs((double)rv[2]), fabs((double)rv[3]))));
            } else {
                rm = fmax(rm, fmax(fabs((double)rv[0]), fabs((double)rv[1])));
            }
        }
    }
    {
        const int64_t first = (int64_t)t0 * NT;  // a block past the last tile has end <= first
        const int64_t start = (end / V) * V > first ? (end / V) * V : first;
        for (int64_t e = start + threadIdx.x; e < end; e += NB) {
            x[e] += a * s[e];
            const T rv = r[e] - a * q[e];
            r[e] = rv;
            rr += (double)rv * (double)rv;
            if (SCALED) {
                const double sg = sig[e];
                if (sg > 0.0) rm = fmax(rm, fabs((double)rv) / sg);
            } else {
                rm = fmax(rm, fabs((double)rv));
            }
        }
    }
    const double bs = block_sum<NB>(rr, lds);
    const double bm = block_max<NB>(rm, lds);
    if (threadIdx.x == 0) {
        red.pb[blockIdx.x] = bs;
        red.pb[gridDim.x + blockIdx.x] = bm;
    }
}

// ---------------------------------------------------------------- true residual

// (A v)[e] for the brick operator; UNIT: unit diagonal (diag unused)
template <class T, class C, int BX, bool UNIT = false>
__device__ inline T brick_apply(const int *__restrict__ nb6, int64_t e, int l, const C *__restrict__ diag,
                                const C *__restrict__ cx, const C *__restrict__ cy, const C *__restrict__ cz,
                                const T *__restrict__ v) {
    using B = BrickShape<BX>;
    constexpr int X = B::X, Y = B::Y;
    const int i = l % X, j = (l / X) % Y, k = l / kWave;
    T acc = UNIT ? v[e] : (T)diag[e] * v[e];
    if (i > 0) {
        acc -= (T)cx[e] * v[e - 1];
    } else {
        const int nb = nb6[0];
        if (nb >= 0) acc -= (T)cx[e] * v[(int64_t)nb * NT + l + (X - 1)];
    }
    if (i < X - 1) {
        acc -= (T)cx[e + 1] * v[e + 1];
    } else {
        const int nb = nb6[1];
        if (nb >= 0) {
            const int64_t m = (int64_t)nb * NT + l - (X - 1);
            acc -= (T)cx[m] * v[m];
        }
    }
    if constexpr (Y > 1) {
        if (j > 0) {
            acc -= (T)cy[e] * v[e - X];
        } else {
            const int nb = nb6[2];
            if (nb >= 0) acc -= (T)cy[e] * v[(int64_t)nb * NT + l + X * (Y - 1)];
        }
        if (j < Y - 1) {
            acc -= (T)cy[e + X] * v[e + X];
        } else {
            const int nb = nb6[3];
            if (nb >= 0) {
                const int64_t m = (int64_t)nb * NT + l - X * (Y - 1);
                acc -= (T)cy[m] * v[m];
            }
        }
    }
    if (k > 0) {
        acc -= (T)cz[e] * v[e - 64];
    } else {
        const int nb = nb6[4];
        if (nb >= 0) acc -= (T)cz[e] * v[(int64_t)nb * NT + l + 448];
    }
    if (k < 7) {
        acc -= (T)cz[e + 64] * v[e + 64];
    } else {
        const int nb = nb6[5];
        if (nb >= 0) {
            const int64_t m = (int64_t)nb * NT + l - 448;
            acc -= (T)cz[m] * v[m];
        }
    }
    return acc;
}

template <int BX>
__global__ __launch_bounds__(NT) void k_residual(BrickOp<double> op, int per, const double *__restrict__ b,
                                                 const double *__restrict__ x, double *__restrict__ r,
                                                 double *__restrict__ part, const double *__restrict__ sig,
                                                 float *__restrict__ rs) {
    __shared__ double lds[NT / kWave];
    int t0, t1;
    chunk_range(op.n_active, per, t0, t1);
    const int l = threadIdx.x;
    double m = 0.0, s2 = 0.0, b2 = 0.0, bm = 0.0;
    for (int slot = t0; slot < t1; ++slot) {
        const int64_t e = (int64_t)slot * NT + l;
        double rv = 0.0, bv = 0.0;
        if (op.diag[e] != 0.0) {
            bv = b[e];
            rv = bv - brick_apply<double, double, BX>(op.nbr + 6 * slot, e, l, op.diag, op.cx, op.cy, op.cz, x);
        }
        if (r) r[e] = rv;
        if (rs) rs[e] = (float)(sig ? sig[e] * rv : rv);
        m = fmax(m, fabs(rv));
        s2 += rv * rv;
        b2 += bv * bv;
        bm = fmax(bm, fabs(bv));
    }
    const double v0 = block_max<NT>(m, lds);
    const double v1 = block_sum<NT>(s2, lds);
    const double v2 = block_sum<NT>(b2, lds);
    const double v3 = block_max<NT>(bm, lds);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = v0;
        part[gridDim.x + blockIdx.x] = v1;
        part[2 * gridDim.x + blockIdx.x] = v2;
        part[3 * gridDim.x + blockIdx.x] = v3;
    }
}

__global__ __launch_bounds__(NT) void k_residual_csr(CsrOp op, int per, const double *__restrict__ b,
                                                     const double *__restrict__ x, double *__restrict__ r,
                                                     double *__restrict__ part) {
    __shared__ double lds[NT / kWave];
    int t0, t1;
    chunk_range(blocks_for(op.rows), per, t0, t1);
    double m = 0.0, s2 = 0.0, b2 = 0.0, bm = 0.0;
    for (int t = t0; t < t1; ++t) {
        const int64_t e = (int64_t)t * NT + threadIdx.x;
        if (e >= op.rows) continue;
        double a = 0.0;
        for (int64_t p = op.rowptr[e]; p < op.rowptr[e + 1]; ++p) a += op.vals[p] * x[op.cols[p]];
        const double bv = b[e];
        const double rv = bv - a;
        if (r) r[e] = rv;
        m = fmax(m, fabs(rv));
        s2 += rv * rv;
        b2 += bv * bv;
        bm = fmax(bm, fabs(bv));
    }
    const double v0 = block_max<NT>(m, lds);
    const double v1 = block_sum<NT>(s2, lds);
    const double v2 = block_sum<NT>(b2, lds);
    const double v3 = block_max<NT>(bm, lds);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = v0;
        part[gridDim.x + blockIdx.x] = v1;
        part[2 * gridDim.x + blockIdx.x] = v2;
        part[3 * gridDim.x + blockIdx.x] = v3;
    }
}

__global__ __launch_bounds__(NF) void k_finish_residual(int nblocks, const double *__restrict__ part,
                                                       double *__restrict__ out) {
    __shared__ double lds[NF / kWave];
    const double v0 = reduce_partials_max<NF>(part, nblocks, lds);
    const double v1 = reduce_partials_sum<NF>(part + nblocks, nblocks, lds);
    const double v2 = reduce_partials_sum<NF>(part + 2 * nblocks, nblocks, lds);
    const double v3 = reduce_partials_max<NF>(part + 3 * nblocks, nblocks, lds);
    if (threadIdx.x == 0) {
        out[0] = v0;
        out[1] = v1;
        out[2] = v2;
        out[3] = v3;
    }
}

// ---------------------------------------------------------------- misc

__global__ __launch_bounds__(NT) void k_add_correction(int64_t n, double *__restrict__ x, const float *__restrict__ d,
                                                       const double *__restrict__ sig) {
    const int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x;
    if (e < n) x[e] += sig ? sig[e] * (double)d[e] : (double)d[e];
}

__global__ __launch_bounds__(NT) void k_unscale(int64_t n, double *__restrict__ x, const double *__restrict__ sig) {
    const int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x;
    if (e < n) x[e] *= sig[e];
}

__global__ __launch_bounds__(NT) void k_csr_diag(CsrOp op, double *__restrict__ diag, double *__restrict__ sig) {
    const int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x;
    if (e < op.rows) {
        double dv = 0.0;
        for (int64_t p = op.rowptr[e]; p < op.rowptr[e + 1]; ++p)
            if (op.cols[p] == e) dv = op.vals[p];
        diag[e] = dv;
        if (sig) sig[e] = dv > 0.0 ? 1.0 / sqrt(dv) : 0.0;
    }
}

// Scaled CSR copy: out[p] = sig[row] * vals[p] * sig[col].
__global__ __launch_bounds__(NT) void k_csr_scale(CsrOp op, const double *__restrict__ sig, double *__restrict__ out) {
    const int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x;
    if (e < op.rows) {
        const double se = sig[e];
        for (int64_t p = op.rowptr[e]; p < op.rowptr[e + 1]; ++p) out[p] = se * op.vals[p] * sig[op.cols[p]];
    }
}

}  // namespace

// ---------------------------------------------------------------- flexible CG on bricks

// FCG(1) (Notay 2000) on FP32 brick vectors of a unit-diagonal operator, the inner
// solve of the learned preconditioner.  Scalars never leave the device: each kernel
// folds the previous one's FP64 partials in one fixed order in every block.
// fst[0] = delta = p.q, fst[1] = alpha, fst[2] = 1 if delta <= 0 or not finite.

// alpha = (p.r) / (p.q) from the direction kernel's partials; x += alpha p, r -= alpha q;
// part2[b] = r.r, part2[gridDim + b] = max|r|.  norm_only: no update, just the norms.
__global__ __launch_bounds__(NB) void k_fcg_axpy(int64_t n, int tiles, int per, float *__restrict__ x,
                                                 float *__restrict__ r, const float *__restrict__ p,
                                                 const float *__restrict__ q, const double *__restrict__ part1,
                                                 double *__restrict__ part2, double *__restrict__ fst, int norm_only,
                                                 const FcgState *__restrict__ fs) {
    __shared__ double lds[NB / kWave];
    __shared__ double bc;
    if (fs && fs->stop) return;
    float a = 0.0f;
    if (!norm_only) {
        const double delta = fold_sum<NB>(part1, gridDim.x, lds, &bc);
        const double pr = fold_sum<NB>(part1 + gridDim.x, gridDim.x, lds, &bc);
        const double alpha = pr / delta;
        const bool bad = !(delta > 0.0) || !(alpha == alpha) || isinf(alpha);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            fst[0] = delta;
            fst[1] = bad ? 0.0 : alpha;
            fst[2] = bad ? 1.0 : 0.0;
        }
        a = bad ? 0.0f : (float)alpha;
    }
    int t0, t1;
    chunk_range(tiles, per, t0, t1);
    const int64_t end = (int64_t)t1 * NT < n ? (int64_t)t1 * NT : n;
    double rr = 0.0, rm = 0.0;
    // four elements per thread and access (n is a multiple of the 512-element tile)
    typedef float f4 __attribute__((ext_vector_type(4)));
    for (int64_t e = (int64_t)t0 * NT + 4 * threadIdx.x; e < end; e += 4 * NB) {
        f4 rv = *reinterpret_cast<const f4 *>(r + e);
        if (!norm_only) {
            const f4 pv = *reinterpret_cast<const f4 *>(p + e), qv = *reinterpret_cast<const f4 *>(q + e);
            f4 xv = *reinterpret_cast<const f4 *>(x + e);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                xv[k] = xv[k] + a * pv[k];
                rv[k] = rv[k] - a * qv[k];
            }
            *reinterpret_cast<f4 *>(x + e) = xv;
            *reinterpret_cast<f4 *>(r + e) = rv;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            rr += (double)rv[k] * (double)rv[k];
            rm = fmax(rm, fabs((double)rv[k]));
        }
    }
    const double bs = block_sum<NB>(rr, lds);
    const double bm = block_max<NB>(rm, lds);
    if (threadIdx.x == 0) {
        part2[blockIdx.x] = bs;
        part2[gridDim.x + blockIdx.x] = bm;
    }
}

__global__ void k_fcg_state_init(FcgState *fs, FcgState v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *fs = v;
}

// ---------------------------------------------------------------- launchers

Grid make_grid(int64_t n, int max_blocks) {
    Grid g;
    g.n = n;
    g.tiles = blocks_for(n);
    if (max_blocks < 1) max_blocks = kMaxBlocks;
    g.per = (g.tiles + max_blocks - 1) / max_blocks;
    if (g.per < 1) g.per = 1;
    g.blocks = (g.tiles + g.per - 1) / g.per;
    if (g.blocks < 1) g.blocks = 1;
    return g;
}

template <class T, class TB>
void launch_cg_init(hipStream_t st, const Grid &g, const TB *rhs, const double *sig, T *x, T *r, T *s, Reduction red,
                    double tol_rel, int64_t max_iter) {
    if (sig)
        hipLaunchKernelGGL((k_cg_init<T, TB, true>), dim3(g.blocks), dim3(NB), 0, st, g.n, g.tiles, g.per, rhs, sig, x,
                           r, s, red);
    else
        hipLaunchKernelGGL((k_cg_init<T, TB, false>), dim3(g.blocks), dim3(NB), 0, st, g.n, g.tiles, g.per, rhs, sig,
                           x, r, s, red);
    hipLaunchKernelGGL(k_finish_init, dim3(1), dim3(NF), 0, st, g.blocks, red, tol_rel, max_iter);
}

// plain launch, or one whose dispatch records ev.start / ev.stop
#define SHK_LAUNCH(ev, kernel, grid, block, st, ...)                                                 \
    do {                                                                                            \
        if ((ev).start)                                                                             \
            hipExtLaunchKernelGGL(kernel, grid, block, 0, st, (ev).start, (ev).stop, 0, __VA_ARGS__); \
        else                                                                                        \
            hipLaunchKernelGGL(kernel, grid, block, 0, st, __VA_ARGS__);                            \
    } while (0)

// the vector streams bypass the caches (non-temporal accesses) when the Krylov working
// set (about six vectors) exceeds the 256 MB Infinity Cache
constexpr int64_t kStreamBytes = 256ll << 20;
template <class T>
bool streaming(const Grid &g) {
    return g.n * 6 * (int64_t)sizeof(T) > kStreamBytes;
}

template <class T, class C, bool NTS>
void launch_dir_spmv_nt(hipStream_t st, const Grid &g, const BrickOp<C> &op, int ilv, const T *r, const T *s_in,
                        T *s_out, T *q, Reduction red, KEvents ev) {
    if (op.diag) {
        if (op.bx == 64)
            SHK_LAUNCH(ev, (k_dir_spmv<T, C, false, 64, NTS, false>), dim3(g.blocks), dim3(NB), st, op, g.per, ilv, r,
                       s_in, s_out, q, red, FcgDir{});
        else
            SHK_LAUNCH(ev, (k_dir_spmv<T, C, false, 8, NTS, false>), dim3(g.blocks), dim3(NB), st, op, g.per, ilv, r,
                       s_in, s_out, q, red, FcgDir{});
    } else {
        if (op.bx == 64)
            SHK_LAUNCH(ev, (k_dir_spmv<T, C, true, 64, NTS, false>), dim3(g.blocks), dim3(NB), st, op, g.per, ilv, r,
                       s_in, s_out, q, red, FcgDir{});
        else
            SHK_LAUNCH(ev, (k_dir_spmv<T, C, true, 8, NTS, false>), dim3(g.blocks), dim3(NB), st, op, g.per, ilv, r,
                       s_in, s_out, q, red, FcgDir{});
    }
}

template <class T, class C, bool NTS>
const void *dir_spmv_kernel(const BrickOp<C> &op) {
    if (op.diag)
        return op.bx == 64 ? (const void *)k_dir_spmv<T, C, false, 64, NTS, false>
                           : (const void *)k_dir_spmv<T, C, false, 8, NTS, false>;
    return op.bx == 64 ? (const void *)k_dir_spmv<T, C, true, 64, NTS, false>
                       : (const void *)k_dir_spmv<T, C, true, 8, NTS, false>;
}

template <class T, class C>
int dir_spmv_resident(const Grid &g, const BrickOp<C> &op) {
    const void *k = streaming<T>(g) ? dir_spmv_kernel<T, C, true>(op) : dir_spmv_kernel<T, C, false>(op);
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, NB, 0) != hipSuccess || per_cu < 1 ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        return g.blocks;
    return per_cu * cus < g.blocks ? per_cu * cus : g.blocks;
}

template <class T, class C>
void launch_dir_spmv(hipStream_t st, const Grid &g, const BrickOp<C> &op, const T *r, const T *s_in, T *s_out, T *q,
                     Reduction red, KEvents ev) {
    const int ilv = streaming<T>(g) ? 1 : 0;
    if (streaming<T>(g))
        launch_dir_spmv_nt<T, C, true>(st, g, op, ilv, r, s_in, s_out, q, red, ev);
    else
        launch_dir_spmv_nt<T, C, false>(st, g, op, ilv, r, s_in, s_out, q, red, ev);
}

void launch_fcg_dir_spmv(hipStream_t st, const Grid &g, const BrickOp<float> &op, const float *z, const float *p_in,
                         float *p_out, float *q, const FcgDir &fd) {
    const int ilv = streaming<float>(g) ? 1 : 0;
    if (ilv)
        hipLaunchKernelGGL((k_dir_spmv<float, float, true, 8, true, true>), dim3(g.blocks), dim3(NB), 0, st, op, g.per,
                           ilv, z, p_in, p_out, q, Reduction{}, fd);
    else
        hipLaunchKernelGGL((k_dir_spmv<float, float, true, 8, false, true>), dim3(g.blocks), dim3(NB), 0, st, op,
                           g.per, ilv, z, p_in, p_out, q, Reduction{}, fd);
}

void launch_dir_spmv_csr(hipStream_t st, const Grid &g, const CsrOp &op, const double *r, const double *s_in,
                         double *s_out, double *q, Reduction red, KEvents ev) {
    SHK_LAUNCH(ev, k_dir_spmv_csr, dim3(g.blocks), dim3(NB), st, op, g.per, r, s_in, s_out, q, red);
}

template <class T>
void launch_axpy_norm(hipStream_t st, const Grid &g, const double *sig, T *x, T *r, const T *s, const T *q,
                      Reduction red, KEvents ev) {
    const bool nts = streaming<T>(g);
    if (sig && nts)
        SHK_LAUNCH(ev, (k_axpy_norm<T, true, true>), dim3(g.blocks), dim3(NB), st, g.n, g.tiles, g.per, sig, x, r, s,
                   q, red);
    else if (sig)
        SHK_LAUNCH(ev, (k_axpy_norm<T, true, false>), dim3(g.blocks), dim3(NB), st, g.n, g.tiles, g.per, sig, x, r, s,
                   q, red);
    else if (nts)
        SHK_LAUNCH(ev, (k_axpy_norm<T, false, true>), dim3(g.blocks), dim3(NB), st, g.n, g.tiles, g.per, sig, x, r, s,
                   q, red);
    else
        SHK_LAUNCH(ev, (k_axpy_norm<T, false, false>), dim3(g.blocks), dim3(NB), st, g.n, g.tiles, g.per, sig, x, r, s,
                   q, red);
}

void launch_residual(hipStream_t st, const Grid &g, const BrickOp<double> &op, const double *b, const double *x,
                     double *r, double *part, double *out, const double *sig, float *rs) {
    auto kern = op.bx == 64 ? k_residual<64> : k_residual<8>;
    hipLaunchKernelGGL(kern, dim3(g.blocks), dim3(NT), 0, st, op, g.per, b, x, r, part, sig, rs);
    hipLaunchKernelGGL(k_finish_residual, dim3(1), dim3(NF), 0, st, g.blocks, part, out);
}

void launch_residual_csr(hipStream_t st, const Grid &g, const CsrOp &op, const double *b, const double *x, double *r,
                         double *part, double *out) {
    hipLaunchKernelGGL(k_residual_csr, dim3(g.blocks), dim3(NT), 0, st, op, g.per, b, x, r, part);
    hipLaunchKernelGGL(k_finish_residual, dim3(1), dim3(NF), 0, st, g.blocks, part, out);
}

void launch_add_correction(hipStream_t st, int64_t n, double *x, const float *d, const double *sig) {
    hipLaunchKernelGGL(k_add_correction, dim3(blocks_for(n)), dim3(NT), 0, st, n, x, d, sig);
}

void launch_unscale(hipStream_t st, int64_t n, double *x, const double *sig) {
    hipLaunchKernelGGL(k_unscale, dim3(blocks_for(n)), dim3(NT), 0, st, n, x, sig);
}

void launch_csr_diag(hipStream_t st, const CsrOp &op, double *diag, double *sig) {
    hipLaunchKernelGGL(k_csr_diag, dim3(blocks_for(op.rows)), dim3(NT), 0, st, op, diag, sig);
}

void launch_csr_scale(hipStream_t st, const CsrOp &op, const double *sig, double *out) {
    hipLaunchKernelGGL(k_csr_scale, dim3(blocks_for(op.rows)), dim3(NT), 0, st, op, sig, out);
}

void launch_fcg_axpy(hipStream_t st, const Grid &g, float *x, float *r, const float *p, const float *q,
                     const double *part1, double *part2, double *fst, bool norm_only, const FcgState *fs) {
    hipLaunchKernelGGL(k_fcg_axpy, dim3(g.blocks), dim3(NB), 0, st, g.n, g.tiles, g.per, x, r, p, q, part1, part2,
                       fst, norm_only ? 1 : 0, fs);
}

void launch_fcg_state_init(hipStream_t st, FcgState *fs, const FcgState &v) {
    hipLaunchKernelGGL(k_fcg_state_init, dim3(1), dim3(64), 0, st, fs, v);
}

template void launch_cg_init<double, double>(hipStream_t, const Grid &, const double *, const double *, double *,
                                             double *, double *, Reduction, double, int64_t);
template void launch_cg_init<float, double>(hipStream_t, const Grid &, const double *, const double *, float *,
                                            float *, float *, Reduction, double, int64_t);
template void launch_cg_init<float, float>(hipStream_t, const Grid &, const float *, const double *, float *, float *,
                                           float *, Reduction, double, int64_t);
template int dir_spmv_resident<double, double>(const Grid &, const BrickOp<double> &);
template int dir_spmv_resident<float, float>(const Grid &, const BrickOp<float> &);
template void launch_dir_spmv<double, double>(hipStream_t, const Grid &, const BrickOp<double> &, const double *,
                                              const double *, double *, double *, Reduction, KEvents);
template void launch_dir_spmv<float, float>(hipStream_t, const Grid &, const BrickOp<float> &, const float *,
                                            const float *, float *, float *, Reduction, KEvents);
template void launch_axpy_norm<double>(hipStream_t, const Grid &, const double *, double *, double *, const double *,
                                       const double *, Reduction, KEvents);
template void launch_axpy_norm<float>(hipStream_t, const Grid &, const double *, float *, float *, const float *,
                                      const float *, Reduction, KEvents);

}  // namespace shk
