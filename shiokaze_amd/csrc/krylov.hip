// krylov.hip -- fused conjugate gradient kernels for gfx950.
//
// One workgroup of 512 threads (8 waves of 64) per 8x8x8 brick / per 512 rows.
// Per iteration:
//   k_spmv_dot    q = A s, partial s.q          -> last block: alpha
//   k_axpy_norm   x += alpha s, r -= alpha q, z = M^-1 r, partials r.z, max|r|
//                                               -> last block: stop test, beta
//   k_update_dir  s = z + beta s
// Dot products are accumulated in FP64 whatever the vector type.
#include "krylov.h"

#include "shk_projection.h"

namespace shk {

namespace {

constexpr int NT = kBrickCells;

__device__ inline int64_t flat_elem(int64_t n) {
    return (int64_t)xcd_slot((int)blockIdx.x, (int)gridDim.x) * NT + threadIdx.x;
}

template <class T>
__device__ inline T absv(T v) {
    return v < T(0) ? -v : v;
}

// ---------------------------------------------------------------- init

template <class T, class C, class TB, bool JAC>
__global__ __launch_bounds__(NT) void k_cg_init(int64_t n, const TB *__restrict__ rhs, const C *__restrict__ diag,
                                                T *__restrict__ x, T *__restrict__ r, T *__restrict__ s, Reduction red,
                                                double tol_rel, int64_t max_iter) {
    __shared__ double lds[NT / kWave];
    const int64_t e = flat_elem(n);
    double rz = 0.0, rm = 0.0;
    if (e < n) {
        const C d = diag[e];
        const T rv = d != C(0) ? (T)rhs[e] : T(0);
        const T z = (JAC && d != C(0)) ? rv / (T)d : rv;
        x[e] = T(0);
        r[e] = rv;
        s[e] = z;
        rz = (double)rv * (double)z;
        rm = fabs((double)rv);
    }
    const double bs = block_sum<NT>(rz, lds);
    const double bm = block_max<NT>(rm, lds);
    if (threadIdx.x == 0) {
        red.pa[blockIdx.x] = bs;
        red.pb[blockIdx.x] = bm;
    }
    if (last_block_arrive(&red.st->ticket[kTicketInit])) {
        const double rho = reduce_partials_sum<NT>(red.pa, gridDim.x, lds);
        const double bmax = reduce_partials_max<NT>(red.pb, gridDim.x, lds);
        if (threadIdx.x == 0) {
            SolveState *st = red.st;
            double tol = tol_rel < 1e-30 ? 1e-30 : tol_rel;  // pcg_solver.h:235
            st->rho = rho;
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
}

// ---------------------------------------------------------------- stencil

template <class T, class C>
__device__ inline T brick_apply(const int *__restrict__ nb6, int64_t e, int l, C dval, const C *__restrict__ cx,
                                const C *__restrict__ cy, const C *__restrict__ cz, const T *__restrict__ v) {
    const int i = l & 7, j = (l >> 3) & 7, k = l >> 6;
    T acc = (T)dval * v[e];
    // -x / +x
    if (i > 0) {
        acc -= (T)cx[e] * v[e - 1];
    } else {
        const int nb = nb6[0];
        if (nb >= 0) acc -= (T)cx[e] * v[(int64_t)nb * NT + l + 7];
    }
    if (i < 7) {
        acc -= (T)cx[e + 1] * v[e + 1];
    } else {
        const int nb = nb6[1];
        if (nb >= 0) {
            const int64_t m = (int64_t)nb * NT + l - 7;
            acc -= (T)cx[m] * v[m];
        }
    }
    // -y / +y
    if (j > 0) {
        acc -= (T)cy[e] * v[e - 8];
    } else {
        const int nb = nb6[2];
        if (nb >= 0) acc -= (T)cy[e] * v[(int64_t)nb * NT + l + 56];
    }
    if (j < 7) {
        acc -= (T)cy[e + 8] * v[e + 8];
    } else {
        const int nb = nb6[3];
        if (nb >= 0) {
            const int64_t m = (int64_t)nb * NT + l - 56;
            acc -= (T)cy[m] * v[m];
        }
    }
    // -z / +z
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

__device__ inline void finish_alpha(const Reduction &red, double *lds) {
    const double sigma = reduce_partials_sum<NT>(red.pa, gridDim.x, lds);
    if (threadIdx.x == 0) {
        SolveState *st = red.st;
        st->sigma = sigma;
        st->alpha = st->rho / sigma;
        if (!(sigma > 0.0)) {
            st->done = 1;
            st->status = (sigma != sigma) ? SHK_NAN : SHK_BREAKDOWN;
        }
    }
}

template <class T, class C>
__global__ __launch_bounds__(NT) void k_spmv_dot(BrickOp<C> op, const T *__restrict__ s, T *__restrict__ q,
                                                 Reduction red) {
    __shared__ double lds[NT / kWave];
    if (red.st->done) return;
    const int slot = xcd_slot((int)blockIdx.x, (int)gridDim.x);
    const int l = threadIdx.x;
    const int64_t e = (int64_t)slot * NT + l;
    const C d = op.diag[e];
    T qv = T(0);
    double prod = 0.0;
    if (d != C(0)) {
        qv = brick_apply<T, C>(op.nbr + 6 * slot, e, l, d, op.cx, op.cy, op.cz, s);
        prod = (double)s[e] * (double)qv;
    }
    q[e] = qv;
    const double bs = block_sum<NT>(prod, lds);
    if (threadIdx.x == 0) red.pa[blockIdx.x] = bs;
    if (last_block_arrive(&red.st->ticket[kTicketAlpha])) finish_alpha(red, lds);
}

__global__ __launch_bounds__(NT) void k_spmv_dot_csr(CsrOp op, const double *__restrict__ s, double *__restrict__ q,
                                                     Reduction red) {
    __shared__ double lds[NT / kWave];
    if (red.st->done) return;
    const int64_t e = flat_elem(op.rows);
    double prod = 0.0;
    if (e < op.rows) {
        double acc = 0.0;
        for (int64_t p = op.rowptr[e]; p < op.rowptr[e + 1]; ++p) acc += op.vals[p] * s[op.cols[p]];
        q[e] = acc;
        prod = s[e] * acc;
    }
    const double bs = block_sum<NT>(prod, lds);
    if (threadIdx.x == 0) red.pa[blockIdx.x] = bs;
    if (last_block_arrive(&red.st->ticket[kTicketAlpha])) finish_alpha(red, lds);
}

// ---------------------------------------------------------------- axpy + norm

template <class T, class C, bool JAC>
__global__ __launch_bounds__(NT) void k_axpy_norm(int64_t n, const C *__restrict__ diag, T *__restrict__ x,
                                                  T *__restrict__ r, const T *__restrict__ s, const T *__restrict__ q,
                                                  Reduction red) {
    __shared__ double lds[NT / kWave];
    if (red.st->done) return;
    const T a = (T)red.st->alpha;
    const int64_t e = flat_elem(n);
    double rz = 0.0, rm = 0.0;
    if (e < n) {
        const C d = diag[e];
        if (d != C(0)) {
            x[e] += a * s[e];
            const T rv = r[e] - a * q[e];
            r[e] = rv;
            const T z = JAC ? rv / (T)d : rv;
            rz = (double)rv * (double)z;
            rm = fabs((double)rv);
        }
    }
    const double bs = block_sum<NT>(rz, lds);
    const double bm = block_max<NT>(rm, lds);
    if (threadIdx.x == 0) {
        red.pa[blockIdx.x] = bs;
        red.pb[blockIdx.x] = bm;
    }
    if (last_block_arrive(&red.st->ticket[kTicketAxpy])) {
        const double rho_new = reduce_partials_sum<NT>(red.pa, gridDim.x, lds);
        const double rn = reduce_partials_max<NT>(red.pb, gridDim.x, lds);
        if (threadIdx.x == 0) {
            SolveState *st = red.st;
            const int64_t it = st->iter + 1;
            st->iter = it;
            st->rnorm = rn;
            if (rn <= st->tol_abs) {
                st->done = 1;
                st->status = SHK_OK;
            } else if (rho_new != rho_new) {
                st->done = 1;
                st->status = SHK_NAN;
            } else if (it >= st->max_iter) {
                st->done = 1;
                st->status = SHK_NOT_CONVERGED;
            } else {
                st->beta = rho_new / st->rho;
                st->rho = rho_new;
            }
        }
    }
}

template <class T, class C, bool JAC>
__global__ __launch_bounds__(NT) void k_update_dir(int64_t n, const C *__restrict__ diag, const T *__restrict__ r,
                                                   T *__restrict__ s, const SolveState *__restrict__ st) {
    if (st->done) return;
    const T b = (T)st->beta;
    const int64_t e = flat_elem(n);
    if (e < n) {
        const C d = diag[e];
        if (d != C(0)) {
            const T rv = r[e];
            const T z = JAC ? rv / (T)d : rv;
            s[e] = z + b * s[e];
        }
    }
}

// ---------------------------------------------------------------- true residual

__device__ inline void finish_residual(const Reduction &red, double *lds, double *out) {
    const double rmax = reduce_partials_max<NT>(red.pa, gridDim.x, lds);
    const double r2 = reduce_partials_sum<NT>(red.pb, gridDim.x, lds);
    if (threadIdx.x == 0) {
        out[0] = rmax;
        out[1] = r2;
    }
}

__global__ __launch_bounds__(NT) void k_residual(BrickOp<double> op, const double *__restrict__ b,
                                                 const double *__restrict__ x, double *__restrict__ r, Reduction red,
                                                 double *__restrict__ b_partial, double *__restrict__ out) {
    __shared__ double lds[NT / kWave];
    const int slot = xcd_slot((int)blockIdx.x, (int)gridDim.x);
    const int l = threadIdx.x;
    const int64_t e = (int64_t)slot * NT + l;
    const double d = op.diag[e];
    double rv = 0.0, bv = 0.0;
    if (d != 0.0) {
        bv = b[e];
        rv = bv - brick_apply<double, double>(op.nbr + 6 * slot, e, l, d, op.cx, op.cy, op.cz, x);
    }
    if (r) r[e] = rv;
    const double m = block_max<NT>(fabs(rv), lds);
    const double s2 = block_sum<NT>(rv * rv, lds);
    const double b2 = block_sum<NT>(bv * bv, lds);
    const double bm = block_max<NT>(fabs(bv), lds);
    if (threadIdx.x == 0) {
        red.pa[blockIdx.x] = m;
        red.pb[blockIdx.x] = s2;
        b_partial[blockIdx.x] = b2;
        b_partial[gridDim.x + blockIdx.x] = bm;
    }
    if (last_block_arrive(&red.st->ticket[kTicketMisc])) {
        finish_residual(red, lds, out);
        const double bb = reduce_partials_sum<NT>(b_partial, gridDim.x, lds);
        const double bx = reduce_partials_max<NT>(b_partial + gridDim.x, gridDim.x, lds);
        if (threadIdx.x == 0) {
            out[2] = bb;
            out[3] = bx;
        }
    }
}

__global__ __launch_bounds__(NT) void k_residual_csr(CsrOp op, const double *__restrict__ b,
                                                     const double *__restrict__ x, double *__restrict__ r,
                                                     Reduction red, double *__restrict__ b_partial,
                                                     double *__restrict__ out) {
    __shared__ double lds[NT / kWave];
    const int64_t e = flat_elem(op.rows);
    double rv = 0.0, bv = 0.0;
    if (e < op.rows) {
        double acc = 0.0;
        for (int64_t p = op.rowptr[e]; p < op.rowptr[e + 1]; ++p) acc += op.vals[p] * x[op.cols[p]];
        bv = b[e];
        rv = bv - acc;
        if (r) r[e] = rv;
    }
    const double m = block_max<NT>(fabs(rv), lds);
    const double s2 = block_sum<NT>(rv * rv, lds);
    const double b2 = block_sum<NT>(bv * bv, lds);
    const double bm = block_max<NT>(fabs(bv), lds);
    if (threadIdx.x == 0) {
        red.pa[blockIdx.x] = m;
        red.pb[blockIdx.x] = s2;
        b_partial[blockIdx.x] = b2;
        b_partial[gridDim.x + blockIdx.x] = bm;
    }
    if (last_block_arrive(&red.st->ticket[kTicketMisc])) {
        finish_residual(red, lds, out);
        const double bb = reduce_partials_sum<NT>(b_partial, gridDim.x, lds);
        const double bx = reduce_partials_max<NT>(b_partial + gridDim.x, gridDim.x, lds);
        if (threadIdx.x == 0) {
            out[2] = bb;
            out[3] = bx;
        }
    }
}

// ---------------------------------------------------------------- misc

__global__ __launch_bounds__(NT) void k_add_correction(int64_t n, double *__restrict__ x, const float *__restrict__ d) {
    const int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x;
    if (e < n) x[e] += (double)d[e];
}

__global__ __launch_bounds__(NT) void k_csr_diag(CsrOp op, double *__restrict__ diag) {
    const int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x;
    if (e < op.rows) {
        double dv = 0.0;
        for (int64_t p = op.rowptr[e]; p < op.rowptr[e + 1]; ++p)
            if (op.cols[p] == e) dv = op.vals[p];
        diag[e] = dv;
    }
}

}  // namespace

// ---------------------------------------------------------------- launchers

template <class T, class C, class TB>
void launch_cg_init(hipStream_t st, int64_t n, const TB *rhs, const C *diag, T *x, T *r, T *s, Reduction red,
                    double tol_rel, int64_t max_iter, bool jacobi) {
    const int g = blocks_for(n);
    if (jacobi)
        hipLaunchKernelGGL((k_cg_init<T, C, TB, true>), dim3(g), dim3(NT), 0, st, n, rhs, diag, x, r, s, red, tol_rel,
                           max_iter);
    else
        hipLaunchKernelGGL((k_cg_init<T, C, TB, false>), dim3(g), dim3(NT), 0, st, n, rhs, diag, x, r, s, red, tol_rel,
                           max_iter);
}

template <class T, class C>
void launch_spmv_dot(hipStream_t st, const BrickOp<C> &op, const T *s, T *q, Reduction red) {
    hipLaunchKernelGGL((k_spmv_dot<T, C>), dim3(op.n_active), dim3(NT), 0, st, op, s, q, red);
}

void launch_spmv_dot_csr(hipStream_t st, const CsrOp &op, const double *s, double *q, Reduction red) {
    hipLaunchKernelGGL(k_spmv_dot_csr, dim3(blocks_for(op.rows)), dim3(NT), 0, st, op, s, q, red);
}

template <class T, class C>
void launch_axpy_norm(hipStream_t st, int64_t n, const C *diag, T *x, T *r, const T *s, const T *q, Reduction red,
                      bool jacobi) {
    const int g = blocks_for(n);
    if (jacobi)
        hipLaunchKernelGGL((k_axpy_norm<T, C, true>), dim3(g), dim3(NT), 0, st, n, diag, x, r, s, q, red);
    else
        hipLaunchKernelGGL((k_axpy_norm<T, C, false>), dim3(g), dim3(NT), 0, st, n, diag, x, r, s, q, red);
}

template <class T, class C>
void launch_update_dir(hipStream_t st, int64_t n, const C *diag, const T *r, T *s, Reduction red, bool jacobi) {
    const int g = blocks_for(n);
    if (jacobi)
        hipLaunchKernelGGL((k_update_dir<T, C, true>), dim3(g), dim3(NT), 0, st, n, diag, r, s, red.st);
    else
        hipLaunchKernelGGL((k_update_dir<T, C, false>), dim3(g), dim3(NT), 0, st, n, diag, r, s, red.st);
}

void launch_residual(hipStream_t st, const BrickOp<double> &op, const double *b, const double *x, double *r,
                     Reduction red, double *out) {
    const int g = op.n_active;
    // pb doubles as the r^2 partials; the b partials follow the a/b partial arrays
    hipLaunchKernelGGL(k_residual, dim3(g), dim3(NT), 0, st, op, b, x, r, red, red.pb + g, out);
}

void launch_residual_csr(hipStream_t st, const CsrOp &op, const double *b, const double *x, double *r, Reduction red,
                         double *out) {
    const int g = blocks_for(op.rows);
    hipLaunchKernelGGL(k_residual_csr, dim3(g), dim3(NT), 0, st, op, b, x, r, red, red.pb + g, out);
}

void launch_add_correction(hipStream_t st, int64_t n, double *x, const float *d) {
    hipLaunchKernelGGL(k_add_correction, dim3(blocks_for(n)), dim3(NT), 0, st, n, x, d);
}

void launch_csr_diag(hipStream_t st, const CsrOp &op, double *diag) {
    hipLaunchKernelGGL(k_csr_diag, dim3(blocks_for(op.rows)), dim3(NT), 0, st, op, diag);
}

template void launch_cg_init<double, double, double>(hipStream_t, int64_t, const double *, const double *, double *,
                                                     double *, double *, Reduction, double, int64_t, bool);
template void launch_cg_init<float, float, double>(hipStream_t, int64_t, const double *, const float *, float *,
                                                   float *, float *, Reduction, double, int64_t, bool);
template void launch_spmv_dot<double, double>(hipStream_t, const BrickOp<double> &, const double *, double *,
                                              Reduction);
template void launch_spmv_dot<float, float>(hipStream_t, const BrickOp<float> &, const float *, float *, Reduction);
template void launch_axpy_norm<double, double>(hipStream_t, int64_t, const double *, double *, double *,
                                               const double *, const double *, Reduction, bool);
template void launch_axpy_norm<float, float>(hipStream_t, int64_t, const float *, float *, float *, const float *,
                                             const float *, Reduction, bool);
template void launch_update_dir<double, double>(hipStream_t, int64_t, const double *, const double *, double *,
                                                Reduction, bool);
template void launch_update_dir<float, float>(hipStream_t, int64_t, const float *, const float *, float *, Reduction,
                                              bool);

}  // namespace shk
