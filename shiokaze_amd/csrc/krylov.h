// krylov.h -- host launchers of the conjugate gradient kernels (krylov.hip).
//
// The iteration is PCGSolver::solve (local/include/pcgsolver/pcg_solver.h:246-295):
// x0 = 0, r = b, z = M^-1 r, s = z, then per iteration
//   q = A s; alpha = rho / s.q; x += alpha s; r -= alpha q;
//   stop if max|r| <= tol * max|b|; z = M^-1 r; rho' = z.r; s = z + (rho'/rho) s.
// M is the identity (the reference's effective preconditioner, pcg_solver.h:379-384)
// or diag(A) (Jacobi).  Each of the three fused kernels publishes one partial per
// workgroup; the last workgroup to finish folds the partials in a fixed order and
// updates the SolveState, so results are deterministic and no host round trip is
// needed between iterations.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "shk_common.h"

namespace shk {

struct Reduction {
    double *pa;       // per-workgroup partials (>= blocks)
    double *pb;
    SolveState *st;   // device state
};

// Operator views.  Brick: vectors hold n_active bricks of 8x8x8 cells, x-fastest
// inside a brick; nbr[6*slot + {-x,+x,-y,+y,-z,+z}] is the neighbour brick's slot
// or -1.  The matrix is diag - sum_faces c_face * neighbour, with c_face the
// coupling of the face stored at its high cell (cx/cy/cz = low face of a cell).
template <class C>
struct BrickOp {
    const int *nbr;
    const C *diag, *cx, *cy, *cz;
    int n_active;
};
struct CsrOp {
    const int64_t *rowptr, *cols;
    const double *vals;
    const double *diag;  // extracted diagonal (Jacobi, masking)
    int64_t rows;
};

inline int blocks_for(int64_t n) { return (int)((n + kBrickCells - 1) / kBrickCells); }

// x = 0, r = rhs, s = M^-1 r; rho = r.s, rnorm0 = max|r|, tol_abs = tol_rel*rnorm0.
template <class T, class C, class TB>
void launch_cg_init(hipStream_t st, int64_t n, const TB *rhs, const C *diag, T *x, T *r, T *s, Reduction red,
                    double tol_rel, int64_t max_iter, bool jacobi);

// One CG iteration is three kernels: q = A s with s.q (spmv_dot), the x/r update
// with r.z and max|r| (axpy_norm) and s = z + beta s (update_dir).
template <class T, class C>
void launch_spmv_dot(hipStream_t st, const BrickOp<C> &op, const T *s, T *q, Reduction red);
void launch_spmv_dot_csr(hipStream_t st, const CsrOp &op, const double *s, double *q, Reduction red);
template <class T, class C>
void launch_axpy_norm(hipStream_t st, int64_t n, const C *diag, T *x, T *r, const T *s, const T *q, Reduction red,
                      bool jacobi);
template <class T, class C>
void launch_update_dir(hipStream_t st, int64_t n, const C *diag, const T *r, T *s, Reduction red, bool jacobi);

// r = b - A x in FP64; out[0] = max|r|, out[1] = sum r^2, out[2] = sum b^2, out[3] = max|b|.
void launch_residual(hipStream_t st, const BrickOp<double> &op, const double *b, const double *x, double *r,
                     Reduction red, double *out);
void launch_residual_csr(hipStream_t st, const CsrOp &op, const double *b, const double *x, double *r, Reduction red,
                         double *out);

// x64 += (double) d32 (refinement correction); optionally r32 = (float) r64.
void launch_add_correction(hipStream_t st, int64_t n, double *x, const float *d);

// diag[row] = A[row,row] of a CSR matrix.
void launch_csr_diag(hipStream_t st, const CsrOp &op, double *diag);

}  // namespace shk
