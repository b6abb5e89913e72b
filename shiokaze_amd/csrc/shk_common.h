// shk_common.h -- device helpers shared by the gfx950 kernels: 64-lane wave
// reductions, block reductions, the deterministic last-block grid reduction and
// the Krylov solve state that lives in device memory.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace shk {

constexpr int kBrick = 8;                 // brick edge of 3D grids (cells)
constexpr int kBrickCells = 512;          // cells per brick, one workgroup per brick
constexpr int kWave = 64;                 // CDNA wavefront width
constexpr int kXcds = 8;                  // MI355X accelerator complex dies

// Brick shapes: 8x8x8 for 3D grids, 64x1x8 for 2D grids (stored as nx x 1 x ny, the
// 2D y axis on z).  Either way a brick holds 512 cells x-fastest and one wave covers
// one z layer of 64 cells, so the stencil kernels keep a z column per lane.
template <int BX>
struct BrickShape {
    static constexpr int X = BX, Y = kWave / BX, Z = kBrickCells / kWave;
    static_assert(X * Y == kWave && Z == 8, "a brick layer is one wave");
    __host__ __device__ static int elem(int i, int j, int k) { return (i % X) + X * ((j % Y) + Y * (k % Z)); }
};

// Krylov state kept on the device so that iterations are launched back to back
// without a host round trip; the host polls `done` every few iterations.
// Each field has one writer kernel (its block 0) and is read only by the other
// kernel of the iteration, so no field is read and written in the same launch.
struct SolveState {
    double rho_k1;    // r.r folded by the direction kernel (read by axpy for alpha)
    double rho_k2;    // the rho the last axpy used (read by the direction kernel for beta)
    double alpha;     // step of the last iteration (axpy)
    double beta;      // direction update of the current iteration (direction kernel)
    double rnorm;     // max |r| after the last update (direction kernel)
    double rnorm0;    // max |rhs| at the start
    double tol_abs;   // stop when rnorm <= tol_abs
    double sigma;     // s.As of the last iteration (axpy)
    int64_t iter;     // completed iterations (axpy)
    int64_t max_iter;
    int32_t done;     // 1 once converged, capped or broken down
    int32_t status;   // shk_status
};

// Device-side control of one inner flexible-CG solve (the learned preconditioner's):
// the stop rules run on the device between an update and the next application of the
// network, the following kernels read the result (a set stop makes them return at
// once) and the host reads it an iteration later.
struct FcgState {
    double scale;     // rms of r (1 if not positive): the network's input scale
    double rn;        // max|r| after the last update
    double target;    // converged once rn <= target
    double best;      // last rn that improved on the one before by 10%
    int64_t it;       // completed iterations
    int64_t best_it;  // the iteration of best
    int64_t cap;      // stop after this many iterations
    int32_t stop;     // 0 running, 1 converged or capped, 2 stalled, broke down or not finite
    int32_t pad;
};

// The stop rules after an update (r2 = r.r, rn = max|r|, bad: p.q <= 0 or a non-finite
// step, so the update was skipped; nunk: unknowns, for the rms).
__host__ __device__ inline void fcg_decide(FcgState &v, double r2, double rn, bool bad, double nunk) {
    if (bad) {  // the preconditioner broke the iteration
        v.stop = 2;
        return;
    }
    v.it += 1;
    if (!(rn == rn)) {
        v.stop = 2;
    } else if (rn < 0.9 * v.best) {
        v.best = rn;
        v.best_it = v.it;
    } else if (v.it - v.best_it >= 20) {  // less than 10% progress in 20 iterations
        v.stop = 2;
    }
    if (v.stop == 0 && rn <= v.target) v.stop = 1;
    if (v.stop == 0 && v.it >= v.cap) v.stop = 1;
    v.rn = rn;
    const double rms = sqrt(r2 / nunk);
    v.scale = rms > 0.0 ? rms : 1.0;
}

// What the first kernel of a network application needs to apply fcg_decide: the
// axpy's 2 x nparts partials (r.r, max|r|), its flags fst (fst[2] != 0: bad), the
// state before (in) and where the state after goes (out, written by block 0, and
// host: pinned host memory the host reads once the iteration's event has completed;
// nullable).
struct FcgCtl {
    const double *part2;
    const double *fst;
    int nparts;
    double nunk;
    const FcgState *in;
    FcgState *out;
    FcgState *host;
};

// Workgroup -> slot mapping that hands each XCD one contiguous range of slots.
// Workgroups are dispatched to the 8 XCDs round-robin by id, and each XCD has its
// own L2, so contiguous slots (spatial neighbours share halo cells) should land
// on the same XCD.  Bijective for any n.
__device__ inline int xcd_slot(int b, int n) {
    const int per = n / kXcds, rem = n % kXcds;
    const int xcd = b % kXcds, idx = b / kXcds;
    return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

// Sum / max over a workgroup of NT threads; the result is valid in thread 0.
// Fixed butterfly + fixed wave order: deterministic.
template <int NT>
__device__ inline double block_sum(double v, double *lds) {
    v = wave_sum(v);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (wave == 0) {
        r = lane < NT / kWave ? lds[lane] : 0.0;
        r = wave_sum(r);
    }
    __syncthreads();
    return r;
}
template <int NT>
__device__ inline double block_max(double v, double *lds) {
    v = wave_max(v);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (wave == 0) {
        r = lane < NT / kWave ? lds[lane] : 0.0;
        r = wave_max(r);
    }
    __syncthreads();
    return r;
}

// Fixed-order reduction of n partials by one workgroup of NT threads (thread t
// folds t, t+NT, ... in order, then the block tree).  Result valid in thread 0.
template <int NT>
__device__ inline double reduce_partials_sum(const double *p, int n, double *lds) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) acc += p[i];
    return block_sum<NT>(acc, lds);
}
template <int NT>
__device__ inline double reduce_partials_max(const double *p, int n, double *lds) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) acc = fmax(acc, p[i]);
    return block_max<NT>(acc, lds);
}

}  // namespace shk
