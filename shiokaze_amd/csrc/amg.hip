// amg.hip -- aggregation multigrid for shk_solve_csr with Preconditioner=mg: CG on an
// arbitrary symmetric CSR matrix, preconditioned by a V-cycle over a hierarchy built
// from the matrix alone.  The counterpart of the reference's "amg" LinSolver
// (src/linsolver/amg.cpp:43-166), which builds amgcl's smoothed-aggregation hierarchy
// on the host for every solve and runs CG with it.
//
// Setup (host, O(nnz) per level, like the reference's, which runs it on every solve; here
// a solve whose matrix has the content of the previous one, by a 128-bit hash computed
// on the device, reuses the hierarchy):
//   - two greedy pairwise aggregation passes per level: rows in index order, an
//     unaggregated row joins its strongest unaggregated strong neighbour
//     (-a_ij >= 0.25 max_k -a_ik, ties to the lower index), else its strongest
//     unaggregated neighbour of any negative coupling, or stays alone; the second pass
//     runs on the Galerkin matrix of the first, so aggregates hold <= 4 rows;
//   - a level whose aggregation keeps more than 80% of its rows is the coarsest (the
//     over-correction alpha would compound over levels that barely coarsen);
//   - the coarse matrix is the Galerkin product P'AP of the piecewise-constant
//     prolongation (exact zeros dropped), until at most kAmgCoarseMax rows remain,
//     whose matrix is inverted densely (FP64 Cholesky).
// V-cycle (device, FP64, 2 kernels per level): the pre-smoothing step from zero
// x = w D^-1 b, its residual and the restriction (sum over the aggregate) in one
// kernel per coarse row; the prolongation with over-correction alpha and the
// post-smoothing step in one kernel per row; the dense coarsest solve in one
// workgroup.  Pre- and post-smoothing are the same symmetric operator, so M is fixed
// and SPD and plain CG applies.  Every sum runs in one fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ctx_internal.h"
#include "krylov.h"

namespace shk {

constexpr int64_t kAmgCoarseMax = 128;
constexpr int kAmgNB = 256;  // threads of the level-0 kernels on the CG grid (NB elsewhere)
constexpr double kAmgStrong = 0.25;
constexpr double kAmgStall = 0.8;
constexpr int kAmgMaxLevels = 40;

struct AmgLevel {
    int64_t n = 0, nnz = 0, nc = 0;  // rows and entries of this level, rows of the next
    // device CSR of this level (level 0: the caller's arrays, int64 columns)
    const int64_t *rowptr = nullptr;
    const int64_t *cols64 = nullptr;
    const int32_t *cols32 = nullptr;
    const double *vals = nullptr;
    DevBuf<int64_t> d_ptr;
    DevBuf<int32_t> d_col;
    DevBuf<double> d_val;
    DevBuf<double> dinv;
    DevBuf<int32_t> agg;            // row -> row of the next level
    DevBuf<int64_t> mptr;           // next level's row -> its rows mem[mptr[I] .. mptr[I+1])
    DevBuf<int32_t> mem;
    DevBuf<double> b, x;            // levels >= 1: the V-cycle's rhs and result
};

struct CsrAmg {
    std::vector<AmgLevel> lv;
    int64_t ncoarse = 0;
    DevBuf<double> cinv;            // dense inverse of the coarsest matrix, row-major
    DevBuf<double> z, pz;           // level-0 result and the r.z partials for CG
    // content of the matrix the hierarchy was built from: rows, entries and two
    // order-independent 64-bit hashes (XOR and sum) of every (row, column, value) entry
    int64_t key_rows = -1, key_nnz = -1;
    unsigned long long key_x = 0, key_s = 0;
    DevBuf<unsigned long long> hash;
};

const double *amg_z(const CsrAmg *h) { return h->z.p; }
double *amg_pz(CsrAmg *h) { return h->pz.p; }

void amg_destroy(CsrAmg *h) {
    if (!h) return;
    for (AmgLevel &L : h->lv) {
        L.d_ptr.release();
        L.d_col.release();
        L.d_val.release();
        L.dinv.release();
        L.agg.release();
        L.mptr.release();
        L.mem.release();
        L.b.release();
        L.x.release();
    }
    h->cinv.release();
    h->z.release();
    h->pz.release();
    h->hash.release();
    delete h;
}

namespace {

struct HostCsr {
    int64_t n = 0;
    std::vector<int64_t> ptr;
    std::vector<int64_t> col;
    std::vector<double> val;
};

// greedy pairwise aggregation; returns the number of aggregates
int64_t pairwise(const HostCsr &A, std::vector<int64_t> &agg) {
    agg.assign(A.n, -1);
    int64_t nc = 0;
    for (int64_t i = 0; i < A.n; ++i) {
        if (agg[i] >= 0) continue;
        double smax = 0.0;
        for (int64_t p = A.ptr[i]; p < A.ptr[i + 1]; ++p)
            if (A.col[p] != i && A.val[p] < 0.0) smax = std::max(smax, -A.val[p]);
        double best = 0.0, weak = 0.0;
        int64_t bj = -1, wj = -1;
        for (int64_t p = A.ptr[i]; p < A.ptr[i + 1]; ++p) {
            const int64_t j = A.col[p];
            const double v = A.val[p];
            if (j == i || !(v < 0.0) || agg[j] >= 0) continue;
            const double s = -v;
            if (s >= kAmgStrong * smax && (s > best || (s == best && j < bj))) {
                best = s;
                bj = j;
            }
            if (s > weak || (s == weak && j < wj)) {
                weak = s;
                wj = j;
            }
        }
        if (bj < 0) bj = wj;
        agg[i] = nc;
        if (bj >= 0) agg[bj] = nc;
        ++nc;
    }
    return nc;
}

// members of each aggregate in ascending row order (counting sort)
void members(const std::vector<int64_t> &agg, int64_t nc, std::vector<int64_t> &mptr, std::vector<int64_t> &mem) {
    mptr.assign(nc + 1, 0);
    for (int64_t a : agg) ++mptr[a + 1];
    for (int64_t I = 0; I < nc; ++I) mptr[I + 1] += mptr[I];
    mem.resize(agg.size());
    std::vector<int64_t> fill(mptr.begin(), mptr.end() - 1);
    for (int64_t i = 0; i < (int64_t)agg.size(); ++i) mem[fill[agg[i]]++] = i;
}

// Galerkin product P'AP of piecewise-constant P (columns sorted, exact zeros dropped)
HostCsr galerkin(const HostCsr &A, const std::vector<int64_t> &agg, int64_t nc) {
    std::vector<int64_t> mptr, mem;
    members(agg, nc, mptr, mem);
    HostCsr C;
    C.n = nc;
    C.ptr.assign(nc + 1, 0);
    std::vector<int64_t> pos(nc, -1);
    std::vector<int64_t> cols;
    std::vector<double> acc;
    for (int64_t I = 0; I < nc; ++I) {
        cols.clear();
        acc.clear();
        for (int64_t m = mptr[I]; m < mptr[I + 1]; ++m) {
            const int64_t i = mem[m];
            for (int64_t p = A.ptr[i]; p < A.ptr[i + 1]; ++p) {
                const int64_t J = agg[A.col[p]];
                if (pos[J] < 0) {
                    pos[J] = (int64_t)cols.size();
                    cols.push_back(J);
                    acc.push_back(0.0);
                }
                acc[pos[J]] += A.val[p];
            }
        }
        std::vector<int64_t> order(cols.size());
        for (size_t k = 0; k < order.size(); ++k) order[k] = (int64_t)k;
        std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return cols[a] < cols[b]; });
        for (int64_t k : order) {
            pos[cols[k]] = -1;
            if (acc[k] != 0.0) {
                C.col.push_back(cols[k]);
                C.val.push_back(acc[k]);
            }
        }
        C.ptr[I + 1] = (int64_t)C.col.size();
    }
    return C;
}

// inverse of a dense SPD matrix (row-major, n x n) through Cholesky; false if not SPD
bool spd_inverse(int64_t n, std::vector<double> &a) {
    std::vector<double> L(a);
    for (int64_t j = 0; j < n; ++j) {
        double d = L[j * n + j];
        for (int64_t k = 0; k < j; ++k) d -= L[j * n + k] * L[j * n + k];
        if (!(d > 0.0)) return false;
        d = std::sqrt(d);
        L[j * n + j] = d;
        for (int64_t i = j + 1; i < n; ++i) {
            double s = L[i * n + j];
            for (int64_t k = 0; k < j; ++k) s -= L[i * n + k] * L[j * n + k];
            L[i * n + j] = s / d;
        }
    }
    std::vector<double> col(n);
    for (int64_t c = 0; c < n; ++c) {
        for (int64_t i = 0; i < n; ++i) {  // L y = e_c
            double s = i == c ? 1.0 : 0.0;
            for (int64_t k = 0; k < i; ++k) s -= L[i * n + k] * col[k];
            col[i] = s / L[i * n + i];
        }
        for (int64_t i = n - 1; i >= 0; --i) {  // L' x = y
            double s = col[i];
            for (int64_t k = i + 1; k < n; ++k) s -= L[k * n + i] * col[k];
            col[i] = s / L[i * n + i];
        }
        for (int64_t i = 0; i < n; ++i) a[i * n + c] = col[i];
    }
    return true;
}

template <class T>
hipError_t upload(DevBuf<T> &dst, const std::vector<T> &src, hipStream_t st) {
    hipError_t e = dst.ensure(std::max<size_t>(src.size(), 1));
    if (e != hipSuccess || src.empty()) return e;
    return hipMemcpyAsync(dst.p, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, st);
}

std::vector<double> inv_diag(const HostCsr &A) {
    std::vector<double> d(A.n, 0.0);
    for (int64_t i = 0; i < A.n; ++i)
        for (int64_t p = A.ptr[i]; p < A.ptr[i + 1]; ++p)
            if (A.col[p] == i) d[i] = A.val[p];
    for (double &v : d) v = v != 0.0 ? 1.0 / v : 0.0;
    return d;
}

// ---------------------------------------------------------------- V-cycle kernels

template <class IDX>
__device__ inline double row_apply_smoothed(const int64_t *__restrict__ ptr, const IDX *__restrict__ col,
                                            const double *__restrict__ val, const double *__restrict__ dinv,
                                            const double *__restrict__ b, double w, int64_t i) {
    // sum_j a_ij (w dinv_j b_j): A applied to the pre-smoothing step from zero
    double a = 0.0;
    for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) {
        const int64_t j = (int64_t)col[p];
        a += val[p] * (w * dinv[j] * b[j]);
    }
    return a;
}

// bc[I] = sum over the rows i of aggregate I of t_i = b_i - (A x)_i, x = w D^-1 b
template <class IDX>
__global__ __launch_bounds__(256) void k_amg_down(int64_t nc, const int64_t *__restrict__ mptr,
                                                  const int32_t *__restrict__ mem, const int64_t *__restrict__ ptr,
                                                  const IDX *__restrict__ col, const double *__restrict__ val,
                                                  const double *__restrict__ dinv, const double *__restrict__ b,
                                                  double w, double *__restrict__ bc, const SolveState *state) {
    if (state && state->done) return;
    const int64_t I = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (I >= nc) return;
    double s = 0.0;
    for (int64_t m = mptr[I]; m < mptr[I + 1]; ++m) {
        const int64_t i = mem[m];
        s += b[i] - row_apply_smoothed(ptr, col, val, dinv, b, w, i);
    }
    bc[I] = s;
}

// the corrected step x'_j = w dinv_j b_j + alpha xc[agg_j] and the post-smoothing
// y_i = x'_i + w dinv_i (b_i - sum_j a_ij x'_j)
template <class IDX>
__device__ inline double amg_up_row(const int64_t *__restrict__ ptr, const IDX *__restrict__ col,
                                    const double *__restrict__ val, const double *__restrict__ dinv,
                                    const int32_t *__restrict__ agg, const double *__restrict__ b,
                                    const double *__restrict__ xc, double w, double alpha, int64_t i) {
    double a = 0.0;
    for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) {
        const int64_t j = (int64_t)col[p];
        a += val[p] * (w * dinv[j] * b[j] + alpha * xc[agg[j]]);
    }
    const double xi = w * dinv[i] * b[i] + alpha * xc[agg[i]];
    return xi + w * dinv[i] * (b[i] - a);
}

template <class IDX>
__global__ __launch_bounds__(256) void k_amg_up(int64_t n, const int64_t *__restrict__ ptr,
                                                const IDX *__restrict__ col, const double *__restrict__ val,
                                                const double *__restrict__ dinv, const int32_t *__restrict__ agg,
                                                const double *__restrict__ b, const double *__restrict__ xc,
                                                double w, double alpha, double *__restrict__ y,
                                                const SolveState *state) {
    if (state && state->done) return;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = amg_up_row(ptr, col, val, dinv, agg, b, xc, w, alpha, i);
}

// level 0 on the CG grid: z = the post-smoothed result, part[block] = r.z over the
// block's rows (the partials the next direction kernel folds for beta)
__global__ __launch_bounds__(kAmgNB) void k_amg_up0(int64_t n, int per, const int64_t *__restrict__ ptr,
                                                const int64_t *__restrict__ col, const double *__restrict__ val,
                                                const double *__restrict__ dinv, const int32_t *__restrict__ agg,
                                                const double *__restrict__ r, const double *__restrict__ xc,
                                                double w, double alpha, double *__restrict__ z,
                                                double *__restrict__ part, const SolveState *state) {
    __shared__ double lds[kAmgNB / kWave];
    if (state && state->done) return;
    // the CG grid's shape (partials = blocks), tiles of kBrickCells rows in block order
    const int tiles = blocks_for(n);
    const int t0 = (int)blockIdx.x * per, t1 = t0 + per < tiles ? t0 + per : tiles;
    const int64_t end = (int64_t)t1 * kBrickCells < n ? (int64_t)t1 * kBrickCells : n;
    double acc = 0.0;
    for (int64_t e = (int64_t)t0 * kBrickCells + threadIdx.x; e < end; e += kAmgNB) {
        const double ze = amg_up_row(ptr, col, val, dinv, agg, r, xc, w, alpha, e);
        z[e] = ze;
        acc += ze * r[e];
    }
    const double s = block_sum<kAmgNB>(acc, lds);
    if (threadIdx.x == 0 && part) part[blockIdx.x] = s;
}

// part[block] = r.z over the block's rows of the CG grid (a one-level hierarchy)
__global__ __launch_bounds__(kAmgNB) void k_amg_partials(int64_t n, int per, const double *__restrict__ r,
                                                     const double *__restrict__ z, double *__restrict__ part,
                                                     const SolveState *state) {
    __shared__ double lds[kAmgNB / kWave];
    if (state && state->done) return;
    // the CG grid's shape (partials = blocks), tiles of kBrickCells rows in block order
    const int tiles = blocks_for(n);
    const int t0 = (int)blockIdx.x * per, t1 = t0 + per < tiles ? t0 + per : tiles;
    const int64_t end = (int64_t)t1 * kBrickCells < n ? (int64_t)t1 * kBrickCells : n;
    double acc = 0.0;
    for (int64_t e = (int64_t)t0 * kBrickCells + threadIdx.x; e < end; e += kAmgNB) acc += z[e] * r[e];
    const double s = block_sum<kAmgNB>(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// x = Cinv b on the coarsest level (one workgroup)
__global__ __launch_bounds__(256) void k_amg_coarse(int64_t nc, const double *__restrict__ cinv,
                                                    const double *__restrict__ b, double *__restrict__ x,
                                                    const SolveState *state) {
    if (state && state->done) return;
    for (int64_t i = threadIdx.x; i < nc; i += blockDim.x) {
        double s = 0.0;
        for (int64_t j = 0; j < nc; ++j) s += cinv[i * nc + j] * b[j];
        x[i] = s;
    }
}

// splitmix64 finaliser
__device__ inline unsigned long long mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// out[0] ^= and out[1] += a hash of every entry (row, column, value bits): both
// combinations are independent of the order the entries are visited in
__global__ __launch_bounds__(256) void k_amg_hash(int64_t rows, const int64_t *__restrict__ rowptr,
                                                  const int64_t *__restrict__ cols, const double *__restrict__ vals,
                                                  unsigned long long *out) {
    unsigned long long hx = 0, hs = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows; i += (int64_t)gridDim.x * blockDim.x) {
        for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            const unsigned long long e =
                mix64(mix64((unsigned long long)i * 0x9e3779b97f4a7c15ull ^ (unsigned long long)cols[p]) ^
                      (unsigned long long)__double_as_longlong(vals[p]));
            hx ^= e;
            hs += e;
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        hx ^= __shfl_xor(hx, o);
        hs += __shfl_xor(hs, o);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        atomicXor(&out[0], hx);
        atomicAdd(&out[1], hs);
    }
}

__global__ void k_amg_hash_clear(unsigned long long *out) {
    out[0] = 0ull;
    out[1] = 0ull;
}

inline unsigned blocks256(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

}  // namespace

// Builds the hierarchy of the CSR matrix (device arrays, int64 columns sorted per row).
int amg_setup(shk_ctx *c, int64_t rows, const int64_t *rowptr, const int64_t *cols, const double *vals) {
    hipStream_t st = c->stream;
    // the hierarchy of the previous solve serves again if the matrix's content is the same
    int64_t nnz0 = 0;
    HIPC(c, hipMemcpyAsync(&nnz0, rowptr + rows, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    if (!c->amg) c->amg = new CsrAmg();
    HIPC(c, c->amg->hash.ensure(2));
    hipLaunchKernelGGL(k_amg_hash_clear, dim3(1), dim3(1), 0, st, c->amg->hash.p);
    hipLaunchKernelGGL(k_amg_hash, dim3((unsigned)std::min<int64_t>(4096, (rows + 255) / 256 + 1)), dim3(256), 0, st,
                       rows, rowptr, cols, vals, c->amg->hash.p);
    HIPC(c, hipGetLastError());
    unsigned long long key[2] = {0, 0};
    HIPC(c, hipMemcpyAsync(key, c->amg->hash.p, sizeof key, hipMemcpyDeviceToHost, st));
    HIPC(c, hipStreamSynchronize(st));
    {
        CsrAmg *h = c->amg;
        if (!h->lv.empty() && h->key_rows == rows && h->key_nnz == nnz0 && h->key_x == key[0] && h->key_s == key[1]) {
            h->lv[0].rowptr = rowptr;  // the same content, maybe in other buffers
            h->lv[0].cols64 = cols;
            h->lv[0].vals = vals;
            return SHK_OK;
        }
    }
    HostCsr A;
    A.n = rows;
    A.ptr.resize(rows + 1);
    HIPC(c, hipMemcpyAsync(A.ptr.data(), rowptr, sizeof(int64_t) * (rows + 1), hipMemcpyDeviceToHost, st));
    HIPC(c, hipStreamSynchronize(st));
    const int64_t nnz = A.ptr[rows];
    if (nnz < 0 || A.ptr[0] != 0) return c->fail(SHK_BAD_ARG, "rowptr must start at 0 and end at nnz >= 0");
    A.col.resize(nnz);
    A.val.resize(nnz);
    HIPC(c, hipMemcpyAsync(A.col.data(), cols, sizeof(int64_t) * nnz, hipMemcpyDeviceToHost, st));
    HIPC(c, hipMemcpyAsync(A.val.data(), vals, sizeof(double) * nnz, hipMemcpyDeviceToHost, st));
    HIPC(c, hipStreamSynchronize(st));
    for (int64_t i = 0; i < rows; ++i)
        for (int64_t p = A.ptr[i]; p < A.ptr[i + 1]; ++p)
            if (A.col[p] < 0 || A.col[p] >= rows) return c->fail(SHK_BAD_ARG, "column index out of range");
    amg_destroy(c->amg);
    c->amg = new CsrAmg();
    CsrAmg *h = c->amg;
    HIPC(c, h->hash.ensure(2));
    HostCsr cur = A;
    for (int l = 0; l < kAmgMaxLevels; ++l) {
        h->lv.emplace_back();
        AmgLevel &L = h->lv.back();
        L.n = cur.n;
        L.nnz = (int64_t)cur.col.size();
        if (l == 0) {
            L.rowptr = rowptr;
            L.cols64 = cols;
            L.vals = vals;
        } else {
            std::vector<int32_t> c32(cur.col.begin(), cur.col.end());
            HIPC(c, upload(L.d_ptr, cur.ptr, st));
            HIPC(c, upload(L.d_col, c32, st));
            HIPC(c, upload(L.d_val, cur.val, st));
            HIPC(c, L.b.ensure((size_t)std::max<int64_t>(cur.n, 1)));
            HIPC(c, L.x.ensure((size_t)std::max<int64_t>(cur.n, 1)));
            L.rowptr = L.d_ptr.p;
            L.cols32 = L.d_col.p;
            L.vals = L.d_val.p;
        }
        HIPC(c, upload(L.dinv, inv_diag(cur), st));
        if (cur.n <= kAmgCoarseMax || l == kAmgMaxLevels - 1) break;
        std::vector<int64_t> a1, a2;
        const int64_t n1 = pairwise(cur, a1);
        const HostCsr A1 = galerkin(cur, a1, n1);
        const int64_t n2 = pairwise(A1, a2);
        std::vector<int64_t> agg(cur.n);
        for (int64_t i = 0; i < cur.n; ++i) agg[i] = a2[a1[i]];
        if ((double)n2 > kAmgStall * (double)cur.n) break;  // barely coarsens: the coarsest level
        std::vector<int64_t> mptr, mem;
        members(agg, n2, mptr, mem);
        L.nc = n2;
        HIPC(c, upload(L.agg, std::vector<int32_t>(agg.begin(), agg.end()), st));
        HIPC(c, upload(L.mptr, mptr, st));
        HIPC(c, upload(L.mem, std::vector<int32_t>(mem.begin(), mem.end()), st));
        // P'AP of the composed aggregation = the second pass's product of the first's
        cur = galerkin(A1, a2, n2);
    }
    // the coarsest level: dense inverse
    AmgLevel &C = h->lv.back();
    if (C.n > 1024)
        return c->fail(SHK_BAD_ARG, "multigrid: the matrix does not coarsen below %lld rows", (long long)C.n);
    h->ncoarse = C.n;
    std::vector<double> dense((size_t)C.n * C.n, 0.0);
    for (int64_t i = 0; i < cur.n; ++i)
        for (int64_t p = cur.ptr[i]; p < cur.ptr[i + 1]; ++p) dense[(size_t)i * C.n + cur.col[p]] += cur.val[p];
    if (!spd_inverse(C.n, dense))
        return c->fail(SHK_BREAKDOWN, "multigrid: the coarsest matrix is not positive definite");
    HIPC(c, upload(h->cinv, dense, st));
    h->key_rows = rows;
    h->key_nnz = nnz0;
    h->key_x = key[0];
    h->key_s = key[1];
    HIPC(c, h->z.ensure((size_t)std::max<int64_t>(rows, 1)));
    HIPC(c, h->pz.ensure((size_t)kMaxBlocks + 64));
    HIPC(c, hipStreamSynchronize(st));
    return SHK_OK;
}

// z = M r (one V-cycle); the last kernel writes the r.z partials for the CG grid g
int amg_vcycle(shk_ctx *c, const double *r, const Grid &g, const SolveState *state) {
    CsrAmg *h = c->amg;
    hipStream_t st = c->stream;
    const double w = c->prm.mg_omega, alpha = c->prm.mg_alpha;
    const int L = (int)h->lv.size() - 1;
    if (L == 0) {
        // a matrix of at most kAmgCoarseMax rows: z = A^-1 r
        hipLaunchKernelGGL(k_amg_coarse, dim3(1), dim3(256), 0, st, h->ncoarse, h->cinv.p, r, h->z.p, state);
        hipLaunchKernelGGL(k_amg_partials, dim3(g.blocks), dim3(kAmgNB), 0, st, g.n, g.per, r, h->z.p, h->pz.p, state);
        HIPC(c, hipGetLastError());
        return SHK_OK;
    }
    // down
    for (int l = 0; l < L; ++l) {
        AmgLevel &F = h->lv[l];
        AmgLevel &Cn = h->lv[l + 1];
        const double *b = l == 0 ? r : F.b.p;
        if (l == 0)
            hipLaunchKernelGGL(k_amg_down<int64_t>, dim3(blocks256(F.nc)), dim3(256), 0, st, F.nc, F.mptr.p,
                               F.mem.p, F.rowptr, F.cols64, F.vals, F.dinv.p, b, w, Cn.b.p, state);
        else
            hipLaunchKernelGGL(k_amg_down<int32_t>, dim3(blocks256(F.nc)), dim3(256), 0, st, F.nc, F.mptr.p,
                               F.mem.p, F.rowptr, F.cols32, F.vals, F.dinv.p, b, w, Cn.b.p, state);
    }
    AmgLevel &Cl = h->lv[L];
    hipLaunchKernelGGL(k_amg_coarse, dim3(1), dim3(256), 0, st, h->ncoarse, h->cinv.p, Cl.b.p, Cl.x.p, state);
    // up
    for (int l = L - 1; l >= 1; --l) {
        AmgLevel &F = h->lv[l];
        hipLaunchKernelGGL(k_amg_up<int32_t>, dim3(blocks256(F.n)), dim3(256), 0, st, F.n, F.rowptr, F.cols32,
                           F.vals, F.dinv.p, F.agg.p, F.b.p, h->lv[l + 1].x.p, w, alpha, F.x.p, state);
    }
    AmgLevel &F0 = h->lv[0];
    hipLaunchKernelGGL(k_amg_up0, dim3(g.blocks), dim3(kAmgNB), 0, st, F0.n, g.per, F0.rowptr, F0.cols64, F0.vals,
                       F0.dinv.p, F0.agg.p, r, h->lv[1].x.p, w, alpha, h->z.p, h->pz.p, state);
    HIPC(c, hipGetLastError());
    return SHK_OK;
}

}  // namespace shk
