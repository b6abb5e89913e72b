// noalloc.hip -- shk_solve_csr_noalloc: the CSR solve of shk_solve_csr with every
// piece of per-call device workspace taken from one file-scope device array instead
// of context buffers that grow with hipMalloc.  For callers that cannot allocate
// device memory during a solve (HIP graph capture, serving loops, harnesses that
// forbid hipMalloc).  The kernels are the ones shk_solve_csr launches.
//
// The array is per device: each GPU this library runs on holds its own copy for the
// life of the process.  Its device address is looked up once per device with
// hipGetSymbolAddress when the first context on that device is created (with the
// device current); host code passes that address, never the array's name, to
// kernels.  Every call on one device shares the array: two calls that overlap on one
// GPU (two threads, two contexts, two streams) race and corrupt each other.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "ctx_internal.h"
#include "krylov.h"

namespace {

constexpr size_t kScratchBytes = (size_t)128 << 20;  // 128 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

constexpr size_t kAlignDoubles = 32;  // 256 bytes
size_t round_up(size_t n) { return (n + kAlignDoubles - 1) / kAlignDoubles * kAlignDoubles; }

// doubles of scratch a CSR solve of `rows` rows and `nnz` entries uses
size_t layout_doubles(int64_t rows, int64_t nnz, bool jac) {
    const size_t parts = round_up((size_t)shk::kMaxBlocks + 64) + round_up(2 * (size_t)shk::kMaxBlocks + 64) +
                         round_up(4 * (size_t)shk::kMaxBlocks + 64);
    return parts + 6 * round_up((size_t)rows) + (jac ? round_up((size_t)std::max<int64_t>(nnz, 1)) : 0);
}

}  // namespace

namespace shk {

size_t scratch_capacity_bytes() { return kScratchBytes; }

// Looks up (once per device, with `device` current) the device address of this
// device's copy of the scratch array.
hipError_t scratch_bind(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

}  // namespace shk

extern "C" {

size_t shk_noalloc_scratch_bytes(void) { return kScratchBytes; }

int shk_solve_csr_noalloc(shk_ctx *c, int64_t rows, const int64_t *rowptr, const int64_t *cols, const double *vals,
                          const double *b, double *x, shk_result *res) {
    using namespace shk;
    if (!c) return SHK_BAD_ARG;
    c->err.clear();
    if (rows < 0 || rows > ((int64_t)1 << 40)) return c->fail(SHK_BAD_ARG, "bad row count");
    if (!rowptr || !cols || !vals || !b || !x) return c->fail(SHK_BAD_ARG, "NULL CSR/vector pointer");
    if (c->prm.host_pointers) return c->fail(SHK_BAD_ARG, "shk_solve_csr_noalloc takes device pointers");
    if (c->prm.precond == 2 || c->prm.precond == 3)
        return c->fail(SHK_BAD_ARG, "shk_solve_csr_noalloc takes Preconditioner=none or jacobi");
    if (c->profile) return c->fail(SHK_BAD_ARG, "shk_solve_csr_noalloc does not run with ProfileKernels=1");
    HIPC(c, hipSetDevice(c->device));
    void *base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_ptr[c->device] : nullptr;
    if (!base) return c->fail(SHK_HIP_ERROR, "no scratch address for device %d", c->device);
    shk_result local{};
    if (!res) res = &local;
    *res = shk_result{};
    res->unknowns = rows;
    res->outer = 1;
    if (rows == 0) return SHK_OK;
    const bool jac = c->prm.precond != 0;
    int64_t nnz = 0;
    if (jac) {
        HIPC(c, hipMemcpyAsync(&nnz, rowptr + rows, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        if (nnz < 0) return c->fail(SHK_BAD_ARG, "rowptr[rows] is negative");
    }
    const size_t need = layout_doubles(rows, nnz, jac) * sizeof(double);
    if (need > kScratchBytes)
        return c->fail(SHK_BAD_ARG,
                       "shk_solve_csr_noalloc: %lld rows and %lld entries need %zu bytes of scratch, more than its "
                       "%zu; use shk_solve_csr",
                       (long long)rows, (long long)nnz, need, kScratchBytes);
    // carve the scratch: partials, then the six vectors, then the scaled values
    double *cur = static_cast<double *>(base);
    auto take = [&](size_t n) {
        double *p = cur;
        cur += round_up(n);
        return p;
    };
    double *pa = take((size_t)kMaxBlocks + 64);
    double *pb = take(2 * (size_t)kMaxBlocks + 64);
    double *pc = take(4 * (size_t)kMaxBlocks + 64);
    double *diag = take((size_t)rows), *sig = take((size_t)rows), *r = take((size_t)rows);
    double *s0 = take((size_t)rows), *s1 = take((size_t)rows), *q = take((size_t)rows);
    const CsrOp op{rowptr, cols, vals, rows};
    CsrOp sop = op;
    const Grid g = make_grid(rows);
    HIPC(c, hipEventRecord(c->ev[0], c->stream));
    launch_csr_diag(c->stream, op, diag, sig);
    if (jac) {
        // the Jacobi-PCG iterates as CG on D^-1/2 A D^-1/2 (scaled copy of the values)
        double *hv = take((size_t)std::max<int64_t>(nnz, 1));
        launch_csr_scale(c->stream, op, sig, hv);
        sop.vals = hv;
    }
    const double *sg = jac ? sig : nullptr;
    const Reduction red{pa, pb, c->state.p};
    launch_cg_init<double, double>(c->stream, g, b, sg, x, r, s0, red, c->prm.residual, c->prm.max_iterations);
    double *S[2] = {s0, s1};
    int m = 0;
    launch_dir_spmv_csr(c->stream, g, sop, r, S[0], S[1], q, red);
    ++m;
    int rc = drive(c, [&]() {
        launch_axpy_norm<double>(c->stream, g, sg, x, r, S[m & 1], q, red);
        launch_dir_spmv_csr(c->stream, g, sop, r, S[m & 1], S[(m + 1) & 1], q, red);
        ++m;
    });
    if (rc) return rc;
    if (jac) launch_unscale(c->stream, rows, x, sig);
    HIPC(c, hipEventRecord(c->ev[1], c->stream));
    res->count = c->h_state->iter;
    res->status = c->h_state->status;
    res->reresid = c->h_state->rnorm0 > 0 ? c->h_state->rnorm / c->h_state->rnorm0 : c->h_state->rnorm;
    launch_residual_csr(c->stream, g, op, b, x, nullptr, pc, c->misc.p);
    HIPC(c, hipGetLastError());
    HIPC(c, hipMemcpyAsync(c->h_misc, c->misc.p, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    const double rr[4] = {c->h_misc[0], c->h_misc[1], c->h_misc[2], c->h_misc[3]};
    fill_true_residual(res, rr);
    float t = 0.0f;
    HIPC(c, hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
    res->ms_solve = t;
    // the context's projection system is untouched (nothing here writes its buffers)
    return res->status;
}

}  // extern "C"
