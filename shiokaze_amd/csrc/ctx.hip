// ctx.hip -- the C-ABI of libshk_projection.so (include/shk_projection.h).
//
// shk_project follows macpressuresolver3::project (src/projection/macpressuresolver3.cpp:54-299):
//   fractions -> unknown labels -> stencil coefficients + rhs -> solve -> pressure
//   scatter -> velocity update,
// with the matrix kept matrix-free on 8x8x8 bricks that hold at least one unknown,
// so the Krylov loop streams only the liquid part of the grid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "krylov.h"
#include "setup.h"
#include "shk_common.h"
#include "shk_fractions.h"
#include "shk_projection.h"
#include "ctx_internal.h"

using namespace shk;

namespace {

thread_local std::string g_create_error;

}  // namespace

namespace {

bool parse_double(const char *s, double &out) {
    if (!s || !*s) return false;
    char *end = nullptr;
    errno = 0;
    const double v = strtod(s, &end);
    if (errno || !end || *end) return false;
    out = v;
    return true;
}
bool parse_int(const char *s, int64_t &out) {
    double v;
    if (!parse_double(s, v) || v != std::floor(v) || std::fabs(v) > 9e15) return false;
    out = (int64_t)v;
    return true;
}
bool parse_bool(const char *s, bool &out) {
    if (!s) return false;
    std::string v(s);
    for (auto &ch : v) ch = (char)tolower(ch);
    if (v == "1" || v == "true" || v == "yes" || v == "on") return out = true, true;
    if (v == "0" || v == "false" || v == "no" || v == "off") return out = false, true;
    return false;
}

int sync_and_time(shk_ctx *c, int a, int b, double &ms) {
    float t = 0.0f;
    HIPC(c, hipEventSynchronize(c->ev[b]));
    HIPC(c, hipEventElapsedTime(&t, c->ev[a], c->ev[b]));
    ms = t;
    return SHK_OK;
}

int ensure_partials(shk_ctx *c) {
    HIPC(c, c->pa.ensure((size_t)kMaxBlocks + 64));
    HIPC(c, c->pb.ensure(2 * (size_t)kMaxBlocks + 64));
    HIPC(c, c->pc.ensure(4 * (size_t)kMaxBlocks + 64));
    return SHK_OK;
}

// Builds the active brick list, slot map and neighbour table from the brick flags.
int build_bricks(shk_ctx *c) {
    const int64_t nb = (int64_t)c->nbx * c->nby * c->nbz;
    c->h_flag.resize(nb);
    HIPC(c, hipMemcpyAsync(c->h_flag.data(), c->brick_flag.p, nb, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    c->h_slot.assign(nb, -1);
    c->h_coord.clear();
    for (int bz = 0; bz < c->nbz; ++bz)
        for (int by = 0; by < c->nby; ++by)
            for (int bx = 0; bx < c->nbx; ++bx) {
                const int64_t b = bx + (int64_t)c->nbx * (by + (int64_t)c->nby * bz);
                if (c->h_flag[b]) {
                    c->h_slot[b] = (int)c->h_coord.size();
                    c->h_coord.push_back(make_int4(bx, by, bz, 0));
                }
            }
    c->n_active = (int)c->h_coord.size();
    c->h_nbr.assign(6 * (size_t)std::max(c->n_active, 1), -1);
    for (int s = 0; s < c->n_active; ++s) {
        const int4 q = c->h_coord[s];
        const int off[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
        for (int f = 0; f < 6; ++f) {
            const int x = q.x + off[f][0], y = q.y + off[f][1], z = q.z + off[f][2];
            if (x < 0 || y < 0 || z < 0 || x >= c->nbx || y >= c->nby || z >= c->nbz) continue;
            c->h_nbr[6 * s + f] = c->h_slot[x + (int64_t)c->nbx * (y + (int64_t)c->nby * z)];
        }
    }
    HIPC(c, c->brick_slot.ensure(nb));
    HIPC(c, hipMemcpyAsync(c->brick_slot.p, c->h_slot.data(), sizeof(int) * nb, hipMemcpyHostToDevice, c->stream));
    if (c->n_active == 0) return SHK_OK;
    const size_t n = (size_t)c->n_active * kBrickCells;
    HIPC(c, c->coord.ensure(c->n_active));
    HIPC(c, c->nbr.ensure(6 * (size_t)c->n_active));
    HIPC(c, hipMemcpyAsync(c->coord.p, c->h_coord.data(), sizeof(int4) * c->n_active, hipMemcpyHostToDevice,
                           c->stream));
    HIPC(c, hipMemcpyAsync(c->nbr.p, c->h_nbr.data(), sizeof(int) * 6 * c->n_active, hipMemcpyHostToDevice,
                           c->stream));
    for (DevBuf<double> *b : {&c->diag64, &c->cx64, &c->cy64, &c->cz64, &c->b64, &c->sig64, &c->x64, &c->r64,
                              &c->s64, &c->t64, &c->q64})
        HIPC(c, b->ensure(n));
    if (c->prm.precond == 1 && c->prm.precision == 0) {
        for (DevBuf<double> *b : {&c->hx64, &c->hy64, &c->hz64}) HIPC(c, b->ensure(n));
        HIPC(c, c->kappa64.ensure(c->n_active));
    }
    if (c->prm.precond == 0 && c->prm.precision == 0) {
        // plain FP64 CG: uniform bricks (one coupling, one diagonal) skip their coefficients
        HIPC(c, c->kappa64.ensure(c->n_active));
        HIPC(c, c->kdiag64.ensure(c->n_active));
    }
    if (c->prm.precision == 1 || c->prm.precond >= 2) {
        for (DevBuf<float> *b : {&c->x32, &c->r32, &c->s32, &c->t32, &c->q32, &c->rs32}) HIPC(c, b->ensure(n));
        if (c->prm.precond == 0) {
            for (DevBuf<float> *b : {&c->diag32, &c->cx32, &c->cy32, &c->cz32}) HIPC(c, b->ensure(n));
            HIPC(c, c->kappa32.ensure(c->n_active));
            HIPC(c, c->kdiag32.ensure(c->n_active));
        } else {
            for (DevBuf<float> *b : {&c->hx32, &c->hy32, &c->hz32}) HIPC(c, b->ensure(n));
            HIPC(c, c->kappa32.ensure(c->n_active));
        }
    }
    c->grid = make_grid((int64_t)c->n_active * kBrickCells);
    // host-side buffers must outlive the async copies above
    HIPC(c, hipStreamSynchronize(c->stream));
    return ensure_partials(c);
}

}  // namespace

namespace shk {

Reduction reduction(shk_ctx *c) { return Reduction{c->pa.p, c->pb.p, c->state.p}; }

int drive_impl(shk_ctx *c, void (*iterate)(void *), void *arg, int interval_override) {
    HIPC(c, hipMemcpyAsync(c->h_state, c->state.p, sizeof(SolveState), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    const int interval = std::max(1, interval_override > 0 ? interval_override : c->prm.check_interval);
    constexpr int K = shk_ctx::kKernels;
    if (c->profile && c->pev.size() < (size_t)(2 * K * interval)) {
        const size_t old = c->pev.size();
        c->pev.resize(2 * K * interval);
        for (size_t i = old; i < c->pev.size(); ++i) HIPC(c, hipEventCreate(&c->pev[i]));
    }
    while (!c->h_state->done) {
        const int64_t it0 = c->h_state->iter;
        c->pcur = 0;
        for (int t = 0; t < interval; ++t) iterate(arg);
        HIPC(c, hipGetLastError());
        HIPC(c, hipMemcpyAsync(c->h_state, c->state.p, sizeof(SolveState), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        if (c->profile && c->pcur == (size_t)(2 * K * interval)) {
            // only launches that did work count: the first `delta` iterations; the
            // direction update of the final iteration returns at once
            const int64_t delta = c->h_state->iter - it0;
            for (int64_t t = 0; t < delta && t < interval; ++t) {
                for (int kk = 0; kk < K; ++kk) {
                    if (kk == K - 1 && c->h_state->done && t == delta - 1) continue;
                    float ms = 0.0f;
                    const size_t l = (size_t)(K * t + kk);
                    HIPC(c, hipEventElapsedTime(&ms, c->pev[2 * l], c->pev[2 * l + 1]));
                    c->pcount[kk] += 1;
                    c->pms[kk] += ms;
                }
            }
        }
    }
    return SHK_OK;
}

BrickOp<double> op64(shk_ctx *c) {
    return BrickOp<double>{c->nbr.p, c->diag64.p, c->cx64.p, c->cy64.p, c->cz64.p, c->n_active, nullptr, c->dims.bx};
}
BrickOp<float> op32(shk_ctx *c) {
    if (c->prm.precond == 0)
        return BrickOp<float>{c->nbr.p, c->diag32.p, c->cx32.p, c->cy32.p, c->cz32.p, c->n_active, c->kappa32.p,
                              c->dims.bx, c->kdiag32.p};
    return BrickOp<float>{c->nbr.p, nullptr, c->hx32.p, c->hy32.p, c->hz32.p, c->n_active, c->kappa32.p, c->dims.bx};
}

int true_residual(shk_ctx *c, const double *x, double *r_out, double out[4], const double *sig, float *rs) {
    launch_residual(c->stream, c->grid, op64(c), c->b64.p, x, r_out, c->pc.p, c->misc.p, sig, rs);
    HIPC(c, hipGetLastError());
    HIPC(c, hipMemcpyAsync(c->h_misc, c->misc.p, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < 4; ++i) out[i] = c->h_misc[i];
    return SHK_OK;
}

void fill_true_residual(shk_result *res, const double rr[4]) {
    res->b_inf = rr[3];
    res->true_relres_inf = rr[3] > 0 ? rr[0] / rr[3] : 0.0;
    res->true_relres_l2 = rr[2] > 0 ? std::sqrt(rr[1] / rr[2]) : 0.0;
}

// FP64 Krylov on the brick stencil: plain CG, or Jacobi PCG as CG on the
// symmetrically scaled system (x = sig * x_scaled).
int solve_fp64(shk_ctx *c, shk_result *res) {
    const bool jac = c->prm.precond != 0;
    const double *sig = jac ? c->sig64.p : nullptr;
    BrickOp<double> op =
        jac ? BrickOp<double>{c->nbr.p, nullptr, c->hx64.p, c->hy64.p, c->hz64.p, c->n_active, c->kappa64.p, c->dims.bx}
            : op64(c);
    if (!jac) {
        op.kappa = c->kappa64.p;
        op.kdiag = c->kdiag64.p;
    }
    // the FP64 stencil holds 3 workgroups per CU, not 4: all kernels of the solve run on
    // that many, so that the stencil's last quarter does not run as a second round
    const Grid g = make_grid(c->grid.n, dir_spmv_resident<double, double>(c->grid, op));
    const Reduction red = reduction(c);
    launch_cg_init<double, double>(c->stream, g, c->b64.p, sig, c->x64.p, c->r64.p, c->s64.p, red, c->prm.residual,
                                   c->prm.max_iterations);
    double *S[2] = {c->s64.p, c->t64.p};
    int m = 0;
    launch_dir_spmv<double, double>(c->stream, g, op, c->r64.p, S[0], S[1], c->q64.p, red);
    ++m;
    int rc = drive(c, [&]() {
        launch_axpy_norm<double>(c->stream, g, sig, c->x64.p, c->r64.p, S[m & 1], c->q64.p, red, c->next_events());
        launch_dir_spmv<double, double>(c->stream, g, op, c->r64.p, S[m & 1], S[(m + 1) & 1], c->q64.p, red, c->next_events());
        ++m;
    });
    if (rc) return rc;
    if (jac) launch_unscale(c->stream, g.n, c->x64.p, c->sig64.p);
    res->count = c->h_state->iter;
    res->status = c->h_state->status;
    res->reresid = c->h_state->rnorm0 > 0 ? c->h_state->rnorm / c->h_state->rnorm0 : c->h_state->rnorm;
    res->outer = 1;
    double rr[4];
    if ((rc = true_residual(c, c->x64.p, nullptr, rr))) return rc;
    fill_true_residual(res, rr);
    return SHK_OK;
}

}  // namespace shk

namespace {

// Mixed precision: FP32 stencil/AXPY/Jacobi inner CG with FP64 dots, FP64 true
// residual and correction between inner solves (iterative refinement).
int solve_mixed(shk_ctx *c, shk_result *res) {
    const Grid &g = c->grid;
    const bool jac = c->prm.precond != 0;
    const double *sig = jac ? c->sig64.p : nullptr;
    HIPC(c, hipMemsetAsync(c->x64.p, 0, sizeof(double) * g.n, c->stream));
    double rr[4];
    int rc;
    // r64 = b - A*0 = b and its scaled FP32 copy, the first inner right-hand side
    if ((rc = true_residual(c, c->x64.p, c->r64.p, rr, sig, c->rs32.p))) return rc;
    const double b_inf = rr[3];
    const double tol_abs = std::max(c->prm.residual, 1e-30) * b_inf;
    double r_inf = rr[0];
    int64_t total = 0;
    int outer = 0;
    int status = SHK_OK;
    const BrickOp<float> op = op32(c);
    const Reduction red = reduction(c);
    while (r_inf > tol_abs && outer < std::max(1, c->prm.outer_max)) {
        double inner_rel = std::max(c->prm.inner_residual, 0.5 * tol_abs / r_inf);
        if (inner_rel > 0.5) inner_rel = 0.5;
        launch_cg_init<float, float>(c->stream, g, c->rs32.p, nullptr, c->x32.p, c->r32.p, c->s32.p, red, inner_rel,
                                     c->prm.max_iterations - total);
        float *S[2] = {c->s32.p, c->t32.p};
        int m = 0;
        launch_dir_spmv<float, float>(c->stream, g, op, c->r32.p, S[0], S[1], c->q32.p, red);
        ++m;
        rc = drive(c, [&]() {
            launch_axpy_norm<float>(c->stream, g, nullptr, c->x32.p, c->r32.p, S[m & 1], c->q32.p, red, c->next_events());
            launch_dir_spmv<float, float>(c->stream, g, op, c->r32.p, S[m & 1], S[(m + 1) & 1], c->q32.p, red, c->next_events());
            ++m;
        });
        if (rc) return rc;
        total += c->h_state->iter;
        ++outer;
        const int st = c->h_state->status;
        launch_add_correction(c->stream, g.n, c->x64.p, c->x32.p, sig);
        if ((rc = true_residual(c, c->x64.p, c->r64.p, rr, sig, c->rs32.p))) return rc;
        r_inf = rr[0];
        if (st == SHK_NAN || st == SHK_BREAKDOWN) {
            status = st;
            break;
        }
        if (total >= c->prm.max_iterations) break;
    }
    if (status == SHK_OK && r_inf > tol_abs) status = SHK_NOT_CONVERGED;
    if (!(r_inf == r_inf)) status = SHK_NAN;
    res->count = total;
    res->outer = outer;
    res->status = status;
    res->reresid = b_inf > 0 ? r_inf / b_inf : 0.0;
    fill_true_residual(res, rr);
    return SHK_OK;
}

int validate_fields(shk_ctx *c, const void *const *ptrs, int n) {
    for (int i = 0; i < n; ++i)
        if (!ptrs[i]) return c->fail(SHK_BAD_ARG, "field pointer %d is NULL", i);
    return SHK_OK;
}

}  // namespace

// ===================================================================== C-ABI

extern "C" {

int shk_abi_version(void) { return SHK_ABI_VERSION; }

const char *shk_build_info(void) {
    return "libshk_projection (gfx950; brick 8^3 matrix-free MAC pressure projection; FP64 + mixed CG; "
           "aggregation-multigrid PCG; learned flexible-CG preconditioner)";
}

const char *shk_last_error(const shk_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

namespace {

// A context over an nx x ny x nz grid with bricks of x extent bx (8: 8x8x8 bricks;
// 64: 64x1x8 bricks, the layout of a 2D grid held as nx x 1 x ny).
int create_impl(int nx, int ny, int nz, int bx, double dx, int device, shk_ctx **out) {
    if (!out) {
        g_create_error = "out is NULL";
        return SHK_BAD_ARG;
    }
    *out = nullptr;
    if (nx <= 0 || ny <= 0 || nz <= 0 || !(dx > 0.0) || (int64_t)nx * ny * nz > (int64_t)1 << 31) {
        g_create_error = "bad shape or dx";
        return SHK_BAD_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        g_create_error = "no HIP device";
        return SHK_NO_DEVICE;
    }
    if (device < 0) {
        if (hipGetDevice(&device) != hipSuccess) device = 0;
    }
    if (device >= ndev) {
        g_create_error = "device index out of range";
        return SHK_BAD_ARG;
    }
    shk_ctx *c = new shk_ctx();
    c->device = device;
    c->dims = Dims{nx, ny, nz, bx};
    c->dx = dx;
    const int ex = bx, ey = kWave / bx, ez = kBrickCells / kWave;
    c->nbx = (nx + ex - 1) / ex;
    c->nby = (ny + ey - 1) / ey;
    c->nbz = (nz + ez - 1) / ez;
    auto bail = [&](hipError_t e, const char *what) {
        g_create_error = std::string(what) + ": " + hipGetErrorString(e);
        shk_destroy(c);
        return e == hipErrorOutOfMemory ? SHK_OUT_OF_MEMORY : SHK_HIP_ERROR;
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return bail(e, "hipSetDevice");
    // the device address of this device's allocation-free solve scratch (noalloc.hip)
    if ((e = shk::scratch_bind(device)) != hipSuccess) return bail(e, "hipGetSymbolAddress(scratch)");
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(e, "hipStreamCreate");
    for (auto &ev : c->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipHostMalloc((void **)&c->h_state, sizeof(SolveState), hipHostMallocDefault)) != hipSuccess)
        return bail(e, "hipHostMalloc");
    if ((e = hipHostMalloc((void **)&c->h_misc, 16 * sizeof(double), hipHostMallocDefault)) != hipSuccess)
        return bail(e, "hipHostMalloc");
    if ((e = c->state.ensure(1)) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = hipMemset(c->state.p, 0, sizeof(SolveState))) != hipSuccess) return bail(e, "hipMemset");
    if ((e = c->misc.ensure(16)) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = c->count.ensure(1)) != hipSuccess) return bail(e, "hipMalloc");
    const int64_t G = (int64_t)nx * ny * nz;
    for (int a = 0; a < 3; ++a) {
        const int64_t nf = face_count(c->dims, a);
        if ((e = c->area[a].ensure(nf)) != hipSuccess) return bail(e, "hipMalloc");
        if ((e = c->rho[a].ensure(nf)) != hipSuccess) return bail(e, "hipMalloc");
    }
    if ((e = c->label.ensure(G)) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = c->pressure.ensure(G)) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = hipMemset(c->pressure.p, 0, sizeof(float) * G)) != hipSuccess) return bail(e, "hipMemset");
    if ((e = c->brick_flag.ensure((size_t)c->nbx * c->nby * c->nbz)) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = c->brick_slot.ensure((size_t)c->nbx * c->nby * c->nbz)) != hipSuccess) return bail(e, "hipMalloc");
    if (bx == 64) {
        // 2D: the y faces of the nx x 1 x ny layout are walls (area 0) with zero
        // velocity; their arrays are set once and never written
        const int64_t nf1 = face_count(c->dims, 1);
        if ((e = hipMemset(c->area[1].p, 0, sizeof(float) * nf1)) != hipSuccess) return bail(e, "hipMemset");
        if ((e = hipMemset(c->rho[1].p, 0, sizeof(float) * nf1)) != hipSuccess) return bail(e, "hipMemset");
        if ((e = c->wall_faces.ensure(nf1)) != hipSuccess) return bail(e, "hipMalloc");
        if ((e = hipMemset(c->wall_faces.p, 0, sizeof(float) * nf1)) != hipSuccess) return bail(e, "hipMemset");
    }
    *out = c;
    return SHK_OK;
}

}  // namespace

int shk_create(int nx, int ny, int nz, double dx, int device, shk_ctx **out) {
    return create_impl(nx, ny, nz, 8, dx, device, out);
}

int shk_create2d(int nx, int ny, double dx, int device, shk_ctx **out) {
    return create_impl(nx, 1, ny, 64, dx, device, out);
}

void shk_destroy(shk_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->unet) unet_destroy(c->unet);
    if (c->mg) mg_destroy(c->mg);
    if (c->amg) amg_destroy(c->amg);
    if (c->redist) redist_destroy(c->redist);
    if (c->liquid) liquid_destroy(c->liquid);
    for (auto &b : c->area) b.release();
    for (auto &b : c->rho) b.release();
    for (auto &b : c->stage) b.release();
    c->curvature.release();
    c->prev64.release();
    c->warm64.release();
    c->wall_faces.release();
    c->label.release();
    c->pressure.release();
    c->brick_flag.release();
    c->brick_slot.release();
    c->count.release();
    c->coord.release();
    c->nbr.release();
    for (DevBuf<double> *b : {&c->diag64, &c->cx64, &c->cy64, &c->cz64, &c->b64, &c->sig64, &c->hx64, &c->hy64,
                              &c->hz64, &c->kappa64, &c->kdiag64, &c->x64, &c->r64, &c->s64, &c->t64, &c->q64, &c->pa, &c->pb,
                              &c->pc, &c->misc})
        b->release();
    for (DevBuf<float> *b : {&c->hx32, &c->hy32, &c->hz32, &c->kappa32, &c->kdiag32, &c->diag32, &c->cx32, &c->cy32, &c->cz32, &c->x32, &c->r32,
                             &c->s32, &c->t32, &c->q32, &c->rs32})
        b->release();
    for (auto &b : c->csr_i) b.release();
    for (auto &b : c->csr_d) b.release();
    c->state.release();
    if (c->h_state) (void)hipHostFree(c->h_state);
    if (c->h_misc) (void)hipHostFree(c->h_misc);
    for (auto &ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto &ev : c->uev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto &ev : c->pev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int shk_configure(shk_ctx *c, const char *key, const char *value) {
    if (!c) return SHK_BAD_ARG;
    if (!key || !value) return c->fail(SHK_BAD_ARG, "key/value is NULL");
    std::string k(key);
    // accept the reference's scoped names ("PCG.Residual", "MacUtility.EpsFluid")
    const size_t dot = k.rfind('.');
    if (dot != std::string::npos) k = k.substr(dot + 1);
    Params &p = c->prm;
    double d;
    int64_t i;
    bool b;
    if (k == "Residual") {
        if (!parse_double(value, d) || !(d >= 0.0)) return c->fail(SHK_BAD_ARG, "bad Residual '%s'", value);
        p.residual = d;
    } else if (k == "MaxIterations") {
        if (!parse_int(value, i) || i < 0) return c->fail(SHK_BAD_ARG, "bad MaxIterations '%s'", value);
        p.max_iterations = i;
    } else if (k == "ModifiedIC") {
        if (!parse_double(value, d)) return c->fail(SHK_BAD_ARG, "bad ModifiedIC '%s'", value);
        p.modified_ic = d;
    } else if (k == "MinDiagRatio") {
        if (!parse_double(value, d)) return c->fail(SHK_BAD_ARG, "bad MinDiagRatio '%s'", value);
        p.min_diag_ratio = d;
    } else if (k == "SecondOrderAccurateFluid") {
        if (!parse_bool(value, b)) return c->fail(SHK_BAD_ARG, "bad SecondOrderAccurateFluid '%s'", value);
        p.so_fluid = b;
    } else if (k == "SecondOrderAccurateSolid") {
        if (!parse_bool(value, b)) return c->fail(SHK_BAD_ARG, "bad SecondOrderAccurateSolid '%s'", value);
        p.so_solid = b;
    } else if (k == "Gain") {
        if (!parse_double(value, d)) return c->fail(SHK_BAD_ARG, "bad Gain '%s'", value);
        p.gain = d;
    } else if (k == "WarmStart") {
        if (!parse_bool(value, b)) return c->fail(SHK_BAD_ARG, "bad WarmStart '%s'", value);
        c->prm.warm_start = b;
    } else if (k == "Report") {
        if (!parse_bool(value, b)) return c->fail(SHK_BAD_ARG, "bad Report '%s'", value);
        p.report = b;
    } else if (k == "IntegrationRate") {
        if (!parse_double(value, d) || !(d > 0.0) || d > 1.0)
            return c->fail(SHK_BAD_ARG, "bad IntegrationRate '%s' (0 < rate <= 1)", value);
        p.redist_rate = d;
    } else if (k == "RedistTemporalScheme") {
        const std::string v(value);
        if (v == "Euler")
            p.redist_rk2 = false;
        else if (v == "RK2")
            p.redist_rk2 = true;
        else
            return c->fail(SHK_BAD_ARG, "RedistTemporalScheme must be Euler or RK2, got '%s'", value);
    } else if (k == "SolidWallExtrapolation") {
        if (!parse_bool(value, b)) return c->fail(SHK_BAD_ARG, "bad SolidWallExtrapolation '%s'", value);
        p.solid_wall_extrap = b;
    } else if (k == "HorizontalSolidWallExtrapolation") {
        if (!parse_bool(value, b)) return c->fail(SHK_BAD_ARG, "bad HorizontalSolidWallExtrapolation '%s'", value);
        p.horizontal_extrap = b;
    } else if (k == "MacCormack") {
        if (!parse_bool(value, b)) return c->fail(SHK_BAD_ARG, "bad MacCormack '%s'", value);
        p.maccormack = b;
    } else if (k == "TrimNarrowBand") {
        if (!parse_int(value, i) || i < 0 || i > 1000) return c->fail(SHK_BAD_ARG, "bad TrimNarrowBand '%s'", value);
        p.trim_narrowband = (int)i;
    } else if (k == "WENO") {
        if (!parse_bool(value, b)) return c->fail(SHK_BAD_ARG, "bad WENO '%s'", value);
        p.weno = b;
    } else if (k == "EpsFluid") {
        if (!parse_double(value, d) || d < 0) return c->fail(SHK_BAD_ARG, "bad EpsFluid '%s'", value);
        p.eps_fluid = d;
    } else if (k == "EpsSolid") {
        if (!parse_double(value, d) || d < 0) return c->fail(SHK_BAD_ARG, "bad EpsSolid '%s'", value);
        p.eps_solid = d;
    } else if (k == "Precision") {
        std::string v(value);
        if (v == "fp64")
            p.precision = 0;
        else if (v == "mixed")
            p.precision = 1;
        else
            return c->fail(SHK_BAD_ARG, "Precision must be fp64 or mixed, got '%s'", value);
    } else if (k == "Preconditioner") {
        std::string v(value);
        if (v == "none")
            p.precond = 0;
        else if (v == "jacobi")
            p.precond = 1;
        else if (v == "unet")
            p.precond = 2;
        else if (v == "mg")
            p.precond = 3;
        else
            return c->fail(SHK_BAD_ARG, "Preconditioner must be none, jacobi, unet or mg, got '%s'", value);
    } else if (k == "OuterMax") {
        if (!parse_int(value, i) || i < 1 || i > 50) return c->fail(SHK_BAD_ARG, "bad OuterMax '%s'", value);
        p.outer_max = (int)i;
    } else if (k == "InnerResidual") {
        if (!parse_double(value, d) || !(d > 0.0) || d >= 1.0)
            return c->fail(SHK_BAD_ARG, "bad InnerResidual '%s'", value);
        p.inner_residual = d;
    } else if (k == "CheckInterval") {
        if (!parse_int(value, i) || i < 1 || i > 4096) return c->fail(SHK_BAD_ARG, "bad CheckInterval '%s'", value);
        p.check_interval = (int)i;
    } else if (k == "HostPointers") {
        if (!parse_bool(value, b)) return c->fail(SHK_BAD_ARG, "bad HostPointers '%s'", value);
        p.host_pointers = b;
    } else if (k == "ProfileKernels") {
        if (!parse_bool(value, b)) return c->fail(SHK_BAD_ARG, "bad ProfileKernels '%s'", value);
        c->profile = b;
        for (int q = 0; q < shk_ctx::kKernels; ++q) {
            c->pcount[q] = 0;
            c->pms[q] = 0.0;
        }
        c->ucount = 0;
        c->ums = 0.0;
    } else if (k == "MgOmega" || k == "MgAlpha") {
        if (!parse_double(value, d) || !(d > 0.0) || !(d < 2.0))
            return c->fail(SHK_BAD_ARG, "bad %s '%s' (0 < value < 2)", k.c_str(), value);
        (k == "MgOmega" ? p.mg_omega : p.mg_alpha) = d;
    } else if (k == "UnetChebyshev") {
        if (!parse_double(value, d) || !(d == 0.0 || (d > 1.0 && d <= 1e6)))
            return c->fail(SHK_BAD_ARG, "bad UnetChebyshev '%s' (0, or a ratio > 1)", value);
        p.unet_cheb = d;
    } else if (k == "UnetSmoothing") {
        if (!parse_int(value, i) || i < 1 || i > 64) return c->fail(SHK_BAD_ARG, "bad UnetSmoothing '%s' (1..64)", value);
        p.unet_smooth = (int)i;
    } else if (k == "UnetWeights") {
        p.unet_weights = value;
        if (c->unet) {
            unet_destroy(c->unet);
            c->unet = nullptr;
        }
    } else {
        return c->fail(SHK_BAD_ARG, "unknown key '%s'", key);
    }
    return SHK_OK;
}

int shk_set_target_volume(shk_ctx *c, double current_volume, double target_volume) {
    if (!c) return SHK_BAD_ARG;
    c->err.clear();
    if (!std::isfinite(current_volume) || !std::isfinite(target_volume))
        return c->fail(SHK_BAD_ARG, "volumes must be finite");
    c->current_volume = current_volume;
    c->target_volume = target_volume;
    return SHK_OK;
}

// Projection state in the field order of macpressuresolver3::serialize
// (macpressuresolver3.cpp:332-344): magic, shape (3 x uint32), dx, target volume,
// current volume, the PI integral y_prev, then a uint64 cell count and the warm-start
// pressure per cell (FP64), the count 0 when there is none.
namespace {
constexpr char kStateMagic[8] = {'S', 'H', 'K', 'P', 'S', 'T', '0', '1'};
constexpr size_t kStateHeader = 8 + 3 * 4 + 4 + 4 * 8 + 8;  // magic, shape, pad, 4 doubles, count
}  // namespace

int shk_state_bytes(const shk_ctx *c, size_t *bytes) {
    if (!c || !bytes) return SHK_BAD_ARG;
    const size_t G = (size_t)c->dims.nx * c->dims.ny * c->dims.nz;
    *bytes = kStateHeader + (c->prev64.p ? G * sizeof(double) : 0);
    return SHK_OK;
}

int shk_save_state(shk_ctx *c, void *buf, size_t capacity) {
    if (!c || !buf) return SHK_BAD_ARG;
    c->err.clear();
    size_t need = 0;
    shk_state_bytes(c, &need);
    if (capacity < need) return c->fail(SHK_BAD_ARG, "state buffer holds %zu bytes, %zu needed", capacity, need);
    HIPC(c, hipSetDevice(c->device));
    unsigned char *p = static_cast<unsigned char *>(buf);
    std::memcpy(p, kStateMagic, 8);
    const uint32_t shape[4] = {(uint32_t)c->dims.nx, (uint32_t)c->dims.ny, (uint32_t)c->dims.nz, 0u};
    std::memcpy(p + 8, shape, sizeof shape);
    const double v[4] = {c->dx, c->target_volume, c->current_volume, c->y_prev};
    std::memcpy(p + 24, v, sizeof v);
    const uint64_t G = c->prev64.p ? (uint64_t)c->dims.nx * c->dims.ny * c->dims.nz : 0;
    std::memcpy(p + 56, &G, 8);
    if (G) {
        HIPC(c, hipMemcpyAsync(p + kStateHeader, c->prev64.p, sizeof(double) * G, hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
    }
    return SHK_OK;
}

int shk_load_state(shk_ctx *c, const void *buf, size_t bytes) {
    if (!c || !buf) return SHK_BAD_ARG;
    c->err.clear();
    const unsigned char *p = static_cast<const unsigned char *>(buf);
    if (bytes < kStateHeader || std::memcmp(p, kStateMagic, 8) != 0)
        return c->fail(SHK_BAD_ARG, "not a projection state (bad magic or size)");
    uint32_t shape[4];
    std::memcpy(shape, p + 8, sizeof shape);
    double v[4];
    std::memcpy(v, p + 24, sizeof v);
    uint64_t G = 0;
    std::memcpy(&G, p + 56, 8);
    if ((int)shape[0] != c->dims.nx || (int)shape[1] != c->dims.ny || (int)shape[2] != c->dims.nz || v[0] != c->dx)
        return c->fail(SHK_BAD_ARG, "the state was saved for another shape or dx");
    const uint64_t cells = (uint64_t)c->dims.nx * c->dims.ny * c->dims.nz;
    if ((G != 0 && G != cells) || bytes != kStateHeader + G * sizeof(double))
        return c->fail(SHK_BAD_ARG, "truncated or inconsistent projection state");
    for (int i = 1; i < 4; ++i)
        if (!std::isfinite(v[i])) return c->fail(SHK_BAD_ARG, "non-finite volume in the state");
    HIPC(c, hipSetDevice(c->device));
    c->target_volume = v[1];
    c->current_volume = v[2];
    c->y_prev = v[3];
    if (G) {
        HIPC(c, c->prev64.ensure(G));
        HIPC(c, hipMemcpyAsync(c->prev64.p, p + kStateHeader, sizeof(double) * G, hipMemcpyHostToDevice, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
    } else {
        c->prev64.release();  // the next warm-started projection starts from zero
    }
    return SHK_OK;
}

int shk_set_solid_density(shk_ctx *c, const float *du, const float *dv, const float *dw) {
    if (!c) return SHK_BAD_ARG;
    c->err.clear();
    const int given = (du != nullptr) + (dv != nullptr) + (dw != nullptr);
    if (c->dims.bx == 64) {
        // 2D: the x and y face arrays, the third NULL
        if (dw || (given != 0 && given != 2))
            return c->fail(SHK_BAD_ARG, "a 2D solid density takes the x and y face arrays (third NULL) or none");
        c->density[0] = du;
        c->density[1] = nullptr;
        c->density[2] = dv;
        return SHK_OK;
    }
    if (given != 0 && given != 3) return c->fail(SHK_BAD_ARG, "solid density needs all three face arrays or none");
    c->density[0] = du;
    c->density[1] = dv;
    c->density[2] = dw;
    return SHK_OK;
}

namespace {

// The whole projection.  user_vel / user_svel are the caller's face arrays per axis;
// a 2D context (held as nx x 1 x ny) passes nullptr for axis 1, which is then the
// context's zero wall faces.
int project_impl(shk_ctx *c, double dt, float *const user_vel[3], const float *const user_svel[3], const float *solid,
                 const float *fluid, double surface_tension, shk_result *res) {
    int rc;
    if (!(dt == dt) || std::isinf(dt)) return c->fail(SHK_BAD_ARG, "dt is not finite");
    if (!std::isfinite(surface_tension)) return c->fail(SHK_BAD_ARG, "surface_tension is not finite");
    const bool two_d = c->dims.bx == 64;
    if (two_d && c->prm.precond == 2)
        return c->fail(SHK_BAD_ARG, "the learned preconditioner is 3D only (2D: none, jacobi, mg)");
    HIPC(c, hipSetDevice(c->device));
    c->velocity_updated = false;
    shk_result local{};
    if (!res) res = &local;
    *res = shk_result{};
    res->precision = c->prm.precision;
    const Dims d = c->dims;
    const int64_t G = (int64_t)d.nx * d.ny * d.nz;
    const int64_t nn = two_d ? (int64_t)(d.nx + 1) * (d.nz + 1) : (int64_t)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
    const int64_t nf[3] = {face_count(d, 0), face_count(d, 1), face_count(d, 2)};
    float *vel[3] = {user_vel[0], user_vel[1], user_vel[2]};
    const float *svel[3] = {user_svel[0], user_svel[1], user_svel[2]};
    const float *dsolid = solid, *dfluid = fluid;
    if (c->prm.host_pointers) {
        // stage host fields: 0..2 u v w, 3..5 us vs ws, 6 solid, 7 fluid
        const float *src[8] = {vel[0], vel[1], vel[2], svel[0], svel[1], svel[2], solid, fluid};
        const int64_t cnt[8] = {nf[0], nf[1], nf[2], nf[0], nf[1], nf[2], nn, G};
        for (int i = 0; i < 8; ++i) {
            if (!src[i]) continue;
            HIPC(c, c->stage[i].ensure(cnt[i]));
            HIPC(c, hipMemcpyAsync(c->stage[i].p, src[i], sizeof(float) * cnt[i], hipMemcpyHostToDevice, c->stream));
        }
        for (int a = 0; a < 3; ++a) {
            if (vel[a]) vel[a] = c->stage[a].p;
            if (svel[a]) svel[a] = c->stage[3 + a].p;
        }
        dsolid = c->stage[6].p;
        dfluid = c->stage[7].p;
    }
    if (two_d) {
        vel[1] = c->wall_faces.p;
        svel[1] = c->wall_faces.p;
    }
    const float *dens[3] = {c->density[0], c->density[1], c->density[2]};
    if (c->prm.host_pointers && dens[0]) {
        for (int a = 0; a < 3; ++a) {
            if (!dens[a]) continue;
            HIPC(c, c->stage[10 + a].ensure(nf[a]));
            HIPC(c, hipMemcpyAsync(c->stage[10 + a].p, dens[a], sizeof(float) * nf[a], hipMemcpyHostToDevice,
                                   c->stream));
            dens[a] = c->stage[10 + a].p;
        }
    }
    float *area[3] = {c->area[0].p, c->area[1].p, c->area[2].p};
    float *rho[3] = {c->rho[0].p, c->rho[1].p, c->rho[2].p};
    res->volume_correct_rhs = 0.0;

    HIPC(c, hipEventRecord(c->ev[0], c->stream));
    launch_fractions(c->stream, d, dsolid, dfluid, c->prm.eps_solid, c->prm.eps_fluid, c->prm.so_solid,
                     c->prm.so_fluid, dens, area, rho);
    if (surface_tension != 0.0) {
        // macpressuresolver3.cpp:95-124: before the labelling, whatever the unknowns
        HIPC(c, c->curvature.ensure(G));
        launch_surface_tension(c->stream, d, dt, c->dx, surface_tension, dfluid, c->curvature.p, rho, vel);
    }
    HIPC(c, hipMemsetAsync(c->brick_flag.p, 0, (size_t)c->nbx * c->nby * c->nbz, c->stream));
    HIPC(c, hipMemsetAsync(c->count.p, 0, sizeof(unsigned long long), c->stream));
    launch_label(c->stream, d, dfluid, area, rho, c->label.p, c->brick_flag.p, c->nbx, c->nby, c->count.p);
    HIPC(c, hipGetLastError());
    unsigned long long nunk = 0;
    HIPC(c, hipMemcpyAsync(&nunk, c->count.p, sizeof nunk, hipMemcpyDeviceToHost, c->stream));
    if ((rc = build_bricks(c))) return rc;
    c->unknowns = (int64_t)nunk;
    res->unknowns = c->unknowns;
    res->active_bricks = c->n_active;
    if (nunk == 0) {
        // macpressuresolver3.cpp:170: nothing is solved; the velocity keeps only the
        // surface-tension jump, if any
        c->have_system = false;
        if (c->prm.host_pointers && surface_tension != 0.0)
            for (int a = 0; a < 3; ++a)
                if (user_vel[a])
                    HIPC(c, hipMemcpyAsync(user_vel[a], vel[a], sizeof(float) * nf[a], hipMemcpyDeviceToHost,
                                           c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        return SHK_OK;
    }
    // volume correction, a PI controller on the relative volume error
    // (macpressuresolver3.cpp:220-236); its integral state advances only when
    // there is a system to correct
    double rhs_correct = 0.0;
    if (c->prm.gain != 0.0 && c->target_volume != 0.0) {
        const double x = (c->current_volume - c->target_volume) / c->target_volume;
        const double y = c->y_prev + x * dt;
        c->y_prev = y;
        const double kp = c->prm.gain * 2.3 / (25.0 * 0.01);
        const double ki = kp * kp / 16.0;
        rhs_correct = -(kp * x + ki * y) / (x + 1.0);
        res->volume_correct_rhs = rhs_correct;
    }
    const bool need32 = c->prm.precision == 1 || c->prm.precond >= 2;
    const bool plain32 = need32 && c->prm.precond == 0;
    launch_coeff(c->stream, d, dt, c->dx, c->n_active, c->coord.p, area, rho, c->label.p, vel, svel, rhs_correct,
                 c->diag64.p,
                 c->cx64.p, c->cy64.p, c->cz64.p, c->b64.p, c->sig64.p, plain32 ? c->diag32.p : nullptr,
                 plain32 ? c->cx32.p : nullptr, plain32 ? c->cy32.p : nullptr, plain32 ? c->cz32.p : nullptr);
    if (c->prm.precond != 0) {
        const bool s64 = c->prm.precision == 0 && c->prm.precond == 1;
        const bool s32 = need32;
        launch_scale_coeff(c->stream, d.bx, c->n_active, c->nbr.p, c->sig64.p, c->cx64.p, c->cy64.p, c->cz64.p,
                           s64 ? c->hx64.p : nullptr, s64 ? c->hy64.p : nullptr, s64 ? c->hz64.p : nullptr,
                           s32 ? c->hx32.p : nullptr, s32 ? c->hy32.p : nullptr, s32 ? c->hz32.p : nullptr);
        if (s64) launch_brick_uniform<double>(c->stream, d.bx, c->n_active, c->nbr.p, c->hx64.p, c->hy64.p, c->hz64.p,
                                              c->kappa64.p);
        if (s32) launch_brick_uniform<float>(c->stream, d.bx, c->n_active, c->nbr.p, c->hx32.p, c->hy32.p, c->hz32.p,
                                             c->kappa32.p);
    } else if (c->prm.precision == 0) {
        launch_brick_uniform<double>(c->stream, d.bx, c->n_active, c->nbr.p, c->cx64.p, c->cy64.p, c->cz64.p,
                                     c->kappa64.p, c->diag64.p, c->kdiag64.p);
    } else {
        launch_brick_uniform<float>(c->stream, d.bx, c->n_active, c->nbr.p, c->cx32.p, c->cy32.p, c->cz32.p,
                                    c->kappa32.p, c->diag32.p, c->kdiag32.p);
    }
    if (c->prm.report) {
        // RCMatrix_utility::report(Lhs) (macpressuresolver3.cpp:236, RCMatrix_utility.h:98-146)
        const int blocks = std::min(c->n_active, 512);
        launch_report(c->stream, d.bx, c->n_active, blocks, c->nbr.p, c->diag64.p, c->cx64.p, c->cy64.p, c->cz64.p,
                      c->pc.p);
        HIPC(c, hipGetLastError());
        std::vector<double> part(8 * (size_t)blocks);
        HIPC(c, hipMemcpyAsync(part.data(), c->pc.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        double nnz = 0.0, rmin = 1e300, rmax = 0.0, dmin = 1e300, dmax = -1e300, ratio = 0.0, nan = 0.0;
        for (int b = 0; b < blocks; ++b) {
            const double *q = &part[8 * (size_t)b];
            if (q[7] == 0.0) continue;  // a block without rows
            nnz += q[0];
            rmin = std::min(rmin, q[1]);
            rmax = std::max(rmax, q[2]);
            dmin = std::min(dmin, q[3]);
            dmax = std::max(dmax, q[4]);
            ratio = std::max(ratio, q[5]);
            nan = std::max(nan, q[6]);
        }
        res->report_valid = 1;
        res->report_nnz = (int64_t)nnz;
        res->report_min_row = (int32_t)rmin;
        res->report_max_row = (int32_t)rmax;
        res->report_min_diag = dmin <= dmax ? dmin : 0.0;
        res->report_max_diag = dmin <= dmax ? dmax : 0.0;
        res->report_max_offdiag_ratio = ratio;
        // one coupling is stored per face and read by both of its rows
        res->report_symmetry_error = 0.0;
        res->report_has_nan = nan != 0.0 ? 1 : 0;
    }
    if (c->prm.warm_start) {
        // macpressuresolver3.cpp:238-247: solve for the change from the previous
        // pressure, rhs' = rhs - A p_prev.  The previous pressure is kept per cell
        // (zero where there was no unknown), not per row number.
        const bool fresh = c->prev64.p == nullptr;
        HIPC(c, c->prev64.ensure(G));
        if (fresh) HIPC(c, hipMemsetAsync(c->prev64.p, 0, sizeof(double) * G, c->stream));
        HIPC(c, c->warm64.ensure((size_t)c->n_active * kBrickCells));
        launch_gather_warm(c->stream, d, c->n_active, c->coord.p, c->label.p, c->prev64.p, c->warm64.p);
        launch_residual(c->stream, c->grid, op64(c), c->b64.p, c->warm64.p, c->r64.p, c->pc.p, c->misc.p);
        HIPC(c, hipMemcpyAsync(c->b64.p, c->r64.p, sizeof(double) * c->grid.n, hipMemcpyDeviceToDevice, c->stream));
    }
    HIPC(c, hipGetLastError());
    HIPC(c, hipEventRecord(c->ev[1], c->stream));

    if (c->prm.precond == 3) {
        rc = solve_mg(c, res);
    } else if (c->prm.precond == 2) {
        rc = unet_solve(c, res);
    } else if (c->prm.precision == 1) {
        rc = solve_mixed(c, res);
    } else {
        rc = solve_fp64(c, res);
    }
    if (rc) return rc;
    // macpressuresolver3.cpp:256-259: the solution is the previous pressure plus the change
    if (c->prm.warm_start) launch_add64(c->stream, c->grid.n, c->x64.p, c->warm64.p);
    c->have_system = true;
    HIPC(c, hipEventRecord(c->ev[2], c->stream));
    launch_scatter(c->stream, d, c->brick_slot.p, c->nbx, c->nby, c->label.p, c->x64.p, c->pressure.p,
                   c->prm.warm_start ? c->prev64.p : nullptr);
    launch_update_velocity(c->stream, d, dt, c->dx, c->pressure.p, area, rho, svel, dens, vel);
    HIPC(c, hipGetLastError());
    c->velocity_updated = true;
    HIPC(c, hipEventRecord(c->ev[3], c->stream));
    if (c->prm.host_pointers) {
        for (int a = 0; a < 3; ++a)
            if (user_vel[a])
                HIPC(c, hipMemcpyAsync(user_vel[a], vel[a], sizeof(float) * nf[a], hipMemcpyDeviceToHost, c->stream));
        c->h_pressure.resize(G);
        HIPC(c, hipMemcpyAsync(c->h_pressure.data(), c->pressure.p, sizeof(float) * G, hipMemcpyDeviceToHost,
                               c->stream));
    }
    HIPC(c, hipStreamSynchronize(c->stream));
    if ((rc = sync_and_time(c, 0, 1, res->ms_setup))) return rc;
    if ((rc = sync_and_time(c, 1, 2, res->ms_solve))) return rc;
    if ((rc = sync_and_time(c, 2, 3, res->ms_update))) return rc;
    return res->status;
}

}  // namespace

int shk_project(shk_ctx *c, double dt, float *u, float *v, float *w, const float *solid, const float *us,
                const float *vs, const float *ws, const float *fluid, double surface_tension, shk_result *res) {
    if (!c) return SHK_BAD_ARG;
    c->err.clear();
    const void *ptrs[8] = {u, v, w, solid, us, vs, ws, fluid};
    int rc = validate_fields(c, ptrs, 8);
    if (rc) return rc;
    if (c->dims.bx != 8) return c->fail(SHK_BAD_ARG, "shk_project needs a 3D context (use shk_project2d)");
    float *vel[3] = {u, v, w};
    const float *svel[3] = {us, vs, ws};
    return project_impl(c, dt, vel, svel, solid, fluid, surface_tension, res);
}

int shk_project2d(shk_ctx *c, double dt, float *u, float *v, const float *solid, const float *us, const float *vs,
                  const float *fluid, double surface_tension, shk_result *res) {
    if (!c) return SHK_BAD_ARG;
    c->err.clear();
    const void *ptrs[6] = {u, v, solid, us, vs, fluid};
    int rc = validate_fields(c, ptrs, 6);
    if (rc) return rc;
    if (c->dims.bx != 64) return c->fail(SHK_BAD_ARG, "shk_project2d needs a 2D context (shk_create2d)");
    float *vel[3] = {u, nullptr, v};
    const float *svel[3] = {us, nullptr, vs};
    return project_impl(c, dt, vel, svel, solid, fluid, surface_tension, res);
}

const float *shk_pressure(const shk_ctx *c) {
    if (!c) return nullptr;
    if (c->prm.host_pointers) return c->h_pressure.empty() ? nullptr : c->h_pressure.data();
    return c->pressure.p;
}

int shk_copy_pressure(shk_ctx *c, float *dst) {
    if (!c || !dst) return SHK_BAD_ARG;
    HIPC(c, hipSetDevice(c->device));
    const int64_t G = (int64_t)c->dims.nx * c->dims.ny * c->dims.nz;
    const hipMemcpyKind kind = c->prm.host_pointers ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    HIPC(c, hipMemcpyAsync(dst, c->pressure.p, sizeof(float) * G, kind, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return SHK_OK;
}

int shk_kernel_stats(shk_ctx *c, const char *kernel, int64_t *launches, double *total_ms) {
    if (!c || !kernel || !launches || !total_ms) return SHK_BAD_ARG;
    const std::string k(kernel);
    static const char *names[shk_ctx::kKernels] = {"axpy_norm", "spmv_dot"};
    if (k == "unet_apply") {
        *launches = c->ucount;
        *total_ms = c->ums;
        return SHK_OK;
    }
    int idx = -1;
    for (int q = 0; q < shk_ctx::kKernels; ++q)
        if (k == names[q]) idx = q;
    if (idx < 0)
        return c->fail(SHK_BAD_ARG, "unknown kernel '%s' (axpy_norm, spmv_dot, unet_apply)", kernel);
    *launches = c->pcount[idx];
    *total_ms = c->pms[idx];
    return SHK_OK;
}

int shk_export_fractions(shk_ctx *c, float *au, float *av, float *aw, float *ru, float *rv, float *rw) {
    if (!c) return SHK_BAD_ARG;
    HIPC(c, hipSetDevice(c->device));
    float *dst[6] = {au, av, aw, ru, rv, rw};
    if (c->dims.bx == 64) {
        // 2D: x faces, then the y faces (held as the z faces of nx x 1 x ny)
        if (aw || rw) return c->fail(SHK_BAD_ARG, "a 2D context exports x and y faces (third arrays NULL)");
        dst[1] = nullptr;
        dst[2] = av;
        dst[4] = nullptr;
        dst[5] = rv;
    }
    const hipMemcpyKind kind = c->prm.host_pointers ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    for (int i = 0; i < 6; ++i) {
        if (!dst[i]) continue;
        const int a = i % 3;
        const float *src = i < 3 ? c->area[a].p : c->rho[a].p;
        HIPC(c, hipMemcpyAsync(dst[i], src, sizeof(float) * face_count(c->dims, a), kind, c->stream));
    }
    HIPC(c, hipStreamSynchronize(c->stream));
    return SHK_OK;
}

int shk_export_system(shk_ctx *c, uint8_t *label, double *diag, double *cx, double *cy, double *cz, double *rhs,
                      double *x) {
    if (!c) return SHK_BAD_ARG;
    if (!c->have_system) return c->fail(SHK_BAD_ARG, "no linear system (no projection with unknowns yet)");
    if (c->prm.host_pointers) return c->fail(SHK_BAD_ARG, "shk_export_system takes device pointers");
    HIPC(c, hipSetDevice(c->device));
    launch_export(c->stream, c->dims, c->brick_slot.p, c->nbx, c->nby, c->label.p, c->diag64.p, c->cx64.p, c->cy64.p,
                  c->cz64.p, c->b64.p, c->x64.p, label, diag, cx, cy, cz, rhs, x);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    return SHK_OK;
}

int shk_divergence(shk_ctx *c, double dt, const float *u, const float *v, const float *w, const float *us,
                   const float *vs, const float *ws, int use_fp64_update, double *max_abs_div) {
    if (!c || !max_abs_div) return SHK_BAD_ARG;
    const bool two_d = c->dims.bx == 64;
    const void *ptrs[6] = {u, v, two_d ? v : w, us, vs, two_d ? vs : ws};
    int rc = validate_fields(c, ptrs, 6);
    if (rc) return rc;
    if (two_d && (w || ws)) return c->fail(SHK_BAD_ARG, "a 2D context takes x and y faces (third arrays NULL)");
    if (!c->have_system) return c->fail(SHK_BAD_ARG, "no linear system (no projection with unknowns yet)");
    if (c->prm.host_pointers) return c->fail(SHK_BAD_ARG, "shk_divergence takes device pointers");
    HIPC(c, hipSetDevice(c->device));
    const int maxb = 4096;
    HIPC(c, c->pa.ensure(maxb));
    float *area[3] = {c->area[0].p, c->area[1].p, c->area[2].p};
    float *rho[3] = {c->rho[0].p, c->rho[1].p, c->rho[2].p};
    const float *vel[3] = {u, v, w};
    const float *svel[3] = {us, vs, ws};
    if (two_d) {
        vel[1] = svel[1] = c->wall_faces.p;
        vel[2] = v;
        svel[2] = vs;
    }
    const int g = launch_divergence(c->stream, c->dims, dt, c->dx, area, rho, c->label.p, vel, svel, use_fp64_update,
                                    c->brick_slot.p, c->nbx, c->nby, c->x64.p, c->pa.p, maxb);
    HIPC(c, hipGetLastError());
    std::vector<double> part(g);
    HIPC(c, hipMemcpyAsync(part.data(), c->pa.p, sizeof(double) * g, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    double m = 0.0;
    for (double p : part) m = std::max(m, p);
    *max_abs_div = m;
    return SHK_OK;
}

}  // extern "C"

namespace {

// shk_solve_csr on device pointers
int solve_csr_device(shk_ctx *c, int64_t rows, const int64_t *rowptr, const int64_t *cols, const double *vals,
                     const double *b, double *x, shk_result *res) {
    shk_result local{};
    if (!res) res = &local;
    *res = shk_result{};
    res->unknowns = rows;
    res->outer = 1;
    if (c->prm.precond == 2)
        return c->fail(SHK_BAD_ARG, "Preconditioner=unet needs the projection's grid; shk_solve_csr takes none, "
                                    "jacobi or mg");
    if (rows == 0) return SHK_OK;
    int rc = ensure_partials(c);
    if (rc) return rc;
    HIPC(c, c->diag64.ensure(rows));
    HIPC(c, c->sig64.ensure(rows));
    HIPC(c, c->r64.ensure(rows));
    HIPC(c, c->s64.ensure(rows));
    HIPC(c, c->t64.ensure(rows));
    HIPC(c, c->q64.ensure(rows));
    const CsrOp op{rowptr, cols, vals, rows};
    const bool jac = c->prm.precond == 1;
    // Preconditioner=mg: CG with the aggregation-multigrid V-cycle, the hierarchy built
    // from this matrix (the reference's amg plugin builds its own per solve too)
    const bool amg = c->prm.precond == 3;
    const Grid g = make_grid(rows);
    HIPC(c, hipEventRecord(c->ev[0], c->stream));
    if (amg && (rc = amg_setup(c, rows, rowptr, cols, vals))) return rc;
    launch_csr_diag(c->stream, op, c->diag64.p, c->sig64.p);
    CsrOp sop = op;
    if (jac) {
        HIPC(c, c->cx64.ensure((size_t)0));
        // scaled copy of the values: the Jacobi-PCG iterates as CG on D^-1/2 A D^-1/2
        int64_t nnz = 0;
        HIPC(c, hipMemcpy(&nnz, rowptr + rows, sizeof(int64_t), hipMemcpyDeviceToHost));
        if (nnz < 0) return c->fail(SHK_BAD_ARG, "rowptr[rows] is negative");
        HIPC(c, c->q64.ensure(std::max<int64_t>(rows, 1)));
        HIPC(c, c->hx64.ensure(std::max<int64_t>(nnz, 1)));
        launch_csr_scale(c->stream, op, c->sig64.p, c->hx64.p);
        sop.vals = c->hx64.p;
    }
    const double *sig = jac ? c->sig64.p : nullptr;
    Reduction red = reduction(c);
    if (amg) red.pz = amg_pz(c->amg);
    // the vector the next direction is formed from: r, or z = M r
    const double *dsrc = amg ? amg_z(c->amg) : c->r64.p;
    launch_cg_init<double, double>(c->stream, g, b, sig, x, c->r64.p, c->s64.p, red, c->prm.residual,
                                   c->prm.max_iterations);
    double *S[2] = {c->s64.p, c->t64.p};
    int m = 0;
    if (amg && (rc = amg_vcycle(c, c->r64.p, g, c->state.p))) return rc;
    launch_dir_spmv_csr(c->stream, g, sop, dsrc, S[0], S[1], c->q64.p, red);
    ++m;
    rc = drive(
        c,
        [&]() {
            launch_axpy_norm<double>(c->stream, g, sig, x, c->r64.p, S[m & 1], c->q64.p, red, c->next_events());
            if (amg) (void)amg_vcycle(c, c->r64.p, g, c->state.p);
            launch_dir_spmv_csr(c->stream, g, sop, dsrc, S[m & 1], S[(m + 1) & 1], c->q64.p, red,
                                c->next_events());
            ++m;
        },
        amg ? std::min(std::max(1, c->prm.check_interval), 4) : 0);
    if (rc) return rc;
    if (jac) launch_unscale(c->stream, rows, x, c->sig64.p);
    HIPC(c, hipEventRecord(c->ev[1], c->stream));
    res->count = c->h_state->iter;
    res->status = c->h_state->status;
    res->reresid = c->h_state->rnorm0 > 0 ? c->h_state->rnorm / c->h_state->rnorm0 : c->h_state->rnorm;
    launch_residual_csr(c->stream, g, op, b, x, nullptr, c->pc.p, c->misc.p);
    HIPC(c, hipGetLastError());
    HIPC(c, hipMemcpyAsync(c->h_misc, c->misc.p, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    double rr[4] = {c->h_misc[0], c->h_misc[1], c->h_misc[2], c->h_misc[3]};
    fill_true_residual(res, rr);
    if ((rc = sync_and_time(c, 0, 1, res->ms_solve))) return rc;
    c->have_system = false;
    return res->status;
}

}  // namespace

extern "C" {

int shk_solve_csr(shk_ctx *c, int64_t rows, const int64_t *rowptr, const int64_t *cols, const double *vals,
                  const double *b, double *x, shk_result *res) {
    if (!c) return SHK_BAD_ARG;
    c->err.clear();
    if (rows < 0 || rows > ((int64_t)1 << 40)) return c->fail(SHK_BAD_ARG, "bad row count");
    if (!rowptr || !cols || !vals || !b || !x) return c->fail(SHK_BAD_ARG, "NULL CSR/vector pointer");
    HIPC(c, hipSetDevice(c->device));
    if (!c->prm.host_pointers) return solve_csr_device(c, rows, rowptr, cols, vals, b, x, res);
    // host arrays: stage the matrix and rhs in HBM, solve, copy x back (the copies are
    // part of the call, as for shk_project with HostPointers=1)
    const int64_t nnz = rowptr[rows];
    if (nnz < 0 || rowptr[0] != 0) return c->fail(SHK_BAD_ARG, "rowptr must start at 0 and end at nnz >= 0");
    HIPC(c, c->csr_i[0].ensure((size_t)rows + 1));
    HIPC(c, c->csr_i[1].ensure((size_t)std::max<int64_t>(nnz, 1)));
    HIPC(c, c->csr_d[0].ensure((size_t)std::max<int64_t>(nnz, 1)));
    HIPC(c, c->csr_d[1].ensure((size_t)std::max<int64_t>(rows, 1)));
    HIPC(c, c->csr_d[2].ensure((size_t)std::max<int64_t>(rows, 1)));
    HIPC(c, hipMemcpyAsync(c->csr_i[0].p, rowptr, sizeof(int64_t) * (rows + 1), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(c->csr_i[1].p, cols, sizeof(int64_t) * nnz, hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(c->csr_d[0].p, vals, sizeof(double) * nnz, hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(c->csr_d[1].p, b, sizeof(double) * rows, hipMemcpyHostToDevice, c->stream));
    const int rc = solve_csr_device(c, rows, c->csr_i[0].p, c->csr_i[1].p, c->csr_d[0].p, c->csr_d[1].p,
                                    c->csr_d[2].p, res);
    if (rc != SHK_OK && rc != SHK_NOT_CONVERGED && rc != SHK_NAN) return rc;
    HIPC(c, hipMemcpyAsync(x, c->csr_d[2].p, sizeof(double) * rows, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return rc;
}

}  // extern "C"
