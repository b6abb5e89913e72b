// ctx_internal.h -- the context behind the C-ABI handle, shared by ctx.hip and
// the learned preconditioner (unet.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "krylov.h"
#include "shk_common.h"
#include "shk_fractions.h"
#include "shk_projection.h"

namespace shk {
struct UnetPrecond;
struct MgHierarchy;
struct LiquidWork;
struct CsrAmg;
struct RedistWork;
}

namespace shk {

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t want) {
        if (want <= n && p) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            n = 0;
            if (e != hipSuccess) return e;
        }
        if (want == 0) want = 1;
        hipError_t e = hipMalloc((void **)&p, sizeof(T) * want);
        if (e == hipSuccess) n = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct Params {
    double residual = 1e-4;           // pcg.cpp:76
    int64_t max_iterations = 30000;   // pcg.cpp:77
    double modified_ic = 0.97;        // pcg.cpp:78
    double min_diag_ratio = 0.25;     // pcg.cpp:79
    bool so_fluid = true;             // macpressuresolver3.cpp:352
    bool so_solid = true;             // macpressuresolver3.cpp:353
    double gain = 1.0;                // macpressuresolver3.cpp:350
    double eps_fluid = 1e-2;          // macutility3.cpp:522
    double eps_solid = 1e-2;          // macutility3.cpp:523
    int precision = 0;                // 0 fp64, 1 mixed
    int precond = 1;                  // 0 none, 1 jacobi, 2 unet, 3 mg
    int outer_max = 3;
    double inner_residual = 1e-6;
    int check_interval = 32;
    bool host_pointers = false;
    std::string unet_weights;
    int unet_smooth = 10;             // smoothing steps after each network application in FCG
    double unet_cheb = 16.0;          // > 1: Chebyshev step weights on [2 / unet_cheb, 2]; 0: omega
    double mg_omega = 0.9;            // multigrid damped-Jacobi weight
    double mg_alpha = 1.5;            // multigrid coarse-correction factor
    bool warm_start = false;          // macpressuresolver3.cpp:354
    bool report = false;              // RCMatrix_utility::report of each assembled system
    bool maccormack = true;           // macadvection3.cpp:281
    bool weno = false;                // macadvection3.cpp:286 (WENO interpolation)
    int trim_narrowband = 1;          // macadvection3.cpp:283
    double redist_rate = 0.75;        // pderedistancer3.cpp:206 (IntegrationRate)
    bool redist_rk2 = false;          // pderedistancer3.cpp:207 (RedistTemporalScheme Euler)
    bool solid_wall_extrap = true;    // gridutility3.cpp:394, 413 (SolidWallExtrapolation)
    bool horizontal_extrap = true;    // gridutility3.cpp:395, 414 (HorizontalSolidWallExtrapolation)
};

}  // namespace shk

struct shk_ctx {
    using Dims = shk::Dims;
    using Params = shk::Params;
    template <class T>
    using DevBuf = shk::DevBuf<T>;
    using SolveState = shk::SolveState;

    int device = 0;
    hipStream_t stream = nullptr;
    Dims dims{0, 0, 0};
    double dx = 0.0;
    int nbx = 0, nby = 0, nbz = 0;
    Params prm;
    std::string err;

    // dense workspace
    DevBuf<float> area[3], rho[3];
    DevBuf<uint8_t> label;
    DevBuf<float> pressure;
    DevBuf<uint8_t> brick_flag;
    DevBuf<int> brick_slot;
    DevBuf<unsigned long long> count;
    // host-pointer staging (0..7 the projection's fields, 8..9 spare, 10..12 solid density)
    DevBuf<float> stage[13];
    // project() branches: surface-tension curvature, solid density, volume
    // correction state (macpressuresolver3.cpp:46-53, 365-371), warm start
    DevBuf<float> curvature;
    const float *density[3] = {nullptr, nullptr, nullptr};
    double current_volume = 0.0, target_volume = 0.0, y_prev = 0.0;
    DevBuf<double> prev64;   // dense FP64 pressure of the last warm-started projection
    DevBuf<double> warm64;   // that pressure gathered onto the current bricks
    DevBuf<float> wall_faces;  // 2D contexts: zero y-face arrays of the nx x 1 x ny layout
    std::vector<float> h_pressure;
    // bricks
    int n_active = 0;
    int64_t unknowns = 0;
    bool have_system = false;
    std::vector<int4> h_coord;
    std::vector<int> h_nbr, h_slot;
    std::vector<uint8_t> h_flag;
    DevBuf<int4> coord;
    DevBuf<int> nbr;
    // unscaled FP64 operator (diag, couplings zero off unknown pairs), rhs, sig = 1/sqrt(diag)
    DevBuf<double> diag64, cx64, cy64, cz64, b64, sig64;
    // couplings of the Jacobi-scaled operator, FP64 and FP32
    DevBuf<double> hx64, hy64, hz64, kappa64;
    DevBuf<float> hx32, hy32, hz32, kappa32;
    // unscaled FP32 operator (mixed precision without preconditioner)
    DevBuf<float> diag32, cx32, cy32, cz32;
    // plain CG: per-brick constant diagonal of uniform bricks (with kappa64 / kappa32)
    DevBuf<double> kdiag64;
    DevBuf<float> kdiag32;
    // Krylov vectors
    DevBuf<double> x64, r64, s64, t64, q64;
    DevBuf<float> x32, r32, s32, t32, q32, rs32;
    DevBuf<double> pa, pb, pc, misc;
    // shk_solve_csr with HostPointers=1: device copies of rowptr, cols and of vals, b, x
    DevBuf<int64_t> csr_i[2];
    DevBuf<double> csr_d[3];
    shk::Grid grid;
    DevBuf<SolveState> state;
    SolveState *h_state = nullptr;
    double *h_misc = nullptr;
    hipEvent_t ev[6] = {};
    // learned preconditioner
    shk::UnetPrecond *unet = nullptr;
    // multigrid preconditioner hierarchy
    shk::MgHierarchy *mg = nullptr;
    // aggregation multigrid of the last shk_solve_csr with Preconditioner=mg
    shk::CsrAmg *amg = nullptr;
    // work arrays of the PDE redistancer (redist.hip)
    shk::RedistWork *redist = nullptr;
    // scratch of the liquid-step kernels (liquid.hip)
    shk::LiquidWork *liquid = nullptr;
    // the last projection ran its velocity update (it had unknowns)
    bool velocity_updated = false;
    // per-kernel timing of the Krylov loop (ProfileKernels=1)
    bool profile = false;
    std::vector<hipEvent_t> pev;
    size_t pcur = 0;
    static constexpr int kKernels = 2;  // launches per Krylov iteration: axpy_norm, dir_spmv
    int64_t pcount[kKernels] = {0, 0};
    double pms[kKernels] = {0.0, 0.0};
    // network applications inside flexible CG (ProfileKernels=1)
    hipEvent_t uev[2] = {};
    int64_t ucount = 0;
    double ums = 0.0;
    // the start/stop events of the next timed Krylov launch (none when not profiling)
    shk::KEvents next_events() {
        if (!profile || pcur + 2 > pev.size()) return shk::KEvents{};
        const shk::KEvents e{pev[pcur], pev[pcur + 1]};
        pcur += 2;
        return e;
    }

    int fail(int code, const char *fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
};

#define HIPC(ctx, expr)                                                                                     \
    do {                                                                                                    \
        hipError_t _e = (expr);                                                                             \
        if (_e != hipSuccess)                                                                               \
            return (ctx)->fail(_e == hipErrorOutOfMemory ? SHK_OUT_OF_MEMORY : SHK_HIP_ERROR, "%s: %s (%s:%d)", \
                               #expr, hipGetErrorString(_e), __FILE__, __LINE__);                           \
    } while (0)

namespace shk {

// Iterates `iterate` CheckInterval times between polls of the device state.
// Runs iterate() in batches of `interval` launches-worth of iterations (default:
// CheckInterval) until the device state says done.
int drive_impl(shk_ctx *c, void (*iterate)(void *), void *arg, int interval);
template <class F>
int drive(shk_ctx *c, F &&f, int interval = 0) {
    return drive_impl(c, [](void *p) { (*static_cast<F *>(p))(); }, &f, interval);
}
Reduction reduction(shk_ctx *c);
BrickOp<double> op64(shk_ctx *c);
BrickOp<float> op32(shk_ctx *c);
// r = b - A x (FP64); out = {max|r|, sum r^2, sum b^2, max|b|}; optionally the
// scaled FP32 copy rs = (float)(sig*r) for the next inner solve
int true_residual(shk_ctx *c, const double *x, double *r_out, double out[4], const double *sig = nullptr,
                  float *rs = nullptr);
void fill_true_residual(shk_result *res, const double rr[4]);
int solve_fp64(shk_ctx *c, shk_result *res);

// allocation-free CSR solve (noalloc.hip): looks up, once per device and with that
// device current, the address of the device's static scratch array
hipError_t scratch_bind(int device);
size_t scratch_capacity_bytes();

// learned preconditioner (unet.hip)
void unet_destroy(UnetPrecond *u);
void mg_destroy(MgHierarchy *h);
// aggregation multigrid for CSR input (amg.hip): setup from the device CSR arrays,
// then z = M r per V-cycle into c->amg (z, and the r.z partials of the CG grid g)
void amg_destroy(CsrAmg *h);
void redist_destroy(RedistWork *w);
int amg_setup(shk_ctx *c, int64_t rows, const int64_t *rowptr, const int64_t *cols, const double *vals);
int amg_vcycle(shk_ctx *c, const double *r, const Grid &g, const SolveState *state);
const double *amg_z(const CsrAmg *h);
double *amg_pz(CsrAmg *h);
void liquid_destroy(LiquidWork *w);
int solve_mg(shk_ctx *c, shk_result *res);
int unet_solve(shk_ctx *c, shk_result *res);
int export_scaled(shk_ctx *c, float *rhat, float *mask, float *hx, float *hy, float *hz, float *xhat);

}  // namespace shk
