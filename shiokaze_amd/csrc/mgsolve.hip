// mgsolve.hip -- Preconditioner=mg: CG preconditioned by the aggregation multigrid
// V-cycle of mg.hip, in the mixed-precision refinement loop (FP32 inner PCG with
// FP64 dots, FP64 true residual and correction).
//
// It takes the place of the reference's "amg" linear-solver plugin
// (src/linsolver/amg.cpp:43-166, smoothed-aggregation AMG from the vendored amgcl):
// an algebraic-multigrid-preconditioned CG behind the same solve() contract, here
// with geometric 2x2x2 aggregates (2x1x2 on a 2D grid) that the brick layout makes
// free to build.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ctx_internal.h"
#include "mg.h"

namespace shk {

struct MgLevelHost {
    MgLevelDev v;
    DevBuf<int> slot, nbr, parent;
    DevBuf<int4> coord;
    DevBuf<float> d, cx, cy, cz, b, x, t;
    std::vector<int> h_slot, h_nbr, h_parent;
    std::vector<int4> h_coord;
};

struct MgHierarchy {
    std::vector<MgLevelHost> lv;
    DevBuf<float> z;          // the preconditioned residual z = M r (level 0 bricks)
    DevBuf<double> pz;        // r.z partials (one per Krylov workgroup)
    DevBuf<double> pcell;     // r.z sums of the post-smoothing blocks (one per 256 cells)
    DevBuf<int> cidx;         // coarsest: active cell indices
    DevBuf<float> cinv;       // coarsest: dense inverse on them
    int ncoarse = 0;
    std::vector<int> key;     // level-0 brick table the host tables were built for
    int key_dims[3] = {0, 0, 0};
};

namespace {

void release(MgLevelHost &L) {
    for (DevBuf<int> *b : {&L.slot, &L.nbr, &L.parent}) b->release();
    L.coord.release();
    for (DevBuf<float> *b : {&L.d, &L.cx, &L.cy, &L.cz, &L.b, &L.x, &L.t}) b->release();
}

const int kOff[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};

void tables(MgLevelHost &L) {
    L.h_nbr.assign(6 * (size_t)std::max(L.v.n_active, 1), -1);
    for (int s = 0; s < L.v.n_active; ++s) {
        const int4 q = L.h_coord[s];
        for (int f = 0; f < 6; ++f) {
            const int x = q.x + kOff[f][0], y = q.y + kOff[f][1], z = q.z + kOff[f][2];
            if (x < 0 || y < 0 || z < 0 || x >= L.v.nbx || y >= L.v.nby || z >= L.v.nbz) continue;
            L.h_nbr[6 * s + f] = L.h_slot[x + (int64_t)L.v.nbx * (y + (int64_t)L.v.nby * z)];
        }
    }
}

// host brick tables of every coarse level (and the parent links), rebuilt only when
// the level-0 brick set changes
int build_tables(shk_ctx *c, MgHierarchy *h) {
    const bool same = h->key == c->h_slot && h->key_dims[0] == c->dims.nx && h->key_dims[1] == c->dims.ny &&
                      h->key_dims[2] == c->dims.nz && !h->lv.empty();
    MgLevelHost &L0 = h->lv.empty() ? (h->lv.emplace_back(), h->lv[0]) : h->lv[0];
    L0.v.nx = c->dims.nx;
    L0.v.ny = c->dims.ny;
    L0.v.nz = c->dims.nz;
    L0.v.nbx = c->nbx;
    L0.v.nby = c->nby;
    L0.v.nbz = c->nbz;
    L0.v.n_active = c->n_active;
    // brick extents (8x8x8, or 64x1x8 on a 2D grid) and the y aggregation factor
    const int bx = c->dims.bx, ex = bx, ey = kWave / bx, ez = kBrickCells / kWave;
    const int ay = ey > 1 ? 2 : 1;
    for (MgLevelHost &L : h->lv) L.v.bx = bx;
    if (same) return SHK_OK;
    for (size_t l = 1; l < h->lv.size(); ++l) release(h->lv[l]);
    h->lv.resize(1);
    L0.h_slot = c->h_slot;
    L0.h_coord = c->h_coord;
    // coarsen until every axis has at most 4 cells (at least one coarse level)
    while (true) {
        const MgLevelHost &F = h->lv.back();
        if (h->lv.size() > 1 && std::max({F.v.nx, F.v.ny, F.v.nz}) <= 4) break;
        MgLevelHost C;
        C.v.bx = bx;
        C.v.nx = (F.v.nx + 1) / 2;
        C.v.ny = (F.v.ny + ay - 1) / ay;
        C.v.nz = (F.v.nz + 1) / 2;
        C.v.nbx = (C.v.nx + ex - 1) / ex;
        C.v.nby = (C.v.ny + ey - 1) / ey;
        C.v.nbz = (C.v.nz + ez - 1) / ez;
        const int64_t nb = (int64_t)C.v.nbx * C.v.nby * C.v.nbz;
        C.h_slot.assign(nb, -1);
        // a coarse brick covers 2 x ay x 2 fine bricks
        for (int cz = 0; cz < C.v.nbz; ++cz)
            for (int cy = 0; cy < C.v.nby; ++cy)
                for (int cx = 0; cx < C.v.nbx; ++cx) {
                    bool any = false;
                    for (int q = 0; q < 4 * ay && !any; ++q) {
                        const int fx = 2 * cx + (q & 1), fy = ay * cy + (ay > 1 ? (q >> 1) & 1 : 0),
                                  fz = 2 * cz + (ay > 1 ? q >> 2 : q >> 1);
                        if (fx < F.v.nbx && fy < F.v.nby && fz < F.v.nbz)
                            any = F.h_slot[fx + (int64_t)F.v.nbx * (fy + (int64_t)F.v.nby * fz)] >= 0;
                    }
                    if (any) {
                        C.h_slot[cx + (int64_t)C.v.nbx * (cy + (int64_t)C.v.nby * cz)] = (int)C.h_coord.size();
                        C.h_coord.push_back(make_int4(cx, cy, cz, 0));
                    }
                }
        C.v.n_active = (int)C.h_coord.size();
        h->lv.push_back(std::move(C));
    }
    for (size_t l = 0; l < h->lv.size(); ++l) {
        MgLevelHost &L = h->lv[l];
        if (l > 0) tables(L);
        if (l + 1 < h->lv.size()) {
            const MgLevelHost &C = h->lv[l + 1];
            L.h_parent.assign(std::max(L.v.n_active, 1), -1);
            for (int s = 0; s < L.v.n_active; ++s) {
                const int4 q = L.h_coord[s];
                L.h_parent[s] = C.h_slot[q.x / 2 + (int64_t)C.v.nbx * (q.y / ay + (int64_t)C.v.nby * (q.z / 2))];
            }
        }
    }
    h->key = c->h_slot;
    h->key_dims[0] = c->dims.nx;
    h->key_dims[1] = c->dims.ny;
    h->key_dims[2] = c->dims.nz;
    // device copies
    hipStream_t st = c->stream;
    for (size_t l = 0; l < h->lv.size(); ++l) {
        MgLevelHost &L = h->lv[l];
        const size_t n = (size_t)std::max(L.v.n_active, 1) * kBrickCells;
        if (l > 0) {
            HIPC(c, L.slot.ensure(L.h_slot.size()));
            HIPC(c, L.coord.ensure(std::max(L.v.n_active, 1)));
            HIPC(c, L.nbr.ensure(L.h_nbr.size()));
            HIPC(c, hipMemcpyAsync(L.slot.p, L.h_slot.data(), sizeof(int) * L.h_slot.size(), hipMemcpyHostToDevice,
                                   st));
            if (L.v.n_active)
                HIPC(c, hipMemcpyAsync(L.coord.p, L.h_coord.data(), sizeof(int4) * L.v.n_active,
                                       hipMemcpyHostToDevice, st));
            HIPC(c, hipMemcpyAsync(L.nbr.p, L.h_nbr.data(), sizeof(int) * L.h_nbr.size(), hipMemcpyHostToDevice, st));
            for (DevBuf<float> *b : {&L.d, &L.cx, &L.cy, &L.cz}) HIPC(c, b->ensure(n));
        } else {
            HIPC(c, L.d.ensure(n));
        }
        if (l > 0)  // level 0 keeps no vectors: its residual is restricted where it is computed
            for (DevBuf<float> *b : {&L.b, &L.x, &L.t}) HIPC(c, b->ensure(n));
        if (!L.h_parent.empty()) {
            HIPC(c, L.parent.ensure(L.h_parent.size()));
            HIPC(c, hipMemcpyAsync(L.parent.p, L.h_parent.data(), sizeof(int) * L.h_parent.size(),
                                   hipMemcpyHostToDevice, st));
        }
    }
    // the host vectors must outlive the asynchronous copies
    HIPC(c, hipStreamSynchronize(st));
    return SHK_OK;
}

// dense inverse (Cholesky, FP64) of the coarsest operator on its active cells;
// the diagonal inverse when that operator is not positive definite
void coarse_inverse(int bx, const std::vector<float> &d, const std::vector<float> &cx, const std::vector<float> &cy,
                    const std::vector<float> &cz, std::vector<int> &idx, std::vector<float> &inv) {
    const int ex = bx, ey = kWave / bx;
    idx.clear();
    for (int e = 0; e < kBrickCells; ++e)
        if (d[e] > 0.0f) idx.push_back(e);
    const int n = (int)idx.size();
    std::vector<int> pos(kBrickCells, -1);
    for (int i = 0; i < n; ++i) pos[idx[i]] = i;
    std::vector<double> A((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        const int e = idx[i];
        A[(size_t)i * n + i] = d[e];
        const int x = e % ex, y = (e / ex) % ey, z = e / kWave;
        const int lo[3] = {x > 0 ? e - 1 : -1, y > 0 ? e - ex : -1, z > 0 ? e - kWave : -1};
        const float cv[3] = {cx[e], cy[e], cz[e]};
        for (int a = 0; a < 3; ++a) {
            if (lo[a] < 0 || pos[lo[a]] < 0 || cv[a] == 0.0f) continue;
            const int j = pos[lo[a]];
            A[(size_t)i * n + j] -= cv[a];
            A[(size_t)j * n + i] -= cv[a];
        }
    }
    // Cholesky A = L L'
    std::vector<double> Lm((size_t)n * n, 0.0);
    bool ok = true;
    for (int j = 0; j < n && ok; ++j) {
        double s = A[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) s -= Lm[(size_t)j * n + k] * Lm[(size_t)j * n + k];
        if (!(s > 0.0)) {
            ok = false;
            break;
        }
        const double ljj = std::sqrt(s);
        Lm[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; ++i) {
            double t = A[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) t -= Lm[(size_t)i * n + k] * Lm[(size_t)j * n + k];
            Lm[(size_t)i * n + j] = t / ljj;
        }
    }
    inv.assign((size_t)n * n, 0.0f);
    if (!ok) {
        for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = (float)(1.0 / A[(size_t)i * n + i]);
        return;
    }
    // columns of A^-1 by forward and back substitution
    std::vector<double> y(n), xcol(n);
    for (int col = 0; col < n; ++col) {
        for (int i = 0; i < n; ++i) {
            double s = (i == col) ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) s -= Lm[(size_t)i * n + k] * y[k];
            y[i] = s / Lm[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = y[i];
            for (int k = i + 1; k < n; ++k) s -= Lm[(size_t)k * n + i] * xcol[k];
            xcol[i] = s / Lm[(size_t)i * n + i];
        }
        for (int i = 0; i < n; ++i) inv[(size_t)i * n + col] = (float)xcol[i];
    }
}

}  // namespace

void mg_destroy(MgHierarchy *h) {
    if (!h) return;
    for (MgLevelHost &L : h->lv) release(L);
    h->z.release();
    h->pz.release();
    h->pcell.release();
    h->cidx.release();
    h->cinv.release();
    delete h;
}

// Builds (or refreshes) the hierarchy for the current system: host tables when the
// brick set changed, the coarse operators every time, and the coarsest inverse.
int mg_setup(shk_ctx *c) {
    if (!c->mg) c->mg = new MgHierarchy();
    MgHierarchy *h = c->mg;
    int rc;
    if ((rc = build_tables(c, h))) return rc;
    hipStream_t st = c->stream;
    MgLevelHost &L0 = h->lv[0];
    const int64_t n0 = (int64_t)c->n_active * kBrickCells;
    HIPC(c, h->z.ensure((size_t)std::max<int64_t>(n0, 1)));
    HIPC(c, h->pz.ensure((size_t)kMaxBlocks + 64));
    HIPC(c, h->pcell.ensure((size_t)std::max<int64_t>((n0 + 255) / 256, 1)));
    launch_mg_mask(st, n0, c->sig64.p, L0.d.p);
    for (size_t l = 0; l < h->lv.size(); ++l) {
        MgLevelHost &L = h->lv[l];
        L.v.slot = l ? L.slot.p : c->brick_slot.p;
        L.v.coord = l ? L.coord.p : c->coord.p;
        L.v.nbr = l ? L.nbr.p : c->nbr.p;
        L.v.parent = L.parent.p;
        L.v.d = L.d.p;
        L.v.cx = l ? L.cx.p : c->hx32.p;
        L.v.cy = l ? L.cy.p : c->hy32.p;
        L.v.cz = l ? L.cz.p : c->hz32.p;
        L.v.b = L.b.p;
        L.v.x = L.x.p;
        L.v.t = L.t.p;
    }
    for (size_t l = 1; l < h->lv.size(); ++l)
        if (h->lv[l].v.n_active) launch_mg_coarsen(st, h->lv[l - 1].v, h->lv[l].v);
    HIPC(c, hipGetLastError());
    // coarsest level: one brick at most
    MgLevelHost &C = h->lv.back();
    std::vector<float> hd(kBrickCells, 0.0f), hx(kBrickCells, 0.0f), hy(kBrickCells, 0.0f), hz(kBrickCells, 0.0f);
    if (C.v.n_active >= 1) {
        HIPC(c, hipMemcpyAsync(hd.data(), C.d.p, sizeof(float) * kBrickCells, hipMemcpyDeviceToHost, st));
        HIPC(c, hipMemcpyAsync(hx.data(), C.cx.p, sizeof(float) * kBrickCells, hipMemcpyDeviceToHost, st));
        HIPC(c, hipMemcpyAsync(hy.data(), C.cy.p, sizeof(float) * kBrickCells, hipMemcpyDeviceToHost, st));
        HIPC(c, hipMemcpyAsync(hz.data(), C.cz.p, sizeof(float) * kBrickCells, hipMemcpyDeviceToHost, st));
    }
    HIPC(c, hipStreamSynchronize(st));
    std::vector<int> idx;
    std::vector<float> inv;
    coarse_inverse(c->dims.bx, hd, hx, hy, hz, idx, inv);
    if ((int)idx.size() > kMgCoarseMax) return c->fail(SHK_HIP_ERROR, "multigrid: coarsest level too large");
    h->ncoarse = (int)idx.size();
    HIPC(c, h->cidx.ensure(std::max<size_t>(idx.size(), 1)));
    HIPC(c, h->cinv.ensure(std::max<size_t>(inv.size(), 1)));
    if (!idx.empty()) {
        HIPC(c, hipMemcpyAsync(h->cidx.p, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, st));
        HIPC(c, hipMemcpyAsync(h->cinv.p, inv.data(), sizeof(float) * inv.size(), hipMemcpyHostToDevice, st));
    }
    HIPC(c, hipStreamSynchronize(st));
    return SHK_OK;
}

// z = M r on level 0 (brick layout).  The last kernel runs on the Krylov grid g and
// writes the r.z partials into part.
void mg_vcycle(shk_ctx *c, const float *r, float *z, const SolveState *state, const Grid *g, double *part) {
    MgHierarchy *h = c->mg;
    hipStream_t st = c->stream;
    const int L = (int)h->lv.size() - 1;
    MgLevelHost &L0 = h->lv[0];
    const float w = (float)c->prm.mg_omega, a = (float)c->prm.mg_alpha;
    const MgOp op0{c->nbr.p, L0.d.p, c->hx32.p, c->hy32.p, c->hz32.p, c->n_active, w, a, c->kappa32.p, c->dims.bx};
    // down: level 1 from the level-0 residual in the pre-smoothing pass, every later
    // level from the residual of the level above, computed where it is restricted
    // (x_L = 0: w = 0)
    launch_mg_pre0(st, op0, r, c->coord.p, L0.parent.p, h->lv[1].v, L > 1 ? w : 0.0f, state);
    for (int l = 1; l < L; ++l) {
        MgLevelHost &F = h->lv[l], &C = h->lv[l + 1];
        if (!C.v.n_active) continue;
        launch_mg_residual_restrict(st, F.v, C.v, l + 1 < L ? w : 0.0f, state);
    }
    MgLevelHost &Cl = h->lv[L];
    if (Cl.v.n_active) launch_mg_coarse(st, h->ncoarse, h->cidx.p, h->cinv.p, Cl.v.b, Cl.v.x, kBrickCells, state);
    // up: the correction from the coarsest level, then the post-smoothing of each level,
    // which also applies the correction of the level above it (level 0's is applied
    // in its post-smoothing)
    if (L > 1 && h->lv[L - 1].v.n_active) launch_mg_prolong(st, h->lv[L - 1].v, Cl.v, a, state);
    for (int l = L - 1; l >= 1; --l) {
        MgLevelHost &F = h->lv[l];
        if (!F.v.n_active) continue;
        launch_mg_smooth(st, F.v, l > 1 ? &h->lv[l - 1].v : nullptr, w, a, state);
        std::swap(F.v.x, F.v.t);
        std::swap(F.x.p, F.t.p);
    }
    launch_mg_post0(st, op0, c->coord.p, L0.parent.p, r, h->lv[1].v.x, z, state, *g, h->pcell.p, part);
}

// Mixed-precision refinement loop with MG-preconditioned inner CG.
int solve_mg(shk_ctx *c, shk_result *res) {
    int rc;
    if ((rc = mg_setup(c))) return rc;
    MgHierarchy *h = c->mg;
    const Grid &g = c->grid;
    const double *sig = c->sig64.p;
    HIPC(c, hipMemsetAsync(c->x64.p, 0, sizeof(double) * g.n, c->stream));
    double rr[4];
    if ((rc = true_residual(c, c->x64.p, c->r64.p, rr, sig, c->rs32.p))) return rc;
    const double b_inf = rr[3];
    const double tol_abs = std::max(c->prm.residual, 1e-30) * b_inf;
    double r_inf = rr[0];
    int64_t total = 0;
    int outer = 0;
    int status = SHK_OK;
    const BrickOp<float> op = op32(c);
    Reduction red = reduction(c);
    red.pz = h->pz.p;
    const int interval = std::min(std::max(1, c->prm.check_interval), 4);
    while (r_inf > tol_abs && outer < std::max(1, c->prm.outer_max)) {
        double inner_rel = std::max(c->prm.inner_residual, 0.5 * tol_abs / r_inf);
        if (inner_rel > 0.5) inner_rel = 0.5;
        launch_cg_init<float, float>(c->stream, g, c->rs32.p, nullptr, c->x32.p, c->r32.p, c->s32.p, red, inner_rel,
                                     c->prm.max_iterations - total);
        float *S[2] = {c->s32.p, c->t32.p};
        int m = 0;
        mg_vcycle(c, c->r32.p, h->z.p, c->state.p, &g, h->pz.p);
        launch_dir_spmv<float, float>(c->stream, g, op, h->z.p, S[0], S[1], c->q32.p, red);
        ++m;
        rc = drive(
            c,
            [&]() {
                launch_axpy_norm<float>(c->stream, g, nullptr, c->x32.p, c->r32.p, S[m & 1], c->q32.p, red,
                                        c->next_events());
                mg_vcycle(c, c->r32.p, h->z.p, c->state.p, &g, h->pz.p);
                launch_dir_spmv<float, float>(c->stream, g, op, h->z.p, S[m & 1], S[(m + 1) & 1], c->q32.p, red,
                                              c->next_events());
                ++m;
            },
            interval);
        if (rc) return rc;
        total += c->h_state->iter;
        ++outer;
        const int stt = c->h_state->status;
        launch_add_correction(c->stream, g.n, c->x64.p, c->x32.p, sig);
        if ((rc = true_residual(c, c->x64.p, c->r64.p, rr, sig, c->rs32.p))) return rc;
        r_inf = rr[0];
        if (stt == SHK_NAN || stt == SHK_BREAKDOWN) {
            status = stt;
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

namespace {

template <int BX>
__global__ void k_dense_to_bricks(int nx, int ny, int nz, const int4 *__restrict__ coord, int n_active,
                                  const float *__restrict__ src, float *__restrict__ dst) {
    using B = BrickShape<BX>;
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= (int64_t)n_active * kBrickCells) return;
    const int s = (int)(e >> 9), l = (int)(e & 511);
    const int4 q = coord[s];
    const int i = q.x * B::X + l % B::X, j = q.y * B::Y + (l / B::X) % B::Y, k = q.z * B::Z + l / kWave;
    dst[e] = (i < nx && j < ny && k < nz) ? src[((int64_t)k * ny + j) * nx + i] : 0.0f;
}

template <int BX>
__global__ void k_bricks_to_dense(int nx, int ny, int nz, const int *__restrict__ slot, int nbx, int nby,
                                  const float *__restrict__ src, float *__restrict__ dst) {
    using B = BrickShape<BX>;
    const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= (int64_t)nx * ny * nz) return;
    const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((int64_t)nx * ny));
    const int s = slot[(i / B::X) + (int64_t)nbx * ((j / B::Y) + (int64_t)nby * (k / B::Z))];
    dst[v] = s >= 0 ? src[(int64_t)s * kBrickCells + B::elem(i, j, k)] : 0.0f;
}

}  // namespace

}  // namespace shk

using namespace shk;

extern "C" int shk_mg_apply(shk_ctx *c, const float *r, float *z) {
    if (!c || !r || !z) return SHK_BAD_ARG;
    if (!c->have_system || !c->mg || c->prm.precond != 3)
        return c->fail(SHK_BAD_ARG, "no multigrid hierarchy (project with Preconditioner=mg first)");
    if (c->prm.host_pointers) return c->fail(SHK_BAD_ARG, "shk_mg_apply takes device pointers");
    HIPC(c, hipSetDevice(c->device));
    const Dims d = c->dims;
    const int64_t nd = (int64_t)d.nx * d.ny * d.nz, nb = (int64_t)c->n_active * kBrickCells;
    hipStream_t st = c->stream;
    const unsigned gb = (unsigned)((nb + 255) / 256), gd = (unsigned)((nd + 255) / 256);
    if (d.bx == 64)
        hipLaunchKernelGGL(k_dense_to_bricks<64>, dim3(gb), dim3(256), 0, st, d.nx, d.ny, d.nz, c->coord.p,
                           c->n_active, r, c->r32.p);
    else
        hipLaunchKernelGGL(k_dense_to_bricks<8>, dim3(gb), dim3(256), 0, st, d.nx, d.ny, d.nz, c->coord.p,
                           c->n_active, r, c->r32.p);
    mg_vcycle(c, c->r32.p, c->mg->z.p, nullptr, &c->grid, c->mg->pz.p);
    if (d.bx == 64)
        hipLaunchKernelGGL(k_bricks_to_dense<64>, dim3(gd), dim3(256), 0, st, d.nx, d.ny, d.nz, c->brick_slot.p,
                           c->nbx, c->nby, c->mg->z.p, z);
    else
        hipLaunchKernelGGL(k_bricks_to_dense<8>, dim3(gd), dim3(256), 0, st, d.nx, d.ny, d.nz, c->brick_slot.p,
                           c->nbx, c->nby, c->mg->z.p, z);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(st));
    return SHK_OK;
}
