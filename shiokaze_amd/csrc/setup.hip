// setup.hip -- per-projection kernels around the Krylov loop:
//   face fractions (with the solid-density blend), surface tension, unknown labels,
//   brick activity, stencil coefficients + rhs (with the volume correction), warm
//   start, pressure scatter, velocity update (gradient subtract), diagnostics.
//
// All floating point here rounds exactly as the reference's CPU code
// (no contraction, IEEE division), so areas, rhos, couplings, rhs, pressure and
// the updated velocity are bit-identical to a faithful CPU evaluation.
#include "setup.h"

#include "shk_common.h"
#include "shk_fractions.h"

namespace shk {

namespace {

constexpr int kNT = 256;

inline unsigned grid_for(int64_t n, int nt = kNT) {
    int64_t g = (n + nt - 1) / nt;
    if (g > (1 << 20)) g = 1 << 20;
    return (unsigned)(g > 0 ? g : 1);
}

// ---------------------------------------------------------------- fractions (a3, a4, a5)

template <int DIM, bool TWO_D>
__global__ void k_fractions(Dims d, const float *__restrict__ solid, const float *__restrict__ fluid,
                            double eps_solid, double eps_fluid, int second_order_solid, int second_order_fluid,
                            const float *__restrict__ density, float *__restrict__ area, float *__restrict__ rho) {
#pragma clang fp contract(off)
    const int fx = d.nx + (DIM == 0), fy = d.ny + (DIM == 1);
    const int64_t n = face_count(d, DIM);
    for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < n; f += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        split_index(f, fx, fy, i, j, k);
        double a = TWO_D ? face_area_2d(d, solid, DIM, i, k, eps_solid) : face_area(d, solid, DIM, i, j, k, eps_solid);
        double r = face_rho(d, fluid, DIM, i, j, k, eps_fluid);
        float af = (float)a, rf = (float)r;
        if (density) {
            // solid density (macpressuresolver3.cpp:75-80): the float product
            // density*area, plus (1 - density) in double, stored as float
            const float dd = density[f];
            const float prod = dd * af;
            af = (float)((double)prod + (1.0 - (double)dd));
        }
        // first-order switches (macpressuresolver3.cpp:83-93): nonzero -> 1
        if (!second_order_solid && af != 0.0f) af = 1.0f;
        if (!second_order_fluid && rf != 0.0f) rf = 1.0f;
        area[f] = af;
        rho[f] = rf;
    }
}

// ---------------------------------------------------------------- labels (a6)

struct Faces {
    const float *a[3];
    const float *r[3];
};

__device__ inline bool is_unknown(const Dims &d, const float *fluid, const Faces &fc, int i, int j, int k) {
    if (!(fluid[cell_index(d, i, j, k)] < 0.0f)) return false;
    for (int nq = 0; nq < 6; ++nq) {
        int qi, qj, qk, fi, fj, fk, dim, sgn;
        neighbour(nq, qi, qj, qk, fi, fj, fk, dim, sgn);
        qi += i;
        qj += j;
        qk += k;
        if (qi < 0 || qi >= d.nx || qj < 0 || qj >= d.ny || qk < 0 || qk >= d.nz) continue;
        if (!(fluid[cell_index(d, qi, qj, qk)] < 0.0f)) continue;
        const int64_t f = face_index(d, dim, i + fi, j + fj, k + fk);
        if (fc.a[dim][f] != 0.0f && fc.r[dim][f] != 0.0f) return true;
    }
    return false;
}

template <int BX>
__global__ void k_label(Dims d, const float *__restrict__ fluid, Faces fc, uint8_t *__restrict__ label,
                        uint8_t *__restrict__ brick_flag, int nbx, int nby, unsigned long long *__restrict__ count) {
    __shared__ unsigned long long wsum[kNT / kWave];
    const int64_t G = (int64_t)d.nx * d.ny * d.nz;
    unsigned long long local = 0;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < G; c += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        split_index(c, d.nx, d.ny, i, j, k);
        const bool u = is_unknown(d, fluid, fc, i, j, k);
        label[c] = u ? 1 : 0;
        if (u) {
            using B = BrickShape<BX>;
            brick_flag[(i / B::X) + (int64_t)nbx * ((j / B::Y) + (int64_t)nby * (k / B::Z))] = 1;
            ++local;
        }
    }
    // integer count (order independent): one atomic per workgroup
    for (int o = kWave / 2; o > 0; o >>= 1) local += __shfl_xor(local, o, kWave);
    if ((threadIdx.x % kWave) == 0) wsum[threadIdx.x / kWave] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kNT / kWave; ++w) t += wsum[w];
        if (t) atomicAdd(count, t);
    }
}

// ---------------------------------------------------------------- coefficients + rhs (a7)

template <int BX>
__global__ void k_coeff(Dims d, double dt, double dx, const int4 *__restrict__ brick_coord, Faces fc,
                        const uint8_t *__restrict__ label, const float *__restrict__ u, const float *__restrict__ v,
                        const float *__restrict__ w, const float *__restrict__ us, const float *__restrict__ vs,
                        const float *__restrict__ ws, double rhs_correct, double *__restrict__ diag,
                        double *__restrict__ cx, double *__restrict__ cy, double *__restrict__ cz,
                        double *__restrict__ rhs, double *__restrict__ sig, float *__restrict__ diag32,
                        float *__restrict__ cx32, float *__restrict__ cy32, float *__restrict__ cz32) {
#pragma clang fp contract(off)
    const int slot = blockIdx.x;
    const int l = threadIdx.x;
    const int4 bc = brick_coord[slot];
    using B = BrickShape<BX>;
    const int i = bc.x * B::X + l % B::X, j = bc.y * B::Y + (l / B::X) % B::Y, k = bc.z * B::Z + l / kWave;
    const int64_t e = (int64_t)slot * kBrickCells + l;
    double cd = 0.0, c0 = 0.0, c1 = 0.0, c2 = 0.0, b = 0.0;
    if (i < d.nx && j < d.ny && k < d.nz && label[cell_index(d, i, j, k)]) {
        const double dx2 = dx * dx;
        // off-diagonal -dt*area/(dx^2*rho) (macpressuresolver3.cpp:199-203) of the
        // cell's low faces, kept only where both cells are unknowns; a neighbour
        // reads its high faces from here
        const int coord[3] = {i, j, k};
        double low[3];
        for (int a = 0; a < 3; ++a) {
            low[a] = 0.0;
            if (coord[a] == 0) continue;
            if (!label[cell_index(d, i - (a == 0), j - (a == 1), k - (a == 2))]) continue;
            const int64_t f = face_index(d, a, i, j, k);
            const double area = fc.a[a][f], rho = fc.r[a][f];
            if (area != 0.0 && rho != 0.0) low[a] = dt * area / (dx2 * rho);
        }
        c0 = low[0];
        c1 = low[1];
        c2 = low[2];
        const float *vel[3] = {u, v, w};
        const float *svel[3] = {us, vs, ws};
        double diagonal = 0.0;
        for (int nq = 0; nq < 6; ++nq) {
            int qi, qj, qk, fi, fj, fk, dim, sgn;
            neighbour(nq, qi, qj, qk, fi, fj, fk, dim, sgn);
            qi += i;
            qj += j;
            qk += k;
            const int64_t f = face_index(d, dim, i + fi, j + fj, k + fk);
            const double area = fc.a[dim][f];
            if (!(qi < 0 || qi >= d.nx || qj < 0 || qj >= d.ny || qk < 0 || qk >= d.nz)) {
                if (area != 0.0) {
                    const double rho = fc.r[dim][f];
                    if (rho != 0.0) diagonal += dt * area / (dx2 * rho);
                }
            }
            b += -sgn * area * (double)vel[dim][f] / dx;
            if (area < 1.0) {
                const double usf = svel[dim][f];
                b += -sgn * (1.0 - area) * usf / dx;
            }
        }
        // volume correction: one constant added to every row (macpressuresolver3.cpp:231-233)
        if (rhs_correct != 0.0) b += rhs_correct;
        cd = diagonal;
    }
    diag[e] = cd;
    cx[e] = c0;
    cy[e] = c1;
    cz[e] = c2;
    rhs[e] = b;
    sig[e] = cd > 0.0 ? 1.0 / sqrt(cd) : 0.0;
    if (diag32) {
        diag32[e] = (float)cd;
        cx32[e] = (float)c0;
        cy32[e] = (float)c1;
        cz32[e] = (float)c2;
    }
}

// Couplings of the Jacobi-scaled system D^-1/2 A D^-1/2: c * sig_cell * sig_low.
template <int BX>
__global__ void k_scale_coeff(const int *__restrict__ nbr, const double *__restrict__ sig,
                              const double *__restrict__ cx, const double *__restrict__ cy,
                              const double *__restrict__ cz, double *__restrict__ hx, double *__restrict__ hy,
                              double *__restrict__ hz, float *__restrict__ hx32, float *__restrict__ hy32,
                              float *__restrict__ hz32) {
    using B = BrickShape<BX>;
    const int slot = blockIdx.x;
    const int l = threadIdx.x;
    const int loc[3] = {l % B::X, (l / B::X) % B::Y, l / kWave};
    const int stride[3] = {1, B::X, kWave};
    const int ext[3] = {B::X, B::Y, B::Z};
    const int64_t e = (int64_t)slot * kBrickCells + l;
    const double *c[3] = {cx, cy, cz};
    double *h[3] = {hx, hy, hz};
    float *h32[3] = {hx32, hy32, hz32};
    const double se = sig[e];
    for (int a = 0; a < 3; ++a) {
        const double ca = c[a][e];
        double v = 0.0;
        if (ca != 0.0) {
            int64_t m;
            if (loc[a] > 0) {
                m = e - stride[a];
            } else {
                const int nb = nbr[6 * slot + 2 * a];
                m = nb >= 0 ? (int64_t)nb * kBrickCells + l + (ext[a] - 1) * stride[a] : -1;
            }
            if (m >= 0) v = ca * se * sig[m];
        }
        if (h[a]) h[a][e] = v;
        if (h32[a]) h32[a][e] = (float)v;
    }
}

// RCMatrix_utility::report (RCMatrix_utility.h:98-146) of the assembled brick system:
// per block (grid-strided over bricks, one thread per brick cell) the row statistics
// of the matrix diag - sum c neighbour, whose rows are the cells with a nonzero
// diagonal.  A row's entries are its diagonal and its nonzero couplings: the three
// stored at the cell (low faces) and the three stored at its +x/+y/+z neighbours
// (high faces).  part[8*block + {0..7}] = {nonzeros, min row, max row, min diag,
// max diag, max |coupling|/diag, NaN present, rows}.
template <int BX>
__global__ __launch_bounds__(kBrickCells) void k_report(int n_active, const int *__restrict__ nbr,
                                                        const double *__restrict__ diag,
                                                        const double *__restrict__ cx,
                                                        const double *__restrict__ cy,
                                                        const double *__restrict__ cz, double *__restrict__ part) {
    using B = BrickShape<BX>;
    __shared__ double red[8][kBrickCells / kWave];
    const int l = threadIdx.x;
    const int loc[3] = {l % B::X, (l / B::X) % B::Y, l / kWave};
    const int stride[3] = {1, B::X, kWave};
    const int ext[3] = {B::X, B::Y, B::Z};
    const double *c[3] = {cx, cy, cz};
    double nnz = 0.0, rmin = 1e300, rmax = 0.0, dmin = 1e300, dmax = -1e300, ratio = 0.0, nan = 0.0, rows = 0.0;
    for (int slot = blockIdx.x; slot < n_active; slot += gridDim.x) {
        const int64_t e = (int64_t)slot * kBrickCells + l;
        const double d = diag[e];
        if (d == 0.0) continue;  // not an unknown (NaN is not == 0 and counts as a row)
        int cnt = 1;
        double off = 0.0;
        bool bad = d != d;
        for (int a = 0; a < 3; ++a) {
            const double lo = c[a][e];
            double hi = 0.0;
            if (loc[a] < ext[a] - 1) {
                hi = c[a][e + stride[a]];
            } else {
                const int nb = nbr[6 * slot + 2 * a + 1];
                if (nb >= 0) hi = c[a][(int64_t)nb * kBrickCells + l - (ext[a] - 1) * stride[a]];
            }
            const double lohi[2] = {lo, hi};
            for (int t = 0; t < 2; ++t) {
                if (lohi[t] != 0.0) {
                    ++cnt;
                    off = fmax(off, fabs(lohi[t]));
                }
                bad = bad || lohi[t] != lohi[t];
            }
        }
        rows += 1.0;
        nnz += cnt;
        rmin = fmin(rmin, (double)cnt);
        rmax = fmax(rmax, (double)cnt);
        if (d == d) {
            dmin = fmin(dmin, d);
            dmax = fmax(dmax, d);
            if (d != 0.0) ratio = fmax(ratio, off / d);
        }
        if (bad) nan = 1.0;
    }
    double v[8] = {nnz, rmin, rmax, dmin, dmax, ratio, nan, rows};
    for (int k = 0; k < 8; ++k) {
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const double w = __shfl_xor(v[k], o);
            v[k] = (k == 0 || k == 7) ? v[k] + w : ((k == 1 || k == 3) ? fmin(v[k], w) : fmax(v[k], w));
        }
    }
    if ((l & (kWave - 1)) == 0)
        for (int k = 0; k < 8; ++k) red[k][l / kWave] = v[k];
    __syncthreads();
    if (l < 8) {
        double a = red[l][0];
        for (int wv = 1; wv < kBrickCells / kWave; ++wv) {
            const double w = red[l][wv];
            a = (l == 0 || l == 7) ? a + w : ((l == 1 || l == 3) ? fmin(a, w) : fmax(a, w));
        }
        part[8 * (int64_t)blockIdx.x + l] = a;
    }
}

// Per-brick constant coupling: kappa[slot] = c when every coupling touching the
// brick's cells (its 3 low faces per cell and the far faces stored in the +x/+y/+z
// neighbour bricks) equals one nonzero value c bit for bit, and (with a diagonal
// array) every diagonal equals one value d, then kdiag[slot] = d; else kappa = 0.
// Such a brick has only unknowns and the stencil needs no coefficient loads there.
template <class C, int BX>
__global__ void k_brick_uniform(const int *__restrict__ nbr, const C *__restrict__ cx, const C *__restrict__ cy,
                                const C *__restrict__ cz, const C *__restrict__ diag, C *__restrict__ kappa,
                                C *__restrict__ kdiag) {
    using B = BrickShape<BX>;
    const int slot = blockIdx.x;
    const int l = threadIdx.x;
    const int loc[3] = {l % B::X, (l / B::X) % B::Y, l / kWave};
    const int stride[3] = {1, B::X, kWave};
    const int ext[3] = {B::X, B::Y, B::Z};
    const C *c[3] = {cx, cy, cz};
    const int64_t e = (int64_t)slot * kBrickCells + l;
    const C k0 = cx[(int64_t)slot * kBrickCells];
    bool ok = k0 != C(0);
    for (int a = 0; a < 3; ++a) {
        if (ext[a] == 1) continue;  // the y axis of a 2D grid has no couplings
        ok = ok && c[a][e] == k0;
        int64_t m;
        if (loc[a] < ext[a] - 1) {
            m = e + stride[a];
        } else {
            const int nb = nbr[6 * slot + 2 * a + 1];
            m = nb >= 0 ? (int64_t)nb * kBrickCells + l - (ext[a] - 1) * stride[a] : -1;
        }
        ok = ok && m >= 0 && c[a][m] == k0;
    }
    const C d0 = diag ? diag[(int64_t)slot * kBrickCells] : C(0);
    if (diag) ok = ok && diag[e] == d0;
    const int all = __syncthreads_and(ok ? 1 : 0);
    if (l == 0) {
        kappa[slot] = all ? k0 : C(0);
        if (kdiag) kdiag[slot] = all ? d0 : C(0);
    }
}

// ---------------------------------------------------------------- scatter + update (a14, a15)

template <int BX>
__device__ inline int64_t brick_elem(const Dims &d, const int *brick_slot, int nbx, int nby, int i, int j, int k) {
    using B = BrickShape<BX>;
    const int s = brick_slot[(i / B::X) + (int64_t)nbx * ((j / B::Y) + (int64_t)nby * (k / B::Z))];
    if (s < 0) return -1;
    return (int64_t)s * kBrickCells + B::elem(i, j, k);
}

template <int BX>
__global__ void k_scatter(Dims d, const int *__restrict__ brick_slot, int nbx, int nby, const uint8_t *__restrict__ label,
                          const double *__restrict__ x, float *__restrict__ pressure, double *__restrict__ prev) {
    const int64_t G = (int64_t)d.nx * d.ny * d.nz;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < G; c += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        split_index(c, d.nx, d.ny, i, j, k);
        double p = 0.0;
        if (label[c]) {
            const int64_t e = brick_elem<BX>(d, brick_slot, nbx, nby, i, j, k);
            if (e >= 0) p = x[e];
        }
        pressure[c] = (float)p;
        if (prev) prev[c] = p;
    }
}

template <int DIM>
__global__ void k_update_velocity(Dims d, double dt, double dx, const float *__restrict__ pressure,
                                  const float *__restrict__ area, const float *__restrict__ rho,
                                  const float *__restrict__ svel, const float *__restrict__ density,
                                  float *__restrict__ vel) {
#pragma clang fp contract(off)
    const int fx = d.nx + (DIM == 0), fy = d.ny + (DIM == 1);
    const int n = DIM == 0 ? d.nx : (DIM == 1 ? d.ny : d.nz);
    const int64_t nf = face_count(d, DIM);
    for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nf; f += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        split_index(f, fx, fy, i, j, k);
        const int pi = DIM == 0 ? i : (DIM == 1 ? j : k);
        const double r = rho[f];
        float nv = 0.0f;
        if (area[f] != 0.0f && r != 0.0) {
            if (pi == 0 || pi == n) {
                nv = svel[f];
            } else {
                const float p1 = pressure[cell_index(d, i, j, k)];
                const float p0 = pressure[cell_index(d, i - (DIM == 0), j - (DIM == 1), k - (DIM == 2))];
                const float dp = p1 - p0;
                const float inc = (float)(dt * (double)dp / (r * dx));
                nv = vel[f] - inc;
            }
            if (density) {
                // macpressuresolver3.cpp:287-293: scale = 1 - (1 - area)*density in
                // double, stored as the float `scale`, times the float velocity
                const float scale = (float)(1.0 - (1.0 - (double)area[f]) * (double)density[f]);
                nv = scale * nv;
            }
        }
        vel[f] = nv;
    }
}

// ---------------------------------------------------------------- surface tension (macpressuresolver3.cpp:95-124)

// Curvature as the 7-point Laplacian of the liquid level set with clamped
// neighbours.  The six neighbour values are summed in float, left to right; the
// centre term and the division by dx^2 are double; the result is stored as float.
__global__ void k_curvature(Dims d, double dx, const float *__restrict__ fluid, float *__restrict__ curv) {
#pragma clang fp contract(off)
    const int64_t G = (int64_t)d.nx * d.ny * d.nz;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < G; c += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        split_index(c, d.nx, d.ny, i, j, k);
        float s = fluid[cell_index(d, clampi(i - 1, 0, d.nx - 1), j, k)];
        s = s + fluid[cell_index(d, clampi(i + 1, 0, d.nx - 1), j, k)];
        s = s + fluid[cell_index(d, i, clampi(j - 1, 0, d.ny - 1), k)];
        s = s + fluid[cell_index(d, i, clampi(j + 1, 0, d.ny - 1), k)];
        s = s + fluid[cell_index(d, i, j, clampi(k - 1, 0, d.nz - 1))];
        s = s + fluid[cell_index(d, i, j, clampi(k + 1, 0, d.nz - 1))];
        const double v = ((double)s - 6.0 * (double)fluid[c]) / (dx * dx);
        curv[c] = (float)v;
    }
}

// The 2D curvature of a grid held as nx x 1 x ny (macpressuresolver2.cpp:93-102):
// the clamped 5-point Laplacian, neighbours summed in float (x-, x+, y-, y+).
__global__ void k_curvature_2d(Dims d, double dx, const float *__restrict__ fluid, float *__restrict__ curv) {
#pragma clang fp contract(off)
    const int64_t G = (int64_t)d.nx * d.nz;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < G; c += (int64_t)gridDim.x * blockDim.x) {
        int i, jj, k;
        split_index(c, d.nx, 1, i, jj, k);
        (void)jj;
        float s = fluid[clampi(i - 1, 0, d.nx - 1) + (int64_t)d.nx * k];
        s = s + fluid[clampi(i + 1, 0, d.nx - 1) + (int64_t)d.nx * k];
        s = s + fluid[i + (int64_t)d.nx * clampi(k - 1, 0, d.nz - 1)];
        s = s + fluid[i + (int64_t)d.nx * clampi(k + 1, 0, d.nz - 1)];
        const double v = ((double)s - 4.0 * (double)fluid[c]) / (dx * dx);
        curv[c] = (float)v;
    }
}

// Ghost-fluid surface-tension jump on faces with 0 < rho < 1: the face curvature
// interpolated with the liquid fraction, added to the face velocity (the increment
// is rounded to float, then added in float as array3::increment does).
template <int DIM>
__global__ void k_surface_tension(Dims d, double dt, double dx, double kappa, const float *__restrict__ fluid,
                                  const float *__restrict__ curv, const float *__restrict__ rho,
                                  float *__restrict__ vel) {
#pragma clang fp contract(off)
    const int fx = d.nx + (DIM == 0), fy = d.ny + (DIM == 1);
    const int64_t nf = face_count(d, DIM);
    for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nf; f += (int64_t)gridDim.x * blockDim.x) {
        const double r = rho[f];
        if (!(r != 0.0 && r < 1.0)) continue;
        int i, j, k;
        split_index(f, fx, fy, i, j, k);
        const int64_t c1 = cell_index(d, clampi(i, 0, d.nx - 1), clampi(j, 0, d.ny - 1), clampi(k, 0, d.nz - 1));
        const int64_t c0 = cell_index(d, clampi(i - (DIM == 0), 0, d.nx - 1), clampi(j - (DIM == 1), 0, d.ny - 1),
                                      clampi(k - (DIM == 2), 0, d.nz - 1));
        const double sgn = fluid[c1] < 0.0f ? -1.0 : 1.0;
        const double theta = sgn < 0 ? 1.0 - r : r;
        const double face_c = theta * (double)curv[c1] + (1.0 - theta) * (double)curv[c0];
        const double inc = -sgn * dt / (dx * r) * kappa * face_c;
        vel[f] = vel[f] + (float)inc;
    }
}

// ---------------------------------------------------------------- warm start (macpressuresolver3.cpp:238-260)

// The previous projection's FP64 pressure (dense) onto the current bricks, zero off
// the unknowns.
template <int BX>
__global__ void k_gather_warm(Dims d, const int4 *__restrict__ brick_coord, const uint8_t *__restrict__ label,
                              const double *__restrict__ prev, double *__restrict__ w) {
    const int slot = blockIdx.x;
    const int l = threadIdx.x;
    const int4 bc = brick_coord[slot];
    using B = BrickShape<BX>;
    const int i = bc.x * B::X + l % B::X, j = bc.y * B::Y + (l / B::X) % B::Y, k = bc.z * B::Z + l / kWave;
    double v = 0.0;
    if (i < d.nx && j < d.ny && k < d.nz) {
        const int64_t c = cell_index(d, i, j, k);
        if (label[c]) v = prev[c];
    }
    w[(int64_t)slot * kBrickCells + l] = v;
}

__global__ void k_add64(int64_t n, double *__restrict__ x, const double *__restrict__ w) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        x[e] = x[e] + w[e];
}

// ---------------------------------------------------------------- diagnostics

template <int BX>
__device__ inline double xval(const Dims &d, const int *brick_slot, int nbx, int nby, const double *x, int i, int j,
                              int k) {
    if (i < 0 || i >= d.nx || j < 0 || j >= d.ny || k < 0 || k >= d.nz) return 0.0;
    const int64_t e = brick_elem<BX>(d, brick_slot, nbx, nby, i, j, k);
    return e >= 0 ? x[e] : 0.0;
}

template <int BX>
__global__ void k_divergence(Dims d, double dt, double dx, Faces fc, const uint8_t *__restrict__ label,
                             const float *__restrict__ u, const float *__restrict__ v, const float *__restrict__ w,
                             const float *__restrict__ us, const float *__restrict__ vs, const float *__restrict__ ws,
                             int use_fp64_update, const int *__restrict__ brick_slot, int nbx, int nby,
                             const double *__restrict__ x, double *__restrict__ partial) {
#pragma clang fp contract(off)
    __shared__ double lds[kNT / kWave];
    const int64_t G = (int64_t)d.nx * d.ny * d.nz;
    const float *vel[3] = {u, v, w};
    const float *svel[3] = {us, vs, ws};
    double best = 0.0;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < G; c += (int64_t)gridDim.x * blockDim.x) {
        if (!label[c]) continue;
        int i, j, k;
        split_index(c, d.nx, d.ny, i, j, k);
        double acc = 0.0;
        for (int nq = 0; nq < 6; ++nq) {
            int qi, qj, qk, fi, fj, fk, dim, sgn;
            neighbour(nq, qi, qj, qk, fi, fj, fk, dim, sgn);
            const int ai = i + fi, aj = j + fj, ak = k + fk;  // face (dim, ai, aj, ak)
            const int64_t f = face_index(d, dim, ai, aj, ak);
            const double area = fc.a[dim][f];
            double uf = vel[dim][f];
            if (use_fp64_update) {
                const double r = fc.r[dim][f];
                const int pi = dim == 0 ? ai : (dim == 1 ? aj : ak);
                const int n = dim == 0 ? d.nx : (dim == 1 ? d.ny : d.nz);
                if (area != 0.0 && r != 0.0) {
                    if (pi == 0 || pi == n) {
                        uf = svel[dim][f];
                    } else {
                        const double p1 = xval<BX>(d, brick_slot, nbx, nby, x, ai, aj, ak);
                        const double p0 = xval<BX>(d, brick_slot, nbx, nby, x, ai - (dim == 0), aj - (dim == 1),
                                               ak - (dim == 2));
                        uf = uf - dt * (p1 - p0) / (r * dx);
                    }
                } else {
                    uf = 0.0;
                }
            }
            acc += -sgn * area * uf / dx;
            if (area < 1.0) acc += -sgn * (1.0 - area) * (double)svel[dim][f] / dx;
        }
        best = fmax(best, fabs(acc));
    }
    const double m = block_max<kNT>(best, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

template <int BX>
__global__ void k_export(Dims d, const int *__restrict__ brick_slot, int nbx, int nby, const uint8_t *__restrict__ label,
                         const double *__restrict__ diag, const double *__restrict__ cx, const double *__restrict__ cy,
                         const double *__restrict__ cz, const double *__restrict__ rhs, const double *__restrict__ x,
                         uint8_t *o_label, double *o_diag, double *o_cx, double *o_cy, double *o_cz, double *o_rhs,
                         double *o_x) {
    const int64_t G = (int64_t)d.nx * d.ny * d.nz;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < G; c += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        split_index(c, d.nx, d.ny, i, j, k);
        const int64_t e = brick_elem<BX>(d, brick_slot, nbx, nby, i, j, k);
        if (o_label) o_label[c] = label[c];
        if (o_diag) o_diag[c] = e >= 0 ? diag[e] : 0.0;
        if (o_cx) o_cx[c] = e >= 0 ? cx[e] : 0.0;
        if (o_cy) o_cy[c] = e >= 0 ? cy[e] : 0.0;
        if (o_cz) o_cz[c] = e >= 0 ? cz[e] : 0.0;
        if (o_rhs) o_rhs[c] = e >= 0 ? rhs[e] : 0.0;
        if (o_x) o_x[c] = e >= 0 ? x[e] : 0.0;
    }
}

}  // namespace

// ---------------------------------------------------------------- launchers

void launch_fractions(hipStream_t st, Dims d, const float *solid, const float *fluid, double eps_solid,
                      double eps_fluid, int so_solid, int so_fluid, const float *const density[3],
                      float *const area[3], float *const rho[3]) {
    if (d.bx == 64) {
        // a 2D grid: x faces and the 2D y faces (held as z faces); the y faces of the
        // nx x 1 x ny layout are walls whose arrays stay zero
        hipLaunchKernelGGL((k_fractions<0, true>), dim3(grid_for(face_count(d, 0))), dim3(kNT), 0, st, d, solid,
                           fluid, eps_solid, eps_fluid, so_solid, so_fluid, density[0], area[0], rho[0]);
        hipLaunchKernelGGL((k_fractions<2, true>), dim3(grid_for(face_count(d, 2))), dim3(kNT), 0, st, d, solid,
                           fluid, eps_solid, eps_fluid, so_solid, so_fluid, density[2], area[2], rho[2]);
        return;
    }
    hipLaunchKernelGGL((k_fractions<0, false>), dim3(grid_for(face_count(d, 0))), dim3(kNT), 0, st, d, solid, fluid,
                       eps_solid, eps_fluid, so_solid, so_fluid, density[0], area[0], rho[0]);
    hipLaunchKernelGGL((k_fractions<1, false>), dim3(grid_for(face_count(d, 1))), dim3(kNT), 0, st, d, solid, fluid,
                       eps_solid, eps_fluid, so_solid, so_fluid, density[1], area[1], rho[1]);
    hipLaunchKernelGGL((k_fractions<2, false>), dim3(grid_for(face_count(d, 2))), dim3(kNT), 0, st, d, solid, fluid,
                       eps_solid, eps_fluid, so_solid, so_fluid, density[2], area[2], rho[2]);
}

void launch_surface_tension(hipStream_t st, Dims d, double dt, double dx, double kappa, const float *fluid,
                            float *curv, float *const rho[3], float *const vel[3]) {
    const int64_t G = (int64_t)d.nx * d.ny * d.nz;
    if (d.bx == 64) {
        hipLaunchKernelGGL(k_curvature_2d, dim3(grid_for(G)), dim3(kNT), 0, st, d, dx, fluid, curv);
    } else {
        hipLaunchKernelGGL(k_curvature, dim3(grid_for(G)), dim3(kNT), 0, st, d, dx, fluid, curv);
    }
    hipLaunchKernelGGL(k_surface_tension<0>, dim3(grid_for(face_count(d, 0))), dim3(kNT), 0, st, d, dt, dx, kappa,
                       fluid, curv, rho[0], vel[0]);
    if (d.bx != 64)
        hipLaunchKernelGGL(k_surface_tension<1>, dim3(grid_for(face_count(d, 1))), dim3(kNT), 0, st, d, dt, dx, kappa,
                           fluid, curv, rho[1], vel[1]);
    hipLaunchKernelGGL(k_surface_tension<2>, dim3(grid_for(face_count(d, 2))), dim3(kNT), 0, st, d, dt, dx, kappa,
                       fluid, curv, rho[2], vel[2]);
}

void launch_gather_warm(hipStream_t st, Dims d, int n_active, const int4 *brick_coord, const uint8_t *label,
                        const double *prev, double *w) {
    if (n_active <= 0) return;
    if (d.bx == 64)
        hipLaunchKernelGGL(k_gather_warm<64>, dim3(n_active), dim3(kBrickCells), 0, st, d, brick_coord, label, prev, w);
    else
        hipLaunchKernelGGL(k_gather_warm<8>, dim3(n_active), dim3(kBrickCells), 0, st, d, brick_coord, label, prev, w);
}

void launch_add64(hipStream_t st, int64_t n, double *x, const double *w) {
    hipLaunchKernelGGL(k_add64, dim3(grid_for(n)), dim3(kNT), 0, st, n, x, w);
}

void launch_label(hipStream_t st, Dims d, const float *fluid, float *const area[3], float *const rho[3], uint8_t *label,
                  uint8_t *brick_flag, int nbx, int nby, unsigned long long *count) {
    Faces fc{{area[0], area[1], area[2]}, {rho[0], rho[1], rho[2]}};
    const int64_t G = (int64_t)d.nx * d.ny * d.nz;
    unsigned g = grid_for(G);
    if (g > 4096) g = 4096;  // grid-stride: bounded number of count atomics
    if (d.bx == 64)
        hipLaunchKernelGGL(k_label<64>, dim3(g), dim3(kNT), 0, st, d, fluid, fc, label, brick_flag, nbx, nby, count);
    else
        hipLaunchKernelGGL(k_label<8>, dim3(g), dim3(kNT), 0, st, d, fluid, fc, label, brick_flag, nbx, nby, count);
}

void launch_coeff(hipStream_t st, Dims d, double dt, double dx, int n_active, const int4 *brick_coord,
                  float *const area[3], float *const rho[3], const uint8_t *label, const float *const vel[3],
                  const float *const svel[3], double rhs_correct, double *diag, double *cx, double *cy, double *cz,
                  double *rhs, double *sig, float *diag32, float *cx32, float *cy32, float *cz32) {
    Faces fc{{area[0], area[1], area[2]}, {rho[0], rho[1], rho[2]}};
    if (n_active <= 0) return;
    auto kern = d.bx == 64 ? k_coeff<64> : k_coeff<8>;
    hipLaunchKernelGGL(kern, dim3(n_active), dim3(kBrickCells), 0, st, d, dt, dx, brick_coord, fc, label, vel[0],
                       vel[1], vel[2], svel[0], svel[1], svel[2], rhs_correct, diag, cx, cy, cz, rhs, sig, diag32,
                       cx32, cy32, cz32);
}

void launch_scale_coeff(hipStream_t st, int bx, int n_active, const int *nbr, const double *sig, const double *cx,
                        const double *cy, const double *cz, double *hx, double *hy, double *hz, float *hx32,
                        float *hy32, float *hz32) {
    if (n_active <= 0) return;
    auto kern = bx == 64 ? k_scale_coeff<64> : k_scale_coeff<8>;
    hipLaunchKernelGGL(kern, dim3(n_active), dim3(kBrickCells), 0, st, nbr, sig, cx, cy, cz, hx, hy, hz, hx32, hy32,
                       hz32);
}

void launch_report(hipStream_t st, int bx, int n_active, int blocks, const int *nbr, const double *diag,
                   const double *cx, const double *cy, const double *cz, double *part) {
    auto kern = bx == 64 ? k_report<64> : k_report<8>;
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBrickCells), 0, st, n_active, nbr, diag, cx, cy, cz, part);
}

template <class C>
void launch_brick_uniform(hipStream_t st, int bx, int n_active, const int *nbr, const C *cx, const C *cy, const C *cz,
                          C *kappa, const C *diag, C *kdiag) {
    if (n_active <= 0) return;
    auto kern = bx == 64 ? k_brick_uniform<C, 64> : k_brick_uniform<C, 8>;
    hipLaunchKernelGGL(kern, dim3(n_active), dim3(kBrickCells), 0, st, nbr, cx, cy, cz, diag, kappa, kdiag);
}
template void launch_brick_uniform<double>(hipStream_t, int, int, const int *, const double *, const double *,
                                           const double *, double *, const double *, double *);
template void launch_brick_uniform<float>(hipStream_t, int, int, const int *, const float *, const float *,
                                          const float *, float *, const float *, float *);

void launch_scatter(hipStream_t st, Dims d, const int *brick_slot, int nbx, int nby, const uint8_t *label,
                    const double *x, float *pressure, double *prev) {
    const int64_t G = (int64_t)d.nx * d.ny * d.nz;
    auto kern = d.bx == 64 ? k_scatter<64> : k_scatter<8>;
    hipLaunchKernelGGL(kern, dim3(grid_for(G)), dim3(kNT), 0, st, d, brick_slot, nbx, nby, label, x, pressure, prev);
}

void launch_update_velocity(hipStream_t st, Dims d, double dt, double dx, const float *pressure, float *const area[3],
                            float *const rho[3], const float *const svel[3], const float *const density[3],
                            float *const vel[3]) {
    hipLaunchKernelGGL(k_update_velocity<0>, dim3(grid_for(face_count(d, 0))), dim3(kNT), 0, st, d, dt, dx, pressure,
                       area[0], rho[0], svel[0], density[0], vel[0]);
    if (d.bx != 64)  // the y faces of a 2D grid's nx x 1 x ny layout are walls
        hipLaunchKernelGGL(k_update_velocity<1>, dim3(grid_for(face_count(d, 1))), dim3(kNT), 0, st, d, dt, dx,
                           pressure, area[1], rho[1], svel[1], density[1], vel[1]);
    hipLaunchKernelGGL(k_update_velocity<2>, dim3(grid_for(face_count(d, 2))), dim3(kNT), 0, st, d, dt, dx, pressure,
                       area[2], rho[2], svel[2], density[2], vel[2]);
}

int launch_divergence(hipStream_t st, Dims d, double dt, double dx, float *const area[3], float *const rho[3],
                      const uint8_t *label, const float *const vel[3], const float *const svel[3], int use_fp64_update,
                      const int *brick_slot, int nbx, int nby, const double *x, double *partial, int max_blocks) {
    Faces fc{{area[0], area[1], area[2]}, {rho[0], rho[1], rho[2]}};
    const int64_t G = (int64_t)d.nx * d.ny * d.nz;
    unsigned g = grid_for(G);
    if ((int)g > max_blocks) g = (unsigned)max_blocks;
    auto kern = d.bx == 64 ? k_divergence<64> : k_divergence<8>;
    hipLaunchKernelGGL(kern, dim3(g), dim3(kNT), 0, st, d, dt, dx, fc, label, vel[0], vel[1], vel[2], svel[0], svel[1],
                       svel[2], use_fp64_update, brick_slot, nbx, nby, x, partial);
    return (int)g;
}

void launch_export(hipStream_t st, Dims d, const int *brick_slot, int nbx, int nby, const uint8_t *label,
                   const double *diag, const double *cx, const double *cy, const double *cz, const double *rhs,
                   const double *x, uint8_t *o_label, double *o_diag, double *o_cx, double *o_cy, double *o_cz,
                   double *o_rhs, double *o_x) {
    const int64_t G = (int64_t)d.nx * d.ny * d.nz;
    auto kern = d.bx == 64 ? k_export<64> : k_export<8>;
    hipLaunchKernelGGL(kern, dim3(grid_for(G)), dim3(kNT), 0, st, d, brick_slot, nbx, nby, label, diag, cx, cy, cz,
                       rhs, x, o_label, o_diag, o_cx, o_cy, o_cz, o_rhs, o_x);
}

}  // namespace shk
