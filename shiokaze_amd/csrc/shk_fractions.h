// shk_fractions.h -- face fraction arithmetic shared by the gfx950 kernels and the
// host scene generator.
//
// Each function rounds exactly like the reference's x86-64 build (no FMA
// contraction, IEEE double division), so the GPU's areas, rhos, couplings and
// rhs are bit-identical to a faithful CPU evaluation.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

namespace shk {

// GFM liquid fraction of the segment between two level set samples
// (include/shiokaze/utility/utility.h:174-182): 1 when both are inside (or one is
// inside and the other exactly 0), 0 when both are outside, else the fraction of
// the segment on the negative side.
__host__ __device__ inline double fraction(double phi0, double phi1) {
#pragma clang fp contract(off)
    if (phi0 * phi1 >= 0.0) {
        if (phi0 < 0.0 || phi1 < 0.0) return 1.0;
        return 0.0;
    }
    const double lo = (phi1 < phi0) ? phi1 : phi0;
    const double gap = fabs(phi1 - phi0);
    return -lo / (gap < DBL_MIN ? DBL_MIN : gap);
}

// Area of {phi < 0} inside the unit square with corner values q00 (0,0),
// q10 (1,0), q11 (1,1), q01 (0,1) (include/shiokaze/utility/utility.h:191-226):
// negative corners and linear edge crossings are visited counter-clockwise and
// the shoelace formula is applied to that point list.
// The shoelace terms are added as the points are generated (each term needs only the
// previous point; the closing term the first and last): the same terms in the same
// order as a sum over the stored point list, without a per-thread point array (which
// a GPU keeps in slow scratch memory when it is indexed at run time).
struct ShoelaceSum {
    double fx = 0.0, fy = 0.0, lx = 0.0, ly = 0.0, sum = 0.0;
    int np = 0;
    __host__ __device__ inline void add(double x, double y) {
#pragma clang fp contract(off)
        if (np == 0) {
            fx = x;
            fy = y;
        } else {
            sum += lx * y - ly * x;
        }
        lx = x;
        ly = y;
        ++np;
    }
    __host__ __device__ inline double close() {
#pragma clang fp contract(off)
        if (np > 0) sum += lx * fy - ly * fx;
        return sum;
    }
};

__host__ __device__ inline double get_area(double q00, double q10, double q11, double q01) {
#pragma clang fp contract(off)
    const double v[4] = {q00, q10, q11, q01};
    const double cx[4] = {0.0, 1.0, 1.0, 0.0};
    const double cy[4] = {0.0, 0.0, 1.0, 1.0};
    ShoelaceSum s;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const int m = (n + 1) & 3;
        if (v[n] < 0.0) s.add(cx[n], cy[n]);
        if (v[n] * v[m] < 0) {
            const double y0 = v[n], y1 = v[m];
            if (y0 - y1 != 0.0) {
                const double a = y0 / (y0 - y1);
                s.add((1.0 - a) * cx[n] + a * cx[m], (1.0 - a) * cy[n] + a * cy[m]);
            }
        }
    }
    return 0.5 * s.close();
}

struct Dims {
    int nx, ny, nz;
    int bx = 8;  // brick x extent: 8 (8x8x8 bricks, 3D) or 64 (64x1x8 bricks, a 2D grid as nx x 1 x ny)
};

// (i, j, k) of the linear index q of an nx x ny x (any) array with 32-bit unsigned
// arithmetic: the grids hold at most 2^31 cells, so every cell, face or node index
// fits in 32 bits, and a 64-bit division per thread would cost several times more
__host__ __device__ inline void split_index(int64_t q, int nx, int ny, int &i, int &j, int &k) {
    const uint32_t u = (uint32_t)q;
    const uint32_t t = u / (uint32_t)nx;
    i = (int)(u - t * (uint32_t)nx);
    k = (int)(t / (uint32_t)ny);
    j = (int)(t - (uint32_t)k * (uint32_t)ny);
}

__host__ __device__ inline int64_t cell_index(const Dims &d, int i, int j, int k) {
    return (int64_t)i + (int64_t)d.nx * ((int64_t)j + (int64_t)d.ny * (int64_t)k);
}
__host__ __device__ inline int64_t node_index(const Dims &d, int i, int j, int k) {
    return (int64_t)i + (int64_t)(d.nx + 1) * ((int64_t)j + (int64_t)(d.ny + 1) * (int64_t)k);
}
__host__ __device__ inline int64_t face_index(const Dims &d, int dim, int i, int j, int k) {
    const int64_t fx = d.nx + (dim == 0), fy = d.ny + (dim == 1);
    return (int64_t)i + fx * ((int64_t)j + fy * (int64_t)k);
}
__host__ __device__ inline int64_t face_count(const Dims &d, int dim) {
    return (int64_t)(d.nx + (dim == 0)) * (d.ny + (dim == 1)) * (d.nz + (dim == 2));
}

// Open solid area of face (dim, i, j, k) (macutility3.cpp:121-147): 0 on the
// domain walls, else 1 - area of the solid part of the face from its four nodal
// solid values, raised to eps_solid when positive but smaller.
__host__ __device__ inline double face_area(const Dims &d, const float *solid, int dim, int i, int j,
                                            int k, double eps_solid) {
#pragma clang fp contract(off)
    const int pi = dim == 0 ? i : (dim == 1 ? j : k);
    const int n = dim == 0 ? d.nx : (dim == 1 ? d.ny : d.nz);
    double area;
    if (pi == 0 || pi == n) {
        area = 0.0;
    } else {
        double q00, q10, q11, q01;
        if (dim == 0) {
            q00 = solid[node_index(d, i, j, k)];
            q10 = solid[node_index(d, i, j + 1, k)];
            q11 = solid[node_index(d, i, j + 1, k + 1)];
            q01 = solid[node_index(d, i, j, k + 1)];
        } else if (dim == 1) {
            q00 = solid[node_index(d, i, j, k)];
            q10 = solid[node_index(d, i + 1, j, k)];
            q11 = solid[node_index(d, i + 1, j, k + 1)];
            q01 = solid[node_index(d, i, j, k + 1)];
        } else {
            q00 = solid[node_index(d, i, j, k)];
            q10 = solid[node_index(d, i + 1, j, k)];
            q11 = solid[node_index(d, i + 1, j + 1, k)];
            q01 = solid[node_index(d, i, j + 1, k)];
        }
        area = 1.0 - get_area(q00, q10, q11, q01);
    }
    if (area != 0.0 && area < eps_solid) area = eps_solid;
    return area;
}

// Open solid area of a face of a 2D grid held as nx x 1 x ny (macutility2.cpp:104-111):
// 0 on the domain walls, else 1 - fraction of the face's two end nodes of the 2D
// nodal solid ((nx+1) x (ny+1), index i + (nx+1) j), raised to eps_solid when
// positive but smaller.  dim is 0 (x faces) or 2 (the 2D y faces).
__host__ __device__ inline double face_area_2d(const Dims &d, const float *solid2, int dim, int i, int k,
                                               double eps_solid) {
#pragma clang fp contract(off)
    const int pi = dim == 0 ? i : k;
    const int n = dim == 0 ? d.nx : d.nz;
    double area;
    if (pi == 0 || pi == n) {
        area = 0.0;
    } else {
        const int64_t a = (int64_t)i + (int64_t)(d.nx + 1) * k;
        const int64_t b = dim == 0 ? a + (d.nx + 1) : a + 1;
        area = 1.0 - fraction(solid2[a], solid2[b]);
    }
    if (area != 0.0 && area < eps_solid) area = eps_solid;
    return area;
}

__host__ __device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Liquid fraction of face (dim, i, j, k) (macutility3.cpp:183-191): fraction of
// the two adjacent cells' level set with indices clamped into the grid
// (shape.h:808-816), raised to eps_fluid when positive but smaller.
__host__ __device__ inline double face_rho(const Dims &d, const float *fluid, int dim, int i, int j,
                                           int k, double eps_fluid) {
#pragma clang fp contract(off)
    const int i0 = clampi(i, 0, d.nx - 1), j0 = clampi(j, 0, d.ny - 1), k0 = clampi(k, 0, d.nz - 1);
    const int i1 = clampi(i - (dim == 0), 0, d.nx - 1), j1 = clampi(j - (dim == 1), 0, d.ny - 1),
              k1 = clampi(k - (dim == 2), 0, d.nz - 1);
    double rho = fraction(fluid[cell_index(d, i0, j0, k0)], fluid[cell_index(d, i1, j1, k1)]);
    if (rho != 0.0 && rho < eps_fluid) rho = eps_fluid;
    return rho;
}

// Neighbour table of macpressuresolver3.cpp:139-141 and 184-187 in its order
// (+x, -x, +y, -y, +z, -z): the neighbour's offset, the face's offset, its axis
// and the outward sign.
__host__ __device__ inline void neighbour(int nq, int &qi, int &qj, int &qk, int &fi, int &fj, int &fk,
                                          int &dim, int &sgn) {
    dim = nq >> 1;
    const int plus = (nq & 1) == 0;
    sgn = plus ? 1 : -1;
    qi = (dim == 0) ? (plus ? 1 : -1) : 0;
    qj = (dim == 1) ? (plus ? 1 : -1) : 0;
    qk = (dim == 2) ? (plus ? 1 : -1) : 0;
    fi = (dim == 0 && plus) ? 1 : 0;
    fj = (dim == 1 && plus) ? 1 : 0;
    fk = (dim == 2 && plus) ? 1 : 0;
}

}  // namespace shk
