// liquid.hip -- the velocity half of the liquid step around the pressure projection
// (src/liquid/macliquid3.cpp:342-391) on gfx950: velocity extrapolation, the solid
// constraint, max |u| for the time step, MacCormack self-advection, gravity and the
// activity left by the projection.
//
// The reference keeps the MAC velocity in sparse arrays with an activity flag per
// face.  Here the faces are dense float arrays in the caller's x-fastest layout with
// one uint8 activity mask per axis (1 = active): HBM holds every face of a 512^3 grid
// several times over, so a dense pass with masked reads replaces the tile bookkeeping,
// and every kernel is one thread per face (or cell), coalesced along x.  An inactive
// face reads as the background value 0, exactly like a sparse array's unset cell.
//
// Every expression rounds as the reference's (Real = float fields, vec3r float
// vectors, float interpolation accumulators; vec.h:208-296, array_interpolator3.h:
// 50-100): fp contraction is off and float quotients are taken in double and rounded
// once, which equals the correctly rounded float quotient.  The results are bit-for-
// bit those of the reference's arithmetic on the same inputs.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <initializer_list>

#include "ctx_internal.h"
#include "shk_common.h"

namespace shk {

struct LiquidWork {
    DevBuf<float> a[3], b[3], lo[3], hi[3];
    DevBuf<uint8_t> band[3];
    DevBuf<unsigned long long> count;
    DevBuf<unsigned int> slots;
    DevBuf<double> part;
    std::vector<double> h_part;
    void release() {
        for (int d = 0; d < 3; ++d) {
            a[d].release();
            b[d].release();
            lo[d].release();
            hi[d].release();
            band[d].release();
        }
        count.release();
        slots.release();
        part.release();
    }
};

void liquid_destroy(LiquidWork *w) {
    if (!w) return;
    w->release();
    delete w;
}

namespace {

constexpr int kNT = 256;
constexpr int kMaxPartials = 2048;

struct FS {  // extents of one face (or cell) array
    int nx, ny, nz;
    __host__ __device__ int64_t n() const { return (int64_t)nx * ny * nz; }
    __host__ __device__ int64_t at(int i, int j, int k) const {
        return (int64_t)i + (int64_t)nx * ((int64_t)j + (int64_t)ny * k);
    }
};

__host__ __device__ inline FS face_fs(const Dims &d, int dim) {
    return FS{d.nx + (dim == 0), d.ny + (dim == 1), d.nz + (dim == 2)};
}

struct Faces {  // the three face arrays of one MAC field and their masks
    float *v[3];
    const uint8_t *m[3];
};
struct CFaces {
    const float *v[3];
    const uint8_t *m[3];
};

// both loads are issued unconditionally (the index is in bounds either way), so a
// masked read costs one memory latency, not two dependent ones
__device__ inline float rd(const float *a, const uint8_t *m, int64_t q) {
    const float v = a[q];
    const uint8_t act = m[q];
    return act ? v : 0.0f;
}

// (i, j, k) of a linear index with 32-bit unsigned arithmetic (shk_create bounds the
// cells by 2^31, so every face array has fewer than 2^32 entries)
// Exact division of a 32-bit index by an extent n without a divide sequence:
// floor(q / n) = floor(m q / 2^64) with m = ceil(2^64 / n) for every 32-bit q and
// 2 <= n < 2^32 (Lemire, Kaser and Kurz, "Faster remainder by direct computation",
// 2019).  m is computed on the host; n = 1 takes q itself.
struct Div32 {
    uint32_t n;
    uint64_t m;
};
inline Div32 div32(int n) { return Div32{(uint32_t)n, n > 1 ? ~0ull / (uint64_t)n + 1 : 0ull}; }
__device__ inline uint32_t quot(uint32_t q, Div32 d) {
    return d.n > 1 ? (uint32_t)__umul64hi(d.m, (uint64_t)q) : q;
}

// the x and y extents of an index space (a face or cell array), for decompose
struct Split {
    Div32 x, y;
};
inline Split split_of(FS s) { return Split{div32(s.nx), div32(s.ny)}; }
struct Split3 {  // the three face arrays
    Split a[3];
};
inline Split cells_split(const Dims &d) { return split_of(FS{d.nx, d.ny, d.nz}); }
inline Split3 faces_split(const Dims &d) {
    return Split3{{split_of(face_fs(d, 0)), split_of(face_fs(d, 1)), split_of(face_fs(d, 2))}};
}

__device__ inline void decompose(int64_t q, Split sp, int &i, int &j, int &k) {
    const uint32_t u = (uint32_t)q;
    const uint32_t t = quot(u, sp.x);
    i = (int)(u - t * sp.x.n);
    k = (int)quot(t, sp.y);
    j = (int)(t - (uint32_t)k * sp.y.n);
}

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

inline unsigned blocks_for(int64_t n) {
    int64_t g = (n + kNT - 1) / kNT;
    return (unsigned)(g > 0 ? g : 1);
}

// XCD-aware block order for the extrapolation stencil.  Workgroup b of a launch goes
// to XCD b % 8, each XCD with its own L2; the z-neighbours of a block's faces lie about
// one plane (nx*ny faces, hundreds of blocks) away, so the launch gives each XCD one
// contiguous range of blocks instead (the advection's short gathers measured slower
// with it and keep the plain order).
// Launch xcd_blocks(n) blocks; a block whose logical index is past the last one exits.
constexpr int kXcds = 8;
inline unsigned xcd_blocks(int64_t n) {
    const int64_t per = ((int64_t)blocks_for(n) + kXcds - 1) / kXcds;
    return (unsigned)(per * kXcds);
}
__device__ inline int64_t xcd_block() {
    const uint32_t per = gridDim.x / kXcds;
    return (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

// array_interpolator3::interpolate_coef (array_interpolator3.h:50-72)
__device__ inline void interp_coef(FS s, const double p[3], int &i, int &j, int &k, double coef[8]) {
#pragma clang fp contract(off)
    const double x = fmax(0.0, fmin(s.nx - 1., p[0]));
    const double y = fmax(0.0, fmin(s.ny - 1., p[1]));
    const double z = fmax(0.0, fmin(s.nz - 1., p[2]));
    i = (int)fmin(x, s.nx - 2.);
    j = (int)fmin(y, s.ny - 2.);
    k = (int)fmin(z, s.nz - 2.);
    coef[0] = (k + 1 - z) * (i + 1 - x) * (j + 1 - y);
    coef[1] = (k + 1 - z) * (x - i) * (j + 1 - y);
    coef[2] = (k + 1 - z) * (i + 1 - x) * (y - j);
    coef[3] = (k + 1 - z) * (x - i) * (y - j);
    coef[4] = (z - k) * (i + 1 - x) * (j + 1 - y);
    coef[5] = (z - k) * (x - i) * (j + 1 - y);
    coef[6] = (z - k) * (i + 1 - x) * (y - j);
    coef[7] = (z - k) * (x - i) * (y - j);
}

__device__ inline int64_t corner(FS s, int i, int j, int k, int n) {
    return s.at(i + (n & 1), j + ((n >> 1) & 1), k + ((n >> 2) & 1));
}

// array_interpolator3::interpolate<float>: float accumulation of the non-zero terms
// (array_interpolator3.h:86-99); m == nullptr reads every value as stored
__device__ inline float interp(const float *a, const uint8_t *m, FS s, const double p[3]) {
#pragma clang fp contract(off)
    int i, j, k;
    double coef[8];
    interp_coef(s, p, i, j, k, coef);
    float value = 0.0f;
#pragma unroll
    for (int n = 0; n < 8; ++n)
        if (coef[n] != 0.0) {
            const int64_t q = corner(s, i, j, k, n);
            const float v = m ? rd(a, m, q) : a[q];
            value = (float)((double)value + (double)v * coef[n]);
        }
    return value;
}

// The 6th-order WENO interpolation at x in [0, 1] of samples f at -2, -1, 0, 1, 2, 3
// (Macdonald and Ruuth, J. Sci. Comput. 35 (2008) 219-240; the reference's WENO::interp6,
// include/shiokaze/math/WENO.h), each operation in the order written there.
__device__ inline double weno6(double x, const double f[6]) {
#pragma clang fp contract(off)
    const double fm2 = f[0], fm1 = f[1], f0 = f[2], f1 = f[3], f2 = f[4], f3 = f[5];
    const double c[3] = {(2.0 - x) * (3.0 - x) / 20.0, (3.0 - x) * (x - -2.0) / 10.0,
                         (x - -2.0) * (x - -1.0) / 20.0};
    const double s[3] = {
        ((814. * (f1 * f1)) + (4326. * (f0 * f0)) + (2976. * (fm1 * fm1)) + (244. * (fm2 * fm2)) - (3579. * f0 * f1) -
         (6927. * f0 * fm1) + (1854. * f0 * fm2) + (2634. * f1 * fm1) - (683. * f1 * fm2) - (1659. * fm1 * fm2)) /
            180.0,
        ((1986. * (f1 * f1)) + (1986. * (f0 * f0)) + (244. * (fm1 * fm1)) + (244. * (f2 * f2)) + (1074. * f0 * f2) -
         (3777. * f0 * f1) - (1269. * f0 * fm1) + (1074. * f1 * fm1) - (1269. * f2 * f1) - (293. * f2 * fm1)) /
            180.0,
        ((814. * (f0 * f0)) + (4326. * (f1 * f1)) + (2976. * (f2 * f2)) + (244. * (f3 * f3)) - (683. * f0 * f3) +
         (2634. * f0 * f2) - (3579. * f0 * f1) - (6927. * f1 * f2) + (1854. * f1 * f3) - (1659. * f2 * f3)) /
            180.0};
    const double p[3] = {
        fm2 + (fm1 - fm2) * (x - -2.0) + (f0 - 2. * fm1 + fm2) * (x - -2.0) * (x - -1.0) / 2.0 +
            (f1 - 3. * f0 + 3. * fm1 - fm2) * (x - -2.0) * (x - -1.0) * (x - 0.0) / 6.0,
        fm1 + (f0 - fm1) * (x - -1.0) + (f1 - 2. * f0 + fm1) * (x - -1.0) * (x - 0.0) / 2.0 +
            (f2 - 3. * f1 + 3. * f0 - fm1) * (x - -1.0) * (x - 0.0) * (x - 1.0) / 6.0,
        f0 + (f1 - f0) * (x - 0.0) + (f2 - 2. * f1 + f0) * (x - 0.0) * (x - 1.0) / 2.0 +
            (f3 - 3. * f2 + 3. * f1 - f0) * (x - 0.0) * (x - 1.0) * (x - 2.0) / 6.0};
    double a[3], sum = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        a[q] = c[q] / (DBL_EPSILON + s[q] * s[q]);
        sum += a[q];
    }
    return a[0] / sum * p[0] + a[1] / sum * p[1] + a[2] / sum * p[2];
}

// WENO3::interpolate(array, p), order 6 (WENO3.h:92-122): the position clamped into the
// array, then interp6 along x for each of the 6 x 6 rows around it, along y, along z,
// every read clamped to the array (m == nullptr reads every value as stored)
__device__ inline double weno3(const float *a, const uint8_t *m, FS s, const double p[3]) {
#pragma clang fp contract(off)
    const double x = fmax(0.0, fmin(s.nx - 1., p[0]));
    const double y = fmax(0.0, fmin(s.ny - 1., p[1]));
    const double z = fmax(0.0, fmin(s.nz - 1., p[2]));
    const int i = (int)fmin(x, s.nx - 2.), j = (int)fmin(y, s.ny - 2.), k = (int)fmin(z, s.nz - 2.);
    double vvv[6];
    for (int kk = 0; kk < 6; ++kk) {
        const int kc = min(max(k + kk - 2, 0), s.nz - 1);
        double vv[6];
        for (int jj = 0; jj < 6; ++jj) {
            const int jc = min(max(j + jj - 2, 0), s.ny - 1);
            double v[6];
#pragma unroll
            for (int ii = 0; ii < 6; ++ii) {
                const int64_t q = s.at(min(max(i + ii - 2, 0), s.nx - 1), jc, kc);
                v[ii] = m ? (double)rd(a, m, q) : (double)a[q];
            }
            vv[jj] = weno6(x - i, v);
        }
        vvv[kk] = weno6(y - j, vv);
    }
    return weno6(z - k, vvv);
}

// macarray3::convert_to_full for one face (macarray3.h:372-393): its own component
// and the clamped 4-face averages of the other two, each stored as float
__device__ inline void full_velocity(const Dims &d, const CFaces &f, int dim, int i, int j, int k, float out[3]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int ud = 0; ud < 3; ++ud) {
        const FS s = face_fs(d, ud);
        if (ud == dim) {
            out[ud] = rd(f.v[ud], f.m[ud], s.at(i, j, k));
            continue;
        }
        double acc = 0.0;
#pragma unroll
        for (int ii = 0; ii < 2; ++ii)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) {
                int q[3] = {i - (dim == 0), j - (dim == 1), k - (dim == 2)};
                q[dim] += ii;
                q[ud] += jj;
                acc += rd(f.v[ud], f.m[ud], s.at(clampi(q[0], 0, s.nx - 1), clampi(q[1], 0, s.ny - 1),
                                                 clampi(q[2], 0, s.nz - 1)));
            }
        acc /= 4.0;
        out[ud] = (float)acc;
    }
}

// ---------------------------------------------------------------- extrapolation

// One dilation round of array_extrapolator3::extrapolate (array_extrapolator3.h:47-76,
// dilate3.h:43-97), in place: an inactive face with active in-bounds 6-neighbours
// takes their float average, summed in the order +x, -x, +y, -y, -z, +z.  The round
// must see only what was active before it (the reference collects every candidate and
// sets them together), so a face activated in round r stores r + 1 in its mask and a
// neighbour counts in round r when its mask is in [1, r]; a face that the same round
// activates reads as inactive whether its write has landed or not, and the values
// read are final.  Only inactive faces do work; active ones read one byte.  A last
// pass turns the round numbers back into 1.  blockIdx.y is the axis.
__global__ __launch_bounds__(kNT) void k_extrapolate_round(Dims d, Split3 sp, Faces f, uint8_t *m0, uint8_t *m1, uint8_t *m2,
                                                            int round, unsigned int *slots) {
#pragma clang fp contract(off)
    const int dim = blockIdx.y;
    const FS s = face_fs(d, dim);
    const int64_t q = xcd_block() * kNT + threadIdx.x;
    uint8_t *m = dim == 0 ? m0 : (dim == 1 ? m1 : m2);
    bool activated = false;
    if (q < s.n()) {
        float *a = f.v[dim];
        int i, j, k;
        decompose(q, sp.a[dim], i, j, k);
        // an active face loads nothing more: the round is bound by the rate of byte
        // loads, not by the latency of the two-step chain
        const uint8_t own = m[q];
        const int di[6] = {1, -1, 0, 0, 0, 0}, dj[6] = {0, 0, 1, -1, 0, 0}, dk[6] = {0, 0, 0, 0, -1, 1};
        int64_t e[6];
        bool use[6];
        bool any = false;
#pragma unroll
        for (int t = 0; t < 6; ++t) {   // the six neighbour masks, loaded together
            const int ii = i + di[t], jj = j + dj[t], kk = k + dk[t];
            const bool in = ii >= 0 && jj >= 0 && kk >= 0 && ii < s.nx && jj < s.ny && kk < s.nz;
            e[t] = s.at(clampi(ii, 0, s.nx - 1), clampi(jj, 0, s.ny - 1), clampi(kk, 0, s.nz - 1));
            use[t] = false;
        }
        if (own == 0) {
#pragma unroll
            for (int t = 0; t < 6; ++t) {
                const int ii = i + di[t], jj = j + dj[t], kk = k + dk[t];
                const bool in = ii >= 0 && jj >= 0 && kk >= 0 && ii < s.nx && jj < s.ny && kk < s.nz;
                const uint8_t nm = m[e[t]];
                use[t] = in && nm != 0 && nm <= round;
                any = any || use[t];
            }
        }
        float sum = 0.0f;
        int weight = 0;
        if (own == 0 && any) {   // values only on the activation shell
#pragma unroll
            for (int t = 0; t < 6; ++t)
                if (use[t]) {
                    sum += a[e[t]];
                    ++weight;
                }
        }
        if (own == 0 && weight) {
            a[q] = (float)((double)sum / (double)weight);
            m[q] = (uint8_t)(round + 1);
            activated = true;
        }
    }
    // the block's count goes to its own slot (rounds run one after another, so no two
    // threads ever update a slot at once): no contended atomics on one counter
    __shared__ unsigned int wave_count[kNT / 64];
    const unsigned long long ballot = __ballot(activated);
    if ((threadIdx.x & 63) == 0) wave_count[threadIdx.x >> 6] = (unsigned int)__popcll(ballot);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int c = 0;
        for (int w = 0; w < kNT / 64; ++w) c += wave_count[w];
        if (c) slots[blockIdx.y * gridDim.x + blockIdx.x] += c;
    }
}

// total of the per-block slots: grid-stride partial sums, one atomic per block
// (*total starts at 0)
__global__ __launch_bounds__(kNT) void k_sum_slots(int64_t n, const unsigned int *slots,
                                                   unsigned long long *total) {
    unsigned long long c = 0;
    for (int64_t q = blockIdx.x * (int64_t)kNT + threadIdx.x; q < n; q += (int64_t)gridDim.x * kNT) c += slots[q];
    __shared__ unsigned long long lds[kNT];
    lds[threadIdx.x] = c;
    __syncthreads();
    for (int o = kNT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) lds[threadIdx.x] += lds[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0 && lds[0]) atomicAdd(total, lds[0]);
}

// masks to 0 / 1 (before the rounds: any nonzero input is active; after them: the
// round tags), four per thread, uchar4 when aligned
__global__ __launch_bounds__(kNT) void k_mask_normalise(Dims d, uint8_t *m0, uint8_t *m1, uint8_t *m2, int vec) {
    const int dim = blockIdx.y;
    const int64_t n = face_fs(d, dim).n();
    const int64_t q4 = (blockIdx.x * (int64_t)kNT + threadIdx.x) * 4;
    if (q4 >= n) return;
    uint8_t *m = dim == 0 ? m0 : (dim == 1 ? m1 : m2);
    if (vec && q4 + 4 <= n) {
        const uchar4 v = *reinterpret_cast<const uchar4 *>(m + q4);
        if (v.x <= 1 && v.y <= 1 && v.z <= 1 && v.w <= 1) return;
        *reinterpret_cast<uchar4 *>(m + q4) = make_uchar4(v.x != 0, v.y != 0, v.z != 0, v.w != 0);
    } else {
        // this thread's own four faces (fewer at the end of the array)
        const int64_t end = q4 + 4 < n ? q4 + 4 : n;
        for (int64_t q = q4; q < end; ++q)
            if (m[q] > 1) m[q] = 1;
    }
}

// ---------------------------------------------------------------- max |u|

// macutility3::compute_max_u (macutility3.cpp:48-60): over cells with all six faces
// active, the squared length of the float cell velocity (float squares summed in
// double, vec.h:272-276); one partial max per block, the host takes the sqrt
__global__ __launch_bounds__(kNT) void k_max_u2(Dims d, Split sp, CFaces f, double *part) {
#pragma clang fp contract(off)
    const int64_t G = (int64_t)d.nx * d.ny * d.nz;
    double best = 0.0;
    for (int64_t c = blockIdx.x * (int64_t)kNT + threadIdx.x; c < G; c += (int64_t)gridDim.x * kNT) {
        int i, j, k;
        decompose(c, sp, i, j, k);
        float lo_v[3], hi_v[3];
        bool valid = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) {   // all twelve loads first
            const FS s = face_fs(d, a);
            const int64_t lo = s.at(i, j, k), hi = s.at(i + (a == 0), j + (a == 1), k + (a == 2));
            lo_v[a] = f.v[a][lo];
            hi_v[a] = f.v[a][hi];
            valid &= (f.m[a][lo] != 0) & (f.m[a][hi] != 0);
        }
        double n2 = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double value = 0.0;
            value += lo_v[a];
            value += hi_v[a];
            const float cv = (float)(0.5 * value);
            n2 += (double)(cv * cv);
        }
        if (valid) best = fmax(best, n2);
    }
    __shared__ double lds[kNT / 64];
    for (int o = 32; o > 0; o >>= 1) best = fmax(best, __shfl_down(best, o));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = lds[0];
        for (int w = 1; w < kNT / 64; ++w) b = fmax(b, lds[w]);
        part[blockIdx.x] = b;
    }
}

// ---------------------------------------------------------------- solid constraint

// macutility3::constrain_velocity (macutility3.cpp:61-92) for the faces of axis DIM:
// an active face inside the solid whose interpolated velocity (from the snapshot
// `save`) points into it loses the normal part; then the y axis's floor and ceiling
// rule.  `solid` is the nodal level set ((nx+1)(ny+1)(nz+1)).
template <int DIM>
__global__ __launch_bounds__(kNT) void k_constrain(Dims d, Split sp, const float *__restrict__ solid, double dx, CFaces save,
                                                   float *__restrict__ vel, const uint8_t *__restrict__ mask) {
#pragma clang fp contract(off)
    const FS s = face_fs(d, DIM);
    const int64_t q = blockIdx.x * (int64_t)kNT + threadIdx.x;
    if (q >= s.n() || !mask[q]) return;
    int i, j, k;
    decompose(q, sp, i, j, k);
    float out = vel[q];
    bool changed = false;
    if (solid) {
        const FS ns{d.nx + 1, d.ny + 1, d.nz + 1};
        const double p[3] = {i + 0.5 * (DIM != 0), j + 0.5 * (DIM != 1), k + 0.5 * (DIM != 2)};
        if (interp(solid, nullptr, ns, p) < 0.0f) {
            // array_derivative3::derivative<float>: the trilinear gradient in index units
            int ci, cj, ck;
            double w[8];
            interp_coef(ns, p, ci, cj, ck, w);  // clamps and cell as derivative_interpolate_coef
            const double x = fmax(0.0, fmin(ns.nx - 1., p[0]));
            const double y = fmax(0.0, fmin(ns.ny - 1., p[1]));
            const double z = fmax(0.0, fmin(ns.nz - 1., p[2]));
            const double gx[8] = {-(ck + 1 - z) * (cj + 1 - y), (ck + 1 - z) * (cj + 1 - y),
                                  -(ck + 1 - z) * (y - cj),     (ck + 1 - z) * (y - cj),
                                  -(z - ck) * (cj + 1 - y),     (z - ck) * (cj + 1 - y),
                                  -(z - ck) * (y - cj),         (z - ck) * (y - cj)};
            const double gy[8] = {-(ck + 1 - z) * (ci + 1 - x), -(ck + 1 - z) * (x - ci),
                                  (ck + 1 - z) * (ci + 1 - x),  (ck + 1 - z) * (x - ci),
                                  -(z - ck) * (ci + 1 - x),     -(z - ck) * (x - ci),
                                  (z - ck) * (ci + 1 - x),      (z - ck) * (x - ci)};
            const double gz[8] = {-(ci + 1 - x) * (cj + 1 - y), -(x - ci) * (cj + 1 - y),
                                  -(ci + 1 - x) * (y - cj),     -(x - ci) * (y - cj),
                                  (ci + 1 - x) * (cj + 1 - y),  (x - ci) * (cj + 1 - y),
                                  (ci + 1 - x) * (y - cj),      (x - ci) * (y - cj)};
            float g[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int n = 0; n < 8; ++n) {
                const double v = (double)solid[corner(ns, ci, cj, ck, n)];
                g[0] = (float)((double)g[0] + gx[n] * v);
                g[1] = (float)((double)g[1] + gy[n] * v);
                g[2] = (float)((double)g[2] + gz[n] * v);
            }
            const double nrm[3] = {(double)g[0] / dx, (double)g[1] / dx, (double)g[2] / dx};
            double n2 = 0.0;
#pragma unroll
            for (int a = 0; a < 3; ++a) n2 += nrm[a] * nrm[a];
            if (n2 != 0.0) {
                double uu[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double pp[3] = {p[0] - 0.5 * (a != 0), p[1] - 0.5 * (a != 1), p[2] - 0.5 * (a != 2)};
                    uu[a] = (double)interp(save.v[a], save.m[a], face_fs(d, a), pp);
                }
                double un = 0.0;
#pragma unroll
                for (int a = 0; a < 3; ++a) un += uu[a] * nrm[a];
                if (un < 0.0) {
                    out = (float)(uu[DIM] - nrm[DIM] * un);
                    changed = true;
                }
            }
        }
    }
    if (DIM == 1) {
        if (j == 0 && out < 0.0f) {
            out = 0.0f;
            changed = true;
        }
        if (j == d.ny && out > 0.0f) {
            out = 0.0f;
            changed = true;
        }
    }
    if (changed) vel[q] = out;
}

// ---------------------------------------------------------------- advection

// Semi-Lagrangian advection of the faces of axis DIM (macadvection3.cpp:70-138) with
// the departure velocity of `src` itself: out = interpolation of src at the departure
// point (the vec3r arithmetic of `dt*u/m_dx` rounds to float twice).  With lo != 0 it
// also keeps the MacCormack limiter's bounds (min and max of the 8 corners of the
// departure point computed in double; the max starts at DBL_MIN, which rounds to
// 0.0f) and the narrow-band flag (fluid at the departed face centre above -trim dx).
// With fin != 0 it is the backward pass and combines the result straight into the
// final MacCormack value (macadvection3.cpp:150-163).
// WENO is a template parameter: the trilinear instantiation keeps the register budget
// (and occupancy) it had without the WENO code
template <int DIM, bool WENO>
__global__ __launch_bounds__(kNT) void k_semi_lagrangian(Dims d, Split sp, double dx, double dt, CFaces src, float *out,
                                                         float *lo, float *hi, uint8_t *band, const float *fluid,
                                                         int trim, const float *fin_vin, const float *fin_lo,
                                                         const float *fin_hi, const uint8_t *fin_band,
                                                         const float *fin_v0, float *fin_out) {
#pragma clang fp contract(off)
    const FS s = face_fs(d, DIM);
    const int64_t q = blockIdx.x * (int64_t)kNT + threadIdx.x;
    if (q >= s.n()) return;
    if (!src.m[DIM][q]) {
        if (out) out[q] = 0.0f;
        if (fin_out) fin_out[q] = 0.0f;
        return;
    }
    int i, j, k;
    decompose(q, sp, i, j, k);
    float uf[3];
    full_velocity(d, src, DIM, i, j, k, uf);
    const bool empty = uf[0] == 0.0f && uf[1] == 0.0f && uf[2] == 0.0f;
    const float here = src.v[DIM][q];
    float value = here;
    // the 8 corner values of the departure cell, kept for the limiter below
    int ci = 0, cj = 0, ck = 0;
    float vals[8];
    if (!empty) {
        const int ijk[3] = {i, j, k};
        double p[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float t = (float)((double)uf[a] * dt);
            t = (float)((double)t / dx);
            p[a] = (double)ijk[a] - (double)t;
        }
        double coef[8];
        interp_coef(s, p, ci, cj, ck, coef);
        float acc = 0.0f;
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            vals[n] = rd(src.v[DIM], src.m[DIM], corner(s, ci, cj, ck, n));
            if (coef[n] != 0.0) acc = (float)((double)acc + (double)vals[n] * coef[n]);
        }
        // WENO=1: the value by WENO3 interpolation (macadvection3.cpp:85-86); the corners
        // above still bound the MacCormack limiter
        value = WENO ? (float)weno3(src.v[DIM], src.m[DIM], s, p) : acc;
    }
    if (out) out[q] = value;
    if (lo) {
        double face_p[3] = {i + 0.5 * (DIM != 0), j + 0.5 * (DIM != 1), k + 0.5 * (DIM != 2)};
        if (!empty) {
            const int ijk[3] = {i, j, k};
            double p[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double disp = dt * (double)uf[a] / dx;
                p[a] = (double)ijk[a] - disp;
                face_p[a] = face_p[a] - disp;
            }
            int di, dj, dk;
            double coef[8];
            interp_coef(s, p, di, dj, dk, coef);
            const bool same = di == ci && dj == cj && dk == ck;   // almost always: reuse the loads
            double mn = DBL_MAX, mx = DBL_MIN;
#pragma unroll
            for (int n = 0; n < 8; ++n) {
                const double v = same ? (double)vals[n] : (double)rd(src.v[DIM], src.m[DIM], corner(s, di, dj, dk, n));
                mn = fmin(mn, v);
                mx = fmax(mx, v);
            }
            lo[q] = (float)mn;
            hi[q] = (float)mx;
        } else {
            lo[q] = here;
            hi[q] = here;
        }
        const double fp[3] = {face_p[0] - 0.5, face_p[1] - 0.5, face_p[2] - 0.5};
        band[q] = (double)interp(fluid, nullptr, FS{d.nx, d.ny, d.nz}, fp) > -dx * (double)trim;
    }
    if (fin_out) {
        float r;
        if (fin_band[q]) {
            r = fin_v0[q];
        } else {
            const double mn = fin_lo[q], mx = fin_hi[q];
            const double vel0 = fin_v0[q];
            const double corr = 0.5 * ((double)fin_vin[q] - (double)value);
            if (vel0 + corr < mn)
                r = (float)mn;
            else if (vel0 + corr > mx)
                r = (float)mx;
            else
                r = (float)(vel0 + corr);
        }
        fin_out[q] = r;
    }
}

// ---------------------------------------------------------------- level-set advection

// macarray3::convert_to_full, cell version (macarray3.h:327-365): the float cell
// velocity of a cell whose six faces are active, else the empty (zero) vector;
// returns whether it moves
__device__ inline bool cell_velocity(const Dims &d, const CFaces &f, int i, int j, int k, double u[3]) {
#pragma clang fp contract(off)
    float lo_v[3], hi_v[3];
    bool valid = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const FS s = face_fs(d, a);
        const int64_t lo = s.at(i, j, k), hi = s.at(i + (a == 0), j + (a == 1), k + (a == 2));
        lo_v[a] = f.v[a][lo];
        hi_v[a] = f.v[a][hi];
        valid &= (f.m[a][lo] != 0) & (f.m[a][hi] != 0);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double value = 0.0;
        value += lo_v[a];
        value += hi_v[a];
        u[a] = valid ? (double)(float)(0.5 * value) : 0.0;
    }
    return u[0] != 0.0 || u[1] != 0.0 || u[2] != 0.0;
}

// MacCormack advection of the active cells of a cell-centred level set by the MAC
// velocity (macadvection3.cpp:178-277, advect_scalar as the level-set tracker calls
// it).  Forward pass (lo != 0): out = semi-Lagrangian value (vec3d departure), the
// limiter's bounds (min from FLT_MAX, max from FLT_MIN) and the narrow-band flag of
// the level set before the step at the departure point.  Backward pass (fin_out !=
// 0): the semi-Lagrangian value of q0 over -dt, combined into the final value.
// Inactive cells of the output are left alone (they keep what they read).
template <bool WENO>
__global__ __launch_bounds__(kNT) void k_advect_cell(Dims d, Split sp, double dx, double dt, const float *q, const uint8_t *mq,
                                                     CFaces vel, const float *fluid, int trim, float *out, float *lo,
                                                     float *hi, uint8_t *band, const float *fin_qin,
                                                     const float *fin_lo, const float *fin_hi,
                                                     const uint8_t *fin_band, const float *fin_q0, float *fin_out,
                                                     int fill_inactive, float background) {
#pragma clang fp contract(off)
    const FS cs{d.nx, d.ny, d.nz};
    const int64_t c = blockIdx.x * (int64_t)kNT + threadIdx.x;
    if (c >= cs.n()) return;
    if (!mq[c]) {
        // q_0 is a new array of the level set's type (macadvection3.cpp:248): outside
        // the band it reads the background value
        if (fill_inactive) out[c] = background;
        return;
    }
    int i, j, k;
    decompose(c, sp, i, j, k);
    double u[3];
    const bool moving = cell_velocity(d, vel, i, j, k, u);
    double p[3] = {(double)i, (double)j, (double)k};
    float value = q[c];
    int ci = 0, cj = 0, ck = 0;
    float vals[8];
    if (moving) {
#pragma unroll
        for (int a = 0; a < 3; ++a) p[a] = p[a] - dt * u[a] / dx;
        double coef[8];
        interp_coef(cs, p, ci, cj, ck, coef);
        float acc = 0.0f;
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            vals[n] = q[corner(cs, ci, cj, ck, n)];
            if (coef[n] != 0.0) acc = (float)((double)acc + (double)vals[n] * coef[n]);
        }
        value = WENO ? (float)weno3(q, nullptr, cs, p) : acc;  // WENO=1 (macadvection3.cpp:193-194)
    }
    if (out) out[c] = value;
    if (lo) {
        if (moving) {
            double mn = FLT_MAX, mx = FLT_MIN;
#pragma unroll
            for (int n = 0; n < 8; ++n) {
                mn = fmin(mn, (double)vals[n]);
                mx = fmax(mx, (double)vals[n]);
            }
            lo[c] = (float)mn;
            hi[c] = (float)mx;
            band[c] = (double)interp(fluid, nullptr, cs, p) > -dx * (double)trim;
        } else {
            lo[c] = q[c];
            hi[c] = q[c];
            band[c] = (double)fluid[c] > -dx * (double)trim;
        }
    }
    if (fin_out) {
        float r;
        if (fin_band[c]) {
            r = fin_q0[c];
        } else {
            const double mn = fin_lo[c], mx = fin_hi[c], q0 = fin_q0[c];
            const double corr = 0.5 * ((double)fin_qin[c] - (double)value);
            if (q0 + corr < mn)
                r = (float)mn;
            else if (q0 + corr > mx)
                r = (float)mx;
            else
                r = (float)(q0 + corr);
        }
        fin_out[c] = r;
    }
}

// ---------------------------------------------------------------- gravity, activity

// `m_velocity += dt*gravity` (macliquid3.cpp:207-208): array3::operator+= touches every
// cell and array3::increment activates an inactive one with background + value
// (array3.h:647-655, 917-921); the vec3d increment is converted to float first
__global__ __launch_bounds__(kNT) void k_gravity(Dims d, float g0, float g1, float g2, Faces f, uint8_t *m0,
                                                 uint8_t *m1, uint8_t *m2, int vec) {
#pragma clang fp contract(off)
    // four faces per thread (float4 / uchar4 of the flat arrays when every array is
    // 16-byte aligned), the last partial group element by element
    const int dim = blockIdx.y;
    const int64_t n = face_fs(d, dim).n();
    const int64_t q4 = (blockIdx.x * (int64_t)kNT + threadIdx.x) * 4;
    if (q4 >= n) return;
    const float add = dim == 0 ? g0 : (dim == 1 ? g1 : g2);
    uint8_t *m = dim == 0 ? m0 : (dim == 1 ? m1 : m2);
    float *v = f.v[dim];
    if (vec && q4 + 4 <= n) {
        float4 val = *reinterpret_cast<const float4 *>(v + q4);
        const uchar4 act = *reinterpret_cast<const uchar4 *>(m + q4);
        val.x = act.x ? val.x + add : 0.0f + add;
        val.y = act.y ? val.y + add : 0.0f + add;
        val.z = act.z ? val.z + add : 0.0f + add;
        val.w = act.w ? val.w + add : 0.0f + add;
        *reinterpret_cast<float4 *>(v + q4) = val;
        *reinterpret_cast<uchar4 *>(m + q4) = make_uchar4(1, 1, 1, 1);
    } else {
        // this thread's own four faces (fewer at the end of the array): every face
        // is incremented exactly once
        const int64_t end = q4 + 4 < n ? q4 + 4 : n;
        for (int64_t q = q4; q < end; ++q) {
            v[q] = m[q] ? v[q] + add : 0.0f + add;
            m[q] = 1;
        }
    }
}

// The projection's set_off (macpressuresolver3.cpp:269-283): an active face stays
// active only where both its area and liquid fractions are non-zero.
__global__ __launch_bounds__(kNT) void k_project_activity(Dims d, const float *a0, const float *a1, const float *a2,
                                                          const float *r0, const float *r1, const float *r2,
                                                          uint8_t *m0, uint8_t *m1, uint8_t *m2, int vec) {
    const int dim = blockIdx.y;
    const int64_t n = face_fs(d, dim).n();
    const int64_t q4 = (blockIdx.x * (int64_t)kNT + threadIdx.x) * 4;
    if (q4 >= n) return;
    const float *a = dim == 0 ? a0 : (dim == 1 ? a1 : a2);
    const float *r = dim == 0 ? r0 : (dim == 1 ? r1 : r2);
    uint8_t *m = dim == 0 ? m0 : (dim == 1 ? m1 : m2);
    if (vec && q4 + 4 <= n) {
        uchar4 act = *reinterpret_cast<const uchar4 *>(m + q4);
        if (!(act.x | act.y | act.z | act.w)) return;
        const float4 av = *reinterpret_cast<const float4 *>(a + q4);
        const float4 rv = *reinterpret_cast<const float4 *>(r + q4);
        act.x = act.x && av.x != 0.0f && rv.x != 0.0f;
        act.y = act.y && av.y != 0.0f && rv.y != 0.0f;
        act.z = act.z && av.z != 0.0f && rv.z != 0.0f;
        act.w = act.w && av.w != 0.0f && rv.w != 0.0f;
        *reinterpret_cast<uchar4 *>(m + q4) = act;
    } else {
        const int64_t end = q4 + 4 < n ? q4 + 4 : n;
        for (int64_t q = q4; q < end; ++q)
            if (m[q] && !(a[q] != 0.0f && r[q] != 0.0f)) m[q] = 0;
    }
}

int64_t max_faces(const Dims &d) {
    int64_t n = 0;
    for (int a = 0; a < 3; ++a) n = face_fs(d, a).n() > n ? face_fs(d, a).n() : n;
    return n;
}

// ---------------------------------------------------------------- level set into solids
// gridutility3::extrapolate_levelset(solid, fluid) (gridutility3.cpp:87-146), in four
// passes over the cells: whether the solid has an active node below 0
// (array_utility3.h:112-122), the extrapolation of the band cells inside the solid from
// a snapshot of phi, one dilation of the band, and combine_levelset with the offset dx
// (:66-86; convert_to_cell :48-65).

__global__ void k_exls_exist(int64_t nn, const float *__restrict__ solid, const uint8_t *__restrict__ sband,
                             unsigned long long *__restrict__ flag) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nn; q += (int64_t)gridDim.x * blockDim.x)
        if ((!sband || sband[q]) && solid[q] < 0.0f) *flag = 1ull;  // every writer stores the same value
}

// array_derivative3::derivative<float> (array_derivative3.h:51-122): the trilinear
// gradient in index units, float accumulation per axis
__device__ inline void derivative_idx(const float *a, FS s, const double p[3], float g[3]) {
#pragma clang fp contract(off)
    int ci, cj, ck;
    double w[8];
    interp_coef(s, p, ci, cj, ck, w);  // clamps and cell as derivative_interpolate_coef
    const double x = fmax(0.0, fmin(s.nx - 1., p[0]));
    const double y = fmax(0.0, fmin(s.ny - 1., p[1]));
    const double z = fmax(0.0, fmin(s.nz - 1., p[2]));
    const double gx[8] = {-(ck + 1 - z) * (cj + 1 - y), (ck + 1 - z) * (cj + 1 - y),
                          -(ck + 1 - z) * (y - cj),     (ck + 1 - z) * (y - cj),
                          -(z - ck) * (cj + 1 - y),     (z - ck) * (cj + 1 - y),
                          -(z - ck) * (y - cj),         (z - ck) * (y - cj)};
    const double gy[8] = {-(ck + 1 - z) * (ci + 1 - x), -(ck + 1 - z) * (x - ci),
                          (ck + 1 - z) * (ci + 1 - x),  (ck + 1 - z) * (x - ci),
                          -(z - ck) * (ci + 1 - x),     -(z - ck) * (x - ci),
                          (z - ck) * (ci + 1 - x),      (z - ck) * (x - ci)};
    const double gz[8] = {-(ci + 1 - x) * (cj + 1 - y), -(x - ci) * (cj + 1 - y),
                          -(ci + 1 - x) * (y - cj),     -(x - ci) * (y - cj),
                          (ci + 1 - x) * (cj + 1 - y),  (x - ci) * (cj + 1 - y),
                          (ci + 1 - x) * (y - cj),      (x - ci) * (y - cj)};
    g[0] = g[1] = g[2] = 0.0f;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const double v = (double)a[corner(s, ci, cj, ck, n)];
        g[0] = (float)((double)g[0] + gx[n] * v);
        g[1] = (float)((double)g[1] + gy[n] * v);
        g[2] = (float)((double)g[2] + gz[n] * v);
    }
}

// vec::normalize (vec.h:311-320) and norm2 (vec.h:272-276)
__device__ inline double norm2_3(const double v[3]) {
#pragma clang fp contract(off)
    double r = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) r += v[d] * v[d];
    return r;
}
__device__ inline void normalize3(double v[3]) {
#pragma clang fp contract(off)
    const double len = sqrt(norm2_3(v));
    if (len != 0.0)
#pragma unroll
        for (int d = 0; d < 3; ++d) v[d] /= len;
}

// pass 2: the band cells inside the solid, from the snapshot `old`
__global__ __launch_bounds__(kNT) void k_exls_wall(Dims d, Split sp, double dx, const float *__restrict__ solid,
                                                   const float *__restrict__ old, const uint8_t *__restrict__ band,
                                                   float *__restrict__ phi, float background, int wall,
                                                   int horizontal) {
#pragma clang fp contract(off)
    const FS s{d.nx, d.ny, d.nz}, ns{d.nx + 1, d.ny + 1, d.nz + 1};
    const int64_t c = blockIdx.x * (int64_t)kNT + threadIdx.x;
    if (c >= s.n() || !band[c]) return;
    int i, j, k;
    decompose(c, sp, i, j, k);
    const double p[3] = {i + 0.5, j + 0.5, k + 0.5};
    const double sl = (double)interp(solid, nullptr, ns, p);
    if (!(sl < 0.0)) return;
    if (!wall) {
        phi[c] = (float)sl;
        return;
    }
    float g[3];
    derivative_idx(solid, ns, p, g);
    double nrm[3] = {(double)g[0], (double)g[1], (double)g[2]};
    normalize3(nrm);
    if (norm2_3(nrm) == 0.0) {
        phi[c] = background;
        return;
    }
    const double t = -sl / dx;
    const double pn[3] = {i + t * nrm[0], j + t * nrm[1], k + t * nrm[2]};
    float value = interp(old, nullptr, s, pn);
    // sin(pi/4) as the double the reference computes (gridutility3.cpp:93)
    const double limit_y = 0.70710678118654746;
    if (horizontal && nrm[1] < limit_y) {
        double h[3] = {nrm[0], 0.0, nrm[2]};
        normalize3(h);
        if (norm2_3(h) != 0.0) {
            const double ph[3] = {i + t * h[0], j + t * h[1], k + t * h[2]};
            const float vh = interp(old, nullptr, s, ph);
            if (vh < value) value = vh;
        }
    }
    phi[c] = value;
}

// pass 3: one dilation; a new cell >= 0 next to an in-bounds cell below 0 becomes dx
__global__ __launch_bounds__(kNT) void k_exls_dilate(Dims d, Split sp, double dx, const float *__restrict__ phi,
                                                     const uint8_t *__restrict__ band, float *__restrict__ phi2,
                                                     uint8_t *__restrict__ band2) {
    const FS s{d.nx, d.ny, d.nz};
    const int64_t c = blockIdx.x * (int64_t)kNT + threadIdx.x;
    if (c >= s.n()) return;
    float v = phi[c];
    uint8_t m = band[c];
    if (!m) {
        int i, j, k;
        decompose(c, sp, i, j, k);
        const int o[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, -1}, {0, 0, 1}};
        bool next = false, below = false;
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            const int ii = i + o[t][0], jj = j + o[t][1], kk = k + o[t][2];
            if (ii < 0 || jj < 0 || kk < 0 || ii >= s.nx || jj >= s.ny || kk >= s.nz) continue;
            const int64_t e = s.at(ii, jj, kk);
            next = next || band[e];
            below = below || phi[e] < 0.0f;
        }
        if (next && v >= 0.0f && below) {
            v = (float)dx;
            m = 1;
        }
    }
    phi2[c] = v;
    band2[c] = m;
}

// pass 4: combine_levelset(solid, fluid, combined, dx); off the union band the cells take
// -/+ background by the sign of the combined value
__global__ __launch_bounds__(kNT) void k_exls_combine(Dims d, Split sp, double dx, const float *__restrict__ solid,
                                                      const uint8_t *__restrict__ sband,
                                                      const float *__restrict__ phi2,
                                                      const uint8_t *__restrict__ band2, float background,
                                                      float *__restrict__ phi, uint8_t *__restrict__ band,
                                                      unsigned long long *__restrict__ count) {
#pragma clang fp contract(off)
    const FS s{d.nx, d.ny, d.nz}, ns{d.nx + 1, d.ny + 1, d.nz + 1};
    const int64_t c = blockIdx.x * (int64_t)kNT + threadIdx.x;
    bool act = false;
    if (c < s.n()) {
        int i, j, k;
        decompose(c, sp, i, j, k);
        double v = 0.0;
        bool any = false;
#pragma unroll
        for (int ii = 0; ii < 2; ++ii)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                for (int kk = 0; kk < 2; ++kk) {
                    const int64_t e = ns.at(i + ii, j + jj, k + kk);
                    v += (double)solid[e];
                    any = any || !sband || sband[e];
                }
        v = v / 8.0;
        const float sc = any ? (float)v : (v < 0.0 ? -background : background);
        const float neg = -(float)dx - sc;
        const float f = phi2[c];
        const float comb = f < neg ? neg : f;
        act = band2[c] || any;
        phi[c] = act ? comb : (comb < 0.0f ? -background : background);
        band[c] = act ? 1 : 0;
    }
    const unsigned long long wave_act = __popcll(__ballot(act));
    if ((threadIdx.x & 63) == 0 && wave_act) atomicAdd(count, wave_act);
}

__global__ void k_exls_fill(unsigned long long *p) { p[0] = 0ull; }

__global__ __launch_bounds__(kNT) void k_count_band(int64_t n, const uint8_t *__restrict__ band,
                                                    unsigned long long *__restrict__ count) {
    const int64_t c = blockIdx.x * (int64_t)kNT + threadIdx.x;
    const bool act = c < n && band[c];
    const unsigned long long wave_act = __popcll(__ballot(act));
    if ((threadIdx.x & 63) == 0 && wave_act) atomicAdd(count, wave_act);
}

bool aligned16(std::initializer_list<const void *> ps) {
    for (const void *p : ps)
        if (reinterpret_cast<uintptr_t>(p) % 16) return false;
    return true;
}

int check3d(shk_ctx *c, const char *fn) {
    if (!c) return SHK_BAD_ARG;
    c->err.clear();
    if (c->dims.bx != 8) return c->fail(SHK_BAD_ARG, "%s needs a 3D context", fn);
    if (c->prm.host_pointers) return c->fail(SHK_BAD_ARG, "%s takes device pointers", fn);
    // trilinear interpolation needs two samples per axis (array_interpolator3.h:54-56)
    if (c->dims.nx < 2 || c->dims.ny < 2 || c->dims.nz < 2)
        return c->fail(SHK_BAD_ARG, "%s needs at least 2 cells per axis", fn);
    if (hipSetDevice(c->device) != hipSuccess) return c->fail(SHK_HIP_ERROR, "hipSetDevice failed");
    return SHK_OK;
}

LiquidWork *work(shk_ctx *c) {
    if (!c->liquid) c->liquid = new LiquidWork();
    return c->liquid;
}

}  // namespace
}  // namespace shk

using namespace shk;

extern "C" {

int shk_extrapolate_velocity(shk_ctx *c, float *u, float *v, float *w, uint8_t *mu, uint8_t *mv, uint8_t *mw,
                             int width, int64_t *activated) {
    int rc;
    if ((rc = check3d(c, "shk_extrapolate_velocity"))) return rc;
    if (!u || !v || !w || !mu || !mv || !mw || width < 0) return c->fail(SHK_BAD_ARG, "null array or negative width");
    // the round number lives in the uint8 mask (round r stores r + 1)
    if (width > 254) return c->fail(SHK_BAD_ARG, "width %d is above 254", width);
    const Dims d = c->dims;
    LiquidWork *lw = work(c);
    const unsigned grid = xcd_blocks(max_faces(d));
    HIPC(c, lw->count.ensure(1));
    HIPC(c, lw->slots.ensure((size_t)grid * 3));
    HIPC(c, hipMemsetAsync(lw->slots.p, 0, sizeof(unsigned int) * (size_t)grid * 3, c->stream));
    const Faces f{{u, v, w}, {mu, mv, mw}};
    const int vec = aligned16({mu, mv, mw}) ? 1 : 0;
    if (width > 0) {
        // any nonzero mask is active: make them 1 so that the rounds can tag theirs
        hipLaunchKernelGGL(k_mask_normalise, dim3(blocks_for((max_faces(d) + 3) / 4), 3), dim3(kNT), 0, c->stream, d,
                           mu, mv, mw, vec);
    }
    for (int r = 1; r <= width; ++r) {
        hipLaunchKernelGGL(k_extrapolate_round, dim3(grid, 3), dim3(kNT), 0, c->stream, d, faces_split(d), f, mu, mv, mw, r,
                           lw->slots.p);
        HIPC(c, hipGetLastError());
    }
    if (width > 0) {
        hipLaunchKernelGGL(k_mask_normalise, dim3(blocks_for((max_faces(d) + 3) / 4), 3), dim3(kNT), 0, c->stream, d,
                           mu, mv, mw, vec);
        HIPC(c, hipGetLastError());
    }
    HIPC(c, hipMemsetAsync(lw->count.p, 0, sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_sum_slots, dim3(256), dim3(kNT), 0, c->stream, (int64_t)grid * 3, lw->slots.p, lw->count.p);
    HIPC(c, hipGetLastError());
    unsigned long long cnt = 0;
    HIPC(c, hipMemcpyAsync(&cnt, lw->count.p, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    if (activated) *activated = (int64_t)cnt;
    return SHK_OK;
}

int shk_extrapolate_levelset(shk_ctx *c, const float *solid, const uint8_t *solid_band, float *phi, uint8_t *band,
                             float background, int64_t *active) {
    int rc;
    if ((rc = check3d(c, "shk_extrapolate_levelset"))) return rc;
    if (!solid || !phi || !band) return c->fail(SHK_BAD_ARG, "null array");
    const Dims d = c->dims;
    const FS s{d.nx, d.ny, d.nz};
    const int64_t n = s.n(), nn = (int64_t)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
    LiquidWork *lw = work(c);
    HIPC(c, lw->count.ensure(2));
    HIPC(c, lw->a[0].ensure((size_t)n));
    HIPC(c, lw->a[1].ensure((size_t)n));
    HIPC(c, lw->band[0].ensure((size_t)n));
    hipStream_t st = c->stream;
    unsigned long long *flag = lw->count.p, *cnt_d = lw->count.p + 1;
    hipLaunchKernelGGL(k_exls_fill, dim3(1), dim3(1), 0, st, flag);
    hipLaunchKernelGGL(k_exls_fill, dim3(1), dim3(1), 0, st, cnt_d);
    hipLaunchKernelGGL(k_exls_exist, dim3(1024), dim3(kNT), 0, st, nn, solid, solid_band, flag);
    HIPC(c, hipGetLastError());
    unsigned long long exist = 0;
    HIPC(c, hipMemcpyAsync(&exist, flag, sizeof exist, hipMemcpyDeviceToHost, st));
    HIPC(c, hipStreamSynchronize(st));
    const Split sp = cells_split(d);
    const unsigned gb = blocks_for(n);
    if (exist) {
        HIPC(c, hipMemcpyAsync(lw->a[0].p, phi, sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_exls_wall, dim3(gb), dim3(kNT), 0, st, d, sp, c->dx, solid, lw->a[0].p, band, phi,
                           background, c->prm.solid_wall_extrap ? 1 : 0, c->prm.horizontal_extrap ? 1 : 0);
        hipLaunchKernelGGL(k_exls_dilate, dim3(gb), dim3(kNT), 0, st, d, sp, c->dx, phi, band, lw->a[1].p,
                           lw->band[0].p);
        hipLaunchKernelGGL(k_exls_combine, dim3(gb), dim3(kNT), 0, st, d, sp, c->dx, solid, solid_band, lw->a[1].p,
                           lw->band[0].p, background, phi, band, cnt_d);
    } else {
        // no solid: nothing changes; count the band
        hipLaunchKernelGGL(k_count_band, dim3(gb), dim3(kNT), 0, st, n, band, cnt_d);
    }
    HIPC(c, hipGetLastError());
    unsigned long long cnt = 0;
    HIPC(c, hipMemcpyAsync(&cnt, cnt_d, sizeof cnt, hipMemcpyDeviceToHost, st));
    HIPC(c, hipStreamSynchronize(st));
    if (active) *active = (int64_t)cnt;
    return SHK_OK;
}

int shk_constrain_velocity(shk_ctx *c, const float *solid, float *u, float *v, float *w, const uint8_t *mu,
                           const uint8_t *mv, const uint8_t *mw) {
    int rc;
    if ((rc = check3d(c, "shk_constrain_velocity"))) return rc;
    if (!u || !v || !w || !mu || !mv || !mw) return c->fail(SHK_BAD_ARG, "null array");
    const Dims d = c->dims;
    LiquidWork *lw = work(c);
    float *user[3] = {u, v, w};
    const uint8_t *um[3] = {mu, mv, mw};
    CFaces save;
    for (int a = 0; a < 3; ++a) {
        const size_t n = (size_t)face_fs(d, a).n();
        HIPC(c, lw->a[a].ensure(n));
        HIPC(c, hipMemcpyAsync(lw->a[a].p, user[a], sizeof(float) * n, hipMemcpyDeviceToDevice, c->stream));
        save.v[a] = lw->a[a].p;
        save.m[a] = um[a];
    }
    hipLaunchKernelGGL(k_constrain<0>, dim3(blocks_for(face_fs(d, 0).n())), dim3(kNT), 0, c->stream, d,
                       split_of(face_fs(d, 0)), solid, c->dx,
                       save, u, mu);
    hipLaunchKernelGGL(k_constrain<1>, dim3(blocks_for(face_fs(d, 1).n())), dim3(kNT), 0, c->stream, d,
                       split_of(face_fs(d, 1)), solid, c->dx,
                       save, v, mv);
    hipLaunchKernelGGL(k_constrain<2>, dim3(blocks_for(face_fs(d, 2).n())), dim3(kNT), 0, c->stream, d,
                       split_of(face_fs(d, 2)), solid, c->dx,
                       save, w, mw);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    return SHK_OK;
}

int shk_compute_max_u(shk_ctx *c, const float *u, const float *v, const float *w, const uint8_t *mu,
                      const uint8_t *mv, const uint8_t *mw, double *max_u) {
    int rc;
    if ((rc = check3d(c, "shk_compute_max_u"))) return rc;
    if (!u || !v || !w || !mu || !mv || !mw || !max_u) return c->fail(SHK_BAD_ARG, "null argument");
    const Dims d = c->dims;
    LiquidWork *lw = work(c);
    const int64_t G = (int64_t)d.nx * d.ny * d.nz;
    const int blocks = (int)((G + kNT - 1) / kNT < kMaxPartials ? (G + kNT - 1) / kNT : kMaxPartials);
    HIPC(c, lw->part.ensure(kMaxPartials));
    lw->h_part.resize(blocks);
    const CFaces f{{u, v, w}, {mu, mv, mw}};
    hipLaunchKernelGGL(k_max_u2, dim3(blocks), dim3(kNT), 0, c->stream, d, cells_split(d), f, lw->part.p);
    HIPC(c, hipGetLastError());
    HIPC(c, hipMemcpyAsync(lw->h_part.data(), lw->part.p, sizeof(double) * blocks, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    double best = 0.0;
    for (double p : lw->h_part) best = p > best ? p : best;
    *max_u = std::sqrt(best);
    return SHK_OK;
}

int shk_advect_velocity(shk_ctx *c, double dt, float *u, float *v, float *w, const uint8_t *mu, const uint8_t *mv,
                        const uint8_t *mw, const float *fluid) {
    int rc;
    if ((rc = check3d(c, "shk_advect_velocity"))) return rc;
    if (!u || !v || !w || !mu || !mv || !mw || !fluid) return c->fail(SHK_BAD_ARG, "null array");
    if (!std::isfinite(dt)) return c->fail(SHK_BAD_ARG, "dt must be finite");
    const Dims d = c->dims;
    LiquidWork *lw = work(c);
    const bool mc = c->prm.maccormack;
    float *user[3] = {u, v, w};
    const uint8_t *um[3] = {mu, mv, mw};
    CFaces vin, v0;
    for (int a = 0; a < 3; ++a) {
        const size_t n = (size_t)face_fs(d, a).n();
        HIPC(c, lw->a[a].ensure(n));
        HIPC(c, hipMemcpyAsync(lw->a[a].p, user[a], sizeof(float) * n, hipMemcpyDeviceToDevice, c->stream));
        vin.v[a] = lw->a[a].p;
        vin.m[a] = um[a];
        if (mc) {
            HIPC(c, lw->b[a].ensure(n));
            HIPC(c, lw->lo[a].ensure(n));
            HIPC(c, lw->hi[a].ensure(n));
            HIPC(c, lw->band[a].ensure(n));
        }
        v0.v[a] = mc ? lw->b[a].p : user[a];
        v0.m[a] = um[a];
    }
    const int trim = c->prm.trim_narrowband;
    const bool weno = c->prm.weno;
#define SHK_SL(DIM, SRC, OUT, LO, HI, BAND, DT, FV, FL, FH, FB, F0, FO)                                             \
    do {                                                                                                            \
        if (weno)                                                                                                   \
            hipLaunchKernelGGL((k_semi_lagrangian<DIM, true>), dim3(blocks_for(face_fs(d, DIM).n())), dim3(kNT), 0, \
                               c->stream, d, split_of(face_fs(d, DIM)), c->dx, DT, SRC, OUT, LO, HI, BAND, fluid, trim, \
                               FV, FL, FH, FB, F0, FO);                                                            \
        else                                                                                                        \
            hipLaunchKernelGGL((k_semi_lagrangian<DIM, false>), dim3(blocks_for(face_fs(d, DIM).n())), dim3(kNT), 0, \
                               c->stream, d, split_of(face_fs(d, DIM)), c->dx, DT, SRC, OUT, LO, HI, BAND, fluid, trim, \
                               FV, FL, FH, FB, F0, FO);                                                            \
    } while (0)
    // forward pass: v0 (the result itself without MacCormack) and the limiter data
    SHK_SL(0, vin, const_cast<float *>(v0.v[0]), mc ? lw->lo[0].p : nullptr, mc ? lw->hi[0].p : nullptr,
           mc ? lw->band[0].p : nullptr, dt, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    SHK_SL(1, vin, const_cast<float *>(v0.v[1]), mc ? lw->lo[1].p : nullptr, mc ? lw->hi[1].p : nullptr,
           mc ? lw->band[1].p : nullptr, dt, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    SHK_SL(2, vin, const_cast<float *>(v0.v[2]), mc ? lw->lo[2].p : nullptr, mc ? lw->hi[2].p : nullptr,
           mc ? lw->band[2].p : nullptr, dt, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (mc) {
        // backward pass of v0 with its own velocity and -dt, combined into the result
        SHK_SL(0, v0, nullptr, nullptr, nullptr, nullptr, -dt, vin.v[0], lw->lo[0].p, lw->hi[0].p, lw->band[0].p,
               v0.v[0], u);
        SHK_SL(1, v0, nullptr, nullptr, nullptr, nullptr, -dt, vin.v[1], lw->lo[1].p, lw->hi[1].p, lw->band[1].p,
               v0.v[1], v);
        SHK_SL(2, v0, nullptr, nullptr, nullptr, nullptr, -dt, vin.v[2], lw->lo[2].p, lw->hi[2].p, lw->band[2].p,
               v0.v[2], w);
    }
#undef SHK_SL
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    return SHK_OK;
}

int shk_advect_levelset(shk_ctx *c, double dt, float *phi, const uint8_t *mphi, const float *u, const float *v,
                        const float *w, const uint8_t *mu, const uint8_t *mv, const uint8_t *mw, float background) {
    int rc;
    if ((rc = check3d(c, "shk_advect_levelset"))) return rc;
    if (!phi || !mphi || !u || !v || !w || !mu || !mv || !mw) return c->fail(SHK_BAD_ARG, "null array");
    if (!std::isfinite(dt)) return c->fail(SHK_BAD_ARG, "dt must be finite");
    const Dims d = c->dims;
    LiquidWork *lw = work(c);
    const bool mc = c->prm.maccormack;
    const int64_t G = (int64_t)d.nx * d.ny * d.nz;
    // scratch: the level set before the step (a[0]), q0 (b[0]), limiter bounds, band
    HIPC(c, lw->a[0].ensure(G));
    HIPC(c, hipMemcpyAsync(lw->a[0].p, phi, sizeof(float) * G, hipMemcpyDeviceToDevice, c->stream));
    if (mc) {
        HIPC(c, lw->b[0].ensure(G));
        HIPC(c, lw->lo[0].ensure(G));
        HIPC(c, lw->hi[0].ensure(G));
        HIPC(c, lw->band[0].ensure(G));
    }
    const CFaces vel{{u, v, w}, {mu, mv, mw}};
    const float *qin = lw->a[0].p;
    auto kern = c->prm.weno ? k_advect_cell<true> : k_advect_cell<false>;
    hipLaunchKernelGGL(kern, dim3(blocks_for(G)), dim3(kNT), 0, c->stream, d, cells_split(d), c->dx, dt, qin, mphi, vel,
                       qin, c->prm.trim_narrowband, mc ? lw->b[0].p : phi, mc ? lw->lo[0].p : nullptr,
                       mc ? lw->hi[0].p : nullptr, mc ? lw->band[0].p : nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, mc ? 1 : 0, background);
    if (mc)
        hipLaunchKernelGGL(kern, dim3(blocks_for(G)), dim3(kNT), 0, c->stream, d, cells_split(d), c->dx, -dt,
                           lw->b[0].p, mphi, vel, qin, c->prm.trim_narrowband, nullptr, nullptr, nullptr, nullptr, qin,
                           lw->lo[0].p, lw->hi[0].p, lw->band[0].p, lw->b[0].p, phi, 0, background);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    return SHK_OK;
}

int shk_add_gravity(shk_ctx *c, double dt, double gx, double gy, double gz, float *u, float *v, float *w,
                    uint8_t *mu, uint8_t *mv, uint8_t *mw) {
    int rc;
    if ((rc = check3d(c, "shk_add_gravity"))) return rc;
    if (!u || !v || !w || !mu || !mv || !mw) return c->fail(SHK_BAD_ARG, "null array");
    const Dims d = c->dims;
    const Faces f{{u, v, w}, {mu, mv, mw}};
    hipLaunchKernelGGL(k_gravity, dim3(blocks_for((max_faces(d) + 3) / 4), 3), dim3(kNT), 0, c->stream, d,
                       (float)(dt * gx), (float)(dt * gy), (float)(dt * gz), f, mu, mv, mw,
                       aligned16({u, v, w, mu, mv, mw}) ? 1 : 0);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    return SHK_OK;
}

int shk_project_activity(shk_ctx *c, uint8_t *mu, uint8_t *mv, uint8_t *mw) {
    int rc;
    if ((rc = check3d(c, "shk_project_activity"))) return rc;
    if (!mu || !mv || !mw) return c->fail(SHK_BAD_ARG, "null mask");
    if (!c->velocity_updated) return SHK_OK;  // no unknowns: the projection left the velocity alone
    const Dims d = c->dims;
    hipLaunchKernelGGL(k_project_activity, dim3(blocks_for((max_faces(d) + 3) / 4), 3), dim3(kNT), 0, c->stream, d,
                       c->area[0].p, c->area[1].p, c->area[2].p, c->rho[0].p, c->rho[1].p, c->rho[2].p, mu, mv, mw,
                       aligned16({mu, mv, mw}) ? 1 : 0);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    return SHK_OK;
}

}  // extern "C"
