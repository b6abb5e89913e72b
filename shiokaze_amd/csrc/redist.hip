// redist.hip -- PDE redistancing of the liquid level set on the GPU: the
// pderedistancer3 plugin's redistance(phi, width) (src/redistancer/pderedistancer3.cpp:
// 39-183) on a dense level set with a per-cell activity mask (the narrow band).
//
// Steps, each a kernel over the dense grid (one thread per cell):
//   1. trim_narrowband (src/utility/gridutility3.cpp:184-216): an active cell stays active
//      iff an active axis neighbour has the opposite sign, decided on the activity before
//      the trim;
//   2. `width` dilations (pderedistancer3.cpp:50-68, with dilate3.h:43-97's rule that a
//      round reads the state before it): a newly active cell takes -dx or +dx by the sign
//      of its first active neighbour in the order -x, +x, -y, +y, -z, +z;
//   3. the initial level set phi0 and its smoothed sign phi0 / sqrt(phi0^2 + dx^2);
//   4. ceil(width / rate) steps of dphi/dt = sgn0 (1 - |grad phi|) with dt = rate * dx
//      (Euler, or Heun's RK2), the gradient from upwind one-sided differences that use
//      the interface distance (utility::fraction of phi0) across a sign change of phi0
//      (pderedistancer3.cpp:86-134); the updates round as the reference's float arrays do:
//      fl(fl(d * fl(dt)) + phi) (array3.h:889-951);
//   5. cells with |phi| > width*dx leave the band, and every cell off the band takes the
//      level-set fill of its side, -width*dx where its value is negative and +width*dx
//      elsewhere (set_as_levelset + flood_fill, pderedistancer3.cpp:176-182).  The dense
//      layout carries each cell's side in its value, which is what the reference's flood
//      fill reconstructs from the band.
// FP contraction is off: every product and sum rounds as written, as in the reference's
// x86-64 build.
#include <hip/hip_runtime.h>

#include <cmath>

#include "ctx_internal.h"
#include "shk_fractions.h"

namespace shk {

struct RedistWork {
    DevBuf<uint8_t> m[2];
    DevBuf<float> p, phi0, sgn, d0, tmp;
    DevBuf<unsigned long long> count;
    void release() {
        for (auto &b : m) b.release();
        p.release();
        phi0.release();
        sgn.release();
        d0.release();
        tmp.release();
        count.release();
    }
};

void redist_destroy(RedistWork *w) {
    if (!w) return;
    w->release();
    delete w;
}

namespace {

struct RShape {
    int nx, ny, nz;
    __host__ __device__ int64_t at(int i, int j, int k) const {
        return (int64_t)i + (int64_t)nx * ((int64_t)j + (int64_t)ny * k);
    }
};

__device__ inline bool decode(const RShape &s, int64_t c, int &i, int &j, int &k) {
    if (c >= (int64_t)s.nx * s.ny * s.nz) return false;
    i = (int)(c % s.nx);
    const int64_t r = c / s.nx;
    j = (int)(r % s.ny);
    k = (int)(r / s.ny);
    return true;
}

// neighbour t of (i, j, k) in the order -x, +x, -y, +y, -z, +z; false if out of bounds
__device__ inline bool neighbour(const RShape &s, int i, int j, int k, int t, int64_t &e) {
    const int a = t >> 1, d = (t & 1) ? 1 : -1;
    const int ii = i + (a == 0) * d, jj = j + (a == 1) * d, kk = k + (a == 2) * d;
    if (ii < 0 || jj < 0 || kk < 0 || ii >= s.nx || jj >= s.ny || kk >= s.nz) return false;
    e = s.at(ii, jj, kk);
    return true;
}

__global__ void k_redist_trim(RShape s, const float *__restrict__ phi, const uint8_t *__restrict__ band,
                              uint8_t *__restrict__ m) {
#pragma clang fp contract(off)
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int i, j, k;
    if (!decode(s, c, i, j, k)) return;
    uint8_t keep = 0;
    if (band[c]) {
        const double p = phi[c];
        for (int t = 0; t < 6 && !keep; ++t) {
            int64_t e;
            if (neighbour(s, i, j, k, t, e) && band[e] && p * (double)phi[e] < 0.0) keep = 1;
        }
    }
    m[c] = keep;
}

__global__ void k_redist_dilate(RShape s, double dx, const float *__restrict__ phi, const uint8_t *__restrict__ m,
                                float *__restrict__ phi_out, uint8_t *__restrict__ m_out) {
#pragma clang fp contract(off)
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int i, j, k;
    if (!decode(s, c, i, j, k)) return;
    float v = phi[c];
    uint8_t a = m[c];
    if (!a) {
        for (int t = 0; t < 6; ++t) {
            int64_t e;
            if (neighbour(s, i, j, k, t, e) && m[e]) {
                v = phi[e] < 0.0f ? (float)fmin(0.0, 0.0 - dx) : (float)fmax(0.0, 0.0 + dx);
                a = 1;
                break;
            }
        }
    }
    phi_out[c] = v;
    m_out[c] = a;
}

__global__ void k_redist_init(RShape s, double dx, const float *__restrict__ phi, const uint8_t *__restrict__ m,
                              float *__restrict__ phi0, float *__restrict__ sgn) {
#pragma clang fp contract(off)
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= (int64_t)s.nx * s.ny * s.nz) return;
    const double v = phi[c];
    phi0[c] = phi[c];
    sgn[c] = m[c] ? (float)(v / sqrt(v * v + dx * dx)) : 0.0f;
}

// sgn0 (1 - |grad phi|) at an active cell (pderedistancer3.cpp:86-134)
__device__ inline float redist_rate(const RShape &s, double dx, const float *__restrict__ phi,
                                    const float *__restrict__ phi0, const float *__restrict__ sgn,
                                    const uint8_t *__restrict__ m, int i, int j, int k, int64_t c) {
#pragma clang fp contract(off)
    const double sgn0 = sgn[c];
    const double p0 = phi0[c];
    const double p = phi[c];
    const int ijk[3] = {i, j, k};
    const int ext[3] = {s.nx, s.ny, s.nz};
    const int64_t stride[3] = {1, (int64_t)s.nx, (int64_t)s.nx * s.ny};
    double g2 = 0.0;
    for (int dim = 0; dim < 3; ++dim) {
        int sel = 0;
        const double tmp = sgn0 * p;
        double pb = 0.0, pf = 0.0, g = 0.0;
        const int64_t cb = c - stride[dim], cf = c + stride[dim];
        if (ijk[dim] > 0 && m[cb]) {
            pb = phi[cb];
            if (sgn0 * pb < tmp) sel = -1;
        }
        if (ijk[dim] < ext[dim] - 1 && m[cf]) {
            pf = phi[cf];
            if (sgn0 * pf < tmp) {
                if (sel == 0)
                    sel = 1;
                else if (sgn0 * pf < sgn0 * pb)
                    sel = 1;
            }
        }
        if (sel == -1) {
            const double frac = fraction(p0, (double)phi0[cb]);
            if (frac == 1.0 || frac == 0.0)
                g = (p - pb) / dx;
            else if (sgn0 < 0.0)
                g = p / (dx * frac);
            else
                g = p / (dx * (1.0 - frac));
        } else if (sel == 1) {
            const double frac = fraction(p0, (double)phi0[cf]);
            if (frac == 1.0 || frac == 0.0)
                g = (pf - p) / dx;
            else if (sgn0 < 0.0)
                g = -p / (dx * frac);
            else
                g = -p / (dx * (1.0 - frac));
        }
        // |grad|^2 accumulated in the order x, y, z (vec3d::len)
        const double gg = g * g;
        g2 = dim == 0 ? gg : g2 + gg;
    }
    return (float)(sgn0 * (1.0 - sqrt(g2)));
}

// Euler: phi_out = fl(fl(d * dt) + phi) on active cells
__global__ void k_redist_euler(RShape s, double dx, float dtf, const float *__restrict__ phi,
                               const float *__restrict__ phi0, const float *__restrict__ sgn,
                               const uint8_t *__restrict__ m, float *__restrict__ phi_out) {
#pragma clang fp contract(off)
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int i, j, k;
    if (!decode(s, c, i, j, k)) return;
    float v = phi[c];
    if (m[c]) {
        const float t = redist_rate(s, dx, phi, phi0, sgn, m, i, j, k, c) * dtf;
        v = t + v;
    }
    phi_out[c] = v;
}

// RK2 first stage: d0 = rate(phi), tmp = fl(fl(d0 * dt) + phi) on active cells
__global__ void k_redist_rk_a(RShape s, double dx, float dtf, const float *__restrict__ phi,
                              const float *__restrict__ phi0, const float *__restrict__ sgn,
                              const uint8_t *__restrict__ m, float *__restrict__ d0, float *__restrict__ tmp) {
#pragma clang fp contract(off)
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int i, j, k;
    if (!decode(s, c, i, j, k)) return;
    float v = phi[c], d = 0.0f;
    if (m[c]) {
        d = redist_rate(s, dx, phi, phi0, sgn, m, i, j, k, c);
        const float t = d * dtf;
        v = t + v;
    }
    d0[c] = d;
    tmp[c] = v;
}

// RK2 second stage: phi_out = fl(phi + fl(fl(d0 + rate(tmp)) * fl(dt / 2)))
__global__ void k_redist_rk_b(RShape s, double dx, float hdtf, const float *__restrict__ phi,
                              const float *__restrict__ tmp, const float *__restrict__ phi0,
                              const float *__restrict__ sgn, const uint8_t *__restrict__ m,
                              const float *__restrict__ d0, float *__restrict__ phi_out) {
#pragma clang fp contract(off)
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int i, j, k;
    if (!decode(s, c, i, j, k)) return;
    float v = phi[c];
    if (m[c]) {
        const float d1 = redist_rate(s, dx, tmp, phi0, sgn, m, i, j, k, c);
        const float sum = d0[c] + d1;
        const float t = sum * hdtf;
        v = v + t;
    }
    phi_out[c] = v;
}

__global__ void k_redist_finish(RShape s, double bw, const float *__restrict__ phi, const uint8_t *__restrict__ m,
                                float *__restrict__ phi_out, uint8_t *__restrict__ band,
                                unsigned long long *__restrict__ count) {
#pragma clang fp contract(off)
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool act = false;
    if (c < (int64_t)s.nx * s.ny * s.nz) {
        float v = phi[c];
        act = m[c] && !(fabs((double)v) > bw);
        if (!act) v = v < 0.0f ? (float)(-bw) : (float)bw;
        phi_out[c] = v;
        band[c] = act ? 1 : 0;
    }
    // one count per wave (integer atomics: the total does not depend on the order)
    const unsigned long long wave_act = __popcll(__ballot(act));
    if ((threadIdx.x & (kWave - 1)) == 0 && wave_act) atomicAdd(count, wave_act);
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v) { *p = v; }

inline unsigned cells_grid(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

}  // namespace shk

using namespace shk;

extern "C" int shk_redistance(shk_ctx *c, float *phi, uint8_t *band, int width, int64_t *active) {
    if (!c) return SHK_BAD_ARG;
    c->err.clear();
    if (!phi || !band) return c->fail(SHK_BAD_ARG, "NULL level set or band pointer");
    if (width < 1 || width > 64) return c->fail(SHK_BAD_ARG, "width must be in 1..64");
    if (c->dims.bx != 8) return c->fail(SHK_BAD_ARG, "shk_redistance needs a 3D context");
    if (c->prm.host_pointers) return c->fail(SHK_BAD_ARG, "shk_redistance takes device pointers");
    HIPC(c, hipSetDevice(c->device));
    const RShape s{c->dims.nx, c->dims.ny, c->dims.nz};
    const int64_t n = (int64_t)s.nx * s.ny * s.nz;
    if (!c->redist) c->redist = new RedistWork();
    RedistWork *w = c->redist;
    for (auto &b : w->m) HIPC(c, b.ensure(n));
    HIPC(c, w->p.ensure(n));
    HIPC(c, w->phi0.ensure(n));
    HIPC(c, w->sgn.ensure(n));
    HIPC(c, w->count.ensure(1));
    const bool rk2 = c->prm.redist_rk2;
    if (rk2) {
        HIPC(c, w->d0.ensure(n));
        HIPC(c, w->tmp.ensure(n));
    }
    hipStream_t st = c->stream;
    const unsigned gb = cells_grid(n);
    const double dx = c->dx, rate = c->prm.redist_rate;
    hipLaunchKernelGGL(k_redist_trim, dim3(gb), dim3(256), 0, st, s, phi, band, w->m[0].p);
    // dilations: ping-pong between (phi, m[0]) and (p, m[1])
    float *P[2] = {phi, w->p.p};
    uint8_t *M[2] = {w->m[0].p, w->m[1].p};
    int cur = 0;
    for (int r = 0; r < width; ++r) {
        hipLaunchKernelGGL(k_redist_dilate, dim3(gb), dim3(256), 0, st, s, dx, P[cur], M[cur], P[cur ^ 1],
                           M[cur ^ 1]);
        cur ^= 1;
    }
    hipLaunchKernelGGL(k_redist_init, dim3(gb), dim3(256), 0, st, s, dx, P[cur], M[cur], w->phi0.p, w->sgn.p);
    const float dtf = (float)(rate * dx), hdtf = (float)(0.5 * (rate * dx));
    const int iters = (int)std::ceil((double)width / rate);
    const uint8_t *m = M[cur];
    for (int it = 0; it < iters; ++it) {
        if (!rk2) {
            hipLaunchKernelGGL(k_redist_euler, dim3(gb), dim3(256), 0, st, s, dx, dtf, P[cur], w->phi0.p, w->sgn.p,
                               m, P[cur ^ 1]);
        } else {
            hipLaunchKernelGGL(k_redist_rk_a, dim3(gb), dim3(256), 0, st, s, dx, dtf, P[cur], w->phi0.p, w->sgn.p,
                               m, w->d0.p, w->tmp.p);
            hipLaunchKernelGGL(k_redist_rk_b, dim3(gb), dim3(256), 0, st, s, dx, hdtf, P[cur], w->tmp.p, w->phi0.p,
                               w->sgn.p, m, w->d0.p, P[cur ^ 1]);
        }
        cur ^= 1;
    }
    // the result into the caller's arrays
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, st, w->count.p, 0ull);
    hipLaunchKernelGGL(k_redist_finish, dim3(gb), dim3(256), 0, st, s, dx * (double)width, P[cur], m, phi, band,
                       w->count.p);
    HIPC(c, hipGetLastError());
    unsigned long long cnt = 0;
    HIPC(c, hipMemcpyAsync(&cnt, w->count.p, sizeof cnt, hipMemcpyDeviceToHost, st));
    HIPC(c, hipStreamSynchronize(st));
    if (active) *active = (int64_t)cnt;
    return SHK_OK;
}
