// scene.hip -- host-side synthetic scene generator (input synthesis for tests and
// benchmarks; needs no GPU).
//
// Scene conventions follow the reference's scene assignment
// (src/utility/macutility3.cpp:314-411): the liquid level set is sampled at cell
// centres and clipped by the solid, phi <- max(phi, -(solid(p) + dx))
// (macutility3.cpp:352-361); the solid level set is sampled at nodes as
// min(1, solid(p)) (macutility3.cpp:405-411); boxes use the max-of-planes
// distance of utility::box (include/shiokaze/utility/utility.h:125-132).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "shk_fractions.h"
#include "shk_projection.h"

namespace {

using shk::Dims;

inline uint64_t mix64(uint64_t z) {
    // SplitMix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
    // number generators", OOPSLA 2014).
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Counter-based uniform stream: value `index` of stream `stream` under `seed`.
struct Stream {
    uint64_t key;
    Stream(uint64_t seed, uint64_t stream) : key(mix64(seed ^ (0xD1B54A32D192ED03ull * (stream + 1)))) {}
    double u01(uint64_t index) const {
        const uint64_t v = mix64(key + 0x9E3779B97F4A7C15ull * (index + 1));
        return (double)(v >> 11) * 0x1.0p-53;
    }
};

enum : uint64_t { STREAM_U = 0, STREAM_V = 1, STREAM_W = 2, STREAM_SOLIDS = 16, STREAM_DROPS = 17 };

struct Solid {
    int kind;  // 0 sphere, 1 box
    double c[3], h[3], r;
};
struct Drop {
    double c[3], r;
};

struct Scene {
    int kind;
    double pool;
    std::vector<Solid> solids;
    std::vector<Drop> drops;

    double solid_sdf(const double p[3]) const {
        double best = 1e300;
        for (const Solid &s : solids) {
            double d;
            if (s.kind == 0) {
                const double dx = p[0] - s.c[0], dy = p[1] - s.c[1], dz = p[2] - s.c[2];
                d = std::sqrt(dx * dx + dy * dy + dz * dz) - s.r;
            } else {
                d = -9999.0;
                for (int a = 0; a < 3; ++a) {
                    d = std::max(d, (s.c[a] - s.h[a]) - p[a]);
                    d = std::max(d, p[a] - (s.c[a] + s.h[a]));
                }
            }
            best = std::min(best, d);
        }
        return best;
    }
    double liquid(const double p[3]) const {
        double v = p[1] - pool;
        for (const Drop &d : drops) {
            const double dx = p[0] - d.c[0], dy = p[1] - d.c[1], dz = p[2] - d.c[2];
            v = std::min(v, std::sqrt(dx * dx + dy * dy + dz * dz) - d.r);
        }
        return v;
    }
};

Scene make_scene(int kind, uint64_t seed) {
    Scene s;
    s.kind = kind;
    if (kind == 1) {
        s.pool = 0.3;
        s.drops.push_back(Drop{{0.5, 0.7, 0.5}, 0.15});
    } else {
        s.pool = 0.4;
        const Stream rs(seed, STREAM_SOLIDS);
        for (int o = 0; o < 20; ++o) {
            Solid so{};
            so.kind = rs.u01(8 * o) < 0.5 ? 0 : 1;
            for (int a = 0; a < 3; ++a) so.c[a] = rs.u01(8 * o + 1 + a);
            so.r = 0.03 + 0.09 * rs.u01(8 * o + 4);
            for (int a = 0; a < 3; ++a) so.h[a] = 0.03 + 0.09 * rs.u01(8 * o + 4 + a);
            s.solids.push_back(so);
        }
        const Stream rd(seed, STREAM_DROPS);
        for (int m = 0; m < 8; ++m) {
            Drop d{};
            for (int a = 0; a < 3; ++a) d.c[a] = rd.u01(4 * m + a);
            d.r = 0.05 + 0.10 * rd.u01(4 * m + 3);
            s.drops.push_back(d);
        }
    }
    return s;
}

template <class F>
void parallel_slabs(int n, F &&f) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    nt = std::min<unsigned>(nt, 32);
    nt = std::min<unsigned>(nt, (unsigned)std::max(n, 1));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            for (int k = (int)t; k < n; k += (int)nt) f(k);
        });
    for (auto &x : th) x.join();
}

// Turn liquid components that have no free-surface face into air: a pure
// Neumann component makes the system singular.  Components are the connected
// sets of unknowns of macpressuresolver3.cpp:134-167 through open faces.
int64_t drop_closed_components(const Dims &d, double dx, const float *solid, float *fluid) {
    const int64_t G = (int64_t)d.nx * d.ny * d.nz;
    // bit a of open[c]: the low face of cell c along axis a is open (area and rho nonzero)
    std::vector<uint8_t> open(G, 0);
    parallel_slabs(d.nz, [&](int k) {
        for (int j = 0; j < d.ny; ++j)
            for (int i = 0; i < d.nx; ++i) {
                uint8_t bits = 0;
                for (int a = 0; a < 3; ++a) {
                    if (shk::face_area(d, solid, a, i, j, k, 1e-2) != 0.0 &&
                        shk::face_rho(d, fluid, a, i, j, k, 1e-2) != 0.0)
                        bits |= (uint8_t)(1u << a);
                }
                open[shk::cell_index(d, i, j, k)] = bits;
            }
    });
    auto liquid = [&](int64_t c) { return fluid[c] < 0.0f; };
    const int64_t sx = 1, sy = d.nx, sz = (int64_t)d.nx * d.ny;
    // neighbour across face nq (reference order) and whether that face is open
    auto visit = [&](int64_t c, int i, int j, int k, int nq, int64_t &nbr, bool &inb, bool &opn) {
        const int a = nq >> 1;
        const bool plus = (nq & 1) == 0;
        const int coord = a == 0 ? i : (a == 1 ? j : k);
        const int n = a == 0 ? d.nx : (a == 1 ? d.ny : d.nz);
        const int64_t st = a == 0 ? sx : (a == 1 ? sy : sz);
        inb = plus ? coord + 1 < n : coord > 0;
        nbr = plus ? c + st : c - st;
        opn = inb && ((open[plus ? nbr : c] >> a) & 1u);
    };
    std::vector<uint8_t> unknown(G, 0);
    parallel_slabs(d.nz, [&](int k) {
        for (int j = 0; j < d.ny; ++j)
            for (int i = 0; i < d.nx; ++i) {
                const int64_t c = shk::cell_index(d, i, j, k);
                if (!liquid(c)) continue;
                for (int nq = 0; nq < 6; ++nq) {
                    int64_t nb;
                    bool inb, opn;
                    visit(c, i, j, k, nq, nb, inb, opn);
                    if (opn && liquid(nb)) {
                        unknown[c] = 1;
                        break;
                    }
                }
            }
    });
    std::vector<uint8_t> seen(G, 0);
    std::vector<int64_t> stack, comp;
    int64_t dropped = 0;
    for (int64_t c0 = 0; c0 < G; ++c0) {
        if (!unknown[c0] || seen[c0]) continue;
        bool free_surface = false;
        comp.clear();
        stack.clear();
        stack.push_back(c0);
        seen[c0] = 1;
        while (!stack.empty()) {
            const int64_t c = stack.back();
            stack.pop_back();
            comp.push_back(c);
            const int i = (int)(c % d.nx), j = (int)((c / d.nx) % d.ny), k = (int)(c / sz);
            for (int nq = 0; nq < 6; ++nq) {
                int64_t nb;
                bool inb, opn;
                visit(c, i, j, k, nq, nb, inb, opn);
                if (!opn) continue;
                if (!liquid(nb)) {
                    free_surface = true;
                } else if (!seen[nb]) {
                    seen[nb] = 1;
                    stack.push_back(nb);
                }
            }
        }
        if (!free_surface) {
            for (int64_t c : comp) fluid[c] = (float)dx;
            dropped += (int64_t)comp.size();
        }
    }
    return dropped;
}

}  // namespace

extern "C" int shk_scene_generate(int kind, int nx, int ny, int nz, double dx, uint64_t seed, float *fluid,
                                  float *solid, float *u, float *v, float *w, float *us, float *vs, float *ws,
                                  int64_t *dropped) {
    if ((kind != 1 && kind != 3) || nx <= 0 || ny <= 0 || nz <= 0 || !(dx > 0.0) || !fluid || !solid || !u ||
        !v || !w || !us || !vs || !ws)
        return SHK_BAD_ARG;
    const Dims d{nx, ny, nz};
    const Scene scene = make_scene(kind, seed);
    // nodal solid level set
    parallel_slabs(nz + 1, [&](int k) {
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                const double p[3] = {dx * i, dx * j, dx * k};
                const double val = scene.solids.empty() ? 1.0 : std::min(1.0, scene.solid_sdf(p));
                solid[shk::node_index(d, i, j, k)] = (float)val;
            }
    });
    // cell-centred liquid level set, clipped by the solid
    parallel_slabs(nz, [&](int k) {
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double p[3] = {dx * (i + 0.5), dx * (j + 0.5), dx * (k + 0.5)};
                double val = scene.liquid(p);
                if (!scene.solids.empty()) val = std::max(val, -(scene.solid_sdf(p) + dx));
                fluid[shk::cell_index(d, i, j, k)] = (float)val;
            }
    });
    // face velocities
    float *vel[3] = {u, v, w};
    float *svel[3] = {us, vs, ws};
    for (int a = 0; a < 3; ++a) {
        const Stream rs(seed, STREAM_U + a);
        const int fx = nx + (a == 0), fy = ny + (a == 1), fz = nz + (a == 2);
        float *out = vel[a];
        parallel_slabs(fz, [&](int k) {
            const int64_t base = (int64_t)k * fx * fy;
            for (int64_t f = base; f < base + (int64_t)fx * fy; ++f) out[f] = (float)(2.0 * rs.u01((uint64_t)f) - 1.0);
        });
        std::memset(svel[a], 0, sizeof(float) * (size_t)((int64_t)fx * fy * fz));
    }
    const int64_t nd = drop_closed_components(d, dx, solid, fluid);
    if (dropped) *dropped = nd;
    return SHK_OK;
}
