// synth.cpp — seeded procedural room: the input of the tests and of the
// benchmark (the reference is driven by TUM sequences or a live PrimeSense,
// MainTum.cc / Main.cc, neither of which exists here).
//
// Scene: a 4 x 3 x 3 m room (inside faces), 6 axis-aligned boxes and 4
// spheres with sizes uniform in 0.2-0.8 m, placed in front of the start pose.
// Depth: analytic ray cast, uint16 at `depth_units_per_m`, Gaussian noise
// sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 m, `dropout` of pixels set to 0.
// RGB: value-noise texture of the 3D hit point (a 4 cm lattice, trilinear)
// times a per-object colour. Labels: object k -> class k + 1, room -> 0.
// Trajectory: smooth sinusoids, at most 1 cm and 0.5 degree per frame; the yaw
// sweeps the camera from the front wall to the right wall over 300 frames.
// Randomness: SplitMix64 streams seeded by (seed) for the scene and by
// (seed, frame) for the per-frame noise; Gaussians by Box-Muller.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/sicp.h"

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct Obj {
    int kind;  // 0 box, 1 sphere
    double c[3], h[3], r;
    double col[3];
};

struct Vec {
    double x, y, z;
};

}  // namespace

struct sicp_synth {
    sicp_synth_config cfg;
    std::vector<Obj> objs;
    double room_lo[3], room_hi[3];
    double pose0_inv[16];
};

namespace {

// Camera -> room pose of frame i.
void room_pose(int i, double* T) {
    const double a = 0.02;
    double px = 0.5 * std::sin(a * i);
    double py = 0.5 + 0.15 * std::sin(2.0 * a * i);
    double pz = -0.5 + 0.2 * (1.0 - std::cos(a * i));
    // yaw sweeps 0 -> 97 degrees over 300 frames at <= 0.5 degree per frame
    double yaw = 0.85 * (1.0 - std::cos(6.283185307179586 * i / 600.0));
    double pitch = -0.25 + 0.1 * std::sin(0.05 * i + 1.0);
    double cy = std::cos(yaw), sy = std::sin(yaw), cp = std::cos(pitch), sp = std::sin(pitch);
    // R = Ry(yaw) * Rx(pitch)
    double R[9] = {cy, sy * sp, sy * cp, 0, cp, -sp, -sy, cy * sp, cy * cp};
    for (int r = 0; r < 3; ++r) {
        T[r * 4 + 0] = R[r * 3 + 0];
        T[r * 4 + 1] = R[r * 3 + 1];
        T[r * 4 + 2] = R[r * 3 + 2];
    }
    T[3] = px;
    T[7] = py;
    T[11] = pz;
    T[12] = T[13] = T[14] = 0;
    T[15] = 1;
}

void rigid_inv(const double* T, double* Ti) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Ti[i * 4 + j] = T[j * 4 + i];
    for (int i = 0; i < 3; ++i)
        Ti[i * 4 + 3] = -(Ti[i * 4 + 0] * T[3] + Ti[i * 4 + 1] * T[7] + Ti[i * 4 + 2] * T[11]);
    Ti[12] = Ti[13] = Ti[14] = 0;
    Ti[15] = 1;
}

void mul4(const double* A, const double* B, double* C) {
    double t[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += A[i * 4 + k] * B[k * 4 + j];
            t[i * 4 + j] = s;
        }
    std::memcpy(C, t, sizeof(t));
}

uint32_t hash3(int x, int y, int z, uint64_t seed) {
    uint64_t h = (uint64_t)(uint32_t)x * 0x9e3779b1ull ^ (uint64_t)(uint32_t)y * 0x85ebca77ull ^
                 (uint64_t)(uint32_t)z * 0xc2b2ae3dull ^ seed;
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    return (uint32_t)h;
}

double value_noise(double x, double y, double z, uint64_t seed) {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    double tx = x - fx, ty = y - fy, tz = z - fz;
    double v[8];
    for (int k = 0; k < 8; ++k)
        v[k] = (hash3(ix + (k & 1), iy + ((k >> 1) & 1), iz + (k >> 2), seed) & 0xffff) / 65535.0;
    double a = v[0] + (v[1] - v[0]) * tx, b = v[2] + (v[3] - v[2]) * tx;
    double c = v[4] + (v[5] - v[4]) * tx, d = v[6] + (v[7] - v[6]) * tx;
    double e = a + (b - a) * ty, f = c + (d - c) * ty;
    return e + (f - e) * tz;
}

// Nearest hit along o + t d (t > 0); returns t, object id (-1 room) and point.
double trace(const sicp_synth* s, const Vec& o, const Vec& d, int* id) {
    double best = 1e30;
    int bid = -1;
    // room interior: exit distance of the slab
    const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    for (int a = 0; a < 3; ++a) {
        if (dd[a] > 1e-12) {
            double t = (s->room_hi[a] - oo[a]) / dd[a];
            if (t > 0 && t < best) best = t;
        } else if (dd[a] < -1e-12) {
            double t = (s->room_lo[a] - oo[a]) / dd[a];
            if (t > 0 && t < best) best = t;
        }
    }
    for (size_t k = 0; k < s->objs.size(); ++k) {
        const Obj& ob = s->objs[k];
        if (ob.kind == 0) {
            double t0 = -1e30, t1 = 1e30;
            bool miss = false;
            for (int a = 0; a < 3 && !miss; ++a) {
                double lo = ob.c[a] - ob.h[a], hi = ob.c[a] + ob.h[a];
                if (std::fabs(dd[a]) < 1e-12) {
                    if (oo[a] < lo || oo[a] > hi) miss = true;
                } else {
                    double ta = (lo - oo[a]) / dd[a], tb = (hi - oo[a]) / dd[a];
                    if (ta > tb) std::swap(ta, tb);
                    if (ta > t0) t0 = ta;
                    if (tb < t1) t1 = tb;
                }
            }
            if (!miss && t0 <= t1 && t0 > 1e-6 && t0 < best) {
                best = t0;
                bid = (int)k;
            }
        } else {
            double ox = oo[0] - ob.c[0], oy = oo[1] - ob.c[1], oz = oo[2] - ob.c[2];
            double A = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
            double B = 2 * (ox * dd[0] + oy * dd[1] + oz * dd[2]);
            double C = ox * ox + oy * oy + oz * oz - ob.r * ob.r;
            double disc = B * B - 4 * A * C;
            if (disc >= 0) {
                double t = (-B - std::sqrt(disc)) / (2 * A);
                if (t > 1e-6 && t < best) {
                    best = t;
                    bid = (int)k;
                }
            }
        }
    }
    *id = bid;
    return best;
}

}  // namespace

extern "C" {

int sicp_synth_create(const sicp_synth_config* cfg, sicp_synth** out) {
    if (!cfg || !out || cfg->cols <= 0 || cfg->rows <= 0) return SICP_ERR_INVALID_ARGUMENT;
    sicp_synth* s = new sicp_synth();
    s->cfg = *cfg;
    s->room_lo[0] = -2.0;
    s->room_hi[0] = 2.0;
    s->room_lo[1] = -1.5;
    s->room_hi[1] = 1.5;
    s->room_lo[2] = -1.5;
    s->room_hi[2] = 1.5;
    SplitMix rng{cfg->seed * 0x632be59bd9b4e019ull + 0x1234567ull};
    for (int k = 0; k < 10; ++k) {
        Obj o{};
        o.kind = k < 6 ? 0 : 1;
        double size[3];
        for (int a = 0; a < 3; ++a) size[a] = 0.2 + 0.6 * rng.uniform();
        // In front of the trajectory: x in [-1.4, 1.4], z in [0.6, 1.2].
        o.c[0] = -1.4 + 2.8 * rng.uniform();
        o.c[2] = 0.6 + 0.6 * rng.uniform();
        if (o.kind == 0) {
            for (int a = 0; a < 3; ++a) o.h[a] = 0.5 * size[a];
            o.c[1] = s->room_hi[1] - o.h[1];  // standing on the floor (y down)
        } else {
            o.r = 0.5 * size[0];
            o.c[1] = -0.6 + 1.6 * rng.uniform();
        }
        for (int a = 0; a < 3; ++a) o.col[a] = 0.3 + 0.7 * rng.uniform();
        s->objs.push_back(o);
    }
    double T0[16];
    room_pose(0, T0);
    rigid_inv(T0, s->pose0_inv);
    *out = s;
    return SICP_OK;
}

int sicp_synth_destroy(sicp_synth* s) {
    delete s;
    return SICP_OK;
}

int sicp_synth_pose(const sicp_synth* s, int frame, double pose[16]) {
    if (!s || !pose) return SICP_ERR_INVALID_ARGUMENT;
    double T[16];
    room_pose(frame, T);
    mul4(s->pose0_inv, T, pose);
    return SICP_OK;
}

int sicp_synth_render(const sicp_synth* s, int frame, uint16_t* depth, uint8_t* rgb,
                      uint8_t* labels) {
    if (!s) return SICP_ERR_INVALID_ARGUMENT;
    const sicp_synth_config& c = s->cfg;
    double T[16];
    room_pose(frame, T);
    SplitMix noise{(c.seed + 1) * 0x9e3779b97f4a7c15ull ^ ((uint64_t)(frame + 1) * 0xd1b54a32d192ed03ull)};
    const uint64_t tex_seed = c.seed * 0x2545f4914f6cdd1dull + 77;
    for (int y = 0; y < c.rows; ++y) {
        for (int x = 0; x < c.cols; ++x) {
            const size_t i = (size_t)y * c.cols + x;
            double dc[3] = {(x - c.cx) / c.fx, (y - c.cy) / c.fy, 1.0};
            Vec o{T[3], T[7], T[11]};
            Vec d{T[0] * dc[0] + T[1] * dc[1] + T[2] * dc[2], T[4] * dc[0] + T[5] * dc[1] + T[6] * dc[2],
                  T[8] * dc[0] + T[9] * dc[1] + T[10] * dc[2]};
            int id;
            double z = trace(s, o, d, &id);  // the ray has unit camera-z, so t = depth
            double px = o.x + z * d.x, py = o.y + z * d.y, pz = o.z + z * d.z;
            double n1 = noise.normal();
            double u = noise.uniform();
            double zn = z;
            if (c.noise) {
                double sigma = 0.0012 + 0.0019 * (z - 0.4) * (z - 0.4);
                zn = z + sigma * n1;
            }
            double q = std::floor(zn * c.depth_units_per_m + 0.5);
            if (q < 0) q = 0;
            if (q > 65535) q = 65535;
            if (u < c.dropout) q = 0;
            if (depth) depth[i] = (uint16_t)q;
            double tex = value_noise(px / 0.04, py / 0.04, pz / 0.04, tex_seed);
            double base[3] = {0.75, 0.72, 0.68};
            if (id >= 0)
                for (int a = 0; a < 3; ++a) base[a] = s->objs[id].col[a];
            if (rgb)
                for (int a = 0; a < 3; ++a) {
                    double v = 255.0 * base[a] * (0.55 + 0.45 * tex);
                    rgb[i * 3 + a] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : (int)v));
                }
            if (labels) labels[i] = (uint8_t)(id >= 0 ? id + 1 : 0);
        }
    }
    return SICP_OK;
}

}  // extern "C"
