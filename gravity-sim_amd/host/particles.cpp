// Particle sets for the drop-in surface: the four generation presets, the
// synthetic disks of the benchmark workloads, and the particle text format.
//
// The presets follow the formulas of the reference generators
// (src/Particles/ParticleSystem.cpp:15-149) operation for operation in FP32,
// but draw from an explicit SplitMix64 stream seeded by the caller instead of
// the time-seeded std::default_random_engine, so every run is reproducible.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "../csrc/common.hpp"
#include "gsim.h"

namespace {

constexpr double kPi = 3.14159265358979323846;  // M_PI

// SplitMix64 (Steele, Lea, Flood 2014): one 64-bit state, one output per call.
class SplitMix64 {
   public:
    explicit SplitMix64(uint64_t seed) : s_(seed) {}
    uint64_t next() {
        uint64_t z = (s_ += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    // Uniform float in [0, 1) with 24 random bits, the range of
    // std::uniform_real_distribution<float>(0, 1).
    float uniform() { return (float)(next() >> 40) * 0x1p-24f; }
    // Uniform double in (0, 1], safe for log().
    double uniform_open() { return ((double)(next() >> 11) + 1.0) * 0x1p-53; }
    // Standard normal by Box-Muller.
    double normal() {
        const double u1 = uniform_open(), u2 = uniform_open();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }

   private:
    uint64_t s_;
};

struct Out {
    float *pos, *vel, *mass;
    void set(size_t i, float px, float py, float vx, float vy, float m) {
        pos[2 * i] = px; pos[2 * i + 1] = py;
        vel[2 * i] = vx; vel[2 * i + 1] = vy;
        mass[i] = m;
    }
};

// ParticleSystem.cpp:15-35
void gen_random(Out o, size_t n, float size, float minm, float maxm, SplitMix64 &rng) {
    for (size_t i = 0; i != n; ++i) {
        const float theta = (float)(rng.uniform() * 2 * kPi);
        const float r = std::sqrt(1 - rng.uniform()) * size;
        const float m = rng.uniform() * (maxm - minm) + minm;
        o.set(i, r * cosf(theta), r * sinf(theta), 0.0f, 0.0f, m);
    }
}

// ParticleSystem.cpp:36-62
void gen_galaxy(Out o, size_t n, float size, float minm, float maxm, float G,
                SplitMix64 &rng) {
    const float avg = (minm + maxm) * 0.5f;
    const float vel = std::sqrt(G * (float)n * avg / (size * size * size));
    for (size_t i = 0; i != n; ++i) {
        const float theta = (float)(rng.uniform() * 2 * kPi);
        const float r = (1 - rng.uniform()) * size;
        const float s = sinf(theta), c = cosf(theta);
        const float m = rng.uniform() * (maxm - minm) + minm;
        o.set(i, r * c, r * s, -s * r * vel, c * r * vel, m);
    }
}

// ParticleSystem.cpp:63-112
void gen_collision(Out o, size_t n, float size, float minm, float maxm, float G,
                   SplitMix64 &rng) {
    const float gs = size / 3;
    const float avg = (minm + maxm) * 0.25f;
    const float vel = std::sqrt(G * (float)n * avg / (gs * gs * gs));
    for (size_t i = 0; i != n; ++i) {
        const float side = i < (n >> 1) ? -1.0f : 1.0f;
        const float theta = (float)(rng.uniform() * 2 * kPi);
        const float r = (1 - rng.uniform()) * gs;
        const float s = sinf(theta), c = cosf(theta);
        const float m = rng.uniform() * (maxm - minm) + minm;
        // first galaxy: x - gs*2, vx + gs*0.1f; second: x + gs*2, vx - gs*0.1f
        const float px = side < 0 ? r * c - gs * 2 : r * c + gs * 2;
        const float vx = side < 0 ? -s * r * vel + gs * 0.1f : -s * r * vel - gs * 0.1f;
        o.set(i, px, r * s, vx, c * r * vel, m);
    }
}

// ParticleSystem.cpp:113-148 (n even)
void gen_symmetric(Out o, size_t n, float size, float minm, float maxm, float G,
                   SplitMix64 &rng) {
    const float gs = size / 3;
    const float avg = (minm + maxm) * 0.25f;
    const float vel = std::sqrt(G * (float)n * avg / (gs * gs * gs));
    for (size_t i = 0; i + 1 < n; i += 2) {
        const float theta = (float)(rng.uniform() * 2 * kPi);
        const float r = (1 - rng.uniform()) * gs;
        const float s = sinf(theta), c = cosf(theta);
        const float m = rng.uniform() * (maxm - minm) + minm;
        o.set(i, r * c - gs * 2, r * s, -s * r * vel + gs * 0.1f, c * r * vel, m);
        o.set(i + 1, -(r * c - gs * 2), r * s, -s * r * vel - gs * 0.1f, -c * r * vel, m);
    }
}

// One exponential disk: surface density ~ exp(-r/h) truncated at rmax, so the
// radius follows Gamma(2, h) (the sum of two exponential draws), redrawn when
// beyond rmax.  Velocities are circular, sqrt(G M(<r) / r) from the sampled
// masses, rotating counter-clockwise, plus a Gaussian of sigma = disp * v_c per
// component, plus the bulk velocity.
void gen_exp_disk(Out o, size_t first, size_t n, float rmax, float h, float cx,
                  float cy, float bulk_vx, float disp, float minm, float maxm,
                  float G, SplitMix64 &rng) {
    std::vector<double> rad(n), ang(n);
    for (size_t k = 0; k < n; ++k) {
        double r;
        do {
            r = -(double)h * std::log(rng.uniform_open() * rng.uniform_open());
        } while (r > rmax);
        rad[k] = r;
        ang[k] = 2.0 * kPi * rng.uniform_open();
        o.mass[first + k] = rng.uniform() * (maxm - minm) + minm;
    }
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), size_t(0));
    std::sort(order.begin(), order.end(),
              [&](size_t a, size_t b) { return rad[a] < rad[b]; });
    double enclosed = 0.0;
    for (size_t q = 0; q < n; ++q) {
        const size_t k = order[q];
        const double r = rad[k];
        const double vc = r > 0 ? std::sqrt((double)G * enclosed / r) : 0.0;
        enclosed += o.mass[first + k];
        const double s = std::sin(ang[k]), c = std::cos(ang[k]);
        double vx = -s * vc, vy = c * vc;
        if (disp > 0) {
            vx += disp * vc * rng.normal();
            vy += disp * vc * rng.normal();
        }
        const size_t i = first + k;
        o.pos[2 * i] = (float)(cx + r * c);
        o.pos[2 * i + 1] = (float)(cy + r * s);
        o.vel[2 * i] = (float)(vx + bulk_vx);
        o.vel[2 * i + 1] = (float)vy;
    }
}

}  // namespace

extern "C" gsim_status gsim_generate(int32_t type, size_t n, float size,
                                     float minm, float maxm, float G,
                                     uint64_t seed, float *pos, float *vel,
                                     float *mass) {
    gsim::clear_error();
    if (!pos || !vel || !mass)
        return gsim::fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_generate: null output array");
    if (n == 0)
        return gsim::fail(GSIM_ERROR_INVALID_ARGUMENT,
                          "The simulation must contain at least one particle!");
    if (!(size > 0) || !std::isfinite(size) || !std::isfinite(minm) || !std::isfinite(maxm))
        return gsim::fail(GSIM_ERROR_INVALID_ARGUMENT,
                          "gsim_generate: generate size and masses must be finite, size > 0");
    SplitMix64 rng(seed);
    Out o{pos, vel, mass};
    switch (type) {
        case GSIM_GENERATE_RANDOM: gen_random(o, n, size, minm, maxm, rng); break;
        case GSIM_GENERATE_GALAXY: gen_galaxy(o, n, size, minm, maxm, G, rng); break;
        case GSIM_GENERATE_GALAXY_COLLISION: gen_collision(o, n, size, minm, maxm, G, rng); break;
        case GSIM_GENERATE_SYMMETRICAL_GALAXY_COLLISION:
            if (n & 1)
                return gsim::fail(GSIM_ERROR_INVALID_ARGUMENT,
                                  "symmetrical-galaxy-collision needs an even particle count");
            gen_symmetric(o, n, size, minm, maxm, G, rng);
            break;
        case GSIM_GENERATE_EXPONENTIAL_DISK:
            gen_exp_disk(o, 0, n, size, size / 5.0f, 0.0f, 0.0f, 0.0f, 0.05f, minm, maxm, G, rng);
            break;
        case GSIM_GENERATE_TWO_EXPONENTIAL_DISKS: {
            // Disk radius = size, scale length = size*7/33, centres at
            // +-size*40/33; the disk at -x moves +2, the one at +x moves -2.
            const size_t half = n / 2;
            const float h = size * 7.0f / 33.0f, c = size * 40.0f / 33.0f;
            gen_exp_disk(o, 0, half, size, h, -c, 0.0f, 2.0f, 0.0f, minm, maxm, G, rng);
            gen_exp_disk(o, half, n - half, size, h, c, 0.0f, -2.0f, 0.0f, minm, maxm, G, rng);
            break;
        }
        default:
            return gsim::fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_generate: unknown type %d", type);
    }
    return GSIM_OK;
}

// ParticleSystem.cpp:462-477: read "%f%f%f%f%f" records until a match fails.
static gsim_status scan_file(const char *path, float *pos, float *vel, float *mass,
                             size_t capacity, size_t *n) {
    if (!path || !n) return gsim::fail(GSIM_ERROR_INVALID_ARGUMENT, "null path or count");
    FILE *f = std::fopen(path, "r");
    if (!f) return gsim::fail(GSIM_ERROR_IO, "Failed to open particle input file! (%s)", path);
    size_t count = 0;
    float p[5];
    while (std::fscanf(f, "%f%f%f%f%f", &p[0], &p[1], &p[2], &p[3], &p[4]) == 5) {
        if (pos && count < capacity) {
            pos[2 * count] = p[0]; pos[2 * count + 1] = p[1];
            vel[2 * count] = p[2]; vel[2 * count + 1] = p[3];
            mass[count] = p[4];
        }
        ++count;
    }
    std::fclose(f);
    if (count == 0)  // ParticleSystem.cpp:483-484
        return gsim::fail(GSIM_ERROR_IO,
                          "The particle input file must contain at least one valid particle!");
    *n = count;
    if (pos && count > capacity)
        return gsim::fail(GSIM_ERROR_CAPACITY, "particle file holds %zu particles, capacity %zu",
                          count, capacity);
    return GSIM_OK;
}

extern "C" gsim_status gsim_count_particles(const char *path, size_t *n) {
    gsim::clear_error();
    return scan_file(path, nullptr, nullptr, nullptr, 0, n);
}

extern "C" gsim_status gsim_load_particles(const char *path, float *pos, float *vel,
                                           float *mass, size_t capacity, size_t *n) {
    gsim::clear_error();
    if (!pos || !vel || !mass)
        return gsim::fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_load_particles: null output array");
    return scan_file(path, pos, vel, mass, capacity, n);
}

// ParticleSystem.cpp:727-750
extern "C" gsim_status gsim_save_particles(const char *path, const float *pos,
                                           const float *vel, const float *mass, size_t n) {
    gsim::clear_error();
    if (!path || !pos || !vel || !mass)
        return gsim::fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_save_particles: null argument");
    FILE *f = std::fopen(path, "w");
    if (!f) return gsim::fail(GSIM_ERROR_IO, "Failed to open particle output file! (%s)", path);
    for (size_t i = 0; i != n; ++i)
        std::fprintf(f, "%.7f %.7f %.7f %.7f %.7f\n", pos[2 * i], pos[2 * i + 1], vel[2 * i],
                     vel[2 * i + 1], mass[i]);
    const bool ok = std::fclose(f) == 0;
    return ok ? GSIM_OK : gsim::fail(GSIM_ERROR_IO, "failed to close %s", path);
}
