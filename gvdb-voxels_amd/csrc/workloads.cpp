// workloads.cpp -- seeded synthetic inputs shared by tests and the benchmark
// (SplitMix64 seeding xoshiro256**).  Host code only; no reference
// counterpart.
#include <math.h>
#include <stdint.h>

#include "gvdbx/gvdbx_c.h"

namespace {

uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Xoshiro256ss {
    uint64_t s[4];
    explicit Xoshiro256ss(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; i++) s[i] = splitmix64(x);
    }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform double in [0,1)
    double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        double u1 = u01(), u2 = u01();
        if (u1 < 1e-300) u1 = 1e-300;
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
};

}  // namespace

extern "C" {

int gvdbx_gen_shell(uint64_t seed, int64_t n, float cx, float cy, float cz, float r0, float r1, float* out) {
    if (n < 0 || !out || !(r1 >= r0) || r0 < 0) return GVDBX_ERR_ARG;
    Xoshiro256ss g(seed);
    double a3 = (double)r0 * r0 * r0, b3 = (double)r1 * r1 * r1;
    for (int64_t i = 0; i < n; i++) {
        double x, y, z, l;
        do {
            x = g.normal();
            y = g.normal();
            z = g.normal();
            l = sqrt(x * x + y * y + z * z);
        } while (l < 1e-12);
        double r = cbrt(a3 + (b3 - a3) * g.u01());
        out[3 * i + 0] = (float)(cx + r * x / l);
        out[3 * i + 1] = (float)(cy + r * y / l);
        out[3 * i + 2] = (float)(cz + r * z / l);
    }
    return GVDBX_OK;
}

int gvdbx_gen_box(uint64_t seed, int64_t n, const float* lo, const float* hi, float jitter, float* out) {
    if (n < 0 || !out || !lo || !hi) return GVDBX_ERR_ARG;
    Xoshiro256ss g(seed);
    for (int64_t i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            double u = lo[a] + (hi[a] - lo[a]) * g.u01();
            if (jitter != 0.0f) u += jitter * (2.0 * g.u01() - 1.0);
            out[3 * i + a] = (float)u;
        }
    return GVDBX_OK;
}

}  // extern "C"
