// bd_rng.hip -- host noise of the BD step (host code only).
//
// MT19937 (Matsumoto & Nishimura 1998; the reference uses their 2002 C
// code, mt.cpp:48-141) and the Marsaglia polar normal sampler of
// utilities.cpp:304-336, which returns the two normals of an accepted pair
// on consecutive calls, each scaled by the stddev of the call that returns
// it.  Draw order of one step: two uniforms every 100th step (linker
// histone bind/unbind, main.cpp:386-388), the rotational noise rr[3n] with
// the per-coordinate scale r_var (main.cpp:479-481), then p[n_D3] ~ N(0,1)
// (main.cpp:484-486).  Unlike the reference (seeded from time(0), state
// not saved) the state can be read and restored for reproducible restarts.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>

#include "../../include/bd_hip.h"

struct bd_rng {
    uint32_t mt[624];
    int idx;
    int cached;
    double n2;
};

namespace {

void seed_state(bd_rng *g, uint32_t s) {
    g->mt[0] = s;
    for (int i = 1; i < 624; i++) g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
    g->cached = 0;
    g->n2 = 0.0;
}

void twist(bd_rng *g) {
    for (int k = 0; k < 624; k++) {
        uint32_t y = (g->mt[k] & 0x80000000u) | (g->mt[(k + 1) % 624] & 0x7fffffffu);
        uint32_t v = g->mt[(k + 397) % 624] ^ (y >> 1);
        g->mt[k] = (y & 1u) ? (v ^ 0x9908b0dfu) : v;
    }
    g->idx = 0;
}

uint32_t next_u32(bd_rng *g) {
    if (g->idx >= 624) twist(g);
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

inline double real1(bd_rng *g) { return next_u32(g) * (1.0 / 4294967295.0); }

double normal(bd_rng *g, double mean, double sd) {
    if (g->cached) {
        g->cached = 0;
        return g->n2 * sd + mean;
    }
    double x, y, r;
    do {
        x = 2.0 * real1(g) - 1;
        y = 2.0 * real1(g) - 1;
        r = x * x + y * y;
    } while (r == 0.0 || r > 1.0);
    double d = std::sqrt(-2.0 * std::log(r) / r);
    g->n2 = y * d;
    g->cached = 1;
    return x * d * sd + mean;
}

}  // namespace

extern "C" {

int bd_rng_create(unsigned int seed, bd_rng **out) {
    if (!out) return BD_E_ARG;
    bd_rng *g = new (std::nothrow) bd_rng;
    if (!g) return BD_E_ARG;
    seed_state(g, seed);
    *out = g;
    return BD_OK;
}

int bd_rng_destroy(bd_rng *g) {
    delete g;
    return BD_OK;
}

unsigned int bd_rng_u32(bd_rng *g) { return next_u32(g); }
double bd_rng_uniform(bd_rng *g) { return real1(g); }
double bd_rng_normal(bd_rng *g, double mean, double stddev) { return normal(g, mean, stddev); }

int bd_rng_step_noise(bd_rng *g, int step, int n3, const double *r_var, int n_D3, double *rr, double *p,
                      double *lh_uniforms) {
    if (!g || (n3 && (!r_var || !rr)) || (n_D3 && !p)) return BD_E_ARG;
    if (step % 100 == 0) {
        double u0 = real1(g), u1 = real1(g);
        if (lh_uniforms) { lh_uniforms[0] = u0; lh_uniforms[1] = u1; }
    }
    for (int i = 0; i < n3; i++) rr[i] = normal(g, 0, r_var[i]);
    for (int i = 0; i < n_D3; i++) p[i] = normal(g, 0, 1);
    return BD_OK;
}

int bd_rng_get_state(bd_rng *g, unsigned int *words, int *idx, int *cached, double *n2) {
    if (!g || !words || !idx || !cached || !n2) return BD_E_ARG;
    std::memcpy(words, g->mt, sizeof g->mt);
    *idx = g->idx;
    *cached = g->cached;
    *n2 = g->n2;
    return BD_OK;
}

int bd_rng_set_state(bd_rng *g, const unsigned int *words, int idx, int cached, double n2) {
    if (!g || !words || idx < 0 || idx > 624) return BD_E_ARG;
    std::memcpy(g->mt, words, sizeof g->mt);
    g->idx = idx;
    g->cached = cached ? 1 : 0;
    g->n2 = n2;
    return BD_OK;
}

}  // extern "C"
