// Debug dumps of the substep's intermediate buffers and the seeded scene initialisers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "sph_host.hpp"

namespace sph {

namespace {

template <class T>
int copy_back(hipStream_t s, std::vector<T> &dst, const void *src, size_t count) {
    dst.resize(count);
    hipError_t e = hipMemcpyAsync(dst.data(), src, count * sizeof(T), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? SPH_OK : hip_fail(e, "debug dump copy");
}

}  // namespace

int dump_stage(const Chain &c, hipStream_t s, int stage, uint32_t p_begin, uint32_t p_end,
               uint32_t t_begin, uint32_t t_end, void *dst, size_t bytes) {
    const size_t n = p_end - p_begin, T = t_end - t_begin;
    size_t need = 0;
    switch (stage) {
        case SPH_DUMP_CELL_KEYS: case SPH_DUMP_SORTED_KEYS: case SPH_DUMP_SORTED_INDEX:
        case SPH_DUMP_DENSITY: case SPH_DUMP_NEAR_DENSITY: need = n * 4; break;
        case SPH_DUMP_START: case SPH_DUMP_END: need = T * 4; break;
        case SPH_DUMP_PRED_POS: case SPH_DUMP_PRED_VEL: case SPH_DUMP_VEL_FORCE: need = n * 16; break;
        default: return fail(SPH_ERR_INVALID_ARG, "unknown dump stage %d", stage);
    }
    if (bytes < need) return fail(SPH_ERR_INVALID_ARG, "dump buffer too small (%zu < %zu)", bytes, need);
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    int rc = SPH_OK;
    // Per-slot intermediates of the per-lane path are in cell order: hash slot of each.
    std::vector<uint32_t> hs;
    if (c.cell_order && (stage == SPH_DUMP_DENSITY || stage == SPH_DUMP_NEAR_DENSITY ||
                         stage == SPH_DUMP_PRED_POS || stage == SPH_DUMP_PRED_VEL ||
                         stage == SPH_DUMP_VEL_FORCE)) {
        if ((rc = copy_back(s, hs, c.b.hslot + p_begin, n))) return rc;
        for (uint32_t &h : hs) h -= p_begin;
    }
    auto at = [&](size_t i) { return hs.empty() ? i : (size_t)hs[i]; };
    switch (stage) {
        case SPH_DUMP_CELL_KEYS: case SPH_DUMP_SORTED_KEYS: {
            std::vector<uint32_t> k;
            const uint32_t *src = stage == SPH_DUMP_CELL_KEYS ? c.b.keys : c.sorted_keys();
            if ((rc = copy_back(s, k, src + p_begin, n))) return rc;
            uint32_t *o = (uint32_t *)dst;
            for (size_t i = 0; i < n; ++i) o[i] = k[i] - t_begin;
            break;
        }
        case SPH_DUMP_SORTED_INDEX: {
            uint32_t *o = (uint32_t *)dst;
            if (!c.sorted_idx()) {
                for (size_t i = 0; i < n; ++i) o[i] = (uint32_t)i;
                break;
            }
            std::vector<uint32_t> v;
            if ((rc = copy_back(s, v, c.sorted_idx() + p_begin, n))) return rc;
            for (size_t i = 0; i < n; ++i) o[i] = v[i] - p_begin;
            break;
        }
        case SPH_DUMP_START: case SPH_DUMP_END: {
            std::vector<int32_t> t;
            if ((rc = copy_back(s, t, c.b.table + t_begin, 2 * T))) return rc;
            int32_t *o = (int32_t *)dst;
            const size_t off = stage == SPH_DUMP_START ? 0 : 1;
            for (size_t i = 0; i < T; ++i) {
                const int32_t v = t[2 * i + off];
                o[i] = v == -1 ? -1 : v - (int32_t)p_begin;
            }
            break;
        }
        case SPH_DUMP_DENSITY: case SPH_DUMP_NEAR_DENSITY: {
            std::vector<float> d;
            if ((rc = copy_back(s, d, c.b.dens + p_begin, 2 * n))) return rc;
            float *o = (float *)dst;
            const size_t off = stage == SPH_DUMP_DENSITY ? 0 : 1;
            for (size_t i = 0; i < n; ++i) o[at(i)] = d[2 * i + off];
            break;
        }
        case SPH_DUMP_PRED_POS: case SPH_DUMP_PRED_VEL: case SPH_DUMP_VEL_FORCE: {
            const float4 *src = stage == SPH_DUMP_PRED_POS ? c.b.pred_pos
                              : stage == SPH_DUMP_PRED_VEL ? c.b.dbg_pred_vel : c.b.dbg_vel_force;
            if (!src) return fail(SPH_ERR_STATE, "stage %d needs debug capture (sph_world_set_debug)", stage);
            std::vector<float> v;
            if ((rc = copy_back(s, v, src + p_begin, 4 * n))) return rc;
            // the predicted position's w is 1 (predictPositions.comp); the buffer's w lane
            // carries the near density once the per-lane density pass has run
            if (stage == SPH_DUMP_PRED_POS)
                for (size_t i = 0; i < n; ++i) v[4 * i + 3] = 1.0f;
            for (size_t i = 0; i < n; ++i) std::memcpy((float *)dst + 4 * at(i), &v[4 * i], 16);
            break;
        }
    }
    return rc;
}

// PCG32 (O'Neill, "PCG: A Family of Simple Fast Space-Efficient Statistically Good
// Algorithms for Random Number Generation", 2014), XSH-RR output, 64-bit LCG state.
struct Pcg32 {
    uint64_t state = 0, inc = 0;
    Pcg32(uint64_t seed, uint64_t seq) {
        inc = (seq << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    double uniform() { return (next() >> 8) * (1.0 / 16777216.0); }          // [0, 1)
    double uniform_open() { return ((next() >> 8) + 1.0) * (1.0 / 16777216.0); }  // (0, 1]
};

// The C library rand() the reference draws its initial block from (randf, renderer.c:506-
// 508, after srand(0), main.c:351), for the two libraries its build can link.
//
// glibc: rand() is random() with the default TYPE_3 state, an additive lagged Fibonacci
// generator r[i] = r[i-3] + r[i-31] (mod 2^32) whose 31 seed words come from the
// Park-Miller LCG r[i] = 16807 r[i-1] mod (2^31 - 1) (seed 0 taken as 1); the first 310
// sums are discarded and each output is r[i] >> 1 (RAND_MAX = 2^31 - 1).
// MSVCRT: s = s * 214013 + 2531011 (mod 2^32), output (s >> 16) & 0x7FFF (RAND_MAX = 32767).
struct LibcRand {
    int kind;
    uint32_t r[34];  // glibc: the last 34 words of the sequence, r[i % 34]
    uint64_t i = 0;
    uint32_t s = 0;  // MSVCRT state
    LibcRand(int kind_, uint32_t seed) : kind(kind_) {
        if (kind == SPH_RNG_MSVCRT) {
            s = seed;
            return;
        }
        int32_t w[344];
        w[0] = (int32_t)(seed == 0 ? 1u : seed);
        for (int k = 1; k < 31; ++k) {
            // 16807 w mod (2^31 - 1) in Schrage's form, as glibc evaluates it (this also
            // fixes the result for seeds >= 2^31, which enter as negative words)
            const int64_t hi = (int64_t)w[k - 1] / 127773, lo = (int64_t)w[k - 1] % 127773;
            int64_t v = 16807 * lo - 2836 * hi;
            if (v < 0) v += 2147483647;
            w[k] = (int32_t)v;
        }
        for (int k = 31; k < 34; ++k) w[k] = w[k - 31];
        for (int k = 34; k < 344; ++k) w[k] = (int32_t)((uint32_t)w[k - 31] + (uint32_t)w[k - 3]);
        for (int k = 0; k < 34; ++k) r[(310 + k) % 34] = (uint32_t)w[310 + k];
        i = 344;
    }
    uint32_t next() {
        if (kind == SPH_RNG_MSVCRT) {
            s = s * 214013u + 2531011u;
            return (s >> 16) & 0x7FFFu;
        }
        const uint32_t v = r[(i - 31) % 34] + r[(i - 3) % 34];
        r[i % 34] = v;
        ++i;
        return v >> 1;
    }
    // randf(): (float)rand() / (RAND_MAX + 1.0f), in float
    float randf() {
        const float scale = kind == SPH_RNG_MSVCRT ? 32768.0f : 2147483648.0f;
        return (float)next() / scale;
    }
};

}  // namespace sph

extern "C" int sph_scene_fill_reference(const sph_params *params, uint32_t n, uint32_t seed,
                                        int rng, uint32_t resets, float *pos4, float *vel4) {
    using namespace sph;
    if (!params || !pos4 || !vel4) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    if (rng != SPH_RNG_GLIBC && rng != SPH_RNG_MSVCRT) return fail(SPH_ERR_INVALID_ARG, "unknown generator %d", rng);
    LibcRand g(rng, seed);
    // each earlier 'R' reset (main.c:85-98) drew 4 numbers per particle from the same stream
    for (uint64_t k = 0; k < 4ull * n * resets; ++k) (void)g.next();
    const float scale[3] = {0.25f, 0.45f, 0.45f};
    const float r = params->particle_radius;
    for (uint32_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            // randf() * (scale * (max - min) - 2 * r) + r + min evaluates in float; the
            // centring term (1 - scale) * 0.5 * (max - min) is double (0.5 is a double
            // literal), so the sum is formed in double and rounded once on assignment
            const float extent = params->max_xyz[a] - params->min_xyz[a];
            const float span = scale[a] * extent - 2.0f * r;
            const float head = g.randf() * span + r + params->min_xyz[a];
            pos4[4 * (size_t)i + a] = (float)((double)head + (double)(1.0f - scale[a]) * 0.5 * (double)extent);
        }
        pos4[4 * (size_t)i + 3] = 0.0f;
        (void)g.randf();  // the `randf() < 0.5` draw of renderer.c:587 (its shift s is 0)
        // initializeVelocitiesBuffer (renderer.c:627-645)
        vel4[4 * (size_t)i + 0] = 0.0f;
        vel4[4 * (size_t)i + 1] = 0.0f;
        vel4[4 * (size_t)i + 2] = -100.0f;
        vel4[4 * (size_t)i + 3] = 0.0f;
    }
    return SPH_OK;
}

extern "C" void sph_params_reference_default(sph_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    // main.c:11-30 and the box of main.c:356-358 (s = 40)
    const float s = 40.0f;
    p->min_xyz[0] = (float)(-3.5 * s); p->min_xyz[1] = (float)(-1.00 * s); p->min_xyz[2] = (float)(-1.00 * s);
    p->max_xyz[0] = (float)(3.5 * s);  p->max_xyz[1] = (float)(1.00 * s);  p->max_xyz[2] = (float)(1.00 * s);
    p->particle_radius = (float)0.3;
    p->smoothing_radius = (float)2.5;
    p->g = (float)100;
    p->particle_mass = 1.0f;
    p->pressure_multiplier = (float)15000;
    p->target_density = (float)1.0;
    p->viscosity_strength = (float)0.0;
    p->near_pressure_multiplier = (float)0;
    p->table_cells = 0;  // T = N (renderer.c:741-754)
    sph_params_derive(p, 0.01, 1, 2.5);
}

extern "C" int sph_scene_fill_block(uint32_t n, uint64_t seed, const float lo[3], const float hi[3],
                                    float vel_sigma, float *pos4, float *vel4) {
    using namespace sph;
    if (!lo || !hi || !pos4 || !vel4) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    for (int a = 0; a < 3; ++a)
        if (!(hi[a] >= lo[a])) return fail(SPH_ERR_INVALID_ARG, "empty block");
    Pcg32 rp(seed, 1), rv(seed, 2);
    const double two_pi = 6.283185307179586;
    for (uint32_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            float x = (float)((double)lo[a] + rp.uniform() * ((double)hi[a] - (double)lo[a]));
            if (x > hi[a]) x = hi[a];
            pos4[4 * (size_t)i + a] = x;
        }
        pos4[4 * (size_t)i + 3] = 0.0f;
        for (int a = 0; a < 3; ++a) {
            float v = 0.0f;
            if (vel_sigma > 0.0f) {
                const double u1 = rv.uniform_open(), u2 = rv.uniform();
                v = (float)((double)vel_sigma * std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2));
            }
            vel4[4 * (size_t)i + a] = v;
        }
        vel4[4 * (size_t)i + 3] = 0.0f;
    }
    return SPH_OK;
}
