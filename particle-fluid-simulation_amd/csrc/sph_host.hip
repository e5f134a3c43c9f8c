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

}  // namespace sph

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
