// Per-run neighbour accumulation shared by every density / force kernel variant, so all of
// them evaluate the same float expressions in the same order (bit-identical results).
#pragma once

#include "sph_device.hpp"

namespace sph {

// calculateDensities.comp:122-128 over one cell run: pp[0..cnt) are the run's predicted
// positions.  smoothingKernel / nearDensitySmoothingKernel are 0 for dist >= R, and
// M + m*0 == M for M >= +0, so those candidates are skipped; d2 > R2_pre implies
// sqrt(d2) >= R, which lets the sqrt be skipped for them too.
__device__ __forceinline__ void density_run(const SceneConst &c, const float4 *__restrict__ pp,
                                            int cnt, const float4 &s, float &M, float &Mn) {
    for (int k = 0; k < cnt; ++k) {
        const float4 p = pp[k];
        const float ex = p.x - s.x, ey = p.y - s.y, ez = p.z - s.z;
        const float d2 = ex * ex + ey * ey + ez * ez;
        if (d2 > c.R2_pre) continue;
        const float dist = sqrtf(d2);
        if (dist >= c.R) continue;
        const float t = c.R - dist;
        M = M + c.m * (t * t);
        Mn = Mn + c.m * (t * t * t);
    }
}

// calculateDensities.comp:133-152
__device__ __forceinline__ float2 density_finish(const SceneConst &c, float M, float Mn) {
    const float density = M / c.volume;
    const float near_density = Mn * c.near_vol_inv;
    return make_float2(fmaxf(0.01f, density), fmaxf(0.01f, near_density));
}

struct ForceSelf {
    float4 s;            // predictedPositions[j]
    float3 u;            // predictedVelocities[j]
    float my_pressure;   // convertDensityToPressure(densities[j])
    float near_pressure; // nearDensities[j] * nearPressureMultiplier
    uint32_t j;          // slot of the particle
};

// applyPressureAndViscosityForce.comp:297-322 over one cell run whose first slot is
// `first`: pp/vv/dd hold the run's predicted positions, velocities and
// (density, near density).
__device__ __forceinline__ void force_run(const SceneConst &c, const float4 *__restrict__ pp,
                                          const float4 *__restrict__ vv,
                                          const float2 *__restrict__ dd, int cnt,
                                          uint32_t first, const ForceSelf &me, float3 &Fp,
                                          float3 &Fv) {
    for (int k = 0; k < cnt; ++k) {
        if (first + (uint32_t)k == me.j) continue;
        const float4 pi = pp[k];
        const float ex = pi.x - me.s.x, ey = pi.y - me.s.y, ez = pi.z - me.s.z;
        const float d2 = ex * ex + ey * ey + ez * ez;
        if (d2 > c.R2) continue;
        const float dist = sqrtf(d2);
        if (fabsf(dist) < 1e-7f) continue;
        const float dirx = ex / dist, diry = ey / dist, dirz = ez / dist;
        const float2 di = dd[k];
        const float slope = dist >= c.R ? 0.0f : (dist - c.R) * c.slope_scale;
        const float shared = ((di.x - c.rho0) * c.k + me.my_pressure) / 2.0f;
        const float a = shared * slope * c.m / di.x;
        Fp.x = Fp.x + a * dirx;
        Fp.y = Fp.y + a * diry;
        Fp.z = Fp.z + a * dirz;
        const float infl = dist >= c.R ? 0.0f : (c.R - dist) * (c.R - dist);
        const float b = c.m * infl / di.x;
        const float4 vi = vv[k];
        Fv.x = Fv.x + b * (vi.x - me.u.x);
        Fv.y = Fv.y + b * (vi.y - me.u.y);
        Fv.z = Fv.z + b * (vi.z - me.u.z);
        const float shared_near = (me.near_pressure + di.y * c.kn) / 2.0f;
        const float x = c.R - dist;
        const float nslope = dist >= c.R ? 0.0f : x * x * c.near_slope_scale;
        const float cn = shared_near * nslope * c.m / di.y;
        Fp.x = Fp.x + cn * dirx;
        Fp.y = Fp.y + cn * diry;
        Fp.z = Fp.z + cn * dirz;
    }
}

// Self terms of the force pass for slot j (applyPressureAndViscosityForce.comp:276-286).
__device__ __forceinline__ ForceSelf force_self(const SceneConst &c, uint32_t j, const float4 &s,
                                                const float4 &P, const float4 &V,
                                                const float2 &myd) {
    ForceSelf me;
    me.s = s;
    float3 ps = make_float3(P.x, P.y, P.z);
    me.u = make_float3(V.x, V.y, V.z);
    kick_drift<false>(c, ps, me.u);
    me.my_pressure = (myd.x - c.rho0) * c.k;
    me.near_pressure = myd.y * c.kn;
    me.j = j;
    return me;
}

// Cell coordinate of the force pass's sample point (samplePointToCellCoord,
// applyPressure...comp:224-228): -1 on every axis when outside the box.
__device__ __forceinline__ int3 force_cell(const SceneConst &c, const float4 &s) {
    const bool outside = s.x < c.minx || s.y < c.miny || s.z < c.minz || s.x > c.maxx ||
                         s.y > c.maxy || s.z > c.maxz;
    if (outside) return make_int3(-1, -1, -1);
    return make_int3(cell_of(s.x, c.minx, c.R), cell_of(s.y, c.miny, c.R), cell_of(s.z, c.minz, c.R));
}

// velocities[j] += vec4((Fp + Fv + 0) / densities[j] * dt, 1) and updateVelocities.
__device__ __forceinline__ void force_finish(const SceneConst &c, const float4 &P, const float4 &V,
                                             float my_density, const float3 &Fp, const float3 &Fv,
                                             float4 &pos_out, float4 &vel_out, float4 &vel_force) {
    float3 v = make_float3(V.x + (Fp.x + Fv.x + 0.0f) / my_density * c.dt,
                           V.y + (Fp.y + Fv.y + 0.0f) / my_density * c.dt,
                           V.z + (Fp.z + Fv.z + 0.0f) / my_density * c.dt);
    vel_force = make_float4(v.x, v.y, v.z, V.w + 1.0f);
    float3 p = make_float3(P.x, P.y, P.z);
    kick_drift<true>(c, p, v);
    pos_out = make_float4(p.x, p.y, p.z, 1.0f);
    vel_out = make_float4(v.x, v.y, v.z, 1.0f);
}

__device__ __forceinline__ uint32_t key_of_cell(const SceneConst &c, int32_t x, int32_t y, int32_t z) {
    return c.table_base + (cell_hash((uint32_t)x, (uint32_t)y, (uint32_t)z) & c.mask);
}

}  // namespace sph
