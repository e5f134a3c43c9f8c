"""Pins of the CPU oracle (oracle/sph_ref32.c, sph_ref64.c) against facts that do not come
from the oracle's own code: known-answer hashes, numpy's stable sort and run detection,
closed forms of the kernels, float64 all-pairs sums, physics invariants, and the FP64
restatement.  The reference ships no tests, fixtures or golden vectors for this path."""
import math

import numpy as np
import pytest

# Known answers of cellCoordToHash (updateCellEntries.comp:36-52) and its masks.
HASH_KAT = [
    ((0, 0, 0), 0x00000000, 0, 0, 0),
    ((1, 0, 0), 0x9D5A7ACD, 15053, 686797, 1735373),
    ((0, 1, 0), 0x2859CC20, 3104, 642080, 1690656),
    ((0, 0, 1), 0x358917CF, 6095, 595919, 595919),
    ((3, 5, 7), 0x03F8DAE0, 6880, 580320, 1628896),
    ((0xFFFFFFFF, 0, 0), 0x43340E54, 3668, 265812, 1314388),
]


def test_hash_fixture_matches_oracle_and_table(orc):
    """tests/golden/hash_kat.json (written by make_hash_kat.py from the shader's steps in
    plain Python integers) agrees with the oracle on every cell and with the table above
    on the cells they share."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hash_kat.json")
    kat = json.load(open(path))
    table = {c: h for c, h, *_ in HASH_KAT}
    assert len(kat) >= len(table)
    for e in kat:
        cell = tuple(e["cell"])
        assert orc.cell_hash(*cell) == e["hash"], cell
        for bits, key in e["key"].items():
            assert e["hash"] & ((1 << int(bits)) - 1) == key
        if cell in table:
            assert table[cell] == e["hash"], cell
    assert all(c in {tuple(e["cell"]) for e in kat} for c in table)


def _params(orc_mod, **kw):
    import sph_hip
    base = dict(particle_radius=0.3, smoothing_radius=0.35, g=9.8, box_min=(0, 0, 0),
                box_max=(10, 6, 6), frame_dt=1 / 120, substeps=1, pressure_multiplier=500.0,
                target_density=55.0, viscosity_strength=0.06, table_cells=0)
    base.update(kw)
    return sph_hip.params_for(**base)


@pytest.mark.parametrize("cell,h,k14,k20,k21", HASH_KAT)
def test_hash_known_answers(orc, cell, h, k14, k20, k21):
    assert orc.cell_hash(*cell) == h
    assert h & ((1 << 14) - 1) == k14
    assert h & ((1 << 20) - 1) == k20
    assert h & ((1 << 21) - 1) == k21


def test_hash_keys_of_positions(orc):
    """Keys of positions = known-answer hash of floor((p - min)/R), masked by T-1."""
    import ctypes as C
    p = _params(orc, table_cells=1 << 14)
    pos = np.zeros((4, 4), np.float32)
    pos[0, :3] = (0.1, 0.1, 0.1)        # cell (0,0,0)
    pos[1, :3] = (0.36, 0.2, 0.3)       # cell (1,0,0)
    pos[2, :3] = (1.1, 1.8, 2.5)        # cell (3,5,7)
    pos[3, :3] = (0.2, 0.4, 0.1)        # cell (0,1,0)
    keys = np.empty(4, np.uint32)
    rp = orc.to_ref(p)
    orc.load().sphref_hash(C.byref(rp), 4, pos.ctypes.data, keys.ctypes.data)
    assert keys.tolist() == [0, 15053, 6880, 3104]


@pytest.mark.parametrize("n,T", [(1, 1), (1000, 1024), (5000, 64), (20000, 1 << 20), (333, 7)])
def test_sort_is_numpy_stable_sort(orc, n, T):
    rng = np.random.default_rng(n + T)
    keys = rng.integers(0, T, n, dtype=np.uint32)
    ok = np.empty(n, np.uint32)
    oi = np.empty(n, np.uint32)
    assert orc.load().sphref_sort(n, T, keys.ctypes.data, ok.ctypes.data, oi.ctypes.data) == 0
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(oi, order.astype(np.uint32))
    assert np.array_equal(ok, keys[order])


@pytest.mark.parametrize("n,T", [(1, 4), (1000, 1024), (5000, 64), (3000, 3000)])
def test_table_matches_run_boundaries(orc, n, T):
    rng = np.random.default_rng(7 * n + T)
    sk = np.sort(rng.integers(0, T, n, dtype=np.uint32))
    st = np.empty(T, np.int32)
    en = np.empty(T, np.int32)
    orc.load().sphref_table(n, T, sk.ctypes.data, st.ctypes.data, en.ctypes.data)
    u, first, count = np.unique(sk, return_index=True, return_counts=True)
    exp_st = np.full(T, -1, np.int32)
    exp_en = np.full(T, -1, np.int32)
    exp_st[u] = first
    exp_en[u] = first + count
    assert np.array_equal(st, exp_st) and np.array_equal(en, exp_en)
    assert int((en[en >= 0] - st[st >= 0]).sum()) == n


def _table_for(orc, params, pos4):
    import ctypes as C
    n = pos4.shape[0]
    T = params.table_cells or n
    rp = orc.to_ref(params)
    rp.table_cells = T
    keys = np.empty(n, np.uint32)
    orc.load().sphref_hash(C.byref(rp), n, pos4.ctypes.data, keys.ctypes.data)
    sk = np.empty(n, np.uint32)
    si = np.empty(n, np.uint32)
    orc.load().sphref_sort(n, T, keys.ctypes.data, sk.ctypes.data, si.ctypes.data)
    pos_sorted = np.ascontiguousarray(pos4[si])
    st = np.empty(T, np.int32)
    en = np.empty(T, np.int32)
    orc.load().sphref_table(n, T, sk.ctypes.data, st.ctypes.data, en.ctypes.data)
    return pos_sorted, sk, st, en, rp


def _density(orc, params, pos4):
    import ctypes as C
    ps, sk, st, en, rp = _table_for(orc, params, pos4)
    n = pos4.shape[0]
    d = np.empty(n, np.float32)
    dn = np.empty(n, np.float32)
    orc.load().sphref_density(C.byref(rp), n, ps.ctypes.data, st.ctypes.data, en.ctypes.data,
                              d.ctypes.data, dn.ctypes.data)
    return ps, sk, st, en, d, dn


def test_density_isolated_particle_closed_form(orc):
    """A lone particle sees only itself: rho = m R^2 / (2 pi R^5/15) = 15/(2 pi R^3),
    rho_near = R^3 * 15/(pi R^6) = 15/(pi R^3)."""
    p = _params(orc, table_cells=1 << 16)
    pos = np.array([[5.0, 3.0, 3.0, 0.0]], np.float32)
    _, _, _, _, d, dn = _density(orc, p, pos)
    R = 0.35
    assert d[0] == pytest.approx(15 / (2 * math.pi * R ** 3), rel=2e-6)
    assert dn[0] == pytest.approx(15 / (math.pi * R ** 3), rel=2e-6)


def test_density_duplicate_keys_are_visited_again(orc):
    """With T = N = 1 every one of the 27 visited cells maps to key 0, so the lone
    particle's own run is summed 27 times (duplicate keys are not de-duplicated,
    calculateDensities.comp:114-131)."""
    p = _params(orc)
    pos = np.array([[5.0, 3.0, 3.0, 0.0]], np.float32)
    _, _, _, _, d, _ = _density(orc, p, pos)
    assert d[0] == pytest.approx(27 * 15 / (2 * math.pi * 0.35 ** 3), rel=2e-6)


def test_density_pair_closed_form(orc):
    p = _params(orc, table_cells=1 << 16)
    dist = 0.21
    pos = np.array([[5.0, 3.0, 3.0, 0], [5.0 + dist, 3.0, 3.0, 0]], np.float32)
    _, _, _, _, d, dn = _density(orc, p, pos)
    R = 0.35
    dd = float(np.float32(5.0 + dist)) - 5.0
    V = 2 * math.pi * R ** 5 / 15
    assert d[0] == pytest.approx((R ** 2 + (R - dd) ** 2) / V, rel=1e-5)
    assert dn[1] == pytest.approx((R ** 3 + (R - dd) ** 3) * 15 / (math.pi * R ** 6), rel=1e-5)


def test_density_all_pairs_float64(orc):
    """Interior particles, a table so large that no two visited cells share a key: the hashed
    27-cell sum equals the float64 all-pairs sum over d < R."""
    p = _params(orc, table_cells=1 << 24)
    rng = np.random.default_rng(5)
    n = 1500
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform((2, 1.5, 1.5), (5, 4.5, 4.5), (n, 3)).astype(np.float32)
    ps, _, _, _, d, dn = _density(orc, p, pos)
    x = ps[:, :3].astype(np.float64)
    R = float(np.float32(0.35))
    diff = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    w = np.where(diff < R, R - diff, 0.0)
    V = float(p.volume)
    rho = (w ** 2).sum(1) / V
    rho_n = (w ** 3).sum(1) * 15 / (3.14159265359 * R ** 6)
    inner = np.all((x > (0.5, 0.5, 0.5)) & (x < (9.5, 5.5, 5.5)), axis=1)
    assert inner.sum() == n
    np.testing.assert_allclose(d, rho, rtol=1e-5)
    np.testing.assert_allclose(dn, rho_n, rtol=1e-5)


def test_density_table_vs_linear_scan(orc):
    """Heavy collisions (T = 64): the start/end lookup visits exactly the particles whose
    key matches, in sorted order, duplicates included (bit-exact vs a linear scan)."""
    p = _params(orc, table_cells=64)
    rng = np.random.default_rng(9)
    n = 600
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform((0, 0, 0), (3, 3, 3), (n, 3)).astype(np.float32)
    ps, sk, _, _, d, dn = _density(orc, p, pos)
    bd, bdn = orc.density_bruteforce(p, ps, sk)
    assert np.array_equal(d.view(np.uint32), bd.view(np.uint32))
    assert np.array_equal(dn.view(np.uint32), bdn.view(np.uint32))


def test_density_sample_point_skips_cells_past_the_wall(orc):
    """calculateDensities.comp:117-118 skips a cell whose offset sample point lies outside
    the box even when a neighbour inside the box lives there: s.x = 9.7, neighbour at 9.85
    (next cell, d = 0.15 < R), s.x + R = 10.05 > max -> only the self term remains."""
    p = _params(orc, table_cells=1 << 16)
    pos = np.array([[9.7, 3.0, 3.0, 0], [9.85, 3.0, 3.0, 0]], np.float32)
    ps, _, _, _, d, _ = _density(orc, p, pos)
    R = 0.35
    i = int(np.argmin(ps[:, 0]))
    assert d[i] == pytest.approx(15 / (2 * math.pi * R ** 3), rel=2e-6)
    # the particle at 9.85 looks back into cell 27 (sample point 9.5 inside) and sees both
    assert d[1 - i] > d[i] * 1.2


def _force(orc, params, pos4, vel4):
    import ctypes as C
    ps, sk, st, en, rp = _table_for(orc, params, pos4)
    n = pos4.shape[0]
    order = np.argsort(_keys(orc, params, pos4), kind="stable")
    vs = np.ascontiguousarray(vel4[order])
    pps = np.empty_like(ps)
    pvs = np.empty_like(vs)
    lib = orc.load()
    lib.sphref_predict(C.byref(rp), n, ps.ctypes.data, vs.ctypes.data, pps.ctypes.data, pvs.ctypes.data)
    d = np.empty(n, np.float32)
    dn = np.empty(n, np.float32)
    lib.sphref_density(C.byref(rp), n, ps.ctypes.data, st.ctypes.data, en.ctypes.data,
                       d.ctypes.data, dn.ctypes.data)
    out = np.empty_like(vs)
    lib.sphref_force(C.byref(rp), n, ps.ctypes.data, pvs.ctypes.data, vs.ctypes.data,
                     d.ctypes.data, dn.ctypes.data, st.ctypes.data, en.ctypes.data, out.ctypes.data)
    return ps, vs, pvs, d, dn, out


def _keys(orc, params, pos4):
    import ctypes as C
    n = pos4.shape[0]
    rp = orc.to_ref(params)
    rp.table_cells = params.table_cells or n
    k = np.empty(n, np.uint32)
    orc.load().sphref_hash(C.byref(rp), n, pos4.ctypes.data, k.ctypes.data)
    return k


def test_force_pair_closed_form(orc):
    """Two particles d apart on x, at rest: equal densities, pressure force along the pair
    axis with (d-R)*15/(pi R^5) slope, equal and opposite; viscosity from v - v* adds
    m (R-d)^2 / rho * (0, 0, g dt) to both."""
    R, k, rho0, dt, g = 0.35, 500.0, 55.0, 1 / 120, 9.8
    p = _params(orc, table_cells=1 << 16)
    dist = 0.2
    pos = np.array([[5.0, 3.0, 3.0, 0], [5.0 + dist, 3.0, 3.0, 0]], np.float32)
    vel = np.zeros((2, 4), np.float32)
    ps, vs, pvs, d, dn, out = _force(orc, p, pos, vel)
    dd = abs(float(ps[1, 0]) - float(ps[0, 0]))
    V = 2 * math.pi * R ** 5 / 15
    rho = (R ** 2 + (R - dd) ** 2) / V
    assert d[0] == pytest.approx(rho, rel=1e-5)
    shared = (rho - rho0) * k
    slope = (dd - R) * 15 / (3.14159265359 * R ** 5)
    a = shared * slope / rho          # magnitude along dir (neighbour - self)
    visc = (R - dd) ** 2 / rho * (g * float(np.float32(dt)))
    left = int(np.argmin(ps[:, 0]))
    right = 1 - left
    dv_left = (out[left, :3].astype(np.float64) - vs[left, :3]) / float(np.float32(dt))
    dv_right = (out[right, :3].astype(np.float64) - vs[right, :3]) / float(np.float32(dt))
    assert dv_left[0] == pytest.approx(a / rho, rel=1e-4)
    assert dv_right[0] == pytest.approx(-a / rho, rel=1e-4)
    assert dv_left[2] == pytest.approx(visc / rho, rel=1e-4)
    assert dv_left[1] == 0.0


def test_force_zero_pressure_at_target_density(orc):
    """rho = rho0 for every particle removes the pressure force; what remains is the
    viscosity pull toward neighbours' (non-predicted) velocities."""
    p = _params(orc, table_cells=1 << 16)
    pos = np.array([[5.0, 3.0, 3.0, 0], [5.2, 3.0, 3.0, 0]], np.float32)
    vel = np.array([[1.0, 0, 0, 0], [-1.0, 0, 0, 0]], np.float32)
    _, _, _, d, _, _ = _force(orc, p, pos, vel)
    p.target_density = float(d[0])
    ps, vs, pvs, d, dn, out = _force(orc, p, pos, vel)
    R = 0.35
    dd = abs(float(ps[1, 0]) - float(ps[0, 0]))
    for j in range(2):
        i = 1 - j
        b = (R - dd) ** 2 / float(d[i])
        expect = vs[j, :3] + b * (vs[i, :3].astype(np.float64) - pvs[j, :3]) / float(d[j]) * float(p.dt)
        np.testing.assert_allclose(out[j, :3], expect, rtol=1e-5, atol=1e-7)


def _near_kernel(R, d):
    """Near-density kernel (R - d)^3 * 15 / (pi R^6), in float64."""
    return (R - d) ** 3 * 15 / (math.pi * R ** 6) if d < R else 0.0


def test_force_near_pressure_pair_closed_form(orc):
    """Near pressure alone (k = 0, k_near = 50), two particles d apart on x at rest: each
    gets k_near * (rho_n_i + rho_n_j) / 2 * (-45 (R - d)^2 / (pi R^6)) * m / rho_n_j / rho_i
    * dt along the pair axis (neighbour minus self), i.e. a repulsion; the y component stays
    0 and z only carries the viscosity of v - v*."""
    R, kn = 0.35, 50.0
    p = _params(orc, table_cells=1 << 16, pressure_multiplier=0.0, near_pressure_multiplier=kn)
    pos = np.array([[5.0, 3.0, 3.0, 0], [5.2, 3.0, 3.0, 0]], np.float32)
    vel = np.zeros((2, 4), np.float32)
    ps, vs, pvs, d, dn, out = _force(orc, p, pos, vel)
    dd = abs(float(ps[1, 0]) - float(ps[0, 0]))
    V = 2 * math.pi * R ** 5 / 15
    rho = (R ** 2 + (R - dd) ** 2) / V
    rho_n = _near_kernel(R, 0.0) + _near_kernel(R, dd)
    assert dn[0] == pytest.approx(rho_n, rel=1e-5) and dn[1] == pytest.approx(rho_n, rel=1e-5)
    c = kn * (rho_n + rho_n) / 2 * (-45 * (R - dd) ** 2 / (math.pi * R ** 6)) / rho_n
    dt = float(p.dt)
    left = int(np.argmin(ps[:, 0]))
    right = 1 - left
    dv_left = out[left, :3].astype(np.float64) - vs[left, :3]
    dv_right = out[right, :3].astype(np.float64) - vs[right, :3]
    assert c < 0
    assert dv_left[0] == pytest.approx(c / rho * dt, rel=1e-4)   # pushed toward -x
    assert dv_right[0] == pytest.approx(-c / rho * dt, rel=1e-4)
    assert dv_left[1] == 0.0 and dv_right[1] == 0.0


def test_force_near_pressure_divides_by_neighbour_near_density(orc):
    """Three particles on x: A at 5.0, B at 5.3, C at 5.36 (C is within R of B only), so the
    near densities differ.  A's velocity change comes from B alone and is divided by B's
    near density (and by A's density): a term divided by the wrong one fails here."""
    R, kn = 0.35, 50.0
    p = _params(orc, table_cells=1 << 16, pressure_multiplier=0.0, near_pressure_multiplier=kn)
    pos = np.array([[5.0, 3.0, 3.0, 0], [5.3, 3.0, 3.0, 0], [5.36, 3.0, 3.0, 0]], np.float32)
    vel = np.zeros((3, 4), np.float32)
    ps, vs, pvs, d, dn, out = _force(orc, p, pos, vel)
    order = np.argsort(ps[:, 0])
    a, b, cc = (int(i) for i in order)
    xa, xb, xc = (float(ps[i, 0]) for i in (a, b, cc))
    V = 2 * math.pi * R ** 5 / 15
    rho_a = (R ** 2 + (R - (xb - xa)) ** 2) / V
    rn = lambda *ds: sum(_near_kernel(R, x) for x in (0.0,) + ds)
    rn_a, rn_b = rn(xb - xa), rn(xb - xa, xc - xb)
    assert dn[a] == pytest.approx(rn_a, rel=1e-5) and dn[b] == pytest.approx(rn_b, rel=1e-5)
    assert xc - xa > R and rn_b > 1.3 * rn_a
    dab = xb - xa
    c = kn * (rn_a + rn_b) / 2 * (-45 * (R - dab) ** 2 / (math.pi * R ** 6)) / rn_b
    dv_a = float(out[a, 0]) - float(vs[a, 0])
    assert dv_a == pytest.approx(c / rho_a * float(p.dt), rel=1e-4)


def test_force_slopes_are_kernel_gradients(orc):
    """The pressure slope is the derivative of the density kernel and the near-pressure
    slope the derivative of the near-density kernel.  Both derivatives are taken here by
    central differences of the oracle's own pair densities at d -/+ h, and the pair force
    with both multipliers on (k = 500, k_near = 50) is rebuilt from them: a wrong constant,
    sign or power in either slope fails."""
    R, k, kn, rho0 = 0.35, 500.0, 50.0, 55.0
    h = 1e-3
    dist = 0.2

    def pair_dens(x):
        pq = _params(orc, table_cells=1 << 16)
        pos = np.array([[5.0, 3.0, 3.0, 0], [5.0 + x, 3.0, 3.0, 0]], np.float32)
        ps, _, _, _, d, dn = _density(orc, pq, pos)
        return abs(float(ps[1, 0]) - float(ps[0, 0])), float(d[0]), float(dn[0])

    x_lo, d_lo, n_lo = pair_dens(dist - h)
    x_hi, d_hi, n_hi = pair_dens(dist + h)
    grad_rho = (d_hi - d_lo) / (x_hi - x_lo)      # m dW/dd
    grad_near = (n_hi - n_lo) / (x_hi - x_lo)     # m dW_near/dd
    p = _params(orc, table_cells=1 << 16, pressure_multiplier=k, near_pressure_multiplier=kn)
    pos = np.array([[5.0, 3.0, 3.0, 0], [5.0 + dist, 3.0, 3.0, 0]], np.float32)
    vel = np.zeros((2, 4), np.float32)
    ps, vs, pvs, d, dn, out = _force(orc, p, pos, vel)
    rho, rho_n = float(d[0]), float(dn[0])
    # with m = 1 the per-pair force is shared pressure * slope / rho_neighbour (twice)
    force = (rho - rho0) * k * grad_rho / rho + kn * rho_n * grad_near / rho_n
    left = int(np.argmin(ps[:, 0]))
    dv = float(out[left, 0]) - float(vs[left, 0])
    assert dv == pytest.approx(force / rho * float(p.dt), rel=2e-3)


def test_fp64_oracle_agrees_with_fp32_near_pressure(orc):
    """The FP32 and FP64 restatements agree at k_near = 50 (the default scene is run at 0)."""
    from sph_hip.workloads import DAM_BREAK_16K as W
    params = W.params()
    params.substeps = 1
    params.near_pressure_multiplier = 50.0
    pos, vel = W.state(4096)
    p32, v32, inter = orc.substep(params, pos, vel)
    o64 = orc.substep64(params, pos, vel, inter)
    dv32 = inter["vel_force"][:, :3].astype(np.float64) - vel[inter["sorted_index"], :3]
    dv64 = o64["vel_force"] - vel[inter["sorted_index"], :3]
    err = np.linalg.norm(dv32 - dv64, axis=1) / np.maximum(o64["dv_scale"], 1e-30)
    assert err.max() < 1e-5
    # and the term matters: the same start at k_near = 0 moves differently
    params.near_pressure_multiplier = 0.0
    _, _, inter0 = orc.substep(params, pos, vel)
    assert np.abs(inter0["vel_force"] - inter["vel_force"]).max() > 1e-3


def test_predict_and_integrate_closed_forms(orc):
    import ctypes as C
    p = _params(orc)
    dt = float(p.dt)
    g = 9.8
    pos = np.array([[5.0, 3.0, 3.0, 0],     # free flight
                    [0.31, 3.0, 3.0, 0],    # hits x-min (margin r = 0.3)
                    [5.0, 3.0, 0.305, 0],   # hits the floor: v.z reflected and * 0.95
                    [9.69, 3.0, 3.0, 0]], np.float32)
    vel = np.array([[1.0, -2.0, 0.5, 0], [-3.0, 0, 0, 0], [0, 0, -2.0, 0], [3.0, 0, 0, 0]], np.float32)
    pp = np.empty_like(pos)
    pv = np.empty_like(vel)
    rp = orc.to_ref(p)
    orc.load().sphref_predict(C.byref(rp), 4, pos.ctypes.data, vel.ctypes.data, pp.ctypes.data, pv.ctypes.data)
    v0 = vel[0, :3].astype(np.float64) + (0, 0, -g * dt)
    np.testing.assert_allclose(pv[0, :3], v0, rtol=1e-6)
    np.testing.assert_allclose(pp[0, :3], pos[0, :3] + v0 * dt, rtol=1e-6)
    assert pp[1, 0] == np.float32(0.3) and pv[1, 0] == 3.0
    assert pp[2, 2] == np.float32(0.3)
    assert pv[2, 2] == pytest.approx((2.0 + g * dt) * 0.95, rel=1e-6)
    assert pp[3, 0] == np.float32(10.0) - np.float32(0.3) and pv[3, 0] == -3.0
    assert np.all(pp[:, 3] == 1.0) and np.all(pv[:, 3] == 1.0)
    # integrate: margin r + 0.5 = 0.8
    ip = pos.copy()
    iv = vel.copy()
    orc.load().sphref_integrate(C.byref(rp), 4, ip.ctypes.data, iv.ctypes.data)
    assert ip[1, 0] == np.float32(0.0) + np.float32(np.float32(0.3) + np.float32(0.5))
    assert ip[3, 0] == np.float32(10.0) - np.float32(np.float32(0.3) + np.float32(0.5))
    np.testing.assert_allclose(ip[0, :3], pos[0, :3] + v0 * dt, rtol=1e-6)


def test_fp64_oracle_agrees_with_fp32(orc):
    """The two restatements are separate code; on the dam-break start they agree to FP32
    precision (a dropped or mis-signed term in either would not)."""
    from sph_hip.workloads import DAM_BREAK_16K as W
    params = W.params()
    params.substeps = 1
    pos, vel = W.state(4096)
    p32, v32, inter = orc.substep(params, pos, vel)
    o64 = orc.substep64(params, pos, vel, inter)
    rel = np.abs(inter["density"] - o64["density"]) / o64["density"]
    assert rel.max() < 1e-5
    dv32 = inter["vel_force"][:, :3].astype(np.float64) - vel[inter["sorted_index"], :3]
    dv64 = o64["vel_force"] - vel[inter["sorted_index"], :3]
    scale = np.abs(dv64).max()
    assert np.abs(dv32 - dv64).max() / scale < 1e-4
    np.testing.assert_allclose(p32[:, :3], o64["pos"], rtol=0, atol=1e-5)


def test_color_ramp_endpoints(orc):
    v = np.zeros((3, 4), np.float32)
    v[1, 0] = 60.0
    v[2, 0] = 0.33 * 60.0 - 1e-3
    c = orc.color(v)
    np.testing.assert_allclose(c[0, :3], (0.3, 0.3, 1.0), atol=1e-6)
    np.testing.assert_allclose(c[1, :3], (1.0, 0.3, 0.3), atol=1e-6)
    np.testing.assert_allclose(c[2, :3], (0.0, 1.0, 1.0), atol=1e-3)
