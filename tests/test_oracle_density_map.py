"""Pins of the oracle's density map (updateDensityMap.comp:114-148) against facts outside
its code: the image size the reference host creates, the closed form of one particle, a
float64 all-pairs sum over the whole particle set, and the shader's double counting when
cells share a table key."""
import ctypes as C
import math

import numpy as np
import pytest


def _params(**kw):
    import sph_hip
    base = dict(particle_radius=0.3, smoothing_radius=0.35, g=9.8, box_min=(0, 0, 0),
                box_max=(3.0, 2.0, 2.5), frame_dt=1 / 120, substeps=1, pressure_multiplier=500.0,
                target_density=55.0, table_cells=1 << 20)
    base.update(kw)
    return sph_hip.params_for(**base)


def _sorted_table(orc, p, pos):
    """Keys, stable sort and start/end table of positions (the integer stages, already
    pinned against numpy in test_oracle_pins.py); returns sorted positions and the table."""
    lib = orc.load()
    rp = orc.to_ref(p)
    n = pos.shape[0]
    T = p.table_cells
    keys = np.empty(n, np.uint32)
    lib.sphref_hash(C.byref(rp), n, pos.ctypes.data, keys.ctypes.data)
    sk = np.empty(n, np.uint32)
    si = np.empty(n, np.uint32)
    assert lib.sphref_sort(n, T, keys.ctypes.data, sk.ctypes.data, si.ctypes.data) == 0
    st = np.empty(T, np.int32)
    en = np.empty(T, np.int32)
    lib.sphref_table(n, T, sk.ctypes.data, st.ctypes.data, en.ctypes.data)
    return np.ascontiguousarray(pos[si]), st, en


def _centres(p, voxel, dims):
    dx, dy, dz = dims
    ax = [np.float32(p.min_xyz[a]) + (np.arange(d, dtype=np.float32) + np.float32(0.5)) * np.float32(voxel)
          for a, d in enumerate(dims)]
    z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return np.stack([x, y, z], axis=-1).astype(np.float64)


def test_dims_are_the_reference_image_size(orc):
    """initializeDensityMap (renderer.c:774-783) for the reference's own box (main.c:356-358)
    and DENSITY_CELL_SIZE 0.5 (main.c:14): a 560 x 160 x 160 image."""
    import sph_hip
    p = _params(box_min=(-140.0, -40.0, -40.0), box_max=(140.0, 40.0, 40.0), smoothing_radius=2.5)
    assert orc.density_map_dims(p, 0.5) == (560, 160, 160)
    assert sph_hip.density_map_dims(p, 0.5) == (560, 160, 160)
    for box, v in (((3.0, 2.0, 2.5), 0.1), ((10.0, 6.0, 6.0), 0.3), ((1.7, 0.9, 2.3), 0.45)):
        q = _params(box_max=box)
        assert sph_hip.density_map_dims(q, v) == orc.density_map_dims(q, v)
        assert orc.density_map_dims(q, v) == tuple(math.ceil(np.float32(b) * np.float32(1 / np.float32(v)))
                                                   for b in box)


def test_dims_reject_bad_voxel():
    import sph_hip
    p = _params()
    for v in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(sph_hip.SphError):
            sph_hip.density_map_dims(p, v)


def test_single_particle_closed_form(orc):
    """One particle: every voxel centre within R gets m (R - d)^2 / volume (float64 closed
    form, float32 tolerance), every other voxel exactly 0."""
    p = _params(particle_mass=1.5)
    pos = np.array([[1.23, 0.77, 1.01, 1.0]], np.float32)
    spos, st, en = _sorted_table(orc, p, pos)
    voxel = 0.1
    got = orc.density_map(p, spos, st, en, voxel)
    c = _centres(p, voxel, orc.density_map_dims(p, voxel))
    d = np.sqrt(((c - pos[0, :3].astype(np.float64)) ** 2).sum(-1))
    R = float(np.float32(p.smoothing_radius))
    want = np.where(d < R, 1.5 * (R - np.minimum(d, R)) ** 2 / float(p.volume), 0.0)
    assert np.count_nonzero(want) > 50
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-7)
    assert np.all(got[want == 0.0] == 0.0)


def test_matches_all_pairs_when_keys_are_distinct(orc):
    """With every cell the map visits (non-negative coordinates up to one past the grid:
    the shader skips negative cells) on its own table key, the 27-cell walk sees each
    particle within R exactly once, so the map equals a float64 sum over all particle
    pairs.  (Cells with negative coordinates do collide structurally: the hash's
    x*p1 ^ y*p2 is unchanged by (x, y) -> (-x, -y) for odd p1, p2, so (-1, -1, z) and
    (1, 1, z) share a key at every table size.)"""
    p = _params()
    R = np.float32(p.smoothing_radius)
    g = [int(np.floor(np.float32(p.max_xyz[a]) / R)) + 1 for a in range(3)]
    keys = {orc.cell_hash(x, y, z) & (p.table_cells - 1)
            for x in range(0, g[0] + 1) for y in range(0, g[1] + 1) for z in range(0, g[2] + 1)}
    assert len(keys) == (g[0] + 1) * (g[1] + 1) * (g[2] + 1)
    assert orc.cell_hash(-1, -1, 3) == orc.cell_hash(1, 1, 3)
    rng = np.random.default_rng(7)
    n = 3000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform((0, 0, 0), (3.0, 2.0, 2.5), (n, 3)).astype(np.float32)
    spos, st, en = _sorted_table(orc, p, pos)
    voxel = 0.13
    got = orc.density_map(p, spos, st, en, voxel)
    c = _centres(p, voxel, orc.density_map_dims(p, voxel)).reshape(-1, 3)
    # a centre outside the box maps to cell (-1,-1,-1) on every axis
    # (samplePointToCellCoord), so only cell (0,0,0) of its 27 passes the bounds check
    outside = np.any(c > np.array([3.0, 2.0, 2.5]), axis=-1)
    assert 0 < outside.sum() < c.shape[0]
    in_origin_cell = np.all(np.floor(pos[:, :3] / R) == 0, axis=-1)
    want = np.zeros(c.shape[0])
    for i in range(n):
        d = np.sqrt(((c - pos[i, :3].astype(np.float64)) ** 2).sum(-1))
        term = np.where(d < float(R), (float(R) - np.minimum(d, float(R))) ** 2, 0.0)
        want += term if in_origin_cell[i] else np.where(outside, 0.0, term)
    want = (want / float(p.volume)).reshape(got.shape)
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-5 * want.max())


def test_shared_key_counts_a_run_once_per_cell(orc):
    """T = 1: every cell maps to key 0, whose run is every particle; a voxel whose 27 cells
    all pass the bounds check therefore counts each particle 27 times (the shader walks the
    same run for every cell)."""
    p = _params(table_cells=1)
    rng = np.random.default_rng(11)
    n = 400
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform((0, 0, 0), (3.0, 2.0, 2.5), (n, 3)).astype(np.float32)
    spos, st, en = _sorted_table(orc, p, pos)
    voxel = 0.25
    got = orc.density_map(p, spos, st, en, voxel)
    p_big = _params()
    spos_b, st_b, en_b = _sorted_table(orc, p_big, pos)
    once = orc.density_map(p_big, spos_b, st_b, en_b, voxel)
    # voxels whose cell has all 27 neighbours at non-negative coordinates
    dims = orc.density_map_dims(p, voxel)
    c = _centres(p, voxel, dims)
    cell = np.floor(c / float(np.float32(p.smoothing_radius)))
    interior = np.all(cell >= 1, axis=-1)
    assert interior.sum() > 100
    np.testing.assert_allclose(got[interior], 27.0 * once[interior], rtol=1e-5, atol=1e-7)
