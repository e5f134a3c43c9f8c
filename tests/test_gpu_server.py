"""Serving loop on the GPU: requests submitted while others run are admitted at frame
boundaries, and every result is bit-identical to the same scene stepped alone (and to
the FP32 oracle); an invalid request is refused without disturbing the others."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bitexact(name, a, b):
    bad = np.flatnonzero(a.view(np.uint32).ravel() != b.view(np.uint32).ravel())
    assert bad.size == 0, f"{name}: {bad.size} mismatches"


def test_continuous_batching_matches_alone(gpu, orc):
    from sph_hip.workloads import mixed_scenes
    specs = mixed_scenes(count=6, seed=40, n_min=400, n_max=4000, extent=(2.0, 9.0))
    steps = [3, 1, 4, 2, 5, 1]
    with gpu.Server(hbm_budget_bytes=1 << 30, max_scenes=4, max_particles=1 << 16,
                    max_table_cells=1 << 17, substeps=3) as srv:
        ids = []
        for i, sp in enumerate(specs):
            p, v = sp.state()
            ids.append(srv.submit(sp.params(), p, v, steps[i]))
            if i == 2:
                time.sleep(0.05)  # later requests arrive while the first ones run
        bad_pos, bad_vel = specs[0].state()
        bad_pos[7, 2] = np.nan
        with pytest.raises(gpu.SphError):
            srv.submit(specs[0].params(), bad_pos, bad_vel, 2)
        results = []
        for rid in ids:
            srv.wait(rid, 120.0)
            results.append(srv.result(rid))
        st = srv.stats()
    assert st["completed"] == len(specs) and st["failed"] == 0
    assert st["frames"] >= max(steps)
    for i, sp in enumerate(specs):
        p, v = sp.state()
        rp, rv = orc.step(sp.params(), p, v, steps[i])
        _bitexact(f"request {i} pos", results[i][0], rp)
        _bitexact(f"request {i} vel", results[i][1], rv)


def test_budget_admission_queues_requests(gpu):
    """A budget that holds one scene at a time still completes every request."""
    from sph_hip.workloads import mixed_scenes
    specs = mixed_scenes(count=3, seed=50, n_min=1000, n_max=1000, extent=(3.0, 4.0))
    # 648 B per particle (136 B of state and slots, the default 128 four-byte list entries)
    one = 1000 * 648 + 1024 * 24 + (16 * 1000 + 4096) * 8  # < two scenes at any grid size
    with gpu.Server(hbm_budget_bytes=one, max_scenes=8, max_particles=1 << 14,
                    max_table_cells=1 << 14, substeps=3) as srv:
        ids = [srv.submit(sp.params(), *sp.state(), 2) for sp in specs]
        for rid in ids:
            srv.wait(rid, 60.0)
        st = srv.stats()
        assert st["completed"] == 3
        for rid in ids:
            srv.result(rid)


def test_serving_stream_full_size_requests(gpu, orc):
    """The bench's serving stream at its request sizes (4,096-262,144 particles, dam-break
    and uniform-fill starts, own box and table per request), submitted together with 2
    frames each: every result is bit-exact against the oracle run on that request alone."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    reqs = bench.request_stream(8, seed=4, rate_hz=10.0)
    total = sum(r[2].shape[0] for r in reqs)
    print(f"serving stream: {len(reqs)} requests, {total} particles", flush=True)
    assert max(r[2].shape[0] for r in reqs) > 100000
    with gpu.Server(hbm_budget_bytes=8 << 30, max_scenes=16, max_particles=total,
                    max_table_cells=sum(r[1].table_cells for r in reqs), substeps=3) as srv:
        ids = [srv.submit(params, pos, vel, 2) for _, params, pos, vel, _ in reqs]
        results = []
        for rid in ids:
            srv.wait(rid, 120.0)
            results.append(srv.result(rid))
        assert srv.stats()["failed"] == 0
    print("oracle", flush=True)
    for i, (_, params, pos, vel, _) in enumerate(reqs):
        rp, rv = orc.step(params, pos, vel, 2)
        _bitexact(f"request {i} pos", results[i][0], rp)
        _bitexact(f"request {i} vel", results[i][1], rv)


def test_served_scenes_with_offset_boxes(gpu, orc):
    """Requests whose boxes start away from the origin (asymmetric, non-integer offsets),
    served together: each result bit-exact against the oracle run on that request alone."""
    from sph_hip.workloads import mixed_scenes
    from test_gpu_batch import OFFSETS, shifted
    specs = mixed_scenes(count=len(OFFSETS), seed=60, n_min=500, n_max=5000, extent=(2.0, 9.0))
    reqs = []
    for sp, off in zip(specs, OFFSETS):
        pos, vel = sp.state()
        p2, q = shifted(gpu, sp.params(), pos, off)
        reqs.append((p2, q, vel))
    with gpu.Server(hbm_budget_bytes=1 << 30, max_scenes=8, max_particles=1 << 16,
                    max_table_cells=1 << 17, substeps=3) as srv:
        ids = [srv.submit(p, q, v, 2 + i) for i, (p, q, v) in enumerate(reqs)]
        results = []
        for rid in ids:
            srv.wait(rid, 120.0)
            results.append(srv.result(rid))
    for i, (p, q, v) in enumerate(reqs):
        rp, rv = orc.step(p, q, v, 2 + i)
        _bitexact(f"request {i} pos", results[i][0], rp)
        _bitexact(f"request {i} vel", results[i][1], rv)


def test_served_resident_32m_beside_the_stream(gpu, orc):
    """Mixed serving at scale on the served path: the 2^25-particle uniform fill of a
    160 x 40 x 80 box (T = N) resident for two frames, submitted first, beside six requests
    of the serving stream that retire after one, two and three frames, under a 200 GB
    budget.  Membership changes twice while the big scene is resident (it keeps slot 0 and
    is not moved) and once after it retires; every result is bit-exact against the oracle
    run on that request alone."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    from sph_hip.workloads import UNIFORM_32M as W
    big_params = W.params(table_cells=0)
    big_pos, big_vel = W.state()
    reqs = bench.request_stream(6, seed=4, rate_hz=10.0)
    steps = [1, 2, 3, 1, 2, 3]
    total = W.n + sum(r[2].shape[0] for r in reqs)
    cells = W.n + sum(r[1].table_cells for r in reqs)
    print(f"served 2^25 + {len(reqs)} requests, {total} particles", flush=True)
    with gpu.Server(hbm_budget_bytes=200 << 30, max_scenes=16, max_particles=total,
                    max_table_cells=cells, substeps=3) as srv:
        big = srv.submit(big_params, big_pos, big_vel, 2)
        ids = [srv.submit(params, pos, vel, steps[i]) for i, (_, params, pos, vel, _) in enumerate(reqs)]
        results = []
        for rid in ids:
            srv.wait(rid, 600.0)
            results.append(srv.result(rid))
        srv.wait(big, 600.0)
        got_big = srv.result(big)
        st = srv.stats()
    assert st["failed"] == 0 and st["completed"] == 7 and st["frames"] >= 3
    print("oracle", flush=True)
    for i, (_, params, pos, vel, _) in enumerate(reqs):
        rp, rv = orc.step(params, pos, vel, steps[i])
        _bitexact(f"request {i} pos", results[i][0], rp)
        _bitexact(f"request {i} vel", results[i][1], rv)
    rp, rv = orc.step(big_params, big_pos, big_vel, 2)
    _bitexact("resident pos", got_big[0], rp)
    _bitexact("resident vel", got_big[1], rv)
