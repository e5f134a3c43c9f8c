"""The bench's closed-form figures and its one-GPU contract, checked without a device:
the whole-chain algorithmic bytes per particle-substep (200 + 16 P + 8 T/N) for the three
table sizes the measurements use, the candidate and neighbour counts of a uniform fill
(n 27 R^3 and n 4/3 pi R^3), the HBM traffic figures it reports against the counter tables
of the profile it names, and the refusal of more than one GPU."""
import math
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, ROOT)
    import bench as b
    return b


@pytest.mark.parametrize("n,T,want", [(1 << 14, 1 << 14, 240.0), (1 << 20, 1 << 20, 256.0),
                                      (1 << 20, 1 << 21, 264.0)])
def test_chain_bytes(bench, n, T, want):
    assert bench.alg_bytes_per_particle_substep(n, T) == want


def test_uniform_fill_work(bench):
    from sph_hip.workloads import UNIFORM_1M, DAM_BREAK_16K
    w = bench.flops_per_particle_substep(UNIFORM_1M)
    dens = (1 << 20) / (40.0 * 20.0 * 20.0)
    assert w["candidates"] == pytest.approx(dens * 27 * 0.35 ** 3)
    assert w["candidates"] == pytest.approx(75.87, abs=0.01)
    assert w["neighbours"] == pytest.approx(dens * 4 / 3 * math.pi * 0.35 ** 3)
    assert w["flops"] == pytest.approx(19 * w["candidates"] + 58 * w["neighbours"])
    # the dam break's particles fill a 4 x 6 x 4 block, not the 10 x 6 x 6 box
    d = bench.flops_per_particle_substep(DAM_BREAK_16K)
    assert d["candidates"] == pytest.approx(16384 / 96.0 * 27 * 0.35 ** 3)


def test_traffic_matches_its_profile(bench):
    """Every (warmup, steps) entry of the bench's per-launch HBM bytes is (2 FETCH_SIZE +
    WRITE_SIZE) KiB of the density and force rows in the profile's table for that run."""
    text = open(os.path.join(ROOT, bench.TRAFFIC_PROFILE)).read()
    rows = {"density": "k_density<OneScene, 4, 128, true, false>",
            "force": "k_force<OneScene, 4, 128, true, false>"}
    assert bench.TRAFFIC_BYTES_PER_LAUNCH
    for (warmup, steps), entry in bench.TRAFFIC_BYTES_PER_LAUNCH.items():
        head = "## Per launch, after %d warmup + %d timed steps" % (warmup, steps)
        assert head in text, head
        section = text.split(head, 1)[1].split("\n## ", 1)[0]
        for kind, kernel in rows.items():
            m = re.search(r"\| `%s` \| (\d+) \| (\d+) \|" % re.escape(kernel), section)
            assert m, (head, kernel)
            fetch, write = int(m.group(1)), int(m.group(2))
            assert entry[kind] == (2 * fetch + write) * 1024, (head, kind)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
