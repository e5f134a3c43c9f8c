"""Benchmark of the SPH step on one MI355X.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload uniform_1m]

A step is one frame = 3 substeps of the full chain (hash, radix sort, reorder + cell
table, density, force + integrate) over the 1,048,576-particle uniform scene
(40x20x20 box, T = 2^21 cells), inputs resident in HBM.  Prints one JSON line with
particle-steps/s, the dominant kernel's HBM roofline fraction (HIP events on the
library's stream) and a CPU baseline (the FP32 C restatement on the host cores, bounded
sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "particle-fluid-simulation_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "particle-steps/s batched vs plain (1M pts, 1 GPU); p50/p99 scene latency"
HBM_PEAK_GBS = 8000.0
FP32_VECTOR_PEAK_TFLOPS = 157.3


def flops_per_particle_substep(workload) -> dict:
    """Floating-point work per particle per substep of the uniform fill, in closed form:
    density ~10 flops per candidate and 8 per in-radius neighbour, force ~9 per candidate
    and 50 per neighbour, with C = n 27 R^3 candidates in the 27 cells and K = n 4/3 pi R^3
    neighbours at the initial fill's number density n = N / volume of the filled block."""
    import math
    vol = 1.0
    for lo, hi in zip(workload.block_lo, workload.block_hi):
        vol *= hi - lo
    dens = workload.n / vol
    R = workload.smoothing_radius
    C = dens * 27.0 * R ** 3
    K = dens * 4.0 / 3.0 * math.pi * R ** 3
    return {"candidates": C, "neighbours": K, "flops": 10 * C + 8 * K + 9 * C + 50 * K}


def density_map_timing(world, voxel_size: float, reps: int = 5) -> dict:
    """updateDensityMap of the world's last substep into device memory (allocated through
    the HIP runtime the library runs on), best of `reps` enqueue + sync; each voxel walks
    the 27 cells around it."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    dims = world.density_map_dims(voxel_size)
    voxels = dims[0] * dims[1] * dims[2]
    dev = C.c_void_p()
    if hip.hipMalloc(C.byref(dev), C.c_size_t(voxels * 4)) != 0:
        raise RuntimeError("hipMalloc of the density map failed")
    try:
        world.density_map_device(voxel_size, dev.value, voxels)
        world.sync()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            world.density_map_device(voxel_size, dev.value, voxels)
            world.sync()
            best = min(best, time.perf_counter() - t0)
    finally:
        hip.hipFree(dev)
    return {"voxel_size": voxel_size, "dims": list(dims), "voxels": voxels, "ms": best * 1e3,
            "voxels_per_s": voxels / best}


def pcie_inclusive(world, frames: int = 20) -> dict:
    """The drop-in loop with the state crossing PCIe, which the boundary's host buffers
    imply for a host without device-memory interop: one frame stepped and the positions
    and velocities downloaded every frame (state uploaded once), and the same with the
    state also uploaded every frame.  Host buffers are ordinary pageable arrays, as a C
    host's would be; each download waits for the step."""
    import numpy as np
    n = world.n
    pos = np.empty((n, 4), np.float32)
    vel = np.empty((n, 4), np.float32)
    down = lambda: world._lib.sph_world_download_state(world._h, pos.ctypes.data, vel.ctypes.data, None)
    if down() != 0:
        raise RuntimeError("download failed")
    t0 = time.perf_counter()
    for _ in range(frames):
        world.step(1)
        if down() != 0:
            raise RuntimeError("download failed")
    t_sd = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(frames):
        world.upload(pos, vel)
        world.step(1)
        if down() != 0:
            raise RuntimeError("download failed")
    t_usd = time.perf_counter() - t0
    # the same step + download into pinned host memory (hipHostMalloc of the runtime the
    # library runs on), which a host can allocate for its readback buffers
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    ptrs = [C.c_void_p(), C.c_void_p()]
    for p in ptrs:
        if hip.hipHostMalloc(C.byref(p), C.c_size_t(n * 16), 0) != 0:
            raise RuntimeError("hipHostMalloc failed")
    lib, h = world._lib, world._h
    try:
        t0 = time.perf_counter()
        for _ in range(frames):
            world.step(1)
            if lib.sph_world_download_state(h, ptrs[0].value, ptrs[1].value, None) != 0:
                raise RuntimeError("download failed")
        t_pin = time.perf_counter() - t0
        # overlapped: frame f's snapshot is copied to the host while frame f + 1 steps
        t0 = time.perf_counter()
        for _ in range(frames):
            world.step(1)
            if lib.sph_world_snapshot_wait(h) != 0 or \
                    lib.sph_world_snapshot_async(h, ptrs[0].value, ptrs[1].value) != 0:
                raise RuntimeError("snapshot failed")
        if lib.sph_world_snapshot_wait(h) != 0:
            raise RuntimeError("snapshot failed")
        t_ovl = time.perf_counter() - t0
    finally:
        lib.sph_world_snapshot_wait(h)  # no copy may still target the buffers freed here
        for p in ptrs:
            hip.hipHostFree(p)
    return {"frames": frames, "bytes_per_frame_each_way": n * 32,
            "step_download_particle_steps_per_s": n * frames / t_sd,
            "step_download_ms_per_step": t_sd * 1e3 / frames,
            "upload_step_download_particle_steps_per_s": n * frames / t_usd,
            "upload_step_download_ms_per_step": t_usd * 1e3 / frames,
            "host_buffers": "pageable",
            "step_download_pinned_particle_steps_per_s": n * frames / t_pin,
            "step_download_pinned_ms_per_step": t_pin * 1e3 / frames,
            "step_snapshot_overlapped_pinned_particle_steps_per_s": n * frames / t_ovl,
            "step_snapshot_overlapped_pinned_ms_per_step": t_ovl * 1e3 / frames}


def copy_bandwidth_gbs(mib: int = 512, reps: int = 5) -> float:
    """Device-to-device copy rate (bytes read + written per second), best of `reps`: the
    measured counterpart of the HBM peak."""
    import torch
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3)
    del src, dst
    return 2.0 * (mib << 20) / best / 1e9


def alg_bytes_per_particle_substep(n: int, table_cells: int) -> float:
    """Algorithmic HBM bytes per particle per substep of the whole chain:
    200 + 16*P + 8*T/N with P = ceil(ceil(log2 T)/8) radix passes."""
    kb = max(1, (table_cells - 1).bit_length())
    passes = (kb + 7) // 8
    return 200.0 + 16.0 * passes + 8.0 * table_cells / n


# Algorithmic bytes per particle of the two neighbour kernels, from the whole-chain
# formula's per-pass terms: density reads the sorted position and velocity (24 B) and
# writes (density, near density) (8 B); force + integrate reads the same 24 B and the
# densities (8 B) and writes the new position and velocity (32 B).  Neighbour reads are
# reuse of data another particle already counts, not algorithmic traffic.
KERNEL_BYTES_PER_PARTICLE = {"density": 24 + 8, "force": 24 + 8 + 32}
KERNEL_NAMES = {"density": "k_density (calculateDensities + force neighbour lists)",
                "force": "k_force (applyPressureAndViscosityForce + updateVelocities)"}

# HBM bytes per launch of the neighbour kernels on uniform_1m (T = 2^21), keyed by the
# (warmup, timed steps) of the run they were measured after: the fill settles and grows
# denser as it runs, so each figure holds for that point of the trajectory only.  Measured
# with rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, on the 30 launches the bench times its
# dominant kernel over (those right after the timed steps), each counter scaled by what it
# reads for device-to-device copies of known size in the same run (FETCH x 2, WRITE x 1).
# A run at any other (warmup, steps) reports no traffic.
TRAFFIC_PROFILE = "profiles/pmc_uniform_1m_traffic.md"
TRAFFIC_KERNELS = ("per-lane, 4 candidates per scan step, 128-lane workgroups, force neighbour lists, "
                   "8-bit radix digits (3 passes)")
TRAFFIC_BYTES_PER_LAUNCH = {
    (5, 20): {"density": (2 * 109309 + 119029) * 1024, "force": (2 * 173590 + 42403) * 1024},
    (5, 100): {"density": (2 * 116664 + 196226) * 1024, "force": (2 * 182953 + 42441) * 1024},
}


def cpu_baseline(workload, table_cells: int, budget_s: float = 12.0) -> dict:
    """Times the FP32 C restatement (OpenMP, all host cores) on whole substeps of the same
    workload until ~budget_s of wall time; reports particle-steps/s (step = frame)."""
    from oracle import oracle as orc
    params = workload.params(table_cells=table_cells)
    pos, vel = workload.state()
    one = type(params).from_buffer_copy(params)
    one.substeps = 1
    substeps = 0
    t0 = time.perf_counter()
    while True:
        pos, vel = orc.step(one, pos, vel, 1)
        substeps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    frames = substeps / workload.substeps
    return {
        "value": workload.n * frames / dt, "unit": "particle-steps/s", "cores": threads,
        "kind": "port",
        "sample": f"{substeps} substeps ({frames:.2f} steps) of {workload.name} "
                  f"(N={workload.n}, T={table_cells}) in {dt:.1f} s; reference-semantics CPU "
                  f"restatement (FP32 C of the reference shaders, OpenMP over particles), not "
                  f"lavapipe, which this image lacks",
    }


def mixed_batch(steps: int, warmup: int, count: int = 256, fitting_only: bool = False) -> dict:
    """Batched vs plain on the mixed workload: `count` scenes (log-uniform 2,048-65,536
    particles, boxes U[4,16]^3, target density U[40,70], seeds 3+s) stepped together in one
    batch, then the same scenes as one world each, stepped one after another (each world's
    steps are enqueued on its own stream, so the runtime may overlap them).  Both are tuned
    after `warmup` + 1 untimed frames.  With `fitting_only`, only the scenes whose resting
    block fits the box at its target density (the others start over-dense)."""
    import torch
    import sph_hip
    from sph_hip.workloads import mixed_scenes

    specs = mixed_scenes(count=count, seed=3)
    if fitting_only:
        specs = [sp for sp in specs if sp.block_fits]
    count = len(specs)
    params = [sp.params() for sp in specs]
    counts = [sp.n for sp in specs]
    states = [sp.state() for sp in specs]
    total = sum(counts)

    b = sph_hip.Batch(params, counts)
    for i, (p, v) in enumerate(states):
        b.upload(i, p, v)
    b.step(warmup)
    b.set_kernel("auto")  # re-tune on the warmed-up state, like the single worlds below
    b.step(1)
    b.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    bricks = sum(b.uses_bricks(i) for i in range(count))
    batch_kernels = b.kernel
    b.close()

    worlds = []
    for i, (p, v) in enumerate(states):
        w = sph_hip.World(params[i], counts[i])
        w.upload(p, v)
        w.step(warmup)
        w.autotune(3)  # tuned on the warmed-up state (restored afterwards)
        w.step(1)
        worlds.append(w)
    for w in worlds:
        w.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for w in worlds:
        w.step(steps)
    for w in worlds:
        w.sync()
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    for w in worlds:
        w.close()
    batched = total * steps / t_batch
    plain = total * steps / t_plain
    return {"scenes": count, "subset": "blocks that fit at the target density" if fitting_only else "all",
            "particles": total, "steps": steps,
            "batched_particle_steps_per_s": batched, "plain_particle_steps_per_s": plain,
            "batched_over_plain": batched / plain, "batched_ms_per_step": t_batch * 1e3 / steps,
            "plain_ms_per_step": t_plain * 1e3 / steps, "scenes_on_brick_kernels": bricks,
            "batched_kernels": batch_kernels}


def request_stream(count: int, seed: int, rate_hz: float, n_range=(4096, 262144),
                   steps_range=(50, 500)):
    """Serving workload: `count` requests with Poisson arrivals at `rate_hz`; each draws
    (numpy PCG64, seed + i) a particle count log-uniform in n_range, a step count uniform in
    steps_range and a dam-break (particles in the lower-x half of the box) or uniform-fill
    start.  Boxes are 2a x a x a sized so the start is at the target density 55: uniform
    fill of the whole box (a = (n / 110)^(1/3)) or a block filling its lower-x half
    (a = (n / 55)^(1/3)); kernel parameters are those of dam_break_16k, T = next
    pow2 >= n."""
    import numpy as np
    import sph_hip
    from sph_hip.workloads import DAM_BREAK_16K as base
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate_hz))
        r = np.random.Generator(np.random.PCG64(seed + 1 + i))
        n = int(round(float(np.exp(r.uniform(np.log(n_range[0]), np.log(n_range[1]))))))
        steps = int(r.integers(steps_range[0], steps_range[1] + 1))
        dam = bool(r.integers(0, 2))
        a = float(np.float32((n / (55.0 if dam else 110.0)) ** (1.0 / 3.0)))
        box = (2 * a, a, a)
        params = sph_hip.params_for(
            particle_radius=base.particle_radius, smoothing_radius=base.smoothing_radius,
            g=base.g, box_min=(0, 0, 0), box_max=box, frame_dt=base.frame_dt,
            substeps=base.substeps, pressure_multiplier=base.pressure_multiplier,
            target_density=base.target_density, viscosity_strength=base.viscosity_strength,
            table_cells=1 << max(0, (n - 1).bit_length()))
        hi = (a if dam else 2 * a, a, a)
        pos, vel = sph_hip.fill_block(n, seed + 1 + i, (0, 0, 0), hi, 0.0)
        out.append((t, params, pos, vel, steps))
    return out


def serving(count: int, rate_hz: float, seed: int = 4, resident: int = 0,
            budget_gb: float = 8.0) -> dict:
    """Submits the request stream at its arrival times to a Server and waits for all; with
    `resident` > 0 a scene of that many particles (uniform fill of a 160x40x80 box) is
    submitted first with enough steps to stay resident for the whole stream."""
    import sph_hip
    from sph_hip.workloads import DAM_BREAK_16K as base
    reqs = request_stream(count, seed, rate_hz)
    max_n = sum(r[2].shape[0] for r in reqs) + resident
    max_t = sum(r[1].table_cells for r in reqs) + (1 << max(0, (resident - 1).bit_length()) if resident else 0)
    srv = sph_hip.Server(hbm_budget_bytes=int(budget_gb * (1 << 30)), max_scenes=min(65535, count + 1),
                         max_particles=max_n, max_table_cells=max_t, substeps=base.substeps)
    big = None
    if resident:
        box = (160.0, 40.0, 80.0)
        params = sph_hip.params_for(
            particle_radius=base.particle_radius, smoothing_radius=base.smoothing_radius,
            g=base.g, box_min=(0, 0, 0), box_max=box, frame_dt=base.frame_dt, substeps=base.substeps,
            pressure_multiplier=base.pressure_multiplier, target_density=base.target_density,
            table_cells=1 << max(0, (resident - 1).bit_length()))
        big_state = sph_hip.fill_block(resident, seed, (0, 0, 0), box, 0.1)
        big_params = params
        big = srv.submit(big_params, *big_state, 1 << 30)
    t0 = time.perf_counter()
    ids = []
    for t, params, pos, vel, steps in reqs:
        delay = t - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(params, pos, vel, steps))
    for rid in ids:
        srv.wait(rid, 600.0)
        srv.result(rid)
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    if big is not None:
        # the resident scene stepped alone as one world for as many frames as the served run
        # made (the fill settles and its neighbour passes slow down as it runs), timing the
        # last 10: the frame time the served frames are compared with
        w = sph_hip.World(big_params, resident)
        w.upload(*big_state)
        t1 = time.perf_counter()
        w.step(max(0, int(st["frames"]) - 10))
        w.sync()
        # mean over the same trajectory as the served frames (its online tuning included)
        st["resident_alone_mean_ms_per_frame"] = ((time.perf_counter() - t1) * 1e3 /
                                                  max(1, int(st["frames"]) - 10))
        st["served_mean_ms_per_frame"] = wall * 1e3 / max(1, int(st["frames"]))
        w.autotune(3)  # on the settled state, like the served chain's fixed configuration
        w.step(1)
        w.sync()
        t1 = time.perf_counter()
        w.step(10)
        w.sync()
        st["resident_alone_ms_per_frame"] = (time.perf_counter() - t1) * 1e3 / 10
        st["resident_alone_frames"] = int(st["frames"])
        w.close()
    st.update({"requests": count, "rate_hz": rate_hz, "wall_s": wall,
               "request_particles": sum(r[2].shape[0] for r in reqs),
               "request_particle_steps": float(sum(r[2].shape[0] * r[4] for r in reqs)),
               "resident_scene_particles": resident})
    return st


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)  # the 1M configuration runs 100 steps
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="uniform_1m", choices=["uniform_1m", "dam_break_16k"])
    ap.add_argument("--table-cells", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--lane-config", type=int, default=None,
                    help="pin one per-lane launch configuration (0-9) instead of autotuning; "
                         "for hardware-counter runs, where counters distort the tuner's timings")
    ap.add_argument("--tile-config", type=int, default=None,
                    help="pin one tiled configuration (0-1: cell tiles staged in LDS) instead of "
                         "autotuning; for hardware-counter runs of the tiled kernels")
    ap.add_argument("--no-tiled-leg", dest="tiled_leg", action="store_false",
                    help="skip timing the tiled kernels (tile configuration 0) beside the tuned step")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--batch-steps", type=int, default=100,
                    help="steps of the mixed 256-scene batched-vs-plain comparison (0 = skip)")
    ap.add_argument("--serve-requests", type=int, default=60,
                    help="requests of the Poisson serving run (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=10.0, help="request arrivals per second")
    ap.add_argument("--serve-resident", type=int, default=0,
                    help="particles of a scene kept resident beside the request stream "
                         "(33554432 = the mixed-serving-at-scale case)")
    ap.add_argument("--serve-budget-gb", type=float, default=8.0)
    ap.add_argument("--no-tolerance-leg", dest="tolerance_leg", action="store_false",
                    help="skip timing the tolerance path (approximate square root / reciprocal)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (scenes shard "
              f"across GPUs with no communication; a single scene would need a halo exchange).",
              file=sys.stderr)
        return 2

    import torch
    import sph_hip
    from sph_hip import workloads as wl

    W = wl.UNIFORM_1M if args.workload == "uniform_1m" else wl.DAM_BREAK_16K
    T = args.table_cells if args.table_cells is not None else (
        1 << 21 if W is wl.UNIFORM_1M else W.n)
    params = W.params(table_cells=T)
    pos, vel = W.state()
    world = sph_hip.World(params, W.n)
    world.upload(pos, vel)
    print(f"[bench] {W.name}: N={W.n} T={T} substeps/step={W.substeps}", file=sys.stderr, flush=True)
    # W untimed warmup steps; the autotuner runs on the state after the first W - 1 (it
    # restores that state afterwards), and the last warmup step builds the tuned graph.
    world.step(max(0, args.warmup - 1))
    world.sync()
    if args.tile_config is not None:
        world.set_tile_config(args.tile_config)
        chosen = world.kernel_config
        print(f"[bench] pinned: {chosen}", file=sys.stderr, flush=True)
    elif args.lane_config is None:
        chosen, tuned_ms = world.autotune(5)
        chosen = world.kernel_config
        print(f"[bench] autotune: {chosen}, {tuned_ms:.3f} ms/substep", file=sys.stderr, flush=True)
    else:
        world.set_lane_config(args.lane_config)
        chosen = world.kernel_config
        print(f"[bench] pinned: {chosen}", file=sys.stderr, flush=True)
    if args.warmup > 0:
        world.step(1)
    world.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    world.step(args.steps)
    world.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = W.n * args.steps / elapsed
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)

    # Dominant kernel, timed with HIP events recorded on the library's stream around every
    # launch of 30 more substeps continuing the timed run.
    pt = world.pass_times(30)
    tile_stats = world.tile_stats()
    variants = {}
    for v in ("brick", "tile", "direct"):  # direct last: keeps the tuned per-lane configuration
        world.set_kernel(v)
        variants[v] = world.pass_times(10)
    dom = "density" if pt["density_ms"] > pt["force_ms"] else "force"
    dom_s = pt[f"{dom}_ms"] * 1e-3
    dom_bytes = W.n * KERNEL_BYTES_PER_PARTICLE[dom]
    dom_gbs = dom_bytes / dom_s / 1e9
    # measured HBM bytes hold for the kernel they were counted on (lane configuration 3,
    # which the tuner picks on this workload) at the same point of the same trajectory
    traffic = None
    if W is wl.UNIFORM_1M and T == 1 << 21 and chosen == TRAFFIC_KERNELS:
        traffic = TRAFFIC_BYTES_PER_LAUNCH.get((args.warmup, args.steps), {}).get(dom)
    # a 64 MiB device-to-device copy: the calibration the HBM byte counters are scaled by
    calib_src = torch.ones(16 << 20, dtype=torch.float32, device="cuda")
    torch.empty_like(calib_src).copy_(calib_src)
    torch.cuda.synchronize()
    b_alg = alg_bytes_per_particle_substep(W.n, T)
    substep_s = ms_per_step * 1e-3 / W.substeps
    chain_gbs = W.n * b_alg / substep_s / 1e9
    copy_gbs = copy_bandwidth_gbs()
    work = flops_per_particle_substep(W)
    tflops = W.n * work["flops"] / substep_s / 1e12

    # The same timed steps through the tolerance path (sph_world_set_precision: approximate
    # square root and reciprocal in the neighbour terms, 1e-5 relative), from the same
    # start, tuned after the same warmup; never the bench's value.
    tol = None
    if args.tolerance_leg:
        wt = sph_hip.World(params, W.n)
        wt.set_precision("tolerance")
        wt.upload(pos, vel)
        wt.step(max(0, args.warmup - 1))
        wt.sync()
        wt.autotune(5)
        if args.warmup > 0:
            wt.step(1)
        wt.sync()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        wt.step(args.steps)
        wt.sync()
        t_tol = time.perf_counter() - t1
        tol = {"value": W.n * args.steps / t_tol, "ms_per_step": t_tol * 1e3 / args.steps,
               "kernels": wt.kernel_config, "pass_ms": wt.pass_times(30),
               "over_exact": (W.n * args.steps / t_tol) / value}
        wt.close()
        print(f"[bench] tolerance path: {tol['ms_per_step']:.3f} ms/step ({tol['kernels']})", file=sys.stderr,
              flush=True)

    # The same timed steps through the tiled kernels (tile configuration 0: cell tiles
    # staged in LDS), from the same start after the same warmup; never the bench's value.
    tiled = None
    if args.tiled_leg and args.tile_config is None:
        wtl = sph_hip.World(params, W.n)
        wtl.set_tile_config(0)
        wtl.upload(pos, vel)
        wtl.step(args.warmup)
        wtl.sync()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        wtl.step(args.steps)
        wtl.sync()
        t_tl = time.perf_counter() - t1
        tiled = {"value": W.n * args.steps / t_tl, "ms_per_step": t_tl * 1e3 / args.steps,
                 "kernels": wtl.kernel_config, "pass_ms": wtl.pass_times(30), "stats": wtl.tile_stats(),
                 "over_tuned": (W.n * args.steps / t_tl) / value}
        wtl.close()
        print(f"[bench] tiled path: {tiled['ms_per_step']:.3f} ms/step ({tiled['kernels']})", file=sys.stderr,
              flush=True)

    # updateDensityMap of the last substep at a voxel of R/3.5 (16M voxels for uniform_1m,
    # the size of the reference's own 560 x 160 x 160 image)
    dmap = density_map_timing(world, 0.1)
    # the same step with the state crossing PCIe every frame (never the bench's value)
    pcie = pcie_inclusive(world)

    # Sanity: the state stays finite and in the box (the integrate clamp keeps every
    # position at least particle_radius + 0.5 inside each wall).
    p_out, v_out = world.download()
    import numpy as np
    ok = bool(np.isfinite(p_out).all() and np.isfinite(v_out).all())
    lo = np.array(params.min_xyz, np.float32)
    hi = np.array(params.max_xyz, np.float32)
    in_box = bool(ok and (p_out[:, :3] >= lo).all() and (p_out[:, :3] <= hi).all())
    # device memory the world holds at the end (its force lists sized to the state)
    memory = {"list_entries": world.list_capacity, "device_bytes": world.device_bytes,
              "bytes_per_particle": world.device_bytes / W.n}
    world.close()

    mixed = mixed_fit = None
    if args.batch_steps > 0:
        print("[bench] mixed 256-scene batch vs plain", file=sys.stderr, flush=True)
        mixed = mixed_batch(args.batch_steps, 3)
        print(f"[bench] batched/plain = {mixed['batched_over_plain']:.2f}", file=sys.stderr, flush=True)
        mixed_fit = mixed_batch(args.batch_steps, 3, fitting_only=True)
        print(f"[bench] scenes whose block fits: batched/plain = {mixed_fit['batched_over_plain']:.2f}",
              file=sys.stderr, flush=True)

    serve = None
    if args.serve_requests > 0:
        print(f"[bench] serving {args.serve_requests} requests at {args.serve_rate}/s", file=sys.stderr,
              flush=True)
        serve = serving(args.serve_requests, args.serve_rate, resident=args.serve_resident,
                        budget_gb=args.serve_budget_gb)
        print(f"[bench] serving p50/p99 scene {serve['scene_p50_ms']:.1f}/{serve['scene_p99_ms']:.1f} ms",
              file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(W, T, args.cpu_budget)

    line = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded PCG32 uniform fill, velocities N(0,0.1))",
        "config": {"workload": W.name, "particles": W.n, "table_cells": T,
                   "substeps_per_step": W.substeps, "box": list(W.box_max),
                   "smoothing_radius": W.smoothing_radius, "path": "plain (one scene per launch)"},
        "roofline": {"bound": "hbm", "achieved": dom_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom_gbs / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "traffic_profile": TRAFFIC_PROFILE if traffic is not None else None,
                     "kernel": KERNEL_NAMES[dom], "avg_launch_ms": pt[f"{dom}_ms"],
                     "alg_bytes_per_launch": dom_bytes},
        "chain_roofline": {"alg_bytes_per_particle_substep": b_alg, "achieved_gbs": chain_gbs,
                           "frac": chain_gbs / HBM_PEAK_GBS, "copy_gbs": copy_gbs,
                           "frac_of_copy": chain_gbs / copy_gbs},
        "chain_flops": {"flops_per_particle_substep": work["flops"],
                        "candidates": work["candidates"], "neighbours": work["neighbours"],
                        "achieved_tflops": tflops, "peak_tflops": FP32_VECTOR_PEAK_TFLOPS,
                        "frac": tflops / FP32_VECTOR_PEAK_TFLOPS},
        "kernels": chosen,
        "pass_ms": pt,
        "pass_ms_by_kernel": variants,
        "tolerance_path": tol,
        "tiled_path": tiled,
        "tile_stats": tile_stats,
        "density_map": dmap,
        "pcie_inclusive": pcie,
        "mixed_batch": mixed,
        "mixed_batch_fitting_scenes": mixed_fit,
        "serving": serve,
        "state_finite": ok,
        "state_in_box": in_box,
        "memory": memory,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
