"""The synthetic workloads the benchmarks and parity tests run (seeded, deterministic).

  dam_break_16k : 16,384 particles seeded uniform-random in a 4x6x4 block at the corner of a
                  10x6x6 box, zero velocity; R 0.35, target density 55, pressure mult 500,
                  viscosity 0.06 (accepted, unused like the reference), gravity 9.8 along -z,
                  frame dt 1/120 with 3 substeps, particle radius 0.3, mass 1, T = N.
  uniform_1m    : 1,048,576 particles uniform over a 40x20x20 box, velocities N(0, 0.1);
                  same kernel parameters; T = 2^21 (throughput) or 2^20 (= N, the
                  reference's table size).
  uniform_32m   : 33,554,432 particles uniform over a 160x40x80 box, velocities N(0, 0.1);
                  same kernel parameters; T = N.  The resident scene of the mixed serving
                  workload.
"""
from __future__ import annotations

from dataclasses import dataclass

from .abi import SphParams, fill_block, params_for


@dataclass(frozen=True)
class Workload:
    name: str
    n: int
    box_max: tuple
    block_lo: tuple
    block_hi: tuple
    vel_sigma: float
    seed: int
    frame_dt: float = 1.0 / 120.0
    substeps: int = 3
    smoothing_radius: float = 0.35
    target_density: float = 55.0
    pressure_multiplier: float = 500.0
    viscosity_strength: float = 0.06
    g: float = 9.8
    particle_radius: float = 0.3
    particle_mass: float = 1.0
    near_pressure_multiplier: float = 0.0

    def params(self, table_cells: int = 0) -> SphParams:
        return params_for(
            particle_radius=self.particle_radius, smoothing_radius=self.smoothing_radius,
            g=self.g, box_min=(0.0, 0.0, 0.0), box_max=self.box_max, frame_dt=self.frame_dt,
            substeps=self.substeps, particle_mass=self.particle_mass,
            pressure_multiplier=self.pressure_multiplier, target_density=self.target_density,
            viscosity_strength=self.viscosity_strength,
            near_pressure_multiplier=self.near_pressure_multiplier, table_cells=table_cells)

    def state(self, n: int | None = None):
        return fill_block(n or self.n, self.seed, self.block_lo, self.block_hi, self.vel_sigma)


DAM_BREAK_16K = Workload("dam_break_16k", 16384, (10.0, 6.0, 6.0), (0.0, 0.0, 0.0),
                         (4.0, 6.0, 4.0), 0.0, 1)
UNIFORM_1M = Workload("uniform_1m", 1 << 20, (40.0, 20.0, 20.0), (0.0, 0.0, 0.0),
                      (40.0, 20.0, 20.0), 0.1, 2)
UNIFORM_32M = Workload("uniform_32m", 1 << 25, (160.0, 40.0, 80.0), (0.0, 0.0, 0.0),
                       (160.0, 40.0, 80.0), 0.1, 5)


@dataclass(frozen=True)
class SceneSpec:
    n: int
    box_max: tuple
    target_density: float
    seed: int

    def params(self, base: Workload = DAM_BREAK_16K) -> SphParams:
        return params_for(
            particle_radius=base.particle_radius, smoothing_radius=base.smoothing_radius,
            g=base.g, box_min=(0.0, 0.0, 0.0), box_max=self.box_max, frame_dt=base.frame_dt,
            substeps=base.substeps, particle_mass=base.particle_mass,
            pressure_multiplier=base.pressure_multiplier, target_density=self.target_density,
            viscosity_strength=base.viscosity_strength,
            near_pressure_multiplier=base.near_pressure_multiplier,
            table_cells=self.table_cells)

    @property
    def table_cells(self) -> int:
        """Next power of two >= n: the key mask T-1 must be all ones below its top bit
        (updateCellEntries.comp:55 requires a power-of-two count), otherwise a scene of
        e.g. 2077 particles would have only 32 distinct keys."""
        return 1 << max(0, (self.n - 1).bit_length())

    @property
    def block_fits(self) -> bool:
        """Whether the resting block at the target density fits under the box height (when
        it does not, state() caps it at the box height and the scene starts over-dense)."""
        bx, by, bz = self.box_max
        return self.n / (self.target_density * bx * by) <= bz

    def state(self):
        """A resting block on the floor of the box at the scene's target density: number
        density n / volume = rho0 (the density estimate of a uniform fill equals its number
        density, since the kernel integrates to one), height n / (rho0 * Lx * Ly) capped at
        the box height; zero velocity.  (A sparse fill far below rho0 is unstable with
        these kernel parameters: pressure pulls isolated pairs together fast enough to blow
        the velocities up.)"""
        bx, by, bz = self.box_max
        h = min(bz, self.n / (self.target_density * bx * by))
        return fill_block(self.n, self.seed, (0.0, 0.0, 0.0), (bx, by, float(h)), 0.0)


def mixed_scenes(count: int = 256, seed: int = 3, n_min: int = 2048, n_max: int = 65536,
                 extent=(4.0, 16.0), rho0=(40.0, 70.0)) -> list:
    """The mixed-size batch: scene s draws from numpy's PCG64 seeded with seed + s a
    particle count log-uniform in [n_min, n_max], box extents uniform in `extent` per axis
    and a target density uniform in `rho0`; its particles start as a resting block at the
    target density on the floor of its box (SceneSpec.state).
    Other parameters are those of dam_break_16k (R 0.35, k 500, g 9.8, 1/120 s x 3)."""
    import numpy as np
    out = []
    for s in range(count):
        rng = np.random.Generator(np.random.PCG64(seed + s))
        n = int(round(float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))))
        box = tuple(float(np.float32(x)) for x in rng.uniform(extent[0], extent[1], 3))
        r0 = float(np.float32(rng.uniform(rho0[0], rho0[1])))
        out.append(SceneSpec(n, box, r0, seed + s))
    return out
