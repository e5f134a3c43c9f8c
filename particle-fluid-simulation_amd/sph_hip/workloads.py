"""The synthetic workloads the benchmarks and parity tests run (seeded, deterministic).

  dam_break_16k : 16,384 particles seeded uniform-random in a 4x6x4 block at the corner of a
                  10x6x6 box, zero velocity; R 0.35, target density 55, pressure mult 500,
                  viscosity 0.06 (accepted, unused like the reference), gravity 9.8 along -z,
                  frame dt 1/120 with 3 substeps, particle radius 0.3, mass 1, T = N.
  uniform_1m    : 1,048,576 particles uniform over a 40x20x20 box, velocities N(0, 0.1);
                  same kernel parameters; T = 2^21 (throughput) or 2^20 (= N, the
                  reference's table size).
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
