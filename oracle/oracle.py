"""ctypes binding of the CPU oracle (oracle/build/libsphoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker / CPU baseline.  The product library never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "libsphoracle.so")


class RefParams(C.Structure):
    _fields_ = [
        ("particle_radius", C.c_float), ("smoothing_radius", C.c_float), ("g", C.c_float),
        ("min_xyz", C.c_float * 3), ("max_xyz", C.c_float * 3), ("dt", C.c_float),
        ("particle_mass", C.c_float), ("pressure_multiplier", C.c_float),
        ("target_density", C.c_float), ("viscosity_strength", C.c_float),
        ("volume", C.c_float), ("near_pressure_multiplier", C.c_float),
        ("table_cells", C.c_uint32), ("substeps", C.c_uint32),
    ]


class RefDerived(C.Structure):
    _fields_ = [("mask", C.c_uint32), ("R2", C.c_float), ("slope_scale", C.c_float),
                ("near_slope_scale", C.c_float), ("near_vol_inv", C.c_float),
                ("integrate_margin", C.c_float)]


class SubstepOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "cell_keys", "sorted_keys", "sorted_idx", "start", "end", "pred_pos", "pred_vel",
        "density", "near_density", "vel_force")]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    P, vp, u32 = C.POINTER, C.c_void_p, C.c_uint32
    sig = {
        "sphref_derive": (None, [P(RefParams), P(RefDerived)]),
        "sphref_cell_hash": (u32, [u32, u32, u32]),
        "sphref_predict": (None, [P(RefParams), u32, vp, vp, vp, vp]),
        "sphref_hash": (None, [P(RefParams), u32, vp, vp]),
        "sphref_sort": (C.c_int, [u32, u32, vp, vp, vp]),
        "sphref_table": (None, [u32, u32, vp, vp, vp]),
        "sphref_density": (None, [P(RefParams), u32, vp, vp, vp, vp, vp]),
        "sphref_force": (None, [P(RefParams), u32, vp, vp, vp, vp, vp, vp, vp, vp]),
        "sphref_integrate": (None, [P(RefParams), u32, vp, vp]),
        "sphref_color": (None, [u32, vp, vp]),
        "sphref_substep": (C.c_int, [P(RefParams), u32, vp, vp, P(SubstepOut)]),
        "sphref_step": (C.c_int, [P(RefParams), u32, vp, vp, u32]),
        "sphref_density_bruteforce": (None, [P(RefParams), u32, vp, vp, vp, vp]),
        "sphref_density_map_dims": (None, [P(RefParams), C.c_float, vp]),
        "sphref_density_map": (None, [P(RefParams), u32, vp, vp, vp, C.c_float, vp]),
        "sphref64_substep": (C.c_int, [P(RefParams), u32, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                       vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def to_ref(params) -> RefParams:
    """Copies any ctypes struct / object with the sph_params field names."""
    r = RefParams()
    for name, _ in RefParams._fields_:
        v = getattr(params, name)
        if name.endswith("_xyz"):
            for a in range(3):
                getattr(r, name)[a] = v[a]
        else:
            setattr(r, name, v)
    return r


def _ptr(a):
    return a.ctypes.data if a is not None else None


def cell_hash(x: int, y: int, z: int) -> int:
    return load().sphref_cell_hash(x & 0xFFFFFFFF, y & 0xFFFFFFFF, z & 0xFFFFFFFF)


def derive(params) -> RefDerived:
    d = RefDerived()
    load().sphref_derive(C.byref(to_ref(params)), C.byref(d))
    return d


def substep(params, pos: np.ndarray, vel: np.ndarray):
    """One FP32 substep; returns (pos, vel, intermediates dict).  Inputs are not modified."""
    lib = load()
    rp = to_ref(params)
    n = pos.shape[0]
    T = rp.table_cells or n
    rp.table_cells = T
    pos = np.ascontiguousarray(pos, np.float32).copy()
    vel = np.ascontiguousarray(vel, np.float32).copy()
    out = {
        "cell_keys": np.empty(n, np.uint32), "sorted_keys": np.empty(n, np.uint32),
        "sorted_index": np.empty(n, np.uint32), "start": np.empty(T, np.int32),
        "end": np.empty(T, np.int32), "pred_pos": np.empty((n, 4), np.float32),
        "pred_vel": np.empty((n, 4), np.float32), "density": np.empty(n, np.float32),
        "near_density": np.empty(n, np.float32), "vel_force": np.empty((n, 4), np.float32),
    }
    so = SubstepOut(_ptr(out["cell_keys"]), _ptr(out["sorted_keys"]), _ptr(out["sorted_index"]),
                    _ptr(out["start"]), _ptr(out["end"]), _ptr(out["pred_pos"]),
                    _ptr(out["pred_vel"]), _ptr(out["density"]), _ptr(out["near_density"]),
                    _ptr(out["vel_force"]))
    rc = lib.sphref_substep(C.byref(rp), n, pos.ctypes.data, vel.ctypes.data, C.byref(so))
    if rc != 0:
        raise RuntimeError("sphref_substep failed")
    return pos, vel, out


def step(params, pos: np.ndarray, vel: np.ndarray, frames: int):
    lib = load()
    rp = to_ref(params)
    n = pos.shape[0]
    rp.table_cells = rp.table_cells or n
    pos = np.ascontiguousarray(pos, np.float32).copy()
    vel = np.ascontiguousarray(vel, np.float32).copy()
    if lib.sphref_step(C.byref(rp), n, pos.ctypes.data, vel.ctypes.data, frames) != 0:
        raise RuntimeError("sphref_step failed")
    return pos, vel


def substep64(params, pos, vel, inter):
    """FP64 oracle of one substep using the FP32 run's integer decisions."""
    lib = load()
    rp = to_ref(params)
    n = pos.shape[0]
    rp.table_cells = rp.table_cells or n
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    o = {k: np.empty(n, np.float64) for k in ("density", "near_density", "dv_scale")}
    for k in ("vel_force", "pos", "vel"):
        o[k] = np.empty((n, 3), np.float64)
    rc = lib.sphref64_substep(C.byref(rp), n, pos.ctypes.data, vel.ctypes.data,
                              inter["sorted_index"].ctypes.data, inter["start"].ctypes.data,
                              inter["end"].ctypes.data, inter["pred_pos"].ctypes.data,
                              o["density"].ctypes.data, o["near_density"].ctypes.data,
                              o["vel_force"].ctypes.data, o["pos"].ctypes.data,
                              o["vel"].ctypes.data, o["dv_scale"].ctypes.data)
    if rc != 0:
        raise RuntimeError("sphref64_substep failed")
    return o


def density_bruteforce(params, pred_pos, sorted_keys):
    lib = load()
    rp = to_ref(params)
    n = pred_pos.shape[0]
    rp.table_cells = rp.table_cells or n
    d = np.empty(n, np.float32)
    dn = np.empty(n, np.float32)
    lib.sphref_density_bruteforce(C.byref(rp), n, np.ascontiguousarray(pred_pos).ctypes.data,
                                  np.ascontiguousarray(sorted_keys).ctypes.data,
                                  d.ctypes.data, dn.ctypes.data)
    return d, dn


def density_map_dims(params, voxel_size: float):
    dims = np.empty(3, np.uint32)
    rp = to_ref(params)
    load().sphref_density_map_dims(C.byref(rp), voxel_size, dims.ctypes.data)
    return tuple(int(d) for d in dims)


def density_map(params, pos, start, end, voxel_size: float) -> np.ndarray:
    """updateDensityMap over positions in sorted order and the start/end table of the
    substep that produced them; returns a float32 array of shape (dz, dy, dx)."""
    lib = load()
    rp = to_ref(params)
    n = pos.shape[0]
    rp.table_cells = rp.table_cells or n
    dx, dy, dz = density_map_dims(params, voxel_size)
    out = np.empty((dz, dy, dx), np.float32)
    pos = np.ascontiguousarray(pos, np.float32)
    lib.sphref_density_map(C.byref(rp), n, pos.ctypes.data,
                           np.ascontiguousarray(start, np.int32).ctypes.data,
                           np.ascontiguousarray(end, np.int32).ctypes.data, voxel_size,
                           out.ctypes.data)
    return out


def color(vel):
    n = vel.shape[0]
    out = np.empty((n, 4), np.float32)
    load().sphref_color(n, np.ascontiguousarray(vel, np.float32).ctypes.data, out.ctypes.data)
    return out
