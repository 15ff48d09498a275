"""dmipy_sim_amd -- the Monte-Carlo diffusion walk of dmipy-sim on AMD MI355X.

Drop-in for the hot path of dmipy_sim (``simulate``, ``simulate_cpmg``,
``simulate_mixture``, the ``Waveform`` builders and the analytic geometries,
dmipy_sim/__init__.py:12-56): the same Python calls, with the walk executed by
hand-written HIP kernels for gfx950 through the C ABI in include/dmipy_walk.h.
"""
from .constants import GAMMA, DEFAULT_SLEW_RATE
from .waveforms import (Waveform, pgse, pgste, ogse, trapezoidal_ogse, b_trapezoidal_ogse, cpmg,
                        calc_b, calc_btensor, btensor_invariants, ste, pte, set_b,
                        rotate_waveform, tile_waveform, fibonacci_directions)
from .geometries import (FreeDiffusion, Box1D, Sphere, Cylinder, PackedCylinders, Ellipsoid,
                         PermeableSlab1D, PermeableShell, MyelinatedCylinder,
                         PackedMyelinatedCylinders, pack_myelinated_cylinders,
                         PackedSpheres, pack_cylinders, pack_spheres, pack_dense)
from .physics import (make_walk_plan, permeable_sub_steps, surface_sub_steps,
                      collision_sub_steps, replay_sub_steps)
from .core import simulate, simulate_cpmg, simulate_mixture, walk, PlannedWalk
from .gpu import check_gpu, gpu_available, free_gpu_memory, gpu_session, list_gpu_processes
from .batch import BatchedWalk
from .trajectories import simulate_trajectories, replay, TrajectorySet, unwrap_periodic
from . import surrogate
from . import autotune
from .surrogate import MLP, SurrogateEnsemble

__all__ = [
    "GAMMA", "DEFAULT_SLEW_RATE", "Waveform", "pgse", "pgste", "ogse", "trapezoidal_ogse",
    "b_trapezoidal_ogse", "cpmg", "calc_b", "calc_btensor", "btensor_invariants", "ste", "pte", "set_b",
    "rotate_waveform", "tile_waveform", "fibonacci_directions", "FreeDiffusion", "Box1D",
    "Sphere", "Cylinder", "PackedCylinders", "PackedSpheres", "Ellipsoid", "PermeableSlab1D",
    "PermeableShell", "MyelinatedCylinder", "PackedMyelinatedCylinders",
    "pack_myelinated_cylinders", "pack_cylinders",
    "pack_spheres", "pack_dense", "make_walk_plan", "permeable_sub_steps",
    "surface_sub_steps", "collision_sub_steps", "replay_sub_steps", "simulate",
    "simulate_cpmg", "simulate_mixture", "walk", "check_gpu", "gpu_available",
    "free_gpu_memory", "gpu_session", "list_gpu_processes", "unwrap_periodic",
    "BatchedWalk", "PlannedWalk", "surrogate", "MLP", "SurrogateEnsemble", "autotune", "simulate_trajectories", "replay",
    "TrajectorySet",
]
