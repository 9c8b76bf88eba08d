"""ismc_amd — MI355X-native integrating-sphere Monte Carlo photon transport.

A drop-in for the pdcook/ISMC propagators (ISMC.py, ISMLMC.py) and for the
lookup-table and minimization inverse drivers, running hand-written HIP
kernels for gfx950 through the C ABI of libismc.so (include/ismc.h).
"""
from .config import (Layer, SampleConfig, Sphere, cap_port_fraction, dual_sphere_slab,
                     mus_from_reduced)
from .engine import EVENTS, TALLIES, Engine, Result, default_engine
from .propagator import ISMC, ISMLMC

__all__ = ["Layer", "SampleConfig", "Sphere", "cap_port_fraction", "dual_sphere_slab",
           "mus_from_reduced", "Engine", "Result", "default_engine", "TALLIES", "EVENTS",
           "ISMC", "ISMLMC"]
