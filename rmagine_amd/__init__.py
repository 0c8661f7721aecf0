"""rmagine_amd: MI355X-native ray-casting simulation (the Rmagine simulate() path).

The compute lives in librmh.so (rmagine_amd/csrc, C ABI in include/rmh.h).  This package
is the Python mirror of Rmagine's Simulator API (rmagine_amd.simulation), its sensor
model and transform types (rmagine_amd.types) and synthetic workloads
(rmagine_amd.scenes).  Importing the package does not touch the GPU.
"""
from . import types  # noqa: F401

__all__ = ["types"]
