"""g4hepem_amd -- MI355X (gfx950) HIP implementation of G4HepEm's run-time
stepping physics behind a C-ABI (include/g4hepem_amd.h).

* ``Tables``   offline builder of G4HepEmData/G4HepEmParameters (reference layout)
* ``Context``  device image of the tables + lookup and stepping entry points
* ``BatchPlan`` many uneven track batches stepped by one graph-capturable launch
* ``tracks``   structure-of-arrays track batches and synthetic workloads
"""
from ._native import (ELECTRON, GAMMA, HEPEM_FP32T, HEPEM_FP64, POSITRON, HepEmError, Tables,  # noqa: F401
                      lib)
from .context import BatchPlan, Context, TrackStack  # noqa: F401

__all__ = ["Tables", "Context", "BatchPlan", "TrackStack", "HepEmError", "lib", "HEPEM_FP64", "HEPEM_FP32T", "ELECTRON", "POSITRON", "GAMMA"]
