"""pfx_amd: an MI355X (gfx950) wavefront path tracer behind the still-render
surface of the reference tracer (`pfx_trace::Trace`, crates/trace/src/gpu.rs).

The product is libpfxhip.so (include/pfxhip.h); this package is its Python
host: ctypes bindings, the `Trace` mirror, scene inputs and the seeded
workload generators.
"""
from .scene import (EQUIRECT, ORTHOGRAPHIC, PERSPECTIVE, Adaptive, Bounces, Camera, Detail, Scene, Sun, material,
                    point_light, spot_light)
from .trace import Converged, Output, Probe, Trace, TraceError, animated_seed

__all__ = [
    "EQUIRECT",
    "ORTHOGRAPHIC",
    "PERSPECTIVE",
    "Adaptive",
    "Bounces",
    "Camera",
    "Converged",
    "Detail",
    "Output",
    "Probe",
    "Scene",
    "Sun",
    "Trace",
    "TraceError",
    "animated_seed",
    "material",
    "point_light",
    "spot_light",
]
