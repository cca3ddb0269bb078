"""MI355X (gfx950) path-tracing core behind swpathtracer's AccelStructure /
Integrator interfaces. The compute path is libspt.so (HIP kernels + C ABI,
include/spt.h); this package is its host-side Python face."""
from . import _lib
from .api import (Bvh, Device, DeviceArray, Integrator, SptError, cook_torrance, ggx_conductor, lambertian,
                  material, orthographic, pinhole, smooth_conductor, smooth_dielectric)

__all__ = ["Bvh", "Device", "DeviceArray", "Integrator", "SptError", "cook_torrance", "ggx_conductor",
           "lambertian", "material", "orthographic", "pinhole", "smooth_conductor", "smooth_dielectric", "_lib"]
