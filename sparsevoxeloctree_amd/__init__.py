"""MI355X-native sparse-voxel-octree hot path: voxelize -> level-by-level octree build -> ESVO
ray marching, as hand-written HIP kernels for gfx950 behind a C ABI (include/svo.h).

The compute lives in lib/libsvo.so; this package is the ctypes binding plus a Python mirror of the
reference's Voxelizer / OctreeBuilder / Octree / OctreeTracer class surface.  Importing the
package fails if the library has not been built: there is no CPU fallback.
"""
from .api import (  # noqa: F401
    HIT_DTYPE,
    BuildPlan,
    Camera,
    Context,
    DeviceArray,
    EnvironmentMap,
    MarchPlan,
    Octree,
    OctreeBuilder,
    OctreeTracer,
    PathTracer,
    Scene,
    Server,
    SvoError,
    Voxelizer,
    autotune,
    bounce_rays,
    build_batch,
    camera_look_at,
    gen_scene,
    load_obj,
    march_batch,
    march_rays,
    march_rays_batch,
    rgba8_snorm,
    rgba8_unorm,
    voxelize_batch,
)
from ._capi import LIB_PATH, exported_symbols  # noqa: F401
