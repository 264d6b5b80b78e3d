"""Python mirror of the reference's hot-path class surface over the C ABI (include/svo.h).

Same class names and method roles as AdamYuan/SparseVoxelOctree's C++ classes, re-based on a
context (one HIP stream) instead of Vulkan command pools; "Cmd*" methods enqueue on the stream:

    Scene.Create                 src/Scene.cpp:387-417 (+ normalisation src/Scene.cpp:90-99)
    Voxelizer.Create / CmdVoxelize / GetLevel / GetVoxelResolution / GetVoxelFragmentCount /
        GetVoxelFragmentList     src/Voxelizer.hpp:43-52
    OctreeBuilder.Create / CmdBuild / GetOctreeRange / GetOctree / GetLevel
                                 src/OctreeBuilder.hpp:36-48
    Octree.Update / GetLevel / GetRange / Empty
                                 src/Octree.hpp:18-33
    Camera (m_position, m_yaw, m_pitch, m_fov, m_aspect_ratio)
                                 src/Camera.hpp:14-17
    OctreeTracer.Create / Resize / CmdBeamRenderPass / CmdDrawPipeline, m_beam_enable
                                 src/OctreeTracer.hpp:15-16,57-70

Batched forms (new; the reference handles one scene at a time): voxelize_batch, build_batch,
MarchPlan / march_batch and march_rays_batch run many scenes per launch.

Errors raise SvoError carrying the library's status code.
"""
import ctypes as C
import math
import os

import numpy as np

from . import _capi
from ._capi import HIT_DTYPE, SvoError, lib, ptr, svo_camera, svo_timing

__all__ = [
    "Context",
    "DeviceArray",
    "Scene",
    "Voxelizer",
    "OctreeBuilder",
    "Octree",
    "Camera",
    "OctreeTracer",
    "gen_scene",
    "camera_look_at",
    "voxelize_batch",
    "build_batch",
    "BuildPlan",
    "MarchPlan",
    "march_batch",
    "march_rays_batch",
    "autotune",
    "HIT_DTYPE",
    "SvoError",
]


class Context:
    """One HIP stream on one device (the reference's device + queue + command pool)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = lib.svo_ctx_create(device, C.byref(h))
        if rc != 0:
            raise SvoError(rc, "svo_ctx_create(%d) failed" % device)
        self.h = h

    def check(self, rc, what=""):
        if rc != 0:
            msg = lib.svo_last_error(self.h)
            raise SvoError(rc, (what + ": " if what else "") + (msg.decode() if msg else ""))

    def close(self):
        if self.h:
            lib.svo_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def synchronize(self):
        self.check(lib.svo_ctx_synchronize(self.h), "synchronize")

    def alloc(self, nbytes):
        return DeviceArray(self, nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        d = DeviceArray(self, arr.nbytes)
        self.check(lib.svo_upload(self.h, d.ptr, ptr(arr), arr.nbytes), "upload")
        return d

    def timing(self):
        t = svo_timing()
        self.check(lib.svo_last_timing(self.h, C.byref(t)), "timing")
        return {n: getattr(t, n) for n, _ in svo_timing._fields_}

    def device_errors(self):
        b = C.c_uint32()
        self.check(lib.svo_ctx_take_device_errors(self.h, C.byref(b)), "device errors")
        return b.value

    def trim(self):
        """Returns the allocator's cached (free) device memory to the device (svo_ctx_trim)."""
        self.check(lib.svo_ctx_trim(self.h), "trim")

    def memory(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.check(lib.svo_ctx_memory(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"in_use": a.value, "cached": b.value, "peak_in_use": c.value}

    def prof_enable(self, on=True):
        self.check(lib.svo_prof_enable(self.h, 1 if on else 0))

    def prof_reset(self):
        self.check(lib.svo_prof_reset(self.h))

    def prof_stats(self):
        n = C.c_int()
        self.check(lib.svo_prof_count(self.h, C.byref(n)))
        out = {}
        for i in range(n.value):
            name, cnt, ms = C.c_char_p(), C.c_uint64(), C.c_double()
            self.check(lib.svo_prof_get(self.h, i, C.byref(name), C.byref(cnt), C.byref(ms)))
            out[name.value.decode()] = (cnt.value, ms.value)
        return out


class DeviceArray:
    """Device memory from the context's allocator."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        ctx.check(lib.svo_malloc(ctx.h, max(self.nbytes, 1), C.byref(p)), "svo_malloc")
        self.ptr = p

    def download(self, dtype, count=None):
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            self.ctx.check(lib.svo_download(self.ctx.h, ptr(out), self.ptr, out.nbytes), "download")
        return out

    def free(self):
        if self.ptr is not None and self.ctx.h:
            lib.svo_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Scene:
    """Triangles resident in HBM.  xyz: (T, 9) float32, rgb: (T,) uint32 (R in the low byte)."""

    def __init__(self, ctx, handle, T):
        self.ctx, self.h, self.T = ctx, handle, T

    @staticmethod
    def Create(ctx, xyz, rgb, normalize=True):
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1)
        rgb = np.ascontiguousarray(rgb, dtype=np.uint32).reshape(-1)
        T = rgb.shape[0]
        assert xyz.shape[0] == 9 * T
        h = C.c_void_p()
        ctx.check(lib.svo_scene_create(ctx.h, ptr(xyz), ptr(rgb), T, 1 if normalize else 0, C.byref(h)), "scene")
        return Scene(ctx, h, T)

    def update(self, xyz, rgb, normalize=True):
        """svo_scene_update: new triangles, the same count (the input of a BuildPlan's next run)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1)
        rgb = np.ascontiguousarray(rgb, dtype=np.uint32).reshape(-1)
        assert rgb.shape[0] == self.T and xyz.shape[0] == 9 * self.T
        self.ctx.check(lib.svo_scene_update(self.ctx.h, self.h, ptr(xyz), ptr(rgb), 1 if normalize else 0),
                       "scene update")

    def GetTriangleCount(self):
        return self.T

    def download(self):
        out = np.empty(9 * self.T, dtype=np.float32)
        self.ctx.check(lib.svo_scene_download(self.ctx.h, self.h, ptr(out)), "scene download")
        return out.reshape(-1, 9)

    def __del__(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib.svo_scene_destroy(self.h)
            self.h = None


class Voxelizer:
    """Voxelizer::Create runs the count pass and allocates the list; CmdVoxelize fills it."""

    def __init__(self, ctx, handle, level, scene=None):
        self.ctx, self.h, self.level, self._scene = ctx, handle, level, scene

    @staticmethod
    def Create(scene, octree_level):
        ctx = scene.ctx
        h = C.c_void_p()
        ctx.check(lib.svo_voxelize_count(ctx.h, scene.h, octree_level, C.byref(h)), "voxelize count")
        return Voxelizer(ctx, h, octree_level, scene)

    @staticmethod
    def FromFragments(ctx, frags, octree_level):
        """A fragment list from the host (uint32 pairs), e.g. read back from the reference."""
        frags = np.ascontiguousarray(frags, dtype=np.uint32).reshape(-1)
        h = C.c_void_p()
        ctx.check(lib.svo_fraglist_create(ctx.h, ptr(frags), frags.shape[0] // 2, octree_level, C.byref(h)), "frags")
        return Voxelizer(ctx, h, octree_level)

    def CmdVoxelize(self):
        self.ctx.check(lib.svo_voxelize_fill(self.ctx.h, self.h), "voxelize fill")

    def GetLevel(self):
        return self.level

    def GetVoxelResolution(self):
        return 1 << self.level

    def GetVoxelFragmentCount(self):
        F = C.c_uint64()
        self.ctx.check(lib.svo_fraglist_info(self.h, None, C.byref(F), None))
        return F.value

    def GetVoxelFragmentList(self):
        """Downloads the fragment list: (F, 2) uint32."""
        F = self.GetVoxelFragmentCount()
        out = np.empty(2 * F, dtype=np.uint32)
        if F:
            self.ctx.check(lib.svo_fraglist_download(self.ctx.h, self.h, ptr(out), F), "fragment download")
        return out.reshape(-1, 2)

    def device_ptr(self):
        p = C.c_void_p()
        self.ctx.check(lib.svo_fraglist_info(self.h, None, None, C.byref(p)))
        return p

    def __del__(self):
        if getattr(self, "h", None) and self.ctx.h and not getattr(self, "_borrowed", False):
            lib.svo_fraglist_destroy(self.h)
            self.h = None


class Octree:
    """Octree::Update takes the builder's buffer and its used range."""

    def __init__(self, ctx=None, handle=None, level=0, W=0):
        self.ctx, self.h, self.level, self.W = ctx, handle, level, W
        self._owner = None

    @staticmethod
    def FromWords(ctx, words, level):
        words = np.ascontiguousarray(words, dtype=np.uint32)
        h = C.c_void_p()
        ctx.check(lib.svo_octree_create(ctx.h, ptr(words), words.shape[0], level, C.byref(h)), "octree create")
        o = Octree(ctx, h, level, words.shape[0])
        o._own = True
        return o

    @staticmethod
    def Load(ctx, path):
        """An octree file written by save() / svo_octree_save."""
        h = C.c_void_p()
        ctx.check(lib.svo_octree_load(ctx.h, str(path).encode(), C.byref(h)), "octree load")
        lvl, W = C.c_uint32(), C.c_uint64()
        lib.svo_octree_info(h, C.byref(lvl), C.byref(W), None)
        o = Octree(ctx, h, lvl.value, W.value)
        o._own = True
        return o

    def save(self, path):
        self.ctx.check(lib.svo_octree_save(self.ctx.h, self.h, str(path).encode()), "octree save")

    def CmdTransferOwnership(self, dst_ctx):
        """Orders work on dst_ctx's stream after this octree's build (svo_octree_transfer;
        Octree::CmdTransferOwnership, src/Octree.cpp:37-43)."""
        self.ctx.check(lib.svo_octree_transfer(self.h, dst_ctx.h), "octree transfer")

    def Update(self, builder):
        W = builder.GetOctreeRange() // 4
        self.ctx, self.h, self.level, self.W = builder.ctx, builder.h, builder.level, W
        self._owner = builder

    def Empty(self):
        return self.h is None

    def GetLevel(self):
        return self.level

    def GetRange(self):
        return self.W * 4

    def words(self):
        out = np.empty(self.W, dtype=np.uint32)
        self.ctx.check(lib.svo_octree_download(self.ctx.h, self.h, ptr(out), self.W), "octree download")
        return out

    def __del__(self):
        if getattr(self, "_own", False) and self.h and self.ctx.h:
            lib.svo_octree_destroy(self.h)
            self.h = None


class OctreeBuilder:
    """CmdBuild enqueues every level pass; GetOctreeRange waits and returns the used bytes."""

    def __init__(self, voxelizer):
        self.ctx, self.voxelizer, self.level = voxelizer.ctx, voxelizer, voxelizer.level
        self.h = None
        self.W = None

    @staticmethod
    def Create(voxelizer):
        return OctreeBuilder(voxelizer)

    def GetLevel(self):
        return self.level

    def CmdBuild(self):
        if self.h is not None:
            lib.svo_octree_destroy(self.h)
        h = C.c_void_p()
        self.ctx.check(lib.svo_build_enqueue(self.ctx.h, self.voxelizer.h, C.byref(h)), "build")
        self.h, self.W = h, None

    def GetOctreeRange(self):
        if self.W is None:
            W = C.c_uint64()
            self.ctx.check(lib.svo_build_finish(self.ctx.h, self.h, C.byref(W)), "build finish")
            self.W = W.value
        return self.W * 4

    def GetOctree(self):
        o = Octree()
        o.Update(self)
        return o

    def CmdTransferOctreeOwnership(self, dst_ctx):
        """Hands the octree being built to dst_ctx's stream: work enqueued there afterwards runs
        after the build, without a host wait (OctreeBuilder::CmdTransferOctreeOwnership,
        src/OctreeBuilder.cpp:215-220; svo_octree_transfer)."""
        self.ctx.check(lib.svo_octree_transfer(self.h, dst_ctx.h), "octree transfer")

    def __del__(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib.svo_octree_destroy(self.h)
            self.h = None


class Camera:
    """Camera::fetch_uniform_data (src/Camera.cpp:77-95)."""

    def __init__(self, position=(1.5, 1.5, 1.5), yaw=0.0, pitch=0.0, fov=math.pi / 3.0, aspect=1280.0 / 720.0):
        self.m_position = tuple(float(x) for x in position)
        self.m_yaw, self.m_pitch, self.m_fov, self.m_aspect_ratio = yaw, pitch, fov, aspect

    def uniform(self):
        cam = svo_camera()
        pos = (C.c_float * 3)(*self.m_position)
        rc = lib.svo_camera_from_yaw_pitch(pos, self.m_yaw, self.m_pitch, self.m_fov, self.m_aspect_ratio, C.byref(cam))
        if rc:
            raise SvoError(rc, "camera")
        return cam

    def uniform_array(self):
        c = self.uniform()
        return np.array(list(c.pos) + list(c.look) + list(c.side) + list(c.up), dtype=np.float32)


def camera_look_at(eye, target, fov=math.pi / 3.0, aspect=1.0):
    """Camera at `eye` (octree space) looking at `target`, via the reference's yaw/pitch form."""
    d = np.asarray(target, dtype=np.float64) - np.asarray(eye, dtype=np.float64)
    d /= np.linalg.norm(d)
    pitch = math.asin(max(-1.0, min(1.0, d[1])))
    yaw = math.atan2(d[0], d[2])
    return Camera(eye, yaw, pitch, fov, aspect)


class OctreeTracer:
    """Primary-ray marching of a whole frame into a device hit buffer (svo_hit records), then the
    reference tracer's shading of it into an RGBA8 image by m_view_type (OctreeTracer,
    src/OctreeTracer.hpp:15-16, shader/octree_tracer.frag:18-47)."""

    kDiffuse, kNormal, kPosition, kIteration = 0, 1, 2, 3  # OctreeTracer::ViewTypes

    def __init__(self, octree, camera):
        self.octree, self.camera = octree, camera
        self.ctx = octree.ctx
        self.m_beam_enable = True
        self.m_view_type = OctreeTracer.kDiffuse
        self.m_light = (5.0, 5.0, 5.0)  # the constant light (kDefaultConstantColor, src/Config.hpp:28)
        self.m_env_map = None  # an EnvironmentMap: misses show its radiance instead
        self.width, self.height = 1280, 720
        self.hits = None
        self.beam = None
        self.image = None

    @staticmethod
    def Create(octree, camera):
        return OctreeTracer(octree, camera)

    def Resize(self, width, height):
        self.width, self.height = int(width), int(height)
        self.hits = None

    def _ensure(self):
        if self.hits is None or self.hits.nbytes != self.width * self.height * 32:
            self.hits = self.ctx.alloc(self.width * self.height * 32)
            self.image = self.ctx.alloc(self.width * self.height * 4)
            bw, bh = (self.width + 7) // 8 + 1, (self.height + 7) // 8 + 1
            self.beam = self.ctx.alloc(bw * bh * 4)

    def CmdBeamRenderPass(self):
        """The beam pre-pass runs inside CmdDrawPipeline when m_beam_enable is set."""

    def CmdDrawPipeline(self):
        self._ensure()
        cam = self.camera.uniform() if isinstance(self.camera, Camera) else self.camera
        self.ctx.check(
            lib.svo_march(
                self.ctx.h,
                self.octree.h,
                C.byref(cam),
                self.width,
                self.height,
                1 if self.m_beam_enable else 0,
                self.hits.ptr,
                self.beam.ptr if self.m_beam_enable else None,
            ),
            "march",
        )
        self.CmdShade()

    def CmdShade(self):
        """Shades the current hit buffer into the RGBA8 image by m_view_type."""
        self._ensure()
        cam = self.camera.uniform() if isinstance(self.camera, Camera) else self.camera
        light = (C.c_float * 3)(*self.m_light)
        env = self.m_env_map
        self.ctx.check(lib.svo_shade(self.ctx.h, self.hits.ptr, C.byref(cam), self.width, self.height,
                                     int(self.m_view_type), light, env.h if env is not None else None,
                                     env.m_rotation if env is not None else 0.0, self.image.ptr), "shade")

    def GetHits(self):
        return self.hits.download(HIT_DTYPE, self.width * self.height).reshape(self.height, self.width)

    def GetImage(self):
        """(height, width) uint32 RGBA8 UNORM texels (R in the low byte)."""
        return self.image.download(np.uint32, self.width * self.height).reshape(self.height, self.width)

    def GetBeam(self):
        bw, bh = (self.width + 7) // 8 + 1, (self.height + 7) // 8 + 1
        return self.beam.download(np.float32, bw * bh).reshape(bh, bw)


class EnvironmentMap:
    """An equirectangular environment map in HBM with its sampling tables (EnvironmentMap,
    src/EnvironmentMap.cpp:55-190): rgba (H, W, 4) or (H, W, 3) float32 linear radiance."""

    def __init__(self, ctx, rgba):
        rgba = np.asarray(rgba, dtype=np.float32)
        if rgba.ndim != 3 or rgba.shape[2] not in (3, 4):
            raise ValueError("environment map must be (H, W, 3|4)")
        if rgba.shape[2] == 3:
            rgba = np.concatenate([rgba, np.zeros(rgba.shape[:2] + (1,), np.float32)], axis=2)
        rgba = np.ascontiguousarray(rgba)
        self.ctx = ctx
        self.m_height, self.m_width = rgba.shape[:2]
        self.m_rotation = 0.0
        self.m_multiplier = 1.0
        self.h = C.c_void_p()
        ctx.check(lib.svo_envmap_create(ctx.h, self.m_width, self.m_height, ptr(rgba), C.byref(self.h)), "envmap")

    @staticmethod
    def Create(ctx, rgba):
        return EnvironmentMap(ctx, rgba)

    def tables(self):
        """(texels (H, W, 4): rgb + pdf, alias (H, W, 2) uint32 {prob, alias}) from the device."""
        n = self.m_width * self.m_height
        t = np.empty(4 * n, np.float32)
        a = np.empty(2 * n, np.uint32)
        self.ctx.check(lib.svo_envmap_download(self.ctx.h, self.h, ptr(t), ptr(a)), "envmap download")
        return t.reshape(self.m_height, self.m_width, 4), a.reshape(self.m_height, self.m_width, 2)

    def close(self):
        if self.h:
            lib.svo_envmap_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass


class PathTracer:
    """Progressive path tracing (PathTracer, src/PathTracer.cpp:176-215; ConstColorPathTrace /
    EnvMapPathTrace, shader/path_tracer.comp): each CmdRender adds sample spp to the running means of
    colour (RGBA32F), first-hit albedo (RGBA8 UNORM) and first-hit normal (RGBA8 SNORM), drawing it
    with the reference's sampler (Sobol state after spp advances, 256x256 blue noise)."""

    def __init__(self, octree, camera):
        self.octree, self.camera = octree, camera
        self.ctx = octree.ctx
        self.m_width, self.m_height = 1280, 720
        self.m_bounce = 4  # kDefaultBounce, src/Config.hpp:25
        self.m_sun_radiance = (5.0, 5.0, 5.0)  # kDefaultConstantColor, src/Config.hpp:28 (uConstColor)
        self.spp = 0
        self.color = self.albedo = self.normal = None
        self.m_env_map = None  # an EnvironmentMap: EnvMapPathTrace instead of the constant colour

    @staticmethod
    def Create(octree, camera):
        return PathTracer(octree, camera)

    def Resize(self, width, height):
        self.m_width, self.m_height = int(width), int(height)
        self.color = None
        self.spp = 0

    def Reset(self):
        """Starts the means again (the reference clears its target images and its Sobol state)."""
        self.spp = 0

    def _ensure(self):
        n = self.m_width * self.m_height
        if self.color is None or self.color.nbytes != n * 16:
            self.color = self.ctx.alloc(n * 16)
            self.albedo, self.normal = self.ctx.alloc(n * 4), self.ctx.alloc(n * 4)
            self.spp = 0

    def CmdRender(self):
        self._ensure()
        cam = self.camera.uniform() if isinstance(self.camera, Camera) else self.camera
        env = self.m_env_map
        if env is not None:
            rc = lib.svo_path_trace_env(self.ctx.h, self.octree.h, C.byref(cam), self.m_width, self.m_height,
                                        self.m_bounce, self.spp, env.h, env.m_rotation, env.m_multiplier,
                                        self.color.ptr, self.albedo.ptr, self.normal.ptr)
        else:
            sky = (C.c_float * 3)(*self.m_sun_radiance)
            rc = lib.svo_path_trace(self.ctx.h, self.octree.h, C.byref(cam), self.m_width, self.m_height,
                                    self.m_bounce, self.spp, sky, self.color.ptr, self.albedo.ptr, self.normal.ptr)
        self.ctx.check(rc, "path trace")
        self.spp += 1  # Sobol::CmdNext after the dispatch (src/PathTracer.cpp:214)

    def GetSampleCount(self):
        return self.spp

    def GetColor(self):
        """(height, width, 4) float32."""
        return self.color.download(np.float32, self.m_width * self.m_height * 4).reshape(self.m_height, self.m_width, 4)

    def GetAlbedo(self):
        """(height, width) uint32 RGBA8 UNORM texels (R in the low byte)."""
        return self.albedo.download(np.uint32, self.m_width * self.m_height).reshape(self.m_height, self.m_width)

    def GetNormal(self):
        """(height, width) uint32 RGBA8 SNORM texels."""
        return self.normal.download(np.uint32, self.m_width * self.m_height).reshape(self.m_height, self.m_width)


def rgba8_unorm(texels):
    """(…, 4) float32 of RGBA8 UNORM texels: byte / 255."""
    b = np.asarray(texels, dtype=np.uint32)[..., None] >> np.array([0, 8, 16, 24], np.uint32)
    return (b & 0xFF).astype(np.float32) / np.float32(255.0)


def rgba8_snorm(texels):
    """(…, 4) float32 of RGBA8 SNORM texels: max(signed byte / 127, -1)."""
    b = np.asarray(texels, dtype=np.uint32)[..., None] >> np.array([0, 8, 16, 24], np.uint32)
    v = (b & 0xFF).astype(np.uint8).view(np.int8).astype(np.float32) / np.float32(127.0)
    return np.maximum(v, np.float32(-1.0))


def bounce_rays(ctx, hits, width, height, spp, cur=1):
    """Bounce `cur` of path-tracer sample spp from a device buffer of width*height primary hits:
    a device buffer of width*height x 6 floats (svo_bounce_rays)."""
    rays = ctx.alloc(width * height * 24)
    ctx.check(lib.svo_bounce_rays(ctx.h, hits.ptr, width, height, spp, cur, rays.ptr), "bounce rays")
    return rays


def march_rays(ctx, octree, rays, variant=_capi.SVO_MARCH_LEAF_ITER):
    """Marches explicit rays (R, 6) float32 in octree space; returns hit records."""
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
    R = rays.shape[0]
    d_rays = ctx.upload(rays)
    d_hits = ctx.alloc(R * 32)
    ctx.check(lib.svo_march_rays(ctx.h, octree.h, d_rays.ptr, R, variant, d_hits.ptr), "march rays")
    return d_hits.download(HIT_DTYPE, R)


def _cam(c):
    return c.uniform() if isinstance(c, Camera) else c


def _handles(objs):
    return (C.c_void_p * max(1, len(objs)))(*[o.h for o in objs])


def voxelize_batch(scenes, levels):
    """svo_voxelize_batch: voxelizes every scene (scene i at levels[i]) in one count launch and one
    fill launch; returns filled Voxelizer objects."""
    n = len(scenes)
    if n == 0:
        return []
    ctx = scenes[0].ctx
    lv = (C.c_uint32 * n)(*[int(x) for x in levels])
    out = (C.c_void_p * n)()
    ctx.check(lib.svo_voxelize_batch(ctx.h, _handles(scenes), lv, n, out), "voxelize batch")
    return [Voxelizer(ctx, C.c_void_p(out[i]), int(levels[i]), scenes[i]) for i in range(n)]


def build_batch(voxelizers, wait=True):
    """svo_build_batch: builds one octree per fragment list with one launch per build pass; returns
    Octree objects (finished when wait is set; otherwise they are ordered on the stream)."""
    n = len(voxelizers)
    if n == 0:
        return []
    ctx = voxelizers[0].ctx
    out = (C.c_void_p * n)()
    ctx.check(lib.svo_build_batch(ctx.h, _handles(voxelizers), n, out), "build batch")
    octs = []
    for i in range(n):
        W = C.c_uint64()
        if wait:
            ctx.check(lib.svo_build_finish(ctx.h, out[i], C.byref(W)), "build finish")
        o = Octree(ctx, C.c_void_p(out[i]), voxelizers[i].level, W.value)
        o._own = True
        octs.append(o)
    return octs


class BuildPlan:
    """svo_build_plan_*: voxelize + build of a fixed set of scenes at fixed levels, with every count
    kept on the device, captured once as a HIP graph and replayed by run() (a steady-state loop:
    Scene.update, run, finish).  Capacities: frag_caps / word_caps per scene, or None to size them
    from one plain build of the scenes as they are now, plus a quarter."""

    def __init__(self, scenes, levels, frag_caps=None, word_caps=None):
        self.ctx = scenes[0].ctx
        self.n = len(scenes)
        self.levels = [int(x) for x in levels]
        lv = (C.c_uint32 * self.n)(*self.levels)
        fc = (C.c_uint64 * self.n)(*[int(x) for x in frag_caps]) if frag_caps is not None else None
        wc = (C.c_uint64 * self.n)(*[int(x) for x in word_caps]) if word_caps is not None else None
        h = C.c_void_p()
        self.ctx.check(lib.svo_build_plan_create(self.ctx.h, _handles(scenes), lv, self.n, fc, wc, C.byref(h)),
                       "build plan")
        self.h = h
        self._scenes = list(scenes)
        self.F = [0] * self.n
        self.W = [0] * self.n

    def use_graph(self, on=True):
        self.ctx.check(lib.svo_build_plan_use_graph(self.h, 1 if on else 0), "build plan graph")

    def run(self):
        self.ctx.check(lib.svo_build_plan_run(self.h), "build plan run")

    def finish(self):
        """Waits for the last run; returns its device ms.  F / W receive the counts the scenes
        needed; raises SvoError (SVO_ERR_OVERFLOW) when one passed its capacity."""
        F = (C.c_uint64 * self.n)()
        W = (C.c_uint64 * self.n)()
        ms = C.c_double()
        rc = lib.svo_build_plan_finish(self.h, F, W, C.byref(ms))
        self.F, self.W = list(F), list(W)
        self.ctx.check(rc, "build plan finish")
        return ms.value

    def fragments(self, i):
        """Scene i's fragment list (owned by the plan)."""
        fl = C.c_void_p()
        self.ctx.check(lib.svo_build_plan_get(self.h, i, C.byref(fl), None), "build plan get")
        v = Voxelizer(self.ctx, fl, self.levels[i], self._scenes[i])
        v._borrowed = True
        v._plan = self
        return v

    def octree(self, i):
        """Scene i's octree (owned by the plan), at the size of the last finished run."""
        o = C.c_void_p()
        self.ctx.check(lib.svo_build_plan_get(self.h, i, None, C.byref(o)), "build plan get")
        lvl, W = C.c_uint32(), C.c_uint64()
        lib.svo_octree_info(o, C.byref(lvl), C.byref(W), None)
        t = Octree(self.ctx, o, lvl.value, W.value)
        t._plan = self
        return t

    def __del__(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib.svo_build_plan_destroy(self.h)
            self.h = None


class MarchPlan:
    """svo_march_plan_*: a reusable batch march of many octrees (one camera each)."""

    def __init__(self, octrees, cameras, width, height, beam=True):
        self.ctx = octrees[0].ctx
        self.n, self.width, self.height = len(octrees), int(width), int(height)
        cams = (svo_camera * self.n)(*[_cam(c) for c in cameras])
        h = C.c_void_p()
        self.ctx.check(
            lib.svo_march_plan_create(
                self.ctx.h, _handles(octrees), cams, self.n, self.width, self.height, 1 if beam else 0, C.byref(h)
            ),
            "march plan",
        )
        self.h = h
        self._octrees = list(octrees)

    def set_cameras(self, cameras):
        cams = (svo_camera * self.n)(*[_cam(c) for c in cameras])
        self.ctx.check(lib.svo_march_plan_set_cameras(self.ctx.h, self.h, cams), "plan cameras")

    def use_graph(self, on=True):
        self.ctx.check(lib.svo_march_plan_use_graph(self.h, 1 if on else 0), "plan graph")

    def set_shape(self, shape):
        """Launch tile: 0 = 16x16, 1 = 8x8, 2 = 16x8, 3 = 32x8, -1 = default / autotuned."""
        self.ctx.check(lib.svo_march_plan_set_shape(self.h, int(shape)), "plan shape")

    def run(self, hits):
        self.ctx.check(lib.svo_march_plan_run(self.ctx.h, self.h, hits.ptr), "plan run")

    def alloc_hits(self):
        return self.ctx.alloc(self.n * self.width * self.height * 32)

    def __del__(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib.svo_march_plan_destroy(self.h)
            self.h = None


def march_batch(octrees, cameras, width, height, beam=True):
    """svo_march_batch: hit records (n, height, width) of every octree's frame."""
    n = len(octrees)
    ctx = octrees[0].ctx
    hits = ctx.alloc(n * width * height * 32)
    cams = (svo_camera * n)(*[_cam(c) for c in cameras])
    ctx.check(
        lib.svo_march_batch(ctx.h, _handles(octrees), cams, n, width, height, 1 if beam else 0, hits.ptr),
        "march batch",
    )
    return hits.download(HIT_DTYPE, n * width * height).reshape(n, height, width)


def march_rays_batch(octrees, ray_sets, variant=_capi.SVO_MARCH_LEAF_ITER):
    """svo_march_rays_batch: ray_sets[i] (R_i, 6) float32 marched against octrees[i], one launch;
    returns one hit array per set."""
    n = len(octrees)
    ctx = octrees[0].ctx
    sets = [np.ascontiguousarray(r, dtype=np.float32).reshape(-1, 6) for r in ray_sets]
    begin = np.zeros(n + 1, dtype=np.uint64)
    begin[1:] = np.cumsum([r.shape[0] for r in sets])
    R = int(begin[-1])
    rays = np.concatenate(sets) if R else np.zeros((1, 6), np.float32)
    d_rays = ctx.upload(rays)
    d_hits = ctx.alloc(max(R, 1) * 32)
    rb = (C.c_uint64 * (n + 1))(*[int(x) for x in begin])
    ctx.check(lib.svo_march_rays_batch(ctx.h, _handles(octrees), n, rb, d_rays.ptr, variant, d_hits.ptr), "ray batch")
    hits = d_hits.download(HIT_DTYPE, R) if R else np.zeros(0, HIT_DTYPE)
    return [hits[int(begin[i]) : int(begin[i + 1])] for i in range(n)]


def autotune(ctx, on=True):
    """svo_autotune; returns the decision report so far."""
    ctx.check(lib.svo_autotune(ctx.h, 1 if on else 0), "autotune")
    n = C.c_uint64()
    lib.svo_autotune_report(ctx.h, None, 0, C.byref(n))
    buf = C.create_string_buffer(n.value + 1)
    lib.svo_autotune_report(ctx.h, buf, n.value + 1, None)
    return buf.value.decode()


class Server:
    """svo_server_*: a serving loop with resident octrees, admission by memory and p50/p99 latency.

    Requests may be submitted from any thread; wait() returns (status, latency_ms).  Scenes, ray and
    hit buffers passed in must stay alive until their request completes."""

    def __init__(self, device=0, budget_bytes=64 << 30, max_octrees=512):
        h = C.c_void_p()
        rc = lib.svo_server_create(device, int(budget_bytes), int(max_octrees), C.byref(h))
        if rc:
            raise SvoError(rc, "server create")
        self.h = h

    def submit_build(self, scene, level):
        rid = C.c_uint64()
        rc = lib.svo_server_submit_build(self.h, scene.h, int(level), C.byref(rid))
        if rc:
            raise SvoError(rc, "submit build")
        return rid.value

    def submit_rays(self, octree_id, rays, R, hits, variant=_capi.SVO_MARCH_LEAF_ITER):
        """rays, hits: DeviceArray (or raw device pointers) of R rays / R hit records."""
        rid = C.c_uint64()
        rp = rays.ptr if isinstance(rays, DeviceArray) else rays
        hp = hits.ptr if isinstance(hits, DeviceArray) else hits
        rc = lib.svo_server_submit_rays(self.h, int(octree_id), rp, int(R), int(variant), hp, C.byref(rid))
        if rc:
            raise SvoError(rc, "submit rays")
        return rid.value

    def wait(self, req_id):
        st, lat = C.c_int(), C.c_double()
        rc = lib.svo_server_wait(self.h, int(req_id), C.byref(st), C.byref(lat))
        if rc:
            raise SvoError(rc, "wait")
        return st.value, lat.value

    def resident(self, octree_id):
        """(W, octree handle) of a resident octree, or None."""
        W, h = C.c_uint64(), C.c_void_p()
        if lib.svo_server_resident(self.h, int(octree_id), C.byref(W), C.byref(h)):
            return None
        return W.value, h

    def stats(self):
        st = _capi.svo_server_stats()
        lib.svo_server_get_stats(self.h, C.byref(st))
        return {n: getattr(st, n) for n, _ in st._fields_}

    def reset_stats(self):
        lib.svo_server_reset_stats(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib.svo_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def load_obj(path):
    """OBJ + MTL file -> (xyz (T, 9) float32 positions as in the file, rgb (T,) uint32 per-triangle
    packUnorm4x8(Kd)), the input of Scene.Create(..., normalize=True) (Scene::load_meshes,
    src/Scene.cpp:36-147)."""
    xyz_p = C.POINTER(C.c_float)()
    rgb_p = C.POINTER(C.c_uint32)()
    T = C.c_uint64()
    msg = C.create_string_buffer(512)
    rc = lib.svo_obj_load(os.fsencode(path), C.byref(xyz_p), C.byref(rgb_p), C.byref(T), msg, len(msg))
    if rc:
        raise SvoError(rc, "load_obj %s: %s" % (path, msg.value.decode(errors="replace")))
    try:
        n = T.value
        xyz = np.ctypeslib.as_array(xyz_p, shape=(n * 9,)).reshape(n, 9).copy() if n else np.zeros((0, 9), np.float32)
        rgb = np.ctypeslib.as_array(rgb_p, shape=(n,)).copy() if n else np.zeros(0, np.uint32)
    finally:
        lib.svo_obj_free(xyz_p, rgb_p)
    return xyz, rgb


def gen_scene(kind, seed, param=0):
    """Seeded synthetic scene: (xyz (T, 9) float32, rgb (T,) uint32)."""
    T = C.c_uint64()
    rc = lib.svo_gen_scene(kind, seed, param, None, None, C.byref(T))
    if rc:
        raise SvoError(rc, "gen_scene")
    xyz = np.empty((T.value, 9), dtype=np.float32)
    rgb = np.empty(T.value, dtype=np.uint32)
    rc = lib.svo_gen_scene(kind, seed, param, ptr(xyz), ptr(rgb), C.byref(T))
    if rc:
        raise SvoError(rc, "gen_scene")
    return xyz, rgb
