"""Scene and detail inputs of the tracer, mirroring `pfx_trace::Scene`
(crates/trace/src/gpu.rs:249-256) and `detail::Detail`
(crates/trace/src/detail.rs:374-398) for the parts the still path uses.

Arrays use the numpy layouts in `pfx_amd._native`, so `Scene.to_c()` hands
their memory to the C ABI without copies.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N


@dataclass
class Camera:
    """gpu::Camera (crates/trace/src/gpu.rs:172-178)."""

    origin: tuple = (0.0, 0.0, 0.0)
    forward: tuple = (0.0, 0.0, 1.0)
    right: tuple = (1.0, 0.0, 0.0)
    up: tuple = (0.0, 1.0, 0.0)

    def to_c(self) -> N.Camera:
        c = N.Camera()
        c.origin[:] = [float(v) for v in self.origin]
        c.forward[:] = [float(v) for v in self.forward]
        c.right[:] = [float(v) for v in self.right]
        c.up[:] = [float(v) for v in self.up]
        return c

    @staticmethod
    def look(origin, target, fov_deg: float, aspect: float, world_up=(0.0, 1.0, 0.0)) -> "Camera":
        """A pinhole camera: forward is unit, right/up span the image half-extents."""
        o = np.asarray(origin, np.float64)
        f = np.asarray(target, np.float64) - o
        f /= np.linalg.norm(f)
        r = np.cross(f, np.asarray(world_up, np.float64))
        r /= np.linalg.norm(r)
        u = np.cross(r, f)
        h = np.tan(np.radians(fov_deg) * 0.5)
        return Camera(tuple(o), tuple(f), tuple(r * h * aspect), tuple(u * h))


@dataclass
class Sun:
    """gpu::Sun (crates/trace/src/gpu.rs:230-235); intensity 0 turns it off."""

    direction: tuple = (0.0, 1.0, 0.0)
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 0.0

    def to_c(self) -> N.Sun:
        s = N.Sun()
        s.direction[:] = [float(v) for v in self.direction]
        s.color[:] = [float(v) for v in self.color]
        s.intensity = float(self.intensity)
        return s


@dataclass
class Bounces:
    """detail::Bounces (crates/trace/src/detail.rs:414-453)."""

    total: int = 16
    diffuse: int = 4
    glossy: int = 8
    transmission: int = 12

    @staticmethod
    def deep(count: int) -> "Bounces":
        return Bounces(count, count, count, count)


ACCUM_COMPENSATED = N.ACCUM_COMPENSATED
ACCUM_PLAIN = N.ACCUM_PLAIN
ACCUM_FP64 = N.ACCUM_FP64
PERSPECTIVE = N.PROJ_PERSPECTIVE
ORTHOGRAPHIC = N.PROJ_ORTHOGRAPHIC
EQUIRECT = N.PROJ_EQUIRECT


@dataclass
class Detail:
    """The parts of detail::Detail the still path reads. The sampler is PCG
    (`Sampler::Pcg`); `accumulate` picks compensated FP32 (default), the
    reference's plain FP32 sum, or FP64."""

    bounces: Bounces = field(default_factory=Bounces)
    lens_shift: tuple = (0.0, 0.0)
    aperture: float = 0.0
    focus_distance: float = 0.0
    sun_radius_deg: float = 0.0
    accumulate: int = ACCUM_COMPENSATED
    wave_paths: int = 0
    projection: int = PERSPECTIVE  # gpu::Projection
    ortho_size: tuple = (0.0, 0.0)  # orthographic view width, height
    emitter_sampling: bool = True  # False: emissive triangles only through BSDF sampling (gpu.rs:781)
    watertight: bool = False  # True: the watertight ray-triangle test instead of the reference's
    left_first: bool = False  # True: closest hits walk the reference's left-first order (default nearest first)
    sampler: str = "pcg"  # "pcg" (Sampler::Pcg) or "sobol" (Sampler::Sobol { samples })
    sobol_samples: int = 256

    def to_c(self) -> N.DetailC:
        d = N.DetailC()
        d.bounces = N.Bounces(self.bounces.total, self.bounces.diffuse, self.bounces.glossy, self.bounces.transmission)
        d.lens_shift[:] = [float(v) for v in self.lens_shift]
        d.aperture = float(self.aperture)
        d.focus_distance = float(self.focus_distance)
        d.sun_radius_deg = float(self.sun_radius_deg)
        d.accumulate = int(self.accumulate)
        d.wave_paths = int(self.wave_paths)
        d.projection = int(self.projection)
        d.ortho_size[:] = [float(v) for v in self.ortho_size]
        d.no_emitter_sampling = 0 if self.emitter_sampling else 1
        d.triangle_test = 1 if self.watertight else 0
        d.walk_order = 1 if self.left_first else 0
        d.sampler = {"pcg": 0, "sobol": 1}[self.sampler]
        d.sobol_samples = int(self.sobol_samples)
        return d


@dataclass
class Adaptive:
    """adaptive::Adaptive (crates/trace/src/adaptive.rs:9-26)."""

    threshold: float = 0.005
    min_samples: int = 32
    max_samples: int = 1024
    growth: float = 1.5

    def to_c(self) -> N.AdaptiveC:
        return N.AdaptiveC(float(self.threshold), int(self.min_samples), int(self.max_samples), float(self.growth))


def _empty(dtype):
    return np.zeros(0, dtype=dtype)


@dataclass
class Scene:
    """A still's geometry, materials, lights and sky.

    triangles:      flat world triangles (bvh::Triangle)
    world_details:  world triangles with vertex normals (Detail::triangles)
    mesh_triangles: triangles of shared meshes, ranged by `meshes`
    instances:      placed meshes with column-major 4x4 transforms
    sky:            (H, W, 4) float32 lat-long environment, or None (black)
    """

    triangles: np.ndarray = field(default_factory=lambda: _empty(N.TRIANGLE))
    materials: np.ndarray = field(default_factory=lambda: _empty(N.MATERIAL))
    world_details: np.ndarray = field(default_factory=lambda: _empty(N.TRIANGLE_DETAIL))
    mesh_triangles: np.ndarray = field(default_factory=lambda: _empty(N.TRIANGLE_DETAIL))
    meshes: np.ndarray = field(default_factory=lambda: _empty(N.MESH))
    instances: np.ndarray = field(default_factory=lambda: _empty(N.INSTANCE))
    lights: np.ndarray = field(default_factory=lambda: _empty(N.LIGHT))
    sky: np.ndarray | None = None
    camera: Camera = field(default_factory=Camera)
    sun: Sun = field(default_factory=Sun)

    def triangle_count(self) -> int:
        return len(self.triangles) + len(self.world_details) + int(
            sum(int(self.meshes[i]["count"]) for i in self.instances["mesh"]) if len(self.instances) else 0
        )

    def to_c(self) -> tuple[N.SceneC, list]:
        """The C struct plus the arrays it points into (keep them alive)."""
        keep = []

        def arr(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return (a.ctypes.data if len(a) else None), len(a)

        s = N.SceneC()
        s.triangles, s.triangle_count = arr(self.triangles, N.TRIANGLE)
        s.world_details, s.world_detail_count = arr(self.world_details, N.TRIANGLE_DETAIL)
        s.mesh_triangles, s.mesh_triangle_count = arr(self.mesh_triangles, N.TRIANGLE_DETAIL)
        s.meshes, s.mesh_count = arr(self.meshes, N.MESH)
        s.instances, s.instance_count = arr(self.instances, N.INSTANCE)
        s.materials, s.material_count = arr(self.materials, N.MATERIAL)
        s.lights, s.light_count = arr(self.lights, N.LIGHT)
        if self.sky is not None and self.sky.size:
            sky = np.ascontiguousarray(self.sky, dtype=np.float32)
            keep.append(sky)
            s.sky = N.Environment(sky.shape[1], sky.shape[0], sky.ctypes.data)
        else:
            s.sky = N.Environment(0, 0, None)
        s.camera = self.camera.to_c()
        s.sun = self.sun.to_c()
        return s, keep


SCENE_MAGIC = b"PFXSCN02"


def save_scene(scene: "Scene", path) -> None:
    """A scene file for non-Python callers (pfxh-render): b"PFXSCN02", nine
    little-endian u32 counts (triangles, world details, mesh triangles,
    meshes, instances, materials, lights, sky width, sky height), the camera
    (12 f32) and sun (7 f32), then the arrays in that order as the C
    structs of include/pfxhip.h lay them out, then the sky texels."""
    sky = scene.sky if scene.sky is not None and scene.sky.size else np.zeros((0, 0, 4), np.float32)
    arrays = [(scene.triangles, N.TRIANGLE), (scene.world_details, N.TRIANGLE_DETAIL),
              (scene.mesh_triangles, N.TRIANGLE_DETAIL), (scene.meshes, N.MESH), (scene.instances, N.INSTANCE),
              (scene.materials, N.MATERIAL), (scene.lights, N.LIGHT)]
    c, sun = scene.camera, scene.sun
    with open(path, "wb") as f:
        f.write(SCENE_MAGIC)
        counts = [len(a) for a, _ in arrays] + [sky.shape[1], sky.shape[0]]
        f.write(np.asarray(counts, "<u4").tobytes())
        f.write(np.asarray(list(c.origin) + list(c.forward) + list(c.right) + list(c.up), "<f4").tobytes())
        f.write(np.asarray(list(sun.direction) + list(sun.color) + [sun.intensity], "<f4").tobytes())
        for a, dt in arrays:
            f.write(np.ascontiguousarray(a, dt).tobytes())
        f.write(np.ascontiguousarray(sky, "<f4").tobytes())


def load_scene(path) -> "Scene":
    data = memoryview(open(path, "rb").read())
    if bytes(data[:8]) != SCENE_MAGIC:
        raise ValueError(f"{path}: not a PFXSCN02 scene file")
    counts = np.frombuffer(data[8:44], "<u4")
    cam = np.frombuffer(data[44:92], "<f4")
    sun = np.frombuffer(data[92:120], "<f4")
    at = 120
    out = []
    for n, dt in zip(counts[:7], (N.TRIANGLE, N.TRIANGLE_DETAIL, N.TRIANGLE_DETAIL, N.MESH, N.INSTANCE, N.MATERIAL,
                                  N.LIGHT)):
        size = int(n) * np.dtype(dt).itemsize
        out.append(np.frombuffer(data[at:at + size], dt).copy())
        at += size
    sw, sh = int(counts[7]), int(counts[8])
    sky = np.frombuffer(data[at:at + sw * sh * 16], "<f4").reshape(sh, sw, 4).copy() if sw * sh else None
    return Scene(triangles=out[0], world_details=out[1], mesh_triangles=out[2], meshes=out[3], instances=out[4],
                 materials=out[5], lights=out[6], sky=sky,
                 camera=Camera(tuple(cam[0:3]), tuple(cam[3:6]), tuple(cam[6:9]), tuple(cam[9:12])),
                 sun=Sun(tuple(sun[0:3]), tuple(sun[3:6]), float(sun[6])))


def material(base=(0.5, 0.5, 0.5), roughness=0.5, metalness=0.0, specular=0.04, clearcoat=0.0,
             clearcoat_roughness=0.05, sheen=0.0, emission=(0.0, 0.0, 0.0), fresnel_power=5.0) -> np.ndarray:
    """One material row; the defaults are `Material::default`
    (crates/materials/src/material.rs:534-575)."""
    m = np.zeros(1, dtype=N.MATERIAL)
    m["base"] = base
    m["roughness"] = roughness
    m["metalness"] = metalness
    m["specular"] = specular
    m["clearcoat"] = clearcoat
    m["clearcoat_roughness"] = clearcoat_roughness
    m["sheen"] = sheen
    m["emission"] = emission
    m["fresnel_power"] = fresnel_power
    return m


def point_light(position, colour, intensity, range_, radius=0.0, shadow=True) -> np.ndarray:
    """LocalLight::point (crates/trace/src/lights.rs:42-52)."""
    l = np.zeros(1, dtype=N.LIGHT)
    l["position"] = position
    l["colour"] = colour
    l["intensity"] = intensity
    l["radius"] = radius
    l["range"] = range_
    l["shadow"] = 1 if shadow else 0
    l["kind"] = N.LIGHT_POINT
    l["direction"] = (0.0, -1.0, 0.0)
    return l


def spot_light(position, colour, intensity, range_, direction, inner_deg, outer_deg, radius=0.0,
               shadow=True) -> np.ndarray:
    """LocalLight with LightShape::Spot (crates/trace/src/lights.rs:13-17)."""
    l = point_light(position, colour, intensity, range_, radius, shadow)
    l["kind"] = N.LIGHT_SPOT
    l["direction"] = direction
    l["inner_deg"] = inner_deg
    l["outer_deg"] = outer_deg
    return l


F32_MAX = float(np.finfo(np.float32).max)


def rect_light(centre, half_u, half_v, colour, intensity, two_sided=False, shadow=True) -> np.ndarray:
    """LocalLight::rect (crates/trace/src/lights.rs:55-76): a rectangle of
    sides 2*half_u and 2*half_v (perpendicular) emitting `colour *
    intensity` watts spread over its area, along half_u x half_v."""
    l = point_light(centre, colour, intensity, F32_MAX, 0.0, shadow)
    l["kind"] = N.LIGHT_RECT
    l["half_u"] = half_u
    l["half_v"] = half_v
    l["two_sided"] = 1 if two_sided else 0
    return l


def disc_light(centre, normal, radius, colour, intensity, two_sided=False, shadow=True) -> np.ndarray:
    """LocalLight::disc (crates/trace/src/lights.rs:78-99)."""
    l = point_light(centre, colour, intensity, F32_MAX, 0.0, shadow)
    l["kind"] = N.LIGHT_DISC
    l["direction"] = normal
    l["disc_radius"] = radius
    l["two_sided"] = 1 if two_sided else 0
    return l
