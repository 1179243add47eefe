"""Staging: indexed meshes and placements into a tracer scene, over the C ABI
(`pfxh_stage_build`), mirroring `pfx_trace::stage::Stage::build`
(crates/trace/src/stage.rs:210-385).

    staged = stage([StageMesh(positions, normals, indices, alpha)],
                   [Placement(0, model, material=1)], materials, camera=cam)
    staged.scene, staged.detail   # a Scene and Detail for Trace.new_detailed
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .scene import Bounces, Camera, Detail, Scene, Sun

ALPHA_CUTOFF = 0.5  # stage.rs:25
IDENTITY = np.eye(4, dtype=np.float32)


class StageMeshC(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("normals", C.c_void_p), ("alpha", C.c_void_p),
                ("vertex_count", C.c_uint32), ("indices", C.c_void_p), ("index_count", C.c_uint32)]


class PlacementC(C.Structure):
    _fields_ = [("mesh", C.c_uint32), ("material", C.c_uint32), ("model", C.c_float * 16),
                ("casts_shadow", C.c_uint32), ("shadow_only", C.c_uint32), ("two_sided", C.c_uint32),
                ("alpha_cutoff", C.c_float)]


class StageC(C.Structure):
    _fields_ = [("meshes", C.c_void_p), ("mesh_count", C.c_uint32), ("placements", C.c_void_p),
                ("placement_count", C.c_uint32), ("materials", C.c_void_p), ("material_count", C.c_uint32),
                ("lights", C.c_void_p), ("light_count", C.c_uint32), ("sky", N.Environment), ("sun", N.Sun),
                ("camera", N.Camera), ("projection", C.c_uint32), ("ortho_size", C.c_float * 2),
                ("lens_shift", C.c_float * 2), ("aperture", C.c_float), ("focus_distance", C.c_float)]


@dataclass
class StageMesh:
    """stage::Mesh (stage.rs:27-36) without uv sets."""

    positions: np.ndarray  # (n, 3)
    normals: np.ndarray  # (n, 3)
    indices: np.ndarray  # (3m,)
    alpha: np.ndarray | None = None  # (n,)


@dataclass
class Placement:
    """stage::Placement (stage.rs:150-195); defaults of Placement::new."""

    mesh: int
    model: np.ndarray = field(default_factory=lambda: IDENTITY.copy())  # column-major 4x4 (columns as rows)
    material: int = 0
    casts_shadow: bool = True
    shadow_only: bool = False
    two_sided: bool = True
    alpha_cutoff: float = ALPHA_CUTOFF


@dataclass
class Staged:
    scene: Scene
    detail: Detail


def stage(meshes, placements, materials, lights=None, sky=None, sun: Sun | None = None,
          camera: Camera | None = None, projection: int = N.PROJ_PERSPECTIVE, ortho_size=(0.0, 0.0),
          lens_shift=(0.0, 0.0), aperture: float = 0.0, focus_distance: float = 1.0) -> Staged:
    """Stage::build; raises ValueError with the reference's message."""
    keep = []

    def arr(a, dtype):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype)
        keep.append(a)
        return a.ctypes.data if a.size else None

    mesh_c = (StageMeshC * max(len(meshes), 1))()
    for k, m in enumerate(meshes):
        pos = np.asarray(m.positions, np.float32).reshape(-1, 3)
        mesh_c[k] = StageMeshC(arr(pos, np.float32), arr(np.asarray(m.normals, np.float32).reshape(-1, 3), np.float32),
                               arr(m.alpha, np.float32), len(pos), arr(np.asarray(m.indices, np.uint32), np.uint32),
                               len(m.indices))
    place_c = (PlacementC * max(len(placements), 1))()
    for k, p in enumerate(placements):
        place_c[k].mesh, place_c[k].material = int(p.mesh), int(p.material)
        place_c[k].model[:] = [float(v) for v in np.asarray(p.model, np.float32).reshape(16)]
        place_c[k].casts_shadow = 1 if p.casts_shadow else 0
        place_c[k].shadow_only = 1 if p.shadow_only else 0
        place_c[k].two_sided = 1 if p.two_sided else 0
        place_c[k].alpha_cutoff = float(p.alpha_cutoff)
    st = StageC()
    st.meshes, st.mesh_count = C.cast(mesh_c, C.c_void_p), len(meshes)
    st.placements, st.placement_count = C.cast(place_c, C.c_void_p), len(placements)
    st.materials, st.material_count = arr(materials, N.MATERIAL), len(materials)
    st.lights, st.light_count = (arr(lights, N.LIGHT), len(lights)) if lights is not None and len(lights) else (None, 0)
    if sky is not None and np.asarray(sky).size:
        s = arr(np.asarray(sky, np.float32), np.float32)
        st.sky = N.Environment(np.asarray(sky).shape[1], np.asarray(sky).shape[0], s)
    else:
        st.sky = N.Environment(0, 0, None)
    st.sun = (sun or Sun()).to_c()
    st.camera = (camera or Camera()).to_c()
    st.projection = int(projection)
    st.ortho_size[:] = [float(v) for v in ortho_size]
    st.lens_shift[:] = [float(v) for v in lens_shift]
    st.aperture, st.focus_distance = float(aperture), float(focus_distance)
    out = C.c_void_p()
    err = N.err_buffer()
    lib = N.lib()
    status = lib.pfxh_stage_build(C.byref(st), C.byref(out), err, len(err))
    if status != N.OK:
        raise ValueError(err.value.decode(errors="replace"))
    try:
        sc = N.SceneC.from_address(lib.pfxh_staged_scene(out))

        def take(ptr, count, dtype):
            if not count:
                return np.zeros(0, dtype)
            raw = (C.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr)
            return np.frombuffer(bytes(raw), dtype).copy()

        tris = take(sc.mesh_triangles, sc.mesh_triangle_count, N.TRIANGLE_DETAIL)
        mesh_rows = take(sc.meshes, sc.mesh_count, N.MESH)
        inst = take(sc.instances, sc.instance_count, N.INSTANCE)
        mats = take(sc.materials, sc.material_count, N.MATERIAL)
        lts = take(sc.lights, sc.light_count, N.LIGHT)
        sky_out = None
        if sc.sky.width:
            sky_out = np.frombuffer(bytes((C.c_char * (sc.sky.width * sc.sky.height * 16)).from_address(sc.sky.texels)),
                                    np.float32).reshape(sc.sky.height, sc.sky.width, 4).copy()
        d = N.DetailC()
        lib.pfxh_staged_detail(out, C.byref(d))
        c = sc.camera
        scene = Scene(materials=mats, mesh_triangles=tris, meshes=mesh_rows, instances=inst, lights=lts, sky=sky_out,
                      camera=Camera(tuple(c.origin), tuple(c.forward), tuple(c.right), tuple(c.up)),
                      sun=Sun(tuple(sc.sun.direction), tuple(sc.sun.color), float(sc.sun.intensity)))
        detail = Detail(bounces=Bounces(d.bounces.total, d.bounces.diffuse, d.bounces.glossy, d.bounces.transmission),
                        lens_shift=tuple(d.lens_shift), aperture=d.aperture, focus_distance=d.focus_distance,
                        projection=d.projection, ortho_size=tuple(d.ortho_size))
        return Staged(scene, detail)
    finally:
        lib.pfxh_staged_free(out)
