"""Scenes of the reference's statistical tracer tests, restated with
triangles (the reference's analytic ellipsoids become icospheres, since
analytic shapes are outside this tracer): the white furnace
(crates/trace/src/gpu.rs:3787-3830), the softbox room (:4153-4213) and the
lamp street (:4316-4378)."""
import numpy as np

from pfx_amd import Camera, Scene, Sun, material, point_light, spot_light
from pfx_amd import _native as N
from pfx_amd.scenes import icosphere, quad


def triangles(verts, mats) -> np.ndarray:
    verts = np.asarray(verts, np.float64)
    out = np.zeros(len(verts), dtype=N.TRIANGLE)
    out["vertices"] = verts.reshape(len(verts), 9).astype(np.float32)
    out["material"] = np.broadcast_to(np.asarray(mats, np.uint32), (len(verts),))
    return out


def sphere(center, radius, subdivisions=3) -> np.ndarray:
    return icosphere(subdivisions) * radius + np.asarray(center, np.float64)


NO_SUN = Sun(direction=(0.0, 1.0, 0.0), color=(0.0, 0.0, 0.0), intensity=0.0)


def furnace() -> Scene:
    """A white Lambertian sphere in front of the camera under a uniform sky of
    radiance 1 (gpu.rs:3788-3818): every pixel's expectation is 1."""
    sky = np.ones((2, 4, 4), np.float32)
    return Scene(
        triangles=triangles(sphere((0.0, 0.0, -3.0), 1.0), 0),
        materials=material(base=(1.0, 1.0, 1.0), specular=0.0, roughness=1.0),
        sky=sky,
        camera=Camera(origin=(0.0, 0.0, 0.0), forward=(0.0, 0.0, -1.0), right=(1.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)),
        sun=NO_SUN,
    )


def softbox_room() -> Scene:
    """A closed room lit by two emissive ceiling quads and a small glowing
    sphere (gpu.rs:4153-4213)."""
    x0, x1, y0, y1, z0, z1 = -2.0, 2.0, 0.0, 3.0, -2.0, 2.0
    faces = [
        [[x0, y0, z0], [x1, y0, z0], [x1, y0, z1], [x0, y0, z1]],
        [[x0, y1, z0], [x0, y1, z1], [x1, y1, z1], [x1, y1, z0]],
        [[x0, y0, z0], [x0, y1, z0], [x1, y1, z0], [x1, y0, z0]],
        [[x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]],
        [[x0, y0, z0], [x0, y0, z1], [x0, y1, z1], [x0, y1, z0]],
        [[x1, y0, z0], [x1, y1, z0], [x1, y1, z1], [x1, y0, z1]],
    ]
    verts = [quad(*f) for f in faces]
    mats = [np.zeros(12, np.uint32)]
    for cx, cy, cz, hx, hz in [(-0.9, 2.6, -0.6, 0.3, 0.2), (1.0, 2.3, 0.5, 0.2, 0.25)]:
        verts.append(quad([cx - hx, cy, cz - hz], [cx + hx, cy, cz - hz], [cx + hx, cy, cz + hz], [cx - hx, cy, cz + hz]))
        mats.append(np.ones(2, np.uint32))
    ball = sphere((-1.6, 2.6, 1.6), 0.12, 2)
    verts.append(ball)
    mats.append(np.full(len(ball), 2, np.uint32))
    tris = triangles(np.concatenate(verts), np.concatenate(mats))
    mat = np.concatenate([
        material(base=(0.7, 0.65, 0.6), specular=0.0, roughness=1.0),
        material(base=(0.0, 0.0, 0.0), specular=0.0, emission=(8.0, 7.5, 7.0)),
        material(base=(0.0, 0.0, 0.0), specular=0.0, emission=(6.0, 3.0, 1.5)),
    ])
    return Scene(
        triangles=tris, materials=mat, sky=None,
        camera=Camera(origin=(0.0, 1.4, 1.9), forward=(0.0, -0.35, -1.0), right=(0.8, 0.0, 0.0), up=(0.0, 0.47, -0.165)),
        sun=NO_SUN,
    )


def lamp_street(count: int) -> tuple:
    """A street lit by `count` point and spot lamps over eight spheres
    (gpu.rs:4316-4378). Returns the scene without lights and the lights."""
    floor = material(base=(0.5, 0.48, 0.45), roughness=0.7)
    verts = [quad([-12.0, 0.0, -6.0], [12.0, 0.0, -6.0], [12.0, 0.0, 70.0], [-12.0, 0.0, 70.0])]
    for k in range(8):
        verts.append(sphere((-2.0 if k % 2 == 0 else 2.5, 0.6, 4.0 + k * 7.5), 0.6, 2))
    tris = triangles(np.concatenate(verts), 0)
    lights = []
    for k in range(count):
        side = -5.0 if k % 2 == 0 else 5.0
        z = (k // 2) * 64.0 / max(count // 2, 1)
        warm = 0.6 + 0.4 * ((k * 7) % 5) / 4.0
        colour = (1.0, warm, warm * 0.8)
        if k % 3 == 0:
            lights.append(spot_light((side, 3.0, z), colour, 6.0, 9.0, (-side * 0.1, -1.0, 0.0), 30.0, 55.0, radius=0.05))
        else:
            lights.append(point_light((side, 3.0, z), colour, 6.0, 9.0, radius=0.05))
    scene = Scene(
        triangles=tris, materials=floor, sky=None,
        camera=Camera(origin=(0.0, 60.0, 32.0), forward=(0.0, -1.0, 0.0), right=(0.2, 0.0, 0.0), up=(0.0, 0.0, -0.6)),
        sun=NO_SUN,
    )
    return scene, np.concatenate(lights) if lights else np.zeros(0, N.LIGHT)


def brightness(color: np.ndarray) -> np.ndarray:
    """gpu.rs:4232-4234: Rec. 709 luma of the RGB channels."""
    c = np.asarray(color, np.float64)
    return 0.2126 * c[..., 0] + 0.7152 * c[..., 1] + 0.0722 * c[..., 2]
