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


# ---- direct light from point and spot lights (gpu.rs:4443-4560, reduced) ---
LIT_FLOOR_BASE = (0.5, 0.45, 0.4)


def lit_floor() -> tuple:
    """A Lambertian floor under six point and spot lights, seen from above
    (the camera of gpu.rs:4477-4482), direct light only: the reference's
    `local_lights_match_the_live_direct_term` without its analytic occluder
    and rect light, which this tracer does not have. Six lights exceed the
    four the reference loops over, so every vertex picks one through the
    light tree. The floor has specular 0 and metalness 0, where the
    reference's BRDF is exactly base / pi (brdf.wgsl:206-220: f0 = 0 gives a
    zero Fresnel term, no specular lobe, full diffuse weight)."""
    floor = quad((-4.0, 0.0, -4.0), (4.0, 0.0, -4.0), (4.0, 0.0, 4.0), (-4.0, 0.0, 4.0))
    mats = material(LIT_FLOOR_BASE, roughness=0.6, specular=0.0)
    lights = np.concatenate([
        point_light((-0.8, 1.2, 0.3), (1.0, 0.9, 0.8), 3.0, 8.0),
        point_light((0.6, 0.7, -0.9), (0.8, 1.0, 0.9), 1.5, 2.5),   # range cuts in: the window shows
        point_light((-1.3, 2.0, -1.1), (0.7, 0.8, 1.0), 6.0, 12.0),
        spot_light((1.2, 1.5, 1.2), (0.9, 1.0, 0.8), 4.0, 10.0, (0.0, -1.0, 0.0), 15.0, 25.0),
        spot_light((-1.0, 1.0, 1.4), (1.0, 0.7, 0.6), 2.5, 6.0, (0.3, -1.0, 0.1), 20.0, 40.0),
        spot_light((0.2, 2.5, -0.2), (0.8, 0.8, 1.0), 8.0, 9.0, (-0.2, -1.0, -0.3), 10.0, 30.0),
    ])
    cam = Camera(origin=(0.0, 4.0, 0.0), forward=(0.0, -1.0, 0.0), right=(0.6, 0.0, 0.0), up=(0.0, 0.0, -0.6))
    scene = Scene(triangles=triangles(floor, 0), materials=mats, camera=cam, sun=NO_SUN, lights=lights)
    return scene, lights


def light_falloff(lights, k, distance):
    """lights.rs falloff (intensity x window^2 / max(d, radius)^2), vectorised;
    tests check it against the oracle's, which the reference's known answers
    pin (lights.rs:578-628)."""
    l = lights[k]
    ratio = distance / l["range"]
    window = np.clip(1.0 - ratio ** 4, 0.0, 1.0)
    out = l["intensity"] * window * window / np.maximum(distance, l["radius"]) ** 2
    return np.where((distance >= l["range"]) | (distance <= 1e-6), 0.0, out)


def light_cone(lights, k, toward):
    """lights.rs cone for a spot (1 for a point), vectorised; `toward` runs
    from the light to the lit point."""
    l = lights[k]
    if l["kind"] != N.LIGHT_SPOT:
        return np.ones(toward.shape[:-1])
    inner, outer = np.cos(np.radians(l["inner_deg"])), np.cos(np.radians(l["outer_deg"]))
    scale = 1.0 / max(inner - outer, 0.001)
    offset = -outer * scale
    axis = np.asarray(l["direction"], np.float64)
    axis = axis / np.linalg.norm(axis)
    t = np.clip((toward / np.linalg.norm(toward, axis=-1, keepdims=True)) @ axis * scale + offset, 0.0, 1.0)
    return t * t


def lit_floor_expected(scene, lights, size, sub=16) -> np.ndarray:
    """Each pixel's expected colour: the direct term summed over the lights,
    (base / pi) cos(theta) colour falloff(d) cone, averaged over a sub x sub
    grid across the pixel (the tracer jitters uniformly inside it)."""
    w, h = size
    cam = scene.camera
    o = np.asarray(cam.origin, np.float64)
    fwd, right, up = (np.asarray(v, np.float64) for v in (cam.forward, cam.right, cam.up))
    s = (np.arange(sub) + 0.5) / sub
    px = (np.arange(w)[:, None] + s[None, :]).ravel()
    py = (np.arange(h)[:, None] + s[None, :]).ravel()
    ndc_x = px / w * 2.0 - 1.0
    ndc_y = py / h * 2.0 - 1.0
    d = fwd[None, None] + right[None, None] * ndc_x[None, :, None] + up[None, None] * (-ndc_y[:, None, None])
    d = d / np.linalg.norm(d, axis=-1, keepdims=True)
    p = o + d * (-o[1] / d[..., 1:2])
    total = np.zeros(p.shape)
    base = np.asarray(LIT_FLOOR_BASE) / np.pi
    for k in range(len(lights)):
        to = np.asarray(lights[k]["position"], np.float64) - p
        dist = np.linalg.norm(to, axis=-1)
        cos = np.clip(to[..., 1] / dist, 0.0, None)
        val = cos * light_falloff(lights, k, dist) * light_cone(lights, k, -to)
        total += val[..., None] * np.asarray(lights[k]["colour"], np.float64) * base
    return total.reshape(h, sub, w, sub, 3).mean(axis=(1, 3))


# ---- direct light with the rect light of gpu.rs:4443-4560, plus a disc ----
RECT = dict(centre=(0.9, 1.0, -0.4), half_u=(0.3, 0.0, 0.0), half_v=(0.0, 0.0, 0.2), colour=(0.8, 0.9, 1.0),
            intensity=2.0)
DISC = dict(centre=(-0.6, 0.9, -1.0), normal=(0.0, -1.0, 0.0), radius=0.25, colour=(1.0, 0.8, 0.6), intensity=1.5)


def lit_floor_area() -> tuple:
    """The reference's `local_lights_match_the_live_direct_term` lights on
    the Lambertian floor: the point light of radius 0.1, the spot light and
    the one-sided rect light facing down (gpu.rs:4491-4509), plus a
    one-sided disc light facing down that casts no shadow (its light is
    sampled without MIS or a visibility test). The analytic occluder sphere
    of the reference's scene is left out (no analytic shapes here). The
    camera looks down on the backs of the area lights, which do not show."""
    from pfx_amd.scene import disc_light, rect_light

    floor = quad((-4.0, 0.0, -4.0), (4.0, 0.0, -4.0), (4.0, 0.0, 4.0), (-4.0, 0.0, 4.0))
    mats = material(LIT_FLOOR_BASE, roughness=0.6, specular=0.0)
    lights = np.concatenate([
        point_light((-0.8, 1.2, 0.3), (1.0, 0.9, 0.8), 3.0, 8.0, radius=0.1),
        spot_light((1.2, 1.5, 1.2), (0.9, 1.0, 0.8), 4.0, 10.0, (0.0, -1.0, 0.0), 15.0, 25.0),
        rect_light(RECT["centre"], RECT["half_u"], RECT["half_v"], RECT["colour"], RECT["intensity"]),
        disc_light(DISC["centre"], DISC["normal"], DISC["radius"], DISC["colour"], DISC["intensity"], shadow=False),
    ])
    cam = Camera(origin=(0.0, 4.0, 0.0), forward=(0.0, -1.0, 0.0), right=(0.6, 0.0, 0.0), up=(0.0, 0.0, -0.6))
    scene = Scene(triangles=triangles(floor, 0), materials=mats, camera=cam, sun=NO_SUN, lights=lights)
    return scene, lights


def _floor_points(scene, size, sub):
    w, h = size
    cam = scene.camera
    o = np.asarray(cam.origin, np.float64)
    fwd, right, up = (np.asarray(v, np.float64) for v in (cam.forward, cam.right, cam.up))
    s = (np.arange(sub) + 0.5) / sub
    px = (np.arange(w)[:, None] + s[None, :]).ravel()
    py = (np.arange(h)[:, None] + s[None, :]).ravel()
    d = fwd[None, None] + right[None, None] * (px / w * 2.0 - 1.0)[None, :, None] \
        + up[None, None] * (-(py / h * 2.0 - 1.0))[:, None, None]
    d = d / np.linalg.norm(d, axis=-1, keepdims=True)
    return o + d * (-o[1] / d[..., 1:2])


def _area_term(p, samples, cos_light, cell_area, radiance):
    """Sum over emitter cells of L cos(floor) cos(light) dA / d^2 at floor
    points p (the floor's normal is +y)."""
    total = np.zeros(p.shape[:-1] + (3,))
    for x, c, a in zip(samples, cos_light, cell_area):
        to = x - p
        d2 = np.einsum("...k,...k->...", to, to)
        l = to / np.sqrt(d2)[..., None]
        cf = np.clip(l[..., 1], 0.0, None)
        cl = np.clip(-(l @ c), 0.0, None)
        total += (cf * cl * a / d2)[..., None] * radiance
    return total


def lit_floor_area_expected(scene, lights, size, sub=4, cells=48) -> np.ndarray:
    """Each pixel's expected colour: base / pi times the point and spot
    terms (as lit_floor_expected) plus the rect and disc integrated over a
    cells x cells grid of their surfaces (the reference's own 48 x 48
    midpoint rule for the rect, gpu.rs:4590-4608; polar cells for the
    disc), averaged over sub x sub points of each pixel."""
    w, h = size
    p = _floor_points(scene, size, sub)
    base = np.asarray(LIT_FLOOR_BASE) / np.pi
    total = np.zeros(p.shape)
    for k in range(2):
        to = np.asarray(lights[k]["position"], np.float64) - p
        dist = np.linalg.norm(to, axis=-1)
        cos = np.clip(to[..., 1] / dist, 0.0, None)
        val = cos * light_falloff(lights, k, dist) * light_cone(lights, k, -to)
        total += val[..., None] * np.asarray(lights[k]["colour"], np.float64)
    # rect: area 4 |u x v|, radiance colour * intensity / area, normal u x v / |u x v|
    hu, hv = np.asarray(RECT["half_u"]), np.asarray(RECT["half_v"])
    n = np.cross(hu, hv)
    area = 4.0 * np.linalg.norm(n)
    n = n / np.linalg.norm(n)
    g = (np.arange(cells) + 0.5) / cells * 2.0 - 1.0
    pts = [np.asarray(RECT["centre"]) + hu * u + hv * v for v in g for u in g]
    total += _area_term(p, pts, [n] * len(pts), [area / cells**2] * len(pts),
                        np.asarray(RECT["colour"]) * RECT["intensity"] / area)
    # disc: polar midpoint cells
    r0, nd = DISC["radius"], np.asarray(DISC["normal"], np.float64)
    nd = nd / np.linalg.norm(nd)
    e1 = np.array([1.0, 0.0, 0.0])
    e2 = np.cross(nd, e1)
    area = np.pi * r0 * r0
    rs = (np.arange(cells) + 0.5) / cells * r0
    phis = (np.arange(2 * cells) + 0.5) / (2 * cells) * 2.0 * np.pi
    pts, cells_a = [], []
    for r in rs:
        for f in phis:
            pts.append(np.asarray(DISC["centre"]) + r * (np.cos(f) * e1 + np.sin(f) * e2))
            cells_a.append(r * (r0 / cells) * (2.0 * np.pi / (2 * cells)))
    total += _area_term(p, pts, [nd] * len(pts), cells_a, np.asarray(DISC["colour"]) * DISC["intensity"] / area)
    total = total * base
    return total.reshape(h, sub, w, sub, 3).mean(axis=(1, 3))
