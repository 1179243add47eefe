"""Seeded synthetic scenes for the still workloads.

box_spheres : a closed box of 6 quads with an emissive ceiling quad and 64
              subdivision-2 icospheres (~20.5k triangles), Lambertian.
terrain     : a 1024x512-quad fBm heightfield (1,048,576 triangles), 256
              instances of 8 random meshes (2k-20k triangles each), 128
              point and spot lights sampled through the light tree, diffuse
              and GGX materials.
Both take their sizes as arguments so tests can build small versions with
the same structure. Everything is drawn from numpy's PCG64 with the seed.
"""
from __future__ import annotations

import math

import numpy as np

from . import _native as N
from .scene import Camera, Scene, material, point_light, spot_light


# ---- geometry helpers ----------------------------------------------------------
def icosphere(subdivisions: int) -> np.ndarray:
    """Unit icosphere triangles, (20 * 4**s, 3, 3) float64."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array(
        [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array(
        [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    tris = v[f]
    for _ in range(subdivisions):
        a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
        ab, bc, ca = a + b, b + c, c + a
        ab /= np.linalg.norm(ab, axis=1, keepdims=True)
        bc /= np.linalg.norm(bc, axis=1, keepdims=True)
        ca /= np.linalg.norm(ca, axis=1, keepdims=True)
        tris = np.concatenate(
            [np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1), np.stack([ca, bc, c], 1), np.stack([ab, bc, ca], 1)]
        )
    return tris


def quad(p0, p1, p2, p3) -> np.ndarray:
    """Two triangles (p0, p1, p2), (p0, p2, p3)."""
    p = np.array([p0, p1, p2, p3], np.float64)
    return np.stack([p[[0, 1, 2]], p[[0, 2, 3]]])


def _triangles(verts: np.ndarray, materials: np.ndarray) -> np.ndarray:
    out = np.zeros(len(verts), dtype=N.TRIANGLE)
    out["vertices"] = verts.reshape(len(verts), 9).astype(np.float32)
    out["material"] = materials.astype(np.uint32)
    return out


def _value_noise(rng: np.random.Generator, x: np.ndarray, z: np.ndarray, cells: int) -> np.ndarray:
    """2-D value noise with smoothstep interpolation on a `cells` periodic lattice."""
    lattice = rng.random((cells, cells))
    xi = np.floor(x).astype(np.int64)
    zi = np.floor(z).astype(np.int64)
    fx, fz = x - xi, z - zi
    ux, uz = fx * fx * (3 - 2 * fx), fz * fz * (3 - 2 * fz)
    x0, x1, z0, z1 = xi % cells, (xi + 1) % cells, zi % cells, (zi + 1) % cells
    a, b, c, d = lattice[x0, z0], lattice[x1, z0], lattice[x0, z1], lattice[x1, z1]
    return (a * (1 - ux) + b * ux) * (1 - uz) + (c * (1 - ux) + d * ux) * uz


def fbm_heights(seed: int, nx: int, nz: int, size_x: float, size_z: float, amplitude: float,
                octaves: int = 5, base_frequency: float = 1.0 / 16.0, lacunarity: float = 2.0,
                gain: float = 0.5) -> np.ndarray:
    """(nz+1, nx+1) vertex heights: fBm of value noise (5 octaves, lacunarity 2, gain 0.5)."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(-size_x / 2, size_x / 2, nx + 1)
    zs = np.linspace(-size_z / 2, size_z / 2, nz + 1)
    X, Z = np.meshgrid(xs, zs)
    h = np.zeros_like(X)
    freq, amp = base_frequency, 1.0
    for _ in range(octaves):
        h += amp * (_value_noise(rng, X * freq + 1000.0, Z * freq + 1000.0, 256) - 0.5)
        freq *= lacunarity
        amp *= gain
    return h * amplitude


def heightfield_triangles(h: np.ndarray, size_x: float, size_z: float) -> np.ndarray:
    """Two triangles per grid quad: (nz*nx*2, 3, 3)."""
    nz, nx = h.shape[0] - 1, h.shape[1] - 1
    xs = np.linspace(-size_x / 2, size_x / 2, nx + 1)
    zs = np.linspace(-size_z / 2, size_z / 2, nz + 1)
    X, Z = np.meshgrid(xs, zs)
    P = np.stack([X, h, Z], -1)
    p00, p10, p01, p11 = P[:-1, :-1], P[:-1, 1:], P[1:, :-1], P[1:, 1:]
    t0 = np.stack([p00, p01, p11], -2)
    t1 = np.stack([p00, p11, p10], -2)
    return np.stack([t0, t1], 2).reshape(-1, 3, 3)


def rock_mesh(rng: np.random.Generator, n: int) -> tuple[np.ndarray, np.ndarray]:
    """A displaced cube-sphere: 12 n^2 triangles with smooth vertex normals."""
    u = np.linspace(-1.0, 1.0, n + 1)
    U, V = np.meshgrid(u, u)
    faces = []
    for axis in range(3):
        for sign in (-1.0, 1.0):
            p = np.zeros(U.shape + (3,))
            p[..., axis] = sign
            p[..., (axis + 1) % 3] = U * sign
            p[..., (axis + 2) % 3] = V
            faces.append(p)
    tris = []
    for p in faces:
        p = p / np.linalg.norm(p, axis=-1, keepdims=True)
        p00, p10, p01, p11 = p[:-1, :-1], p[:-1, 1:], p[1:, :-1], p[1:, 1:]
        tris.append(np.stack([np.stack([p00, p10, p11], -2), np.stack([p00, p11, p01], -2)], 2).reshape(-1, 3, 3))
    unit = np.concatenate(tris)
    # smooth radial displacement from a few random lobes
    dirs = rng.normal(size=(6, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    weights = rng.uniform(0.05, 0.25, 6)
    radius = 1.0 + np.einsum("tvk,lk->tvl", unit, dirs).clip(0) ** 2 @ weights
    verts = unit * radius[..., None]
    # vertex normals: the unit direction perturbed by the displacement gradient is
    # approximated by averaging face normals over shared positions
    fn = np.cross(verts[:, 1] - verts[:, 0], verts[:, 2] - verts[:, 0])
    key = np.round(unit.reshape(-1, 3) * 1e5).astype(np.int64)
    _, inv = np.unique(key, axis=0, return_inverse=True)
    acc = np.zeros((inv.max() + 1, 3))
    np.add.at(acc, inv.reshape(-1), np.repeat(fn, 3, axis=0))
    acc /= np.linalg.norm(acc, axis=1, keepdims=True)
    normals = acc[inv.reshape(-1)].reshape(verts.shape)
    return verts, normals


def transform(position, yaw: float, scale: float) -> np.ndarray:
    """Column-major 4x4: translate * rotate_y(yaw) * scale."""
    c, s = math.cos(yaw), math.sin(yaw)
    m = np.zeros((4, 4), np.float64)  # m[column][row]
    m[0, 0], m[0, 2] = c * scale, -s * scale
    m[1, 1] = scale
    m[2, 0], m[2, 2] = s * scale, c * scale
    m[3, :3] = position
    m[3, 3] = 1.0
    return m.astype(np.float32).reshape(16)


# ---- workloads -------------------------------------------------------------------
def box_spheres(seed: int = 1, spheres: int = 64, subdivisions: int = 2, width: int = 512, height: int = 512) -> Scene:
    """Closed unit box, emissive ceiling quad (radiance 10), `spheres`
    icospheres with centres U(box), radii U[0.05, 0.2], albedo U[0.2, 0.8]^3."""
    rng = np.random.default_rng(seed)
    walls = np.concatenate(
        [
            quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)),
            quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)),
            quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0)),
            quad((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)),
            quad((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0)),
            quad((0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)),
        ]
    )
    light = quad((0.35, 0.995, 0.35), (0.65, 0.995, 0.35), (0.65, 0.995, 0.65), (0.35, 0.995, 0.65))
    unit = icosphere(subdivisions)
    centres = rng.random((spheres, 3))
    radii = rng.uniform(0.05, 0.2, spheres)
    albedo = rng.uniform(0.2, 0.8, (spheres, 3))
    balls = (unit[None] * radii[:, None, None, None] + centres[:, None, None, :]).reshape(-1, 3, 3)
    mats = [material((0.75, 0.75, 0.75), roughness=1.0, specular=0.0),
            material((0.0, 0.0, 0.0), roughness=1.0, specular=0.0, emission=(10.0, 10.0, 10.0))]
    mats += [material(tuple(a), roughness=1.0, specular=0.0) for a in albedo]
    ids = np.concatenate(
        [np.zeros(len(walls)), np.ones(len(light)), 2 + np.repeat(np.arange(spheres), len(unit))]
    )
    verts = np.concatenate([walls, light, balls])
    # the eye sits just inside the front wall, at the first point of a fixed
    # search pattern that no sphere's circumscribed ball contains
    eye = (0.5, 0.5, 0.001)
    for r in np.linspace(0.0, 0.45, 10):
        for a in np.linspace(0.0, 2.0 * math.pi, 12, endpoint=False):
            cand = np.array([0.5 + r * math.cos(a), 0.5 + r * math.sin(a), 0.001])
            if np.all(np.linalg.norm(centres - cand, axis=1) > radii * 1.001):
                eye = tuple(cand)
                break
        else:
            continue
        break
    camera = Camera.look(eye, (eye[0], eye[1], 1.0), 70.0, width / height)
    return Scene(triangles=_triangles(verts, ids), materials=np.concatenate(mats), camera=camera)


def terrain(seed: int = 2, grid=(1024, 512), size=(128.0, 64.0), amplitude: float = 10.0, meshes: int = 8,
            mesh_tris=(2000, 20000), instances: int = 256, lights: int = 128, width: int = 3840,
            height: int = 2160, light_range: float = 40.0) -> Scene:
    """fBm heightfield, instanced random meshes, point and spot lights."""
    rng = np.random.default_rng(seed)
    nx, nz = grid
    sx, sz = size
    h = fbm_heights(seed * 7919 + 1, nx, nz, sx, sz, amplitude)
    ground = heightfield_triangles(h, sx, sz)
    # materials: 8 terrain patches + one per mesh; even diffuse, odd GGX
    mats = []
    for k in range(8 + meshes):
        r = float(rng.uniform(0.05, 1.0))
        if k % 2 == 0:
            mats.append(material(tuple(rng.uniform(0.2, 0.8, 3)), roughness=r, specular=0.04))
        else:
            mats.append(material(tuple(rng.uniform(0.5, 0.95, 3)), roughness=r, metalness=1.0, specular=0.04))
    qz, qx = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    patch = ((qx * 4 // max(nx, 1)) + 4 * ((qz * 2) // max(nz, 1))) % 8
    tri_mat = np.repeat(patch.reshape(-1), 2)
    scene = Scene(triangles=_triangles(ground, tri_mat), materials=np.concatenate(mats))
    # meshes
    lo, hi = mesh_tris
    mesh_rows, mesh_ranges, first = [], [], 0
    for m in range(meshes):
        target = int(rng.integers(lo, hi + 1))
        n = max(2, int(round((target / 12.0) ** 0.5)))
        v, nrm = rock_mesh(rng, n)
        rows = np.zeros(len(v), dtype=N.TRIANGLE_DETAIL)
        rows["vertices"] = v.reshape(len(v), 9).astype(np.float32)
        rows["normals"] = nrm.reshape(len(v), 9).astype(np.float32)
        rows["material"] = 8 + m
        mesh_rows.append(rows)
        mesh_ranges.append((first, len(v)))
        first += len(v)
    scene.mesh_triangles = np.concatenate(mesh_rows) if mesh_rows else scene.mesh_triangles
    scene.meshes = np.array(mesh_ranges, dtype=N.MESH) if mesh_ranges else scene.meshes

    def ground_height(x, z):
        fx = (x + sx / 2) / sx * nx
        fz = (z + sz / 2) / sz * nz
        ix = np.clip(np.floor(fx).astype(int), 0, nx - 1)
        iz = np.clip(np.floor(fz).astype(int), 0, nz - 1)
        tx, tz = fx - ix, fz - iz
        return ((h[iz, ix] * (1 - tx) + h[iz, ix + 1] * tx) * (1 - tz)
                + (h[iz + 1, ix] * (1 - tx) + h[iz + 1, ix + 1] * tx) * tz)

    inst = np.zeros(instances if meshes else 0, dtype=N.INSTANCE)
    for i in range(len(inst)):
        x = float(rng.uniform(-0.47, 0.47) * sx)
        z = float(rng.uniform(-0.47, 0.47) * sz)
        scale = float(rng.uniform(0.5, 3.0))
        yaw = float(rng.uniform(0.0, 2.0 * math.pi))
        y = float(ground_height(x, z)) + 0.4 * scale
        inst[i]["mesh"] = int(rng.integers(0, meshes))
        inst[i]["word"] = 0
        inst[i]["transform"] = transform((x, y, z), yaw, scale)
    scene.instances = inst
    # lights: alternate point / spot, log-uniform intensity in [1, 1000]
    rows = []
    for k in range(lights):
        x = float(rng.uniform(-0.5, 0.5) * sx)
        z = float(rng.uniform(-0.5, 0.5) * sz)
        y = float(ground_height(x, z)) + float(rng.uniform(0.5, 10.0))
        intensity = float(math.exp(rng.uniform(math.log(1.0), math.log(1000.0))))
        colour = tuple(rng.uniform(0.6, 1.0, 3))
        radius = float(rng.uniform(0.0, 0.25))
        if k % 2 == 0:
            rows.append(point_light((x, y, z), colour, intensity, light_range, radius=radius))
        else:
            d = np.array([rng.uniform(-0.5, 0.5), -1.0, rng.uniform(-0.5, 0.5)])
            outer = float(rng.uniform(15.0, 60.0))
            rows.append(spot_light((x, y, z), colour, intensity, light_range, tuple(d / np.linalg.norm(d)),
                                   0.8 * outer, outer, radius=radius))
    scene.lights = np.concatenate(rows) if rows else scene.lights
    top = float(h.max())
    scene.camera = Camera.look((0.0, top + 0.3 * sz * 0.4 + 6.0, -0.72 * sz), (0.0, 0.0, 0.1 * sz), 55.0,
                               width / height)
    return scene


def sphere_terrain(seed: int, tris: int = 100_000, lights: int = 8, width: int = 512, height: int = 512,
                   size: float = 64.0) -> Scene:
    """A seeded sphere-and-terrain mix of about `tris` triangles: an fBm
    heightfield holding about 60% of them (world triangles), the rest as
    instances of a subdivision-3 icosphere (1,280 triangles each), diffuse
    and GGX materials, `lights` point and spot lights (log-uniform
    intensities over [1, 1000]), a camera looking down across the terrain."""
    rng = np.random.default_rng(seed)
    ground_tris = max(2, int(tris * 0.6))
    n = max(1, int(round((ground_tris / 2) ** 0.5)))
    h = fbm_heights(seed * 7919 + 3, n, n, size, size, size * 0.1)
    ground = heightfield_triangles(h, size, size)
    mats = []
    for k in range(6):
        r = float(rng.uniform(0.05, 1.0))
        if k % 2 == 0:
            mats.append(material(tuple(rng.uniform(0.2, 0.8, 3)), roughness=r, specular=0.04))
        else:
            mats.append(material(tuple(rng.uniform(0.5, 0.95, 3)), roughness=r, metalness=1.0, specular=0.04))
    qz, qx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    tri_mat = np.repeat(((qx * 3 // n) + 3 * ((qz * 2) // n)) % 4, 2)
    scene = Scene(triangles=_triangles(ground, tri_mat), materials=np.concatenate(mats))
    unit = icosphere(3)
    rows = np.zeros(len(unit), dtype=N.TRIANGLE_DETAIL)
    rows["vertices"] = unit.reshape(len(unit), 9).astype(np.float32)
    rows["normals"] = unit.reshape(len(unit), 9).astype(np.float32)  # unit sphere: normal = position
    rows["material"] = 4
    scene.mesh_triangles = rows
    scene.meshes = np.array([(0, len(unit))], dtype=N.MESH)
    spheres = max(0, (tris - len(ground)) // len(unit))
    inst = np.zeros(spheres, dtype=N.INSTANCE)
    top = float(h.max())
    for i in range(spheres):
        x, z = rng.uniform(-0.45, 0.45, 2) * size
        ix = int(np.clip((x / size + 0.5) * n, 0, n))
        iz = int(np.clip((z / size + 0.5) * n, 0, n))
        radius = float(rng.uniform(0.3, 2.0))
        inst[i]["mesh"] = 0
        inst[i]["word"] = 0
        inst[i]["transform"] = transform((x, float(h[iz, ix]) + radius, z), 0.0, radius)
    scene.instances = inst
    light_rows = []
    for k in range(lights):
        x, z = rng.uniform(-0.5, 0.5, 2) * size
        y = top + float(rng.uniform(0.5, 8.0))
        intensity = float(math.exp(rng.uniform(0.0, math.log(1000.0))))
        colour = tuple(rng.uniform(0.6, 1.0, 3))
        if k % 2 == 0:
            light_rows.append(point_light((x, y, z), colour, intensity, size * 0.6, radius=0.1))
        else:
            d = np.array([rng.uniform(-0.5, 0.5), -1.0, rng.uniform(-0.5, 0.5)])
            outer = float(rng.uniform(15.0, 60.0))
            light_rows.append(spot_light((x, y, z), colour, intensity, size * 0.6, tuple(d / np.linalg.norm(d)),
                                         0.8 * outer, outer, radius=0.1))
    scene.lights = np.concatenate(light_rows) if light_rows else scene.lights
    yaw = float(rng.uniform(0, 2 * math.pi))
    eye = (0.55 * size * math.cos(yaw), top + size * 0.15, 0.55 * size * math.sin(yaw))
    scene.camera = Camera.look(eye, (0.0, 0.0, 0.0), 55.0, width / height)
    return scene


def batch_mix(seed: int = 3, stills: int = 32, sizes=(256, 512, 1024, 2048), tris=(10_000, 500_000),
              lights=(1, 64), spps=(4, 16, 64)) -> list[tuple[Scene, int, int, int, int]]:
    """Uneven stills for a batch: per still a square size drawn from `sizes`,
    a sphere-and-terrain scene of log-uniform triangle count in `tris` with
    a light count drawn from `lights`, spp drawn from `spps` and a seed.
    Returns (scene, width, height, spp, seed) tuples."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(stills):
        side = int(rng.choice(sizes))
        t = int(math.exp(rng.uniform(math.log(tris[0]), math.log(tris[1]))))
        nl = int(rng.integers(lights[0], lights[1] + 1))
        spp = int(rng.choice(spps))
        scene = sphere_terrain(seed * 1000 + k, tris=t, lights=nl, width=side, height=side)
        out.append((scene, side, side, spp, int(rng.integers(1, 2**31))))
    return out
