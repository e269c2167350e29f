"""Procedural builds of the benchmark scenes, through the asset pipeline.

Cornell box (36 triangles)
    Room x in [-1, 1], y in [0, 2], z in [-1, 1], open towards +z, five walls
    facing inwards (white floor/ceiling/back 0.73, red (0.65, 0.05, 0.05)
    left, green (0.12, 0.45, 0.15) right); a tall 0.6 x 1.2 x 0.6 box turned
    +17 degrees and a 0.6^3 box turned -17 degrees about y; a 0.5 x 0.5
    ceiling light at y = 1.98 facing -y with radiance 15.  Materials are
    base-only: specular weight 0 (the specular lobe vanishes), specular
    roughness 1.  Same extents as the reference's scenes/cornell_box.gltf
    after its root Z-up to Y-up rotation.

Sphere field (4 096 258 triangles)
    800 icospheres (4 subdivisions: 2562 vertices, 5120 triangles), each a
    distinct mesh, centres U[0, 100]^3, uniform random rotations, scales
    log-uniform in [0.5, 4]; a 2-triangle ground quad at y = 0 over
    [-50, 150]^2; 256 emissive triangles (side U[1, 3], centres U[0, 100]^3,
    facing -y, luminance log-uniform in [1, 100]).  Sphere materials: base
    colour U[0.05, 0.95]^3, specular roughness U[0, 1], metalness
    Bernoulli(0.3), specular IOR 1.5, coat weight Bernoulli(0.2).

Emissive meshes are inserted first so every light triangle's global index is
below 65535: the reference's visibility hash is stored mod 65536
(src/asset/mod.rs:186) but checked mod 65535 (integrate.glsl:184), which
would make next-event estimation miss lights placed later in the index
buffer.  Random draws come from numpy's PCG64 generator with the given seed,
in a fixed order.
"""
from __future__ import annotations

import math
from typing import Tuple

import numpy as np

from .api import AssetScene, Camera, Material


def _quad(center, normal, u_axis, a, b):
    c, n, u = (np.asarray(x, dtype=np.float64) for x in (center, normal, u_axis))
    v = np.cross(n, u)
    p = np.array([c - a * u - b * v, c + a * u - b * v, c + a * u + b * v, c - a * u + b * v])
    return p, n, u


def _box_faces(size, rotation_y, center):
    """12 outward-facing triangles of a box (24 vertices, flat normals)."""
    half = np.asarray(size, dtype=np.float64) / 2
    cs, sn = math.cos(rotation_y), math.sin(rotation_y)
    rot = np.array([[cs, 0, sn], [0, 1, 0], [-sn, 0, cs]])
    positions, normals, tangents = [], [], []
    for axis in range(3):
        for sign in (1.0, -1.0):
            n = np.zeros(3)
            n[axis] = sign
            u = np.zeros(3)
            u[(axis + 1) % 3] = 1.0
            w = (axis + 2) % 3
            c = n * half[axis]
            a, b = half[(axis + 1) % 3], half[w]
            p, _, _ = _quad(c, n, u, a, b)
            positions.append(p @ rot.T + np.asarray(center))
            normals.append(np.tile(n @ rot.T, (4, 1)))
            tangents.append(np.tile(np.append(u @ rot.T, 1.0), (4, 1)))
    return np.concatenate(positions), np.concatenate(normals), np.concatenate(tangents)


def _quad_indices(count):
    base = np.arange(count, dtype=np.uint32) * 4
    return np.stack([base, base + 1, base + 2, base, base + 2, base + 3], axis=1).reshape(-1)


def _insert_flat(scene: AssetScene, positions, normals, tangents, indices) -> int:
    uv = np.tile(np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float32), (positions.shape[0] // 4, 1))
    return scene.insert_mesh(positions, uv, indices, normals, tangents)


def cornell_box(variant: int = 0) -> Tuple[AssetScene, Camera]:
    """The Cornell box (36 triangles, 2 emissive).  variant > 0 turns the two
    boxes by a further 11 degrees per step and swaps the wall colours on odd
    variants (distinct resident scenes for the serving mix)."""
    scene = AssetScene()
    base_only = dict(specular_weight=0.0, specular_roughness=1.0)
    light = scene.insert_material(Material(base_color=(0.73, 0.73, 0.73), emission_luminance=15.0,
                                           emission_color=(1.0, 1.0, 1.0), **base_only))
    white = scene.insert_material(Material(base_color=(0.73, 0.73, 0.73), **base_only))
    red = scene.insert_material(Material(base_color=(0.65, 0.05, 0.05), **base_only))
    green = scene.insert_material(Material(base_color=(0.12, 0.45, 0.15), **base_only))
    if variant % 2 == 1:
        red, green = green, red
    turn = math.radians(11.0 * variant)

    def wall(center, normal, u, a, b, material):
        p, n, t = _quad(center, normal, u, a, b)
        mesh = _insert_flat(scene, p, np.tile(n, (4, 1)), np.tile(np.append(t, 1.0), (4, 1)), _quad_indices(1))
        scene.add_model(scene.add_node(), mesh, material)

    # The light first: its triangles get global indices 0 and 1.
    wall((0, 1.98, 0), (0, -1, 0), (1, 0, 0), 0.25, 0.25, light)
    wall((0, 0, 0), (0, 1, 0), (1, 0, 0), 1, 1, white)      # floor
    wall((0, 2, 0), (0, -1, 0), (1, 0, 0), 1, 1, white)     # ceiling
    wall((0, 1, -1), (0, 0, 1), (1, 0, 0), 1, 1, white)     # back
    wall((-1, 1, 0), (1, 0, 0), (0, 0, 1), 1, 1, red)       # left
    wall((1, 1, 0), (-1, 0, 0), (0, 0, 1), 1, 1, green)     # right
    for size, angle, center in (((0.6, 1.2, 0.6), math.radians(17.0) + turn, (-0.33, 0.6, -0.3)),
                                ((0.6, 0.6, 0.6), math.radians(-17.0) - turn, (0.33, 0.3, 0.3))):
        p, n, t = _box_faces(size, angle, center)
        mesh = _insert_flat(scene, p, n, t, _quad_indices(6))
        scene.add_model(scene.add_node(), mesh, white)
    camera = Camera(position=(0.0, 1.0, 3.4), yaw=math.pi / 2, pitch=0.0, fov=0.7)
    return scene, camera


def textured_cornell_box(size: int = 64, block_compression: bool = True) -> Tuple[AssetScene, Camera]:
    """The Cornell box with textured floor and back wall: a size x size
    checker base colour, roughness, IOR (a float channel), metalness stripes
    and a bumpy normal map on the floor; a colour gradient on the back wall.
    Each face maps [0, 1]^2, so the floor repeats its texture through the
    sampler's wrap when scaled (uv of the floor quad x 3).  Colour and normal
    textures are block-compressed unless block_compression is False."""
    scene = AssetScene(block_compression=block_compression)
    base_only = dict(specular_weight=0.0, specular_roughness=1.0)
    n = size
    y, x = np.mgrid[0:n, 0:n].astype(np.float32) / n
    checker = ((np.floor(x * 8) + np.floor(y * 8)) % 2)[..., None]
    floor_color = (0.15 + 0.6 * checker * np.array([0.9, 0.8, 0.6], np.float32) +
                   0.1 * (1 - checker) * np.array([0.2, 0.5, 0.9], np.float32)).astype(np.float32)
    roughness = (0.2 + 0.6 * np.abs(np.sin(6.0 * x) * np.cos(5.0 * y))).astype(np.float32)
    ior = (1.3 + 0.6 * y).astype(np.float32)
    metal = (np.floor(x * 4) % 2 * (y > 0.5)).astype(np.float32)
    bump = np.stack([0.25 * np.sin(12.0 * x), 0.25 * np.cos(10.0 * y), np.ones_like(x)], axis=-1)
    bump = (bump / np.linalg.norm(bump, axis=-1, keepdims=True)).astype(np.float32)
    gradient = np.stack([0.2 + 0.6 * x, 0.3 + 0.4 * y, 0.7 - 0.5 * x], axis=-1).astype(np.float32)

    light = scene.insert_material(Material(base_color=(0.73, 0.73, 0.73), emission_luminance=15.0,
                                           emission_color=(1.0, 1.0, 1.0), **base_only))
    white = scene.insert_material(Material(base_color=(0.73, 0.73, 0.73), **base_only))
    red = scene.insert_material(Material(base_color=(0.65, 0.05, 0.05), **base_only))
    green = scene.insert_material(Material(base_color=(0.12, 0.45, 0.15), **base_only))
    floor = scene.insert_material(Material(specular_weight=1.0),
                                  {"base_color": floor_color, "specular_roughness": roughness, "specular_ior": ior,
                                   "base_metalness": metal, "geometry_normal": bump})
    back = scene.insert_material(Material(**base_only), {"base_color": gradient})

    def wall(center, normal, u, a, b, material, uv_scale=1.0):
        p, nrm, t = _quad(center, normal, u, a, b)
        uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float32) * uv_scale
        mesh = scene.insert_mesh(p, uv, _quad_indices(1), np.tile(nrm, (4, 1)), np.tile(np.append(t, 1.0), (4, 1)))
        scene.add_model(scene.add_node(), mesh, material)

    wall((0, 1.98, 0), (0, -1, 0), (1, 0, 0), 0.25, 0.25, light)
    wall((0, 0, 0), (0, 1, 0), (1, 0, 0), 1, 1, floor, uv_scale=3.0)
    wall((0, 2, 0), (0, -1, 0), (1, 0, 0), 1, 1, white)
    wall((0, 1, -1), (0, 0, 1), (1, 0, 0), 1, 1, back)
    wall((-1, 1, 0), (1, 0, 0), (0, 0, 1), 1, 1, red)
    wall((1, 1, 0), (-1, 0, 0), (0, 0, 1), 1, 1, green)
    for size, angle, center in (((0.6, 1.2, 0.6), math.radians(17.0), (-0.33, 0.6, -0.3)),
                                ((0.6, 0.6, 0.6), math.radians(-17.0), (0.33, 0.3, 0.3))):
        p, nrm, t = _box_faces(size, angle, center)
        mesh = _insert_flat(scene, p, nrm, t, _quad_indices(6))
        scene.add_model(scene.add_node(), mesh, white)
    camera = Camera(position=(0.0, 1.0, 3.4), yaw=math.pi / 2, pitch=0.0, fov=0.7)
    return scene, camera


def icosphere(subdivisions: int = 4):
    """Unit icosphere: (positions, triangle indices), CCW seen from outside."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        new_faces = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            new_faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new_faces
    return np.array(verts), np.array(faces, dtype=np.uint32)


def _sphere_attributes(positions):
    n = positions / np.linalg.norm(positions, axis=1, keepdims=True)
    up = np.array([0.0, 1.0, 0.0])
    t = np.cross(up, n)
    length = np.linalg.norm(t, axis=1, keepdims=True)
    t = np.where(length > 1e-6, t / np.maximum(length, 1e-12), np.array([1.0, 0.0, 0.0]))
    tangents = np.concatenate([t, np.ones((t.shape[0], 1))], axis=1)
    uv = np.stack([np.arctan2(n[:, 2], n[:, 0]) / (2 * math.pi) + 0.5, np.arccos(np.clip(n[:, 1], -1, 1)) / math.pi],
                  axis=1)
    return n, tangents, uv


def _random_rotation(rng):
    """Uniform rotation matrix (Shoemake 1992, from a uniform unit quaternion)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def sphere_field(seed: int = 0, spheres: int = 800, lights: int = 256, subdivisions: int = 4
                 ) -> Tuple[AssetScene, Camera]:
    rng = np.random.default_rng(seed)
    scene = AssetScene()
    # Emissive triangles first (global triangle indices 0 .. lights-1).
    for _ in range(lights):
        side = rng.uniform(1.0, 3.0)
        center = rng.uniform(0.0, 100.0, 3)
        theta0 = rng.uniform(0.0, 2 * math.pi)
        luminance = math.exp(rng.uniform(math.log(1.0), math.log(100.0)))
        r = side / math.sqrt(3.0)
        p = np.array([center + r * np.array([math.cos(theta0 + 2 * math.pi * k / 3), 0.0,
                                             math.sin(theta0 + 2 * math.pi * k / 3)]) for k in range(3)])
        normals = np.tile([0.0, -1.0, 0.0], (3, 1))
        tangents = np.tile([1.0, 0.0, 0.0, 1.0], (3, 1))
        uv = np.array([[0, 0], [1, 0], [0, 1]], dtype=np.float32)
        mesh = scene.insert_mesh(p, uv, np.array([0, 1, 2], dtype=np.uint32), normals, tangents)
        material = scene.insert_material(Material(base_color=(0.8, 0.8, 0.8), emission_luminance=luminance,
                                                  specular_roughness=0.5))
        scene.add_model(scene.add_node(), mesh, material)
    # Ground.
    p, n, t = _quad((50.0, 0.0, 50.0), (0, 1, 0), (1, 0, 0), 100.0, 100.0)
    ground = _insert_flat(scene, p, np.tile(n, (4, 1)), np.tile(np.append(t, 1.0), (4, 1)), _quad_indices(1))
    ground_material = scene.insert_material(Material(base_color=(0.5, 0.5, 0.5), specular_roughness=0.5))
    scene.add_model(scene.add_node(), ground, ground_material)
    # Spheres: distinct meshes of the same unit geometry, placed by nodes.
    unit, faces = icosphere(subdivisions)
    normals, tangents, uv = _sphere_attributes(unit)
    indices = faces.reshape(-1)
    for _ in range(spheres):
        center = rng.uniform(0.0, 100.0, 3)
        rotation = _random_rotation(rng)
        scale = math.exp(rng.uniform(math.log(0.5), math.log(4.0)))
        base_color = rng.uniform(0.05, 0.95, 3)
        roughness = rng.uniform(0.0, 1.0)
        metalness = 1.0 if rng.random() < 0.3 else 0.0
        coat = 1.0 if rng.random() < 0.2 else 0.0
        mesh = scene.insert_mesh(unit, uv, indices, normals, tangents)
        material = scene.insert_material(Material(base_color=tuple(base_color), specular_roughness=roughness,
                                                  base_metalness=metalness, specular_ior=1.5, coat_weight=coat))
        transform = np.eye(4)
        transform[:3, :3] = rotation * scale
        transform[:3, 3] = center
        scene.add_model(scene.add_node(transform), mesh, material)
    camera = Camera.looking_at((-25.0, 60.0, -25.0), (55.0, 5.0, 55.0), fov=math.pi / 3)
    return scene, camera


def torus(major: float, minor: float, nu: int, nv: int):
    """Torus around y: (positions, normals, tangents, uv, indices), CCW seen
    from outside, 2 nu nv triangles."""
    u = np.arange(nu) * (2 * math.pi / nu)
    v = np.arange(nv) * (2 * math.pi / nv)
    uu, vv = np.meshgrid(u, v, indexing="ij")
    cu, su, cv, sv = np.cos(uu), np.sin(uu), np.cos(vv), np.sin(vv)
    ring = major + minor * cv
    pos = np.stack([ring * cu, minor * sv, ring * su], axis=-1).reshape(-1, 3)
    nrm = np.stack([cv * cu, sv, cv * su], axis=-1).reshape(-1, 3)
    tan = np.stack([-su, np.zeros_like(su), cu], axis=-1).reshape(-1, 3)
    tan = np.concatenate([tan, np.ones((tan.shape[0], 1))], axis=1)
    uv = np.stack([uu / (2 * math.pi), vv / (2 * math.pi)], axis=-1).reshape(-1, 2)
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    idx = np.stack([a, d, c, a, c, b], axis=-1).reshape(-1).astype(np.uint32)
    return pos, nrm, tan, uv, idx


_ICOSPHERES = {}


def displaced_icosphere(subdivisions: int, rng: np.random.Generator, amplitude: float = 0.15):
    """Icosphere with smooth radial displacement 1 + amplitude * sum of three
    random sinusoids; normals from the displaced faces."""
    if subdivisions not in _ICOSPHERES:
        _ICOSPHERES[subdivisions] = icosphere(subdivisions)
    unit, faces = _ICOSPHERES[subdivisions]
    k = rng.normal(size=(3, 3)) * 3.0
    phase = rng.uniform(0, 2 * math.pi, 3)
    r = 1.0 + amplitude / 3.0 * np.sin(unit @ k.T + phase).sum(axis=1)
    pos = unit * r[:, None]
    fn = np.cross(pos[faces[:, 1]] - pos[faces[:, 0]], pos[faces[:, 2]] - pos[faces[:, 0]])
    nrm = np.zeros_like(pos)
    for c in range(3):
        np.add.at(nrm, faces[:, c], fn)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    _, tangents, uv = _sphere_attributes(pos)
    t = tangents[:, :3] - nrm * (tangents[:, :3] * nrm).sum(axis=1, keepdims=True)
    bad = np.linalg.norm(t, axis=1) < 1e-6
    t[bad] = np.cross(nrm[bad], [0.0, 0.0, 1.0])
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    tan = np.concatenate([t, np.ones((t.shape[0], 1))], axis=1)
    return pos, nrm, tan, uv, faces.reshape(-1)


def instanced_field(seed: int = 0, meshes: int = 64, instances: int = 12000, lights: int = 16,
                    min_triangles: int = 5000, max_triangles: int = 200000, extent: float = 200.0
                    ) -> Tuple[AssetScene, Camera]:
    """The instanced benchmark scene: `meshes` distinct meshes (even: displaced
    icospheres with 20 * 4^s triangles, s in 4..6; odd: tori with ~T
    triangles) whose target sizes T are log-uniform in [min, max],
    `instances` instances of them (mesh uniform, uniform random rotation,
    scale log-uniform in [0.5, 3], position U[0, extent]^3), a ground quad,
    and `lights` instances of one 256-triangle emissive torus (4096 emissive
    triangles for 16 lights, luminance log-uniform in [2, 50]).  Seed 0:
    3 215 512 distinct triangles, 596 951 286 instanced (12 017 instances);
    meshes average ~50k triangles, so 12 000 uniformly chosen instances carry
    far more than 64M triangles."""
    rng = np.random.default_rng(seed)
    scene = AssetScene()
    # Lights first (global triangle indices 0 .. 255 for the emissive mesh).
    lp, ln, lt, luv, li = torus(1.0, 0.3, 16, 8)
    light_mesh = scene.insert_mesh(lp, luv, li, ln, lt)
    for _ in range(lights):
        lum = math.exp(rng.uniform(math.log(2.0), math.log(50.0)))
        mat = scene.insert_material(Material(base_color=(0.8, 0.8, 0.8), emission_luminance=lum,
                                             specular_roughness=0.5))
        transform = np.eye(4)
        transform[:3, :3] = _random_rotation(rng) * math.exp(rng.uniform(math.log(1.0), math.log(4.0)))
        transform[:3, 3] = rng.uniform(0.0, extent, 3)
        scene.add_model(scene.add_node(transform), light_mesh, mat)
    p, n, t = _quad((extent / 2, 0.0, extent / 2), (0, 1, 0), (1, 0, 0), extent, extent)
    ground = _insert_flat(scene, p, np.tile(n, (4, 1)), np.tile(np.append(t, 1.0), (4, 1)), _quad_indices(1))
    scene.add_model(scene.add_node(), ground, scene.insert_material(Material(base_color=(0.5, 0.5, 0.5),
                                                                              specular_roughness=0.5)))
    mesh_ids, mesh_mats, mesh_tris = [], [], []
    for k in range(meshes):
        target = math.exp(rng.uniform(math.log(min_triangles), math.log(max_triangles)))
        if k % 2 == 0:
            s = int(min(6, max(4, round(math.log(target / 20.0, 4)))))
            pos, nrm, tan, uv, idx = displaced_icosphere(s, rng)
        else:
            nu = max(8, int(round(math.sqrt(target))))
            nv = max(4, int(round(target / (2 * nu))))
            pos, nrm, tan, uv, idx = torus(1.0, 0.35, nu, nv)
        mesh_ids.append(scene.insert_mesh(pos, uv, idx, nrm, tan))
        mesh_tris.append(idx.size // 3)
        mesh_mats.append(scene.insert_material(Material(
            base_color=tuple(rng.uniform(0.05, 0.95, 3)), specular_roughness=float(rng.uniform(0.0, 1.0)),
            base_metalness=1.0 if rng.random() < 0.3 else 0.0, specular_ior=1.5,
            coat_weight=1.0 if rng.random() < 0.2 else 0.0)))
    for _ in range(instances):
        m = int(rng.integers(meshes))
        transform = np.eye(4)
        transform[:3, :3] = _random_rotation(rng) * math.exp(rng.uniform(math.log(0.5), math.log(3.0)))
        transform[:3, 3] = rng.uniform(0.0, extent, 3)
        scene.add_model(scene.add_node(transform), mesh_ids[m], mesh_mats[m])
    camera = Camera.looking_at((-0.25 * extent, 0.6 * extent, -0.25 * extent), (0.55 * extent, 0.05 * extent,
                                                                                  0.55 * extent), fov=math.pi / 3)
    return scene, camera
