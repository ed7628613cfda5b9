"""Seeded synthetic workloads (triangle soups and ray batches).

Every vertex is rounded to FP32 and stored as FP64, so the FP64 reference
arithmetic and the FP32 device arrays see the same geometry.  Generators use
numpy's PCG64 ``default_rng(seed)``, so a seed fixes the workload on any host.
"""
from __future__ import annotations

import math

import numpy as np

T_MIN_CAMERA = 0.001  # render.go:108 (camera rays)
T_MIN_SCATTER = 0.0001  # render.go:216 (scattered rays)


def _unit_vectors(rng: np.random.Generator, shape) -> np.ndarray:
    v = rng.standard_normal((*shape, 3))
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    return v


def random_triangles(n: int, seed: int = 1, mean_edge: float = 0.005) -> np.ndarray:
    """Triangles with centroids ~U[0,1]^3 and vertices at centroid + u*l, u
    uniform on the sphere and l ~ Exp(mean_edge).  Returns (n, 9) float64."""
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3))
    u = _unit_vectors(rng, (n, 3))
    length = rng.exponential(mean_edge, (n, 3, 1))
    v = c[:, None, :] + u * length
    return v.astype(np.float32).astype(np.float64).reshape(n, 9)


def incoherent_rays(n: int, seed: int = 2, tmin: float = T_MIN_CAMERA, tmax: float = math.inf) -> np.ndarray:
    """Origins ~U[0,1]^3, directions uniform on the sphere.  (n, 8) float32:
    origin.xyz, tmin, direction.xyz, tmax."""
    rng = np.random.default_rng(seed)
    r = np.empty((n, 8), np.float32)
    r[:, 0:3] = rng.random((n, 3))
    r[:, 3] = tmin
    r[:, 4:7] = _unit_vectors(rng, (n,))
    r[:, 7] = tmax
    return r


class Camera:
    """Pinhole camera of camera.go:53-97 in FP64 (vertical field of view in degrees)."""

    def __init__(self, aspect: float, fov: float, look_from, look_at, up):
        self.viewport_height = 2.0 * math.tan(math.radians(fov) / 2)
        self.viewport_width = aspect * self.viewport_height
        lf = np.asarray(look_from, np.float64)
        la = np.asarray(look_at, np.float64)
        upv = np.asarray(up, np.float64)
        w = lf - la
        w = w / np.linalg.norm(w)
        u = np.cross(upv, w)
        u = u / np.linalg.norm(u)
        v = np.cross(w, u)
        self.origin = lf
        self.horizontal = u * self.viewport_width
        self.vertical = v * self.viewport_height
        self.lower_left = lf - self.horizontal * 0.5 - self.vertical * 0.5 - w

    def directions(self, s: np.ndarray, t: np.ndarray) -> np.ndarray:
        """castRayReuse (camera.go:95-97): LL + s*H + t*V - origin."""
        return (self.lower_left[None, :] + s[:, None] * self.horizontal[None, :]
                + t[:, None] * self.vertical[None, :] - self.origin[None, :])


def camera_rays(cam: Camera, width: int, height: int, tmin: float = T_MIN_CAMERA) -> np.ndarray:
    """One ray through each pixel centre; s = (x + 0.5)/(W - 1), t = (y + 0.5)/(H - 1)
    (the reference divides by W - 1 and H - 1, render.go:201-205).  Row-major,
    y = 0 at the bottom.  (W*H, 8) float32."""
    y, x = np.meshgrid(np.arange(height, dtype=np.float64), np.arange(width, dtype=np.float64), indexing="ij")
    s = ((x + 0.5) / (width - 1)).ravel()
    t = ((y + 0.5) / (height - 1)).ravel()
    d = cam.directions(s, t)
    r = np.empty((width * height, 8), np.float32)
    r[:, 0:3] = cam.origin
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = math.inf
    return r


def sphere_prims(centres, radii) -> np.ndarray:
    """Analytic spheres (NewSphere, primitive.go:31-40) as primitive slots:
    centre, radius, then NaN as the fifth value (the tag), zeros after."""
    c = np.asarray(centres, np.float64).reshape(-1, 3)
    r = np.broadcast_to(np.asarray(radii, np.float64), (len(c),))
    out = np.zeros((len(c), 9), np.float64)
    out[:, :3] = c
    out[:, 3] = r
    out[:, 4] = np.nan
    return out


def mixed_scene(n_tris: int = 20_000, n_spheres: int = 400, seed: int = 81) -> np.ndarray:
    """Random triangles in [0, 1]^3 with analytic spheres (radius 0.002 .. 0.03)
    interleaved in caller order; FP32-exact."""
    rng = np.random.default_rng(seed)
    tris = random_triangles(n_tris, seed=seed)
    sph = sphere_prims(rng.uniform(0, 1, (n_spheres, 3)), rng.uniform(0.002, 0.03, n_spheres))
    pos = np.concatenate([tris, sph])[rng.permutation(n_tris + n_spheres)]
    return pos.astype(np.float32).astype(np.float64)


def camera_cfg2(width: int = 4096, height: int = 2048) -> Camera:
    """Pinhole camera outside the unit cube looking at its centre; the cube
    fills most of the frame."""
    return Camera(width / height, 60.0, (0.5, 0.5, -0.6), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0))


def coherent_rays_cfg2(width: int = 4096, height: int = 2048) -> np.ndarray:
    """Primary rays of camera_cfg2."""
    return camera_rays(camera_cfg2(width, height), width, height)


MAT_DIFFUSE, MAT_LIGHT, MAT_REFLECTIVE, MAT_REFRACTIVE = 0, 1, 2, 3


def _quad(p0, p1, p2, p3) -> np.ndarray:
    return np.array([[*p0, *p1, *p2], [*p0, *p2, *p3]], np.float64)


def uv_sphere(center, radius: float, slices: int, rings: int, jitter: float = 0.0,
              rng: np.random.Generator | None = None) -> np.ndarray:
    """UV sphere with `slices` segments around and `rings` - 1 latitude bands
    (2*slices*(rings - 2) triangles); every vertex (poles included, shared
    vertices once) is moved radially by U(-jitter, jitter).  (n, 9) float64."""
    c = np.asarray(center, np.float64)
    bands = rings - 1
    theta = np.linspace(0.0, math.pi, bands + 1)
    phi = np.linspace(0.0, 2 * math.pi, slices + 1)[:-1]
    grid = np.empty((bands + 1, slices, 3))
    grid[..., 0] = np.sin(theta)[:, None] * np.cos(phi)[None, :]
    grid[..., 1] = np.cos(theta)[:, None] * np.ones_like(phi)[None, :]
    grid[..., 2] = np.sin(theta)[:, None] * np.sin(phi)[None, :]
    r = np.full((bands + 1, slices), radius)
    if jitter and rng is not None:
        r = r + rng.uniform(-jitter, jitter, r.shape)
        r[0, :] = r[0, 0]  # one pole vertex
        r[-1, :] = r[-1, 0]
    verts = c + grid * r[..., None]
    s = np.arange(slices)
    s1 = (s + 1) % slices
    parts = [np.concatenate([verts[0, s], verts[1, s1], verts[1, s]], axis=1)]  # cap at the top pole
    if bands > 2:
        a, bb = verts[1:bands - 1, s], verts[1:bands - 1, s1]
        cc, d = verts[2:bands, s1], verts[2:bands, s]
        quads = np.stack([np.concatenate([a, bb, cc], axis=-1), np.concatenate([a, cc, d], axis=-1)], axis=2)
        parts.append(quads.reshape(-1, 9))
    if bands > 1:
        b = bands - 1
        parts.append(np.concatenate([verts[b, s], verts[b, s1], verts[b + 1, s1]], axis=1))  # bottom cap
    return np.ascontiguousarray(np.concatenate(parts), np.float64)


def displaced_sphere(ntri: int, seed: int, center=(0.0, 0.0, 0.0), radius: float = 1.0,
                     jitter: float = 0.01) -> np.ndarray:
    """A UV sphere of about `ntri` triangles (slices ~ sqrt(ntri)) with seeded
    radial vertex jitter U(-jitter, jitter) * radius; FP32-exact vertices."""
    rng = np.random.default_rng(seed)
    slices = max(3, int(round(math.sqrt(ntri))))
    rings = max(3, int(round(ntri / (2 * slices))) + 2)
    pos = uv_sphere(center, radius, slices, rings, jitter=jitter * radius, rng=rng)
    return pos.astype(np.float32).astype(np.float64)


def batch_meshes(count: int = 64, lo: int = 1000, hi: int = 500_000, seed: int = 11):
    """`count` displaced spheres with triangle counts log-uniform in [lo, hi],
    each at its own seeded centre in [0, 10]^3 and radius in [0.5, 2]."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(lo), math.log(hi), count)).astype(np.int64)
    centres = rng.uniform(0.0, 10.0, (count, 3))
    radii = rng.uniform(0.5, 2.0, count)
    return [displaced_sphere(int(n), seed + 1 + i, centres[i], float(radii[i])) for i, n in enumerate(sizes)]


def box_scene(seed: int = 1, slices: int = 100, rings: int = 102, width: int = 640, height: int = 360):
    """Five-wall box [-1,1]^3 open towards +z (10 triangles: white floor,
    ceiling and back 0.75, red left (0.75, 0.25, 0.25), blue right
    (0.25, 0.25, 0.75)), a 2-triangle area light under the ceiling with
    emission 15, and a UV sphere of 2*slices*(rings-2) triangles with radial
    jitter U(-0.01, 0.01) (seeded).  Vertices are FP32-exact.  Returns a dict
    with pos, mat_idx, materials, camera, width, height."""
    rng = np.random.default_rng(seed)
    L, H_ = -1.0, 1.0
    walls = [
        (_quad((L, L, H_), (H_, L, H_), (H_, L, L), (L, L, L)), 0),   # floor
        (_quad((L, H_, L), (H_, H_, L), (H_, H_, H_), (L, H_, H_)), 0),  # ceiling
        (_quad((L, L, L), (H_, L, L), (H_, H_, L), (L, H_, L)), 0),   # back
        (_quad((L, L, H_), (L, L, L), (L, H_, L), (L, H_, H_)), 1),   # left (red)
        (_quad((H_, L, L), (H_, L, H_), (H_, H_, H_), (H_, H_, L)), 2),  # right (blue)
    ]
    y = 0.98
    light = _quad((-0.3, y, -0.3), (0.3, y, -0.3), (0.3, y, 0.3), (-0.3, y, 0.3))
    sphere = uv_sphere((0.0, -0.45, 0.0), 0.55, slices, rings, jitter=0.01, rng=rng)
    pos = np.concatenate([w for w, _ in walls] + [light, sphere])
    mat_idx = np.concatenate([np.full(2, m, np.int32) for _, m in walls]
                             + [np.full(2, 3, np.int32), np.full(len(sphere), 0, np.int32)])
    materials = np.array([
        [MAT_DIFFUSE, 0.75, 0.75, 0.75],
        [MAT_DIFFUSE, 0.75, 0.25, 0.25],
        [MAT_DIFFUSE, 0.25, 0.25, 0.75],
        [MAT_LIGHT, 15.0, 15.0, 15.0],
    ], np.float32)
    cam = Camera(width / height, 45.0, (0.0, 0.0, 3.4), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    pos = pos.astype(np.float32).astype(np.float64)
    return {"pos": pos, "mat_idx": mat_idx, "materials": materials, "camera": cam, "width": width,
            "height": height}


def specular_box_scene(seed: int = 1, slices: int = 48, rings: int = 26, width: int = 320, height: int = 180):
    """The box and area light of box_scene with two smooth UV spheres instead of
    one: a mirror (Reflective, albedo 0.9, Diffusion 0.05) on the left and a
    glass ball (Refractive, albedo 0.95, Ratio 1.5) on the right, both with
    analytic vertex normals (CornellBox-like, cornellbox.go:5-78).  Vertices
    are FP32-exact.  Returns pos, normals, mat_idx, materials8 (kind, r, g, b,
    param, 0, 0, 0), camera, width, height."""
    base = box_scene(seed=seed, slices=3, rings=3, width=width, height=height)
    walls = base["pos"][:12]
    wall_mat = base["mat_idx"][:12]
    c1, c2, r = np.array([-0.45, -0.55, -0.1]), np.array([0.45, -0.55, 0.25]), 0.42
    s1 = uv_sphere(tuple(c1), r, slices, rings).astype(np.float32).astype(np.float64)
    s2 = uv_sphere(tuple(c2), r, slices, rings).astype(np.float32).astype(np.float64)
    pos = np.concatenate([walls, s1, s2])

    def sphere_normals(tris, c):
        v = tris.reshape(-1, 3, 3) - c
        return (v / np.linalg.norm(v, axis=2, keepdims=True)).reshape(-1, 9)

    wn = np.empty((12, 9))
    for k in range(12):  # geometric normal of each wall / light triangle at its three vertices
        t = walls[k].reshape(3, 3)
        g = np.cross(t[1] - t[0], t[2] - t[0])
        wn[k] = np.tile(g / np.linalg.norm(g), 3)
    normals = np.concatenate([wn, sphere_normals(s1, c1), sphere_normals(s2, c2)])
    mat_idx = np.concatenate([wall_mat, np.full(len(s1), 4, np.int32), np.full(len(s2), 5, np.int32)])
    materials = np.zeros((6, 8), np.float32)
    materials[:4, :4] = base["materials"]
    materials[4, :5] = [MAT_REFLECTIVE, 0.9, 0.9, 0.9, 0.05]
    materials[5, :5] = [MAT_REFRACTIVE, 0.95, 0.95, 0.95, 1.5]
    return {"pos": pos, "normals": normals, "mat_idx": mat_idx, "materials": materials, "camera": base["camera"],
            "width": width, "height": height}


def adjoint_scene(ntri: int = 200_000, seed: int = 7, width: int = 512, height: int = 512):
    """The five-wall box and area light of box_scene (12 triangles) with four
    displaced spheres sharing the remaining ~ntri triangles, a ground-truth
    per-triangle RGB albedo ~U(0.1, 0.9) (seed) and per-triangle emission
    (15 on the light, 0 elsewhere).  Vertices are FP32-exact.  Returns a dict
    with pos, mat_idx, materials, albedo (n, 3) float32, emission (n, 3)
    float32, camera, width, height."""
    base = box_scene(seed=seed, slices=3, rings=3, width=width, height=height)
    walls = base["pos"][:12]
    wall_mi = base["mat_idx"][:12]
    per = max(8, (ntri - 12) // 4)
    centres = [(-0.45, -0.6, -0.35), (0.45, -0.6, -0.3), (-0.4, -0.55, 0.45), (0.42, -0.62, 0.45)]
    radii = [0.38, 0.36, 0.33, 0.34]
    spheres = [displaced_sphere(per, seed + 10 + i, centres[i], radii[i]) for i in range(4)]
    pos = np.concatenate([walls] + spheres).astype(np.float32).astype(np.float64)
    n = len(pos)
    mat_idx = np.concatenate([wall_mi, np.zeros(n - 12, np.int32)]).astype(np.int32)
    rng = np.random.default_rng(seed)
    albedo = rng.uniform(0.1, 0.9, (n, 3)).astype(np.float32)
    emission = np.zeros((n, 3), np.float32)
    emission[mat_idx == 3] = 15.0
    return {"pos": pos, "mat_idx": mat_idx, "materials": base["materials"], "albedo": albedo,
            "emission": emission, "camera": base["camera"], "width": width, "height": height}


def fbm_heightfield(n: int = 2048, octaves: int = 8, seed: int = 5, base_freq: int = 4,
                    gain: float = 0.5) -> np.ndarray:
    """n x n fractional Brownian motion: `octaves` layers of smoothstep-interpolated
    value noise on lattices of base_freq * 2^o cells, amplitude gain^o (seeded).
    float32, scaled to [0, 0.35]."""
    rng = np.random.default_rng(seed)
    h = np.zeros((n, n), np.float64)
    amp, freq = 1.0, base_freq
    for _ in range(octaves):
        lat = rng.standard_normal((freq + 1, freq + 1))
        x = np.linspace(0.0, freq, n)
        i = np.minimum(x.astype(np.int64), freq - 1)
        f = x - i
        f = f * f * (3.0 - 2.0 * f)
        a = lat[i][:, i] * (1 - f)[None, :] + lat[i][:, i + 1] * f[None, :]
        b = lat[i + 1][:, i] * (1 - f)[None, :] + lat[i + 1][:, i + 1] * f[None, :]
        h += amp * (a * (1 - f)[:, None] + b * f[:, None])
        amp *= gain
        freq *= 2
    h = (h - h.min()) / (h.max() - h.min()) * 0.35
    return h.astype(np.float32)


def terrain_scene(n: int = 2048, octaves: int = 8, seed: int = 5, n_spheres: int = 5000,
                  width: int = 1920, height: int = 1080):
    """The large configuration: an n x n fBm heightfield over [-1, 1]^2 (2 (n-1)^2
    triangles, first in the buffer) plus n_spheres copies of one 400-triangle
    sphere (radius 0.006 .. 0.02, seeded positions resting on the terrain).
    Terrain triangles carry a ground-truth albedo ~U(0.1, 0.9) (seed); spheres
    are diffuse 0.8.  Vertices are FP32-exact.  Returns pos, mat_idx,
    materials, albedo, emission, camera, n_terrain, width, height."""
    hf = fbm_heightfield(n, octaves, seed).astype(np.float64)
    g = np.linspace(-1.0, 1.0, n)
    X, Z = np.meshgrid(g, g)
    V = np.stack([X, hf, Z], axis=-1)  # [row (z), col (x)]
    a, b = V[:-1, :-1], V[:-1, 1:]
    c, d = V[1:, 1:], V[1:, :-1]
    t1 = np.concatenate([a, d, c], axis=-1).reshape(-1, 9)
    t2 = np.concatenate([a, c, b], axis=-1).reshape(-1, 9)
    terrain = np.stack([t1, t2], axis=1).reshape(-1, 9)
    rng = np.random.default_rng(seed + 1)
    unit = uv_sphere((0.0, 0.0, 0.0), 1.0, 20, 12)  # 400 triangles
    r = rng.uniform(0.006, 0.02, n_spheres)
    px = rng.uniform(-0.95, 0.95, n_spheres)
    pz = rng.uniform(-0.95, 0.95, n_spheres)
    col = np.clip(np.round((px + 1.0) / 2.0 * (n - 1)).astype(np.int64), 0, n - 1)
    row = np.clip(np.round((pz + 1.0) / 2.0 * (n - 1)).astype(np.int64), 0, n - 1)
    py = hf[row, col] + r
    centre = np.stack([px, py, pz], axis=1)
    spheres = (unit.reshape(1, -1, 3, 3) * r[:, None, None, None] + centre[:, None, None, :]).reshape(-1, 9)
    pos = np.concatenate([terrain, spheres]).astype(np.float32).astype(np.float64)
    n_terrain = len(terrain)
    mat_idx = np.concatenate([np.zeros(n_terrain, np.int32), np.ones(len(spheres), np.int32)])
    materials = np.array([[MAT_DIFFUSE, 0.5, 0.5, 0.5], [MAT_DIFFUSE, 0.8, 0.8, 0.8]], np.float32)
    albedo = np.empty((len(pos), 3), np.float32)
    albedo[:n_terrain] = np.random.default_rng(seed + 2).uniform(0.1, 0.9, (n_terrain, 3))
    albedo[n_terrain:] = 0.8
    emission = np.zeros((len(pos), 3), np.float32)
    cam = Camera(width / height, 40.0, (0.0, 1.1, 2.3), (0.0, 0.1, 0.0), (0.0, 1.0, 0.0))
    return {"pos": pos, "mat_idx": mat_idx, "materials": materials, "albedo": albedo, "emission": emission,
            "camera": cam, "n_terrain": n_terrain, "width": width, "height": height}


def terrain_scene_graph(n: int = 2048, octaves: int = 8, seed: int = 5, n_spheres: int = 5000):
    """The large configuration as a scene graph (SceneNode tree under
    Scene.root): node 0 holds the terrain mesh (identity), nodes 1..n_spheres
    each place the one 400-triangle unit sphere mesh with a scale-and-translate
    matrix.  Same sphere radii and positions as terrain_scene (before its FP32
    rounding).  Returns meshes, mesh materials, parent, node_mesh, matrices
    (n_nodes, 16)."""
    hf = fbm_heightfield(n, octaves, seed).astype(np.float64)
    g = np.linspace(-1.0, 1.0, n)
    X, Z = np.meshgrid(g, g)
    V = np.stack([X, hf, Z], axis=-1)
    a, b = V[:-1, :-1], V[:-1, 1:]
    c, d = V[1:, 1:], V[1:, :-1]
    t1 = np.concatenate([a, d, c], axis=-1).reshape(-1, 9)
    t2 = np.concatenate([a, c, b], axis=-1).reshape(-1, 9)
    terrain = np.stack([t1, t2], axis=1).reshape(-1, 9)
    rng = np.random.default_rng(seed + 1)
    unit = uv_sphere((0.0, 0.0, 0.0), 1.0, 20, 12)
    r = rng.uniform(0.006, 0.02, n_spheres)
    px = rng.uniform(-0.95, 0.95, n_spheres)
    pz = rng.uniform(-0.95, 0.95, n_spheres)
    col = np.clip(np.round((px + 1.0) / 2.0 * (n - 1)).astype(np.int64), 0, n - 1)
    row = np.clip(np.round((pz + 1.0) / 2.0 * (n - 1)).astype(np.int64), 0, n - 1)
    py = hf[row, col] + r
    mats = np.zeros((n_spheres + 1, 4, 4))
    mats[:] = np.eye(4)
    mats[1:, 0, 0] = mats[1:, 1, 1] = mats[1:, 2, 2] = r
    mats[1:, 0, 3], mats[1:, 1, 3], mats[1:, 2, 3] = px, py, pz
    parent = np.full(n_spheres + 1, -1, np.int32)
    node_mesh = np.concatenate([[0], np.ones(n_spheres, np.int32)]).astype(np.int32)
    return [terrain, unit], [0, 1], parent, node_mesh, mats.reshape(-1, 16)


def parse_obj(path: str) -> tuple[np.ndarray, np.ndarray | None]:
    """OBJ triangles as obj.go:19-106 reads them: 'v', 'vn', 'f' (fan
    triangulation, negative indices relative to the current count, normals only
    when every corner has one).  Returns (n, 9) positions and (n, 9) normals or
    None when some face lacks normals (those get geometric normals later)."""
    verts = [np.zeros(3)]
    norms = [np.zeros(3)]
    pos, nrm, has_n = [], [], []
    with open(path) as f:
        for line in f:
            fields = line.split()
            if not fields:
                continue
            key, vals = fields[0], fields[1:]
            if key == "v":
                verts.append(np.array([float(x) for x in vals[:3]]))
            elif key == "vn":
                norms.append(np.array([float(x) for x in vals[:3]]))
            elif key == "f":
                vi, ni = [], []
                for a in vals:
                    parts = a.split("/")
                    k = int(parts[0])
                    vi.append(len(verts) + k if k < 0 else k)
                    if len(parts) >= 3:
                        m = int(parts[2])
                        ni.append(len(norms) + m if m < 0 else m)
                for i in range(1, len(vi) - 1):
                    idx = [vi[0], vi[i], vi[i + 1]]
                    pos.append(np.concatenate([verts[j] for j in idx]))
                    if len(ni) == len(vals):
                        nidx = [ni[0], ni[i], ni[i + 1]]
                        nrm.append(np.concatenate([norms[j] for j in nidx]))
                        has_n.append(True)
                    else:
                        nrm.append(np.zeros(9))
                        has_n.append(False)
    p = np.array(pos, np.float64).reshape(-1, 9)
    n = np.array(nrm, np.float64).reshape(-1, 9)
    return p, (n if all(has_n) else None)
