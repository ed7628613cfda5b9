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


def coherent_rays_cfg2(width: int = 4096, height: int = 2048) -> np.ndarray:
    """Primary rays of a pinhole camera outside the unit cube looking at its
    centre; the cube fills most of the frame."""
    cam = Camera(width / height, 60.0, (0.5, 0.5, -0.6), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0))
    return camera_rays(cam, width, height)


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
