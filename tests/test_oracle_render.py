"""Pins the oracle's path tracer (CPU only) with cases whose answers are
fixed by the mathematics, not by the oracle's own formulas:
Philox known-answer vectors, a furnace test, the set of values a path can
take under the reference's depth rule, and direct views of emitters."""
import numpy as np

from trace_amd import scenes

A = np.float32(0.6)


def test_philox_known_answers(oracle):
    # Random123 known-answer vectors for philox4x32-10
    assert oracle.philox((0, 0, 0, 0), (0, 0)) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0)) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def sphere_scene():
    pos = scenes.uv_sphere((0, 0, 0), 1.0, 24, 14).astype(np.float32).astype(np.float64)
    mat_idx = np.zeros(len(pos), np.int32)
    mats = np.array([[scenes.MAT_DIFFUSE, A, A, A]], np.float32)
    cam = scenes.Camera(2.0, 60.0, (0, 0, 3.0), (0, 0, 0), (0, 1, 0))
    return pos, mat_idx, mats, cam


def test_furnace_convex_sphere(oracle):
    """Geometric normals on a convex polyhedron: the scatter direction n + w
    never re-enters it, so under a white sky every hit pixel is exactly the
    albedo and every miss pixel exactly 1, sample by sample."""
    pos, mi, mats, cam = sphere_scene()
    t = oracle.Tree(pos)
    s, q, last, paths = oracle.render(t, mi, mats, cam, 64, 32, 4, 5, 1, seed=7, path_values=True)
    vals = np.unique(paths)
    assert set(vals.tolist()) <= {float(A), 1.0}
    hit_frac = np.mean(paths[..., 0] == float(A))
    assert 0.1 < hit_frac < 0.9
    assert np.array_equal(last, paths[-1])
    assert np.allclose(s, paths.sum(axis=0)) and np.allclose(q, (paths ** 2).sum(axis=0))


def test_path_values_follow_depth_rule(oracle):
    """All-diffuse grey box (albedo 0.5, no emitter), white miss: a path that
    leaves after scattering at vertex K <= max_depth is worth 0.5^(K+1); one
    that reaches vertex max_depth + 1 is worth 0; a camera miss is worth 1.
    A shader that let one more bounce through would produce 0.5^(max_depth+2)."""
    sc = scenes.box_scene(seed=3, slices=16, rings=10, width=48, height=27)
    mats = np.array([[scenes.MAT_DIFFUSE, 0.5, 0.5, 0.5]] * 4, np.float32)
    t = oracle.Tree(sc["pos"])
    max_depth = 3
    _, _, _, paths = oracle.render(t, sc["mat_idx"], mats, sc["camera"], 48, 27, 4, max_depth, 1, seed=5,
                                   path_values=True)
    allowed = {0.0, 1.0} | {0.5 ** (k + 1) for k in range(max_depth + 1)}
    vals = set(np.unique(paths).tolist())
    assert vals <= allowed, vals - allowed
    assert 0.5 ** (max_depth + 1) in vals and 0.0 in vals


def test_light_seen_directly(oracle):
    """A camera ray that hits an emitter returns its emission (vertex 0)."""
    h = 0.25
    quad = np.array([[-h, -h, 0, h, -h, 0, h, h, 0], [-h, -h, 0, h, h, 0, -h, h, 0]], np.float64)
    mats = np.array([[scenes.MAT_LIGHT, 15.0, 7.5, 3.0]], np.float32)
    cam = scenes.Camera(1.0, 30.0, (0, 0, 3.0), (0, 0, 0), (0, 1, 0))
    t = oracle.Tree(quad)
    s, q, last = oracle.render(t, np.zeros(2, np.int32), mats, cam, 16, 16, 2, 5, 0, seed=1)
    # the 30 degree view at distance 3 spans 1.6 units: the 0.5-unit quad covers the centre only
    assert np.array_equal(s[8, 8], [30.0, 15.0, 6.0])
    assert np.array_equal(s[0, 0], [0.0, 0.0, 0.0])


def test_sky_miss_shader(oracle):
    """SkyMissShader: straight up gives the blue end, straight down the white."""
    quad = np.array([[100, 100, 100, 101, 100, 100, 100, 101, 100]], np.float64)  # out of view
    mats = np.array([[scenes.MAT_DIFFUSE, 0.5, 0.5, 0.5]], np.float32)
    t = oracle.Tree(quad)
    up = scenes.Camera(1.0, 1e-3, (0, 0, 0), (0, 1, 0), (0, 0, 1))
    s, _, _ = oracle.render(t, np.zeros(1, np.int32), mats, up, 2, 2, 1, 1, 2, seed=1)
    assert np.allclose(s[0, 0], [0.25, 0.35, 0.5], atol=1e-6)
    down = scenes.Camera(1.0, 1e-3, (0, 0, 0), (0, -1, 0), (0, 0, 1))
    s, _, _ = oracle.render(t, np.zeros(1, np.int32), mats, down, 2, 2, 1, 1, 2, seed=1)
    assert np.allclose(s[0, 0], [0.8, 0.8, 0.8], atol=1e-6)
