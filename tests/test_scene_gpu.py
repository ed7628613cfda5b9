"""GPU parity of the scene compile (Scene.collectTracables, scene.go:95-151,
primitive.go:152-170) against the FP64 oracle: flattened positions bit-exact,
normals equal to the oracle's FP64 normals rounded to FP32, material indices
equal; the compiled mesh builds and renders like a mesh made from the
oracle's flattened triangles."""
import numpy as np
import pytest

from trace_amd import _lib, scenes
from trace_amd.device import Bvh, Mesh, Scene, render

pytestmark = pytest.mark.gpu


def _affine(rng, identity=False):
    m = np.eye(4)
    if not identity:
        q, _ = np.linalg.qr(rng.normal(0, 1, (3, 3)))
        m[:3, :3] = q * rng.uniform(0.5, 2.0)
        m[:3, 3] = rng.uniform(-1, 1, 3)
    return m.reshape(-1)


def _graph(seed, n_nodes=12, n_mesh=4):
    rng = np.random.default_rng(seed)
    parent = [-1] + [int(rng.integers(-1, k)) for k in range(1, n_nodes)]
    node_mesh = [int(rng.integers(-1, n_mesh)) for _ in range(n_nodes)]
    mats = np.stack([_affine(rng, identity=rng.uniform() < 0.3) for _ in range(n_nodes)])
    return parent, node_mesh, mats


def _meshes():
    ms = [scenes.uv_sphere((0, 0, 0), 0.5, 12, 8), scenes.random_triangles(300, seed=71),
          scenes.displaced_sphere(500, seed=72), scenes.random_triangles(40, seed=73, mean_edge=0.1)]
    rng = np.random.default_rng(74)
    nrm = [rng.normal(0, 1, (len(m), 9)) for m in ms]
    return ms, nrm


MATERIALS = np.array([[scenes.MAT_DIFFUSE, 0.7, 0.6, 0.5], [scenes.MAT_DIFFUSE, 0.3, 0.8, 0.2],
                      [scenes.MAT_LIGHT, 4.0, 4.0, 4.0]], np.float32)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_compile_matches_oracle(ctx, seed):
    from oracle.pyoracle import scene_flatten

    print(f"scene compile seed={seed}", flush=True)
    ms, nrm = _meshes()
    mm = [0, 1, 0, 2]
    parent, node_mesh, mats = _graph(seed)
    sc = Scene(ctx, ms, mm, MATERIALS, normals=nrm)
    mesh = sc.compile(parent, node_mesh, mats)
    opos, onrm, omat = scene_flatten(ms, mm, parent, node_mesh, mats, normals=nrm)
    assert mesh.ntri == len(opos) > 0
    assert np.array_equal(mesh.export(), opos)
    gn, gm = mesh.export_shading()
    assert np.array_equal(gn, onrm.astype(np.float32))
    assert np.array_equal(gm, omat)
    # recompile in place with other matrices (an animation frame)
    _, _, mats2 = _graph(seed + 100)
    sc.compile(parent, node_mesh, mats2, mesh=mesh)
    opos2, onrm2, _ = scene_flatten(ms, mm, parent, node_mesh, mats2, normals=nrm)
    assert np.array_equal(mesh.export(), opos2)
    assert np.array_equal(mesh.export_shading()[0], onrm2.astype(np.float32))


def test_compiled_mesh_builds_and_renders(ctx, oracle):
    from oracle.pyoracle import scene_flatten

    print("scene compile render", flush=True)
    ms, _ = _meshes()
    mm = [0, 1, 0, 2]
    parent, node_mesh, mats = _graph(5)
    sc = Scene(ctx, ms, mm, MATERIALS)  # geometric normals
    mesh = sc.compile(parent, node_mesh, mats)
    opos, _, omat = scene_flatten(ms, mm, parent, node_mesh, mats)
    bvh = Bvh.lbvh(ctx, mesh)
    keys, perm = bvh.sorted()
    okeys, operm = oracle.Tree(opos).sorted()
    assert np.array_equal(keys, okeys) and np.array_equal(perm, operm)
    ref = Mesh(ctx, opos)
    ref.set_shading(omat, MATERIALS)
    assert np.array_equal(mesh.export_shading()[0], ref.export_shading()[0])
    rbvh = Bvh.lbvh(ctx, ref)
    lo, hi = opos.reshape(-1, 3).min(0), opos.reshape(-1, 3).max(0)
    c = (lo + hi) / 2
    cam = scenes.Camera(1.0, 50.0, c + [0.0, 0.0, 2.5 * float(np.max(hi - lo))], c, (0, 1, 0))
    a = render(ctx, mesh, bvh, cam, 48, 48, 4, 3, _lib.MISS_SKY, seed=3)
    b = render(ctx, ref, rbvh, cam, 48, 48, 4, 3, _lib.MISS_SKY, seed=3)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert a[0].mean() > 0


def test_compile_rejects_bad_graphs(ctx):
    ms, _ = _meshes()
    sc = Scene(ctx, ms, [0, 1, 0, 2], MATERIALS)
    eye = np.eye(4).reshape(1, 16)
    with pytest.raises(_lib.TraceError):
        sc.compile([0], [0], eye)  # a node cannot be its own parent
    with pytest.raises(_lib.TraceError):
        sc.compile([-1], [7], eye)  # no mesh 7
    with pytest.raises(_lib.TraceError):
        sc.compile([-1], [-1], eye)  # no triangles
    mesh = sc.compile([-1], [0], eye)
    with pytest.raises(_lib.TraceError):
        sc.compile([-1], [1], eye, mesh=mesh)  # another triangle count
