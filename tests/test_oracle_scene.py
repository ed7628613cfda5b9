"""Pins the oracle's Scene.collectTracables restatement (scene.go:95-151,
primitive.go:152-170, math.go:144-320) with closed forms:

- composition order: t = node x parent, so a scale below a translation
  moves x to 2 (x + 1), not 2 x + 1;
- a mesh whose own node matrix is the identity stays untransformed, even
  below a translated parent (the reference compares the node's matrix, not
  the composed one), while its children still inherit the parent;
- normals: a translation leaves them unchanged, a uniform scale s divides
  them by s (the inverse transpose, not renormalised), and for a general
  linear map they equal inv(A^T) n to FP64 accuracy;
- output order: a node's mesh, then its children in index order."""
import numpy as np

from trace_amd import scenes


def _translate(x, y, z):
    m = np.eye(4)
    m[:3, 3] = (x, y, z)
    return m.reshape(-1)


def _scale(s):
    m = np.eye(4) * s
    m[3, 3] = 1
    return m.reshape(-1)


TRI = np.array([[0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.5, 0.25, 0.0]])
NRM = np.array([[0.0, 0.0, 1.0, 0.5, 0.5, 0.5, 1.0, 0.0, 0.0]])


def test_composition_order_and_identity_quirk():
    from oracle.pyoracle import scene_flatten

    # node 0: translate (1, 0, 0), no mesh; node 1 (child of 0): scale 2, mesh 0;
    # node 2 (child of 0): identity, mesh 0; node 3 (child of 2): scale 2, mesh 0
    parent = [-1, 0, 0, 2]
    node_mesh = [-1, 0, 0, 0]
    mats = np.stack([_translate(1, 0, 0), _scale(2), np.eye(4).reshape(-1), _scale(2)])
    pos, nrm, mat = scene_flatten([TRI], [3], parent, node_mesh, mats, normals=[NRM])
    assert pos.shape == (3, 9) and mat.tolist() == [3, 3, 3]
    v = TRI.reshape(3, 3)
    assert np.array_equal(pos[0].reshape(3, 3), 2 * (v + [1, 0, 0]))  # node 1: scale after translate
    assert np.array_equal(pos[1], TRI[0])                              # node 2: identity -> raw
    assert np.array_equal(pos[2].reshape(3, 3), 2 * (v + [1, 0, 0]))  # node 3: inherits the parent
    n = NRM.reshape(3, 3)
    assert np.array_equal(nrm[0].reshape(3, 3), n / 2)  # scale 2: normals halved (translation drops out)
    assert np.array_equal(nrm[1], NRM[0])


def test_translation_keeps_normals_and_order_is_preorder():
    from oracle.pyoracle import scene_flatten

    a = scenes.uv_sphere((0, 0, 0), 1.0, 6, 4)
    b = scenes.random_triangles(5, seed=3)
    na = np.tile([0.0, 1.0, 0.0], (len(a), 3))
    nb = np.tile([1.0, 0.0, 0.0], (len(b), 3))
    # root children 0 and 1; node 2 is a child of node 0 (comes before node 1)
    parent = [-1, -1, 0]
    node_mesh = [0, 1, 1]
    mats = np.stack([_translate(0.5, -2, 3), _translate(1, 1, 1), _translate(-4, 0, 0)])
    pos, nrm, mat = scene_flatten([a, b], [0, 1], parent, node_mesh, mats, normals=[na, nb])
    la, lb = len(a), len(b)
    assert mat.tolist() == [0] * la + [1] * lb + [1] * lb
    assert np.array_equal(pos[:la].reshape(-1, 3), a.reshape(-1, 3) + [0.5, -2, 3])
    # node 2: t = T(-4, 0, 0) x T(0.5, -2, 3)
    assert np.allclose(pos[la:la + lb].reshape(-1, 3), b.reshape(-1, 3) + [-3.5, -2, 3], rtol=0, atol=1e-15)
    assert np.array_equal(pos[la + lb:].reshape(-1, 3), b.reshape(-1, 3) + [1, 1, 1])
    assert np.array_equal(nrm[:la], na) and np.array_equal(nrm[la:], np.concatenate([nb, nb]))


def test_general_linear_map_normals():
    from oracle.pyoracle import scene_flatten

    rng = np.random.default_rng(7)
    A = np.eye(4)
    A[:3, :3] = rng.normal(0, 1, (3, 3))
    A[:3, 3] = rng.normal(0, 1, 3)
    tri = rng.normal(0, 1, (4, 9))
    nrm_in = rng.normal(0, 1, (4, 9))
    pos, nrm, _ = scene_flatten([tri], [0], [-1], [0], A.reshape(1, 16), normals=[nrm_in])
    want_p = (tri.reshape(-1, 3) @ A[:3, :3].T + A[:3, 3]).reshape(-1, 9)
    assert np.allclose(pos, want_p, rtol=0, atol=1e-13)
    want_n = (nrm_in.reshape(-1, 3) @ np.linalg.inv(A[:3, :3].T).T).reshape(-1, 9)
    assert np.allclose(nrm, want_n, rtol=1e-12, atol=1e-12)
