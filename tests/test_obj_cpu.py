"""OBJ models on the CPU: the library's loader (rumi_obj_load) against the
oracle's restatement of Model::LoadObjectFile (src/Shape/model.cpp:72-338),
bit for bit, on the reference's own model files and on edge cases; the
oracle pinned by closed forms (hand-derived triangles, exact transforms, a
true inverse) and by an independent plain OBJ reading of the reference's
files.  No GPU calls."""
import gzip
import os
import re

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = os.path.join(HERE, "golden", "models")


@pytest.fixture(scope="module")
def model_files(tmp_path_factory):
    d = tmp_path_factory.mktemp("obj")
    out = {}
    for name in ("cube.obj", "edge.obj", "teapot.obj.gz", "dragon.obj.gz"):
        src = os.path.join(MODELS, name)
        key = name.split(".")[0]
        if name.endswith(".gz"):
            dst = d / name[:-3]
            dst.write_bytes(gzip.open(src).read())
            out[key] = str(dst)
        else:
            out[key] = src
    return out


def scenes_module():
    import importlib
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "rumi-renderer_amd"))
    return importlib.import_module("scenes")


def transforms():
    sc = scenes_module()
    return {"identity": None,
            "dragon": sc.model_transform(0.3),
            "teapot": sc.model_transform(0.65),
            "rotated": sc.mat_mul(sc.mat_mul(sc.mat_rotation_y(0.7), sc.mat_rotation_x(-0.4)),
                                  sc.mat_mul(sc.mat_translation(0.25, -1.5, 2.0),
                                             sc.mat_scaling(1.5, 0.5, 2.0)))}


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


@pytest.mark.parametrize("model", ["cube", "edge", "teapot", "dragon"])
@pytest.mark.parametrize("xf", ["identity", "dragon", "teapot", "rotated"])
def test_library_loader_equals_oracle_bit_for_bit(rumi, oracle, model_files, model, xf):
    t = transforms()[xf]
    m = rumi.load_obj(model_files[model], t)
    ox, on, oc = oracle.load_obj(model_files[model], t)
    n = m["idx"].shape[0]
    assert n == ox.shape[0] > 0
    np.testing.assert_array_equal(m["counts"], oc)
    np.testing.assert_array_equal(bits(m["xyz"].reshape(n, 9)), bits(ox))
    np.testing.assert_array_equal(bits(m["normals"].reshape(n, 9)), bits(on))
    np.testing.assert_array_equal(m["idx"].reshape(-1), np.arange(3 * n))


def test_edge_file_triangles_by_hand(oracle, model_files):
    # f 1//1 2//1 3//1 ; f -4//-2 -3//-2 -1//-2 (relative to the totals:
    # 4 vertices, 2 normals) ; the quad's first three corners with the zero
    # normal (face normal instead) ; f 2//1 3//1 4//1 after s/g/usemtl
    x, n, counts = oracle.load_obj(model_files["edge"])
    np.testing.assert_array_equal(counts, [4, 1, 2, 4])
    V = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    want_v = np.stack([V[[0, 1, 2]], V[[0, 1, 3]], V[[0, 2, 3]], V[[1, 2, 3]]]).reshape(4, 9)
    z = [0, 0, 1]
    want_n = np.float32([z * 3, z * 3, [1, 0, 0] * 3, z * 3])
    np.testing.assert_array_equal(x, want_v)
    np.testing.assert_array_equal(n, want_n)
    # a non-uniform scale and a translation: vertices exactly, normals
    # through the transposed inverse (diag(1/2, 1, 1)) then normalised
    T = np.float32([[2, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 5], [0, 0, 0, 1]])
    x2, n2, _ = oracle.load_obj(model_files["edge"], T)
    np.testing.assert_array_equal(x2.reshape(-1, 3), want_v.reshape(-1, 3) * [2, 1, 1] + [0, 0, 5])
    np.testing.assert_array_equal(n2, want_n)


def test_normals_through_the_transposed_inverse(oracle, tmp_path):
    # one triangle, one normal (1, 1, 0)/sqrt(2) under diag(2, 1, 1):
    # (M^-1)^T n = (1/2, 1, 0)/sqrt(2), normalised (1, 2, 0)/sqrt(5)
    p = tmp_path / "n.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0.70710678 0.70710678 0\nf 1//1 2//1 3//1\n")
    T = np.diag(np.float32([2, 1, 1, 1]))
    _, n, _ = oracle.load_obj(str(p), T)
    np.testing.assert_allclose(n.reshape(3, 3), np.tile([1, 2, 0], (3, 1)) / np.sqrt(5), atol=2e-7)


def test_inverse_is_an_inverse(oracle):
    rng = np.random.default_rng(7)
    for _ in range(20):
        a = rng.normal(size=(4, 4)).astype(np.float32) + 4 * np.eye(4, dtype=np.float32)
        inv = oracle.mat4_inverse(a)
        np.testing.assert_allclose(inv.astype(np.float64) @ a.astype(np.float64), np.eye(4), atol=1e-5)
    d = np.diag(np.float32([2, 4, 0.5, 1]))
    np.testing.assert_array_equal(oracle.mat4_inverse(d), np.diag(np.float32([0.5, 0.25, 2, 1])))


def plain_obj(path):
    """A plain OBJ reading (positive v/vt/vn indices, triangles) for files
    the reference's reader handles the standard way."""
    v, vn, faces = [], [], []
    for line in open(path, newline=""):
        line = line.rstrip("\r\n")
        if line.startswith("v "):
            v.append([float(t) for t in line.split()[1:4]])
        elif line.startswith("vn "):
            vn.append([float(t) for t in line.split()[1:4]])
        elif line.startswith("f "):
            faces.append([[int(x) - 1 if x else -1 for x in re.split("/", c)] + [-1, -1]
                          for c in line.split()[1:4]])
    return np.float32(v), np.float64(vn), faces


@pytest.mark.parametrize("model,tris", [("cube", 12), ("teapot", 2304), ("dragon", 100000)])
def test_reference_models_read_plainly(oracle, model_files, model, tris):
    x, n, counts = oracle.load_obj(model_files[model])
    assert x.shape[0] == tris == counts[3]
    v, vn, faces = plain_obj(model_files[model])
    vi = np.array([[c[0] for c in f] for f in faces])
    np.testing.assert_array_equal(x.reshape(-1, 3, 3), v[vi])
    if len(vn):
        ni = np.array([[c[2] for c in f] for f in faces])
        want = vn[ni]
        want /= np.linalg.norm(want, axis=-1, keepdims=True)
        np.testing.assert_allclose(n.reshape(-1, 3, 3), want, atol=3e-7)
    else:
        p = x.reshape(-1, 3, 3).astype(np.float64)
        c = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        c /= np.linalg.norm(c, axis=1, keepdims=True)
        np.testing.assert_allclose(n.reshape(-1, 3, 3), np.repeat(c[:, None], 3, 1), atol=2e-4)
        assert (np.abs(np.linalg.norm(n.reshape(-1, 3).astype(np.float64), axis=1) - 1) < 1e-6).all()


@pytest.mark.parametrize("text", ["v 0 0 0\nv 1 0 0\nf 1 2\n",
                                  "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1/1/1/1 2 3\n",
                                  "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 9\n",
                                  "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1 2 3\n"])
def test_undefined_faces_are_rejected(rumi, oracle, tmp_path, text):
    p = tmp_path / "bad.obj"
    p.write_text(text)
    with pytest.raises(ValueError):
        oracle.load_obj(str(p))
    with pytest.raises(rumi.RumiError):
        rumi.load_obj(str(p))


def test_missing_file_is_an_error(rumi, tmp_path):
    with pytest.raises(rumi.RumiError):
        rumi.load_obj(str(tmp_path / "none.obj"))
