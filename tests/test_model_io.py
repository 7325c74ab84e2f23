"""Model files in the reference's PLY layout (R/scene/gaussian_model.py:257-394,
801-819): known-answer parsing, round trips, the ASCII twin, and a model the
reference saved (tests/golden/mito1024wveg_iter1000.ply)."""
import os
import struct

import numpy as np
import pytest

from gvc import model_io as mio

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "mito1024wveg_iter1000.ply")


def random_params(rng, n):
    return mio.GaussianParams(means=rng.uniform(-1, 2, (n, 3)).astype(np.float32),
                              values=rng.normal(size=(n, 1)).astype(np.float32),
                              weights=rng.normal(size=(n, 1)).astype(np.float32),
                              scales=rng.normal(-5, 1, (n, 3)).astype(np.float32),
                              rotations=rng.normal(size=(n, 4)).astype(np.float32))


def same(a, b):
    for f in ("means", "values", "weights", "scales", "rotations"):
        x, y = getattr(a, f), getattr(b, f)
        assert x.shape == y.shape and x.dtype == np.float32
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f


def test_binary_round_trip_and_header(tmp_path):
    p = random_params(np.random.default_rng(0), 257)
    p.values[3, 0] = np.inf          # saved models can hold infinite logits
    p.rotations[4] = 0.0
    path = str(tmp_path / "m" / "point_cloud.ply")
    mio.write_ply(path, p)
    raw = open(path, "rb").read()
    header = ("ply\nformat binary_little_endian 1.0\nelement vertex 257\n"
              + "".join(f"property float {a}\n" for a in mio.ATTRIBUTES) + "end_header\n").encode()
    assert raw.startswith(header)
    assert len(raw) == len(header) + 257 * 12 * 4
    # first row, little-endian float32 in the reference's column order
    row0 = struct.unpack("<12f", raw[len(header):len(header) + 48])
    expect = np.concatenate([p.means[0], p.values[0], p.weights[0], p.scales[0], p.rotations[0]])
    assert np.array_equal(np.array(row0, dtype=np.float32), expect)
    same(mio.read_ply(path), p)


def test_ascii_twin(tmp_path):
    p = random_params(np.random.default_rng(1), 40)
    path = str(tmp_path / "point_cloud.ply")
    mio.write_ply(path, p)
    out = mio.write_ply_ascii(path)
    assert out == str(tmp_path / "point_cloud_ascii.ply")
    lines = open(out).read().splitlines()
    assert lines[:3] == ["ply", "format ascii 1.0", "element vertex 40"]
    assert lines[3:15] == [f"property float {a}" for a in mio.ATTRIBUTES] and lines[15] == "end_header"
    # each value printed as Python prints a float32: parses back to the same float32
    first = [float(w) for w in lines[16].split()]
    assert np.array_equal(np.array(first, dtype=np.float32), np.concatenate(
        [p.means[0], p.values[0], p.weights[0], p.scales[0], p.rotations[0]]))
    same(mio.read_ply(out), p)


def test_known_answer_permuted_properties(tmp_path):
    # properties out of order, suffixes unsorted, one double and one int column,
    # a comment, and a trailing element that is ignored
    props = [("rot_2", "float", "f"), ("x", "float", "f"), ("scale_1", "double", "d"), ("value", "float", "f"),
             ("rot_0", "float", "f"), ("y", "float", "f"), ("z", "float", "f"), ("scale_0", "float", "f"),
             ("weight", "float", "f"), ("rot_3", "float", "f"), ("scale_2", "float", "f"), ("rot_1", "float", "f"),
             ("flags", "int", "i")]
    rows = [dict(x=1.0, y=2.0, z=3.0, value=0.5, weight=-1.0, scale_0=-4.0, scale_1=-5.25, scale_2=-6.0,
                 rot_0=1.0, rot_1=0.0, rot_2=0.25, rot_3=-0.5, flags=7),
            dict(x=-1.5, y=0.0, z=8.0, value=-2.0, weight=3.0, scale_0=-1.0, scale_1=-2.0, scale_2=-3.0,
                 rot_0=0.0, rot_1=1.0, rot_2=0.0, rot_3=0.0, flags=-1)]
    head = "ply\nformat binary_little_endian 1.0\ncomment made by hand\nelement vertex 2\n"
    head += "".join(f"property {t} {n}\n" for n, t, _ in props)
    head += "element face 1\nproperty list uchar int vertex_indices\nend_header\n"
    body = b"".join(struct.pack("<" + "".join(c for _, _, c in props), *[r[n] for n, _, _ in props]) for r in rows)
    path = str(tmp_path / "hand.ply")
    open(path, "wb").write(head.encode() + body + b"\x03\x00\x00\x00\x00")
    with pytest.raises(ValueError, match="list"):
        mio.read_ply(path)   # list properties anywhere in the header are refused
    head = head.replace("element face 1\nproperty list uchar int vertex_indices\n", "")
    open(path, "wb").write(head.encode() + body)
    p = mio.read_ply(path)
    assert p.count == 2
    assert np.array_equal(p.means, np.array([[1, 2, 3], [-1.5, 0, 8]], np.float32))
    assert np.array_equal(p.values[:, 0], np.array([0.5, -2.0], np.float32))
    assert np.array_equal(p.weights[:, 0], np.array([-1.0, 3.0], np.float32))
    assert np.array_equal(p.scales, np.array([[-4, -5.25, -6], [-1, -2, -3]], np.float32))
    assert np.array_equal(p.rotations, np.array([[1, 0, 0.25, -0.5], [0, 1, 0, 0]], np.float32))


@pytest.mark.parametrize("text, match", [
    ("ply\nformat binary_big_endian 1.0\nelement vertex 0\nend_header\n", "binary_big_endian"),
    ("plx\n", "not a PLY"),
    ("ply\nformat ascii 1.0\nelement face 0\nend_header\n", "vertex"),
    ("ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\n", "end_header"),
])
def test_refused_files(tmp_path, text, match):
    path = str(tmp_path / "bad.ply")
    open(path, "wb").write(text.encode())
    with pytest.raises(ValueError, match=match):
        mio.read_ply_table(path)


def test_short_file_refused(tmp_path):
    p = random_params(np.random.default_rng(2), 10)
    path = str(tmp_path / "m.ply")
    mio.write_ply(path, p)
    open(path, "r+b").truncate(os.path.getsize(path) - 4)
    with pytest.raises(ValueError, match="shorter"):
        mio.read_ply(path)


def test_reference_saved_model():
    names, table = mio.read_ply_table(GOLDEN)
    assert names == mio.ATTRIBUTES
    size = os.path.getsize(GOLDEN)
    header_len = open(GOLDEN, "rb").read().index(b"end_header\n") + len(b"end_header\n")
    assert size == header_len + table.shape[0] * 48
    p = mio.read_ply(GOLDEN)
    assert p.count == 5258
    assert np.all(np.isfinite(p.means)) and np.all(np.isfinite(p.scales)) and np.all(np.isfinite(p.rotations))
    assert np.all(np.isfinite(p.weights))
    assert np.all(p.scales < 0)                      # raw log-scales
    n = np.linalg.norm(p.rotations, axis=1)
    assert np.all(n > 0.5) and np.all(n < 1.5)       # raw, unnormalised quaternions
    assert np.isposinf(p.values).sum() >= 1          # a saturated value logit
