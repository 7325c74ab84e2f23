"""Gaussian model files in the reference's PLY layout.

The reference keeps a model as one ``vertex`` element of float32 properties
``x y z value weight scale_0 scale_1 scale_2 rot_0 rot_1 rot_2 rot_3``
(R/scene/gaussian_model.py:257-263): ``save_ply`` writes the raw parameters
(means, value and weight logits, log-scales, unnormalised quaternions; :265-285),
``save_ply_activated`` the activated ones (:287-303), ``convert_ply_to_ascii`` an
ASCII twin (:801-819), and ``load_ply`` reads the raw parameters back, taking the
``scale_*`` and ``rot*`` properties in the order of their numeric suffix
(:312-394).  This module reads and writes the same files without plyfile.

Reading parses the header and the numbers and executes nothing from the file.
Accepted: ``format binary_little_endian 1.0`` or ``format ascii 1.0``; one
``vertex`` element (other elements only after it); scalar properties of the
standard PLY types, in any order.  List properties and big-endian files raise
ValueError.
"""
import os
from dataclasses import dataclass

import numpy as np

ATTRIBUTES = ["x", "y", "z", "value", "weight", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]

_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1",
    "short": "i2", "int16": "i2", "ushort": "u2", "uint16": "u2",
    "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}


@dataclass
class GaussianParams:
    """Per-Gaussian arrays: means [N,3], values [N,1], weights [N,1],
    scales [N,3], rotations [N,4] (raw or activated, as the file holds them)."""
    means: np.ndarray
    values: np.ndarray
    weights: np.ndarray
    scales: np.ndarray
    rotations: np.ndarray

    @property
    def count(self):
        return int(self.means.shape[0])


def _parse_header(f):
    first = f.readline()
    if first.strip() != b"ply":
        raise ValueError("not a PLY file")
    fmt = None
    elements = []  # [name, count, [(prop name, numpy type)]]
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header has no end_header")
        words = line.decode("ascii", errors="strict").split()
        if not words or words[0] in ("comment", "obj_info"):
            continue
        if words[0] == "end_header":
            break
        if words[0] == "format":
            if len(words) != 3 or words[2] != "1.0":
                raise ValueError(f"unsupported PLY format line: {line!r}")
            fmt = words[1]
            if fmt not in ("binary_little_endian", "ascii"):
                raise ValueError(f"unsupported PLY format {fmt}")
        elif words[0] == "element":
            if len(words) != 3:
                raise ValueError(f"bad element line: {line!r}")
            elements.append([words[1], int(words[2]), []])
        elif words[0] == "property":
            if not elements:
                raise ValueError("property before any element")
            if words[1] == "list":
                raise ValueError("list properties are not supported")
            if len(words) != 3 or words[1] not in _TYPES:
                raise ValueError(f"bad property line: {line!r}")
            elements[-1][2].append((words[2], _TYPES[words[1]]))
        else:
            raise ValueError(f"unexpected PLY header line: {line!r}")
    if fmt is None:
        raise ValueError("PLY header has no format line")
    if not elements or elements[0][0] != "vertex":
        raise ValueError("the first PLY element must be 'vertex'")
    return fmt, elements[0]


def read_ply_table(path):
    """(property names, structured array of the vertex element)."""
    with open(path, "rb") as f:
        fmt, (_, n, props) = _parse_header(f)
        names = [p for p, _ in props]
        if len(set(names)) != len(names):
            raise ValueError("duplicate PLY property names")
        if fmt == "binary_little_endian":
            dt = np.dtype([(p, "<" + t) for p, t in props])
            raw = f.read(n * dt.itemsize)
            if len(raw) != n * dt.itemsize:
                raise ValueError("PLY file is shorter than its header says")
            table = np.frombuffer(raw, dtype=dt, count=n).copy()
        else:
            dt = np.dtype([(p, t) for p, t in props])
            table = np.empty(n, dtype=dt)
            for i in range(n):
                words = f.readline().split()
                if len(words) != len(props):
                    raise ValueError(f"PLY row {i} has {len(words)} values, expected {len(props)}")
                for (p, t), w in zip(props, words):
                    table[p][i] = float(w) if t[0] == "f" else int(w)
    return names, table


def _indexed(names, prefix):
    found = [p for p in names if p.startswith(prefix)]
    return sorted(found, key=lambda p: int(p.split("_")[-1]))


def read_ply(path):
    """The Gaussian parameters of a model file, as float32 arrays
    (R/scene/gaussian_model.py:312-394)."""
    names, t = read_ply_table(path)
    for p in ("x", "y", "z", "value", "weight"):
        if p not in names:
            raise ValueError(f"PLY vertex element has no '{p}' property")
    scale_names = _indexed(names, "scale_")
    rot_names = _indexed(names, "rot")
    if len(scale_names) != 3 or len(rot_names) != 4:
        raise ValueError("expected scale_0..2 and rot_0..3 properties")
    col = lambda ps: np.stack([np.asarray(t[p], dtype=np.float32) for p in ps], axis=1)
    return GaussianParams(means=col(["x", "y", "z"]), values=col(["value"]), weights=col(["weight"]),
                          scales=col(scale_names), rotations=col(rot_names))


def _rows(params):
    cols = [params.means, params.values, params.weights, params.scales, params.rotations]
    cols = [np.asarray(c, dtype=np.float32).reshape(params.count, -1) for c in cols]
    data = np.concatenate(cols, axis=1)
    if data.shape[1] != len(ATTRIBUTES):
        raise ValueError("expected means [N,3], values [N,1], weights [N,1], scales [N,3], rotations [N,4]")
    return data


def _header(n, fmt):
    lines = ["ply", f"format {fmt} 1.0", f"element vertex {n}"]
    lines += [f"property float {p}" for p in ATTRIBUTES]
    lines.append("end_header")
    return ("\n".join(lines) + "\n").encode("ascii")


def write_ply(path, params):
    """Binary little-endian file in the reference's layout (save_ply /
    save_ply_activated, R/scene/gaussian_model.py:265-303); the caller chooses
    raw or activated parameters."""
    data = _rows(params)
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    with open(path, "wb") as f:
        f.write(_header(data.shape[0], "binary_little_endian"))
        f.write(np.ascontiguousarray(data, dtype="<f4").tobytes())


def write_ply_ascii(binary_path):
    """The ASCII twin of a binary model file, ``<name>_ascii.ply`` beside it, as
    convert_ply_to_ascii writes it (R/scene/gaussian_model.py:801-819): every
    property declared float, each value as Python prints it.  Returns its path."""
    names, t = read_ply_table(binary_path)
    out = binary_path.replace(".ply", "_ascii.ply")
    with open(out, "w") as f:
        f.write("ply\n")
        f.write("format ascii 1.0\n")
        f.write(f"element vertex {t.shape[0]}\n")
        for p in names:
            f.write(f"property float {p}\n")
        f.write("end_header\n")
        for row in t:
            f.write(" ".join(str(v) for v in row) + "\n")
    return out


def to_torch(params, device):
    """The arrays as float32 device tensors (means, values, weights, scales,
    rotations)."""
    import torch
    return tuple(torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float32).to(device)
                 for a in (params.means, params.values, params.weights, params.scales, params.rotations))
