"""Scene-file oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, as the checker of tt_scene_unpack / tt_scene_pack and of
diff_tritracer.sceneio; the product path never does.

The reference has no scene file (it passes scenes as tensors to its autograd Function,
diff_tritracer/__init__.py:54-139); the format is this project's specification, restated
here with numpy structured dtypes:

    binary little-endian PLY
    element vertex V   float x, y, z (+ any other scalar properties)
    element face P     list uchar int vertex_indices, float opacity, float sigma,
                       float f_dc_0..2, float f_rest_0..3(M-1)-1 (+ other scalars, any order)
    shs[i][0][c] = f_dc_c;  shs[i][j][c] = f_rest_{c (M-1) + j - 1} for j >= 1
    vertices[3 i + k] = vertex[vertex_indices[i][k]]; npts = 3; cumsum = 3 i
    faces with a count other than 3 or an index outside [0, V) are "bad": their vertices read 0

parity: pinned by tests/test_oracle_plyio.py (files assembled field by field with struct,
round trips, the written bytes of a hand-assembled two-triangle scene).
"""
from __future__ import annotations

import numpy as np

_NP = {"char": "i1", "uchar": "u1", "int8": "i1", "uint8": "u1", "short": "<i2", "ushort": "<u2", "int16": "<i2",
       "uint16": "<u2", "int": "<i4", "uint": "<u4", "int32": "<i4", "uint32": "<u4", "float": "<f4",
       "float32": "<f4", "double": "<f8", "float64": "<f8"}


def _header(buf):
    end = buf.index(b"end_header\n") + len(b"end_header\n")
    elements = []
    for line in buf[:end].decode("ascii").splitlines():
        w = line.split()
        if w and w[0] == "element":
            elements.append([w[1], int(w[2]), []])
        elif w and w[0] == "property":
            if w[1] == "list":
                elements[-1][2].append((w[4], "u1", 0))   # count byte, then three int32
                elements[-1][2].append((w[4] + "__idx", "<i4", 3))
            else:
                elements[-1][2].append((w[2], _NP[w[1]], 0))
    return end, elements


def read(buf: bytes):
    """dict(vertices [3P, 3], npts, cumsum, opacity [P, 1], sigma [P, 1], shs [P, M, 3], M, bad_faces)"""
    pos, elements = _header(buf)
    arrays = {}
    for name, count, props in elements:
        dt = np.dtype([(n, t, (k,)) if k else (n, t) for n, t, k in props])
        arrays[name] = np.frombuffer(buf, dt, count, pos)
        pos += count * dt.itemsize
    v, f = arrays["vertex"], arrays["face"]
    V, P = len(v), len(f)
    names = f.dtype.names
    lst = "vertex_indices" if "vertex_indices" in names else "vertex_index"
    idx = f[lst + "__idx"].astype(np.int64)
    ok = (f[lst] == 3) & (idx >= 0).all(1) & (idx < V).all(1)
    xyz = np.stack([v["x"], v["y"], v["z"]], 1).astype(np.float32)
    verts = np.zeros((P, 3, 3), np.float32)
    verts[ok] = xyz[idx[ok]]
    n_rest = sum(1 for n in names if n.startswith("f_rest_"))
    M = 1 + n_rest // 3
    col = lambda n: f[n].astype(np.float32) if n in names else np.zeros(P, np.float32)
    shs = np.zeros((P, M, 3), np.float32)
    for c in range(3):
        shs[:, 0, c] = col(f"f_dc_{c}")
        for j in range(1, M):
            shs[:, j, c] = col(f"f_rest_{c * (M - 1) + j - 1}")
    return dict(vertices=verts.reshape(-1, 3), npts=np.full(P, 3, np.int32), cumsum=(3 * np.arange(P)).astype(np.int32),
                opacity=col("opacity").reshape(-1, 1), sigma=col("sigma").reshape(-1, 1), shs=shs, M=M,
                bad_faces=int((~ok).sum()))


def write(vertices, cumsum, opacity, sigma, shs) -> bytes:
    """The bytes of the canonical file (header text as diff_tritracer.sceneio writes it)."""
    P, M = len(cumsum), shs.shape[1]
    head = ["ply", "format binary_little_endian 1.0",
            "comment triangle splats: per-face opacity, sigma, SH (f_dc_c, f_rest_{c (M-1) + j - 1})",
            f"element vertex {3 * P}", "property float x", "property float y", "property float z",
            f"element face {P}", "property list uchar int vertex_indices", "property float opacity",
            "property float sigma"] + [f"property float f_dc_{c}" for c in range(3)] + \
           [f"property float f_rest_{k}" for k in range(3 * (M - 1))] + ["end_header"]
    verts = np.asarray(vertices, np.float32).reshape(-1, 3)
    rows = verts[np.asarray(cumsum, np.int64)[:, None] + np.arange(3)[None, :]].reshape(-1, 3)
    dt = np.dtype([("n", "u1"), ("idx", "<i4", (3,)), ("opacity", "<f4"), ("sigma", "<f4"), ("dc", "<f4", (3,)),
                   ("rest", "<f4", (3 * (M - 1),))])
    f = np.zeros(P, dt)
    f["n"] = 3
    f["idx"] = 3 * np.arange(P)[:, None] + np.arange(3)[None, :]
    f["opacity"] = np.asarray(opacity, np.float32).reshape(-1)
    f["sigma"] = np.asarray(sigma, np.float32).reshape(-1)
    s = np.asarray(shs, np.float32).reshape(P, M, 3)
    f["dc"] = s[:, 0, :]
    f["rest"] = s[:, 1:, :].transpose(0, 2, 1).reshape(P, -1)  # channel-major
    return ("\n".join(head) + "\n").encode("ascii") + rows.astype("<f4").tobytes() + f.tobytes()
