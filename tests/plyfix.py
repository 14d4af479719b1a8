"""Scene files assembled field by field with struct (independent of the oracle's and the
product's PLY code), with the arrays they must decode to."""
import struct

import numpy as np


def shared_vertex_file():
    """4 shared vertices with extra properties (uchar red, float nx), 3 faces with the
    properties in a scrambled order, an extra int property, M = 4 (9 f_rest), and one face
    whose index lies outside [0, V). Returns (bytes, expected dict)."""
    V = [(0.0, 0.0, 0.0), (1.0, 0.0, 0.5), (0.0, 1.0, -0.25), (1.5, 1.25, 2.0)]
    faces = [(0, 1, 2), (1, 3, 2), (2, 3, 7)]
    M = 4
    rng = np.random.default_rng(3)
    sig = rng.random(3).astype(np.float32)
    opa = rng.random(3).astype(np.float32)
    dc = rng.standard_normal((3, 3)).astype(np.float32)
    rest = rng.standard_normal((3, 9)).astype(np.float32)
    head = ["ply", "format binary_little_endian 1.0", "comment hand-assembled", "element vertex 4",
            "property float x", "property uchar red", "property float y", "property float z", "property float nx",
            "element face 3", "property float sigma", "property list uchar int vertex_indices",
            "property float f_dc_2", "property float f_dc_0", "property float f_dc_1", "property float opacity",
            "property int foo"] + [f"property float f_rest_{k}" for k in range(9)] + ["end_header"]
    body = b""
    for i, (x, y, z) in enumerate(V):
        body += struct.pack("<fBfff", x, 200 + i, y, z, -1.0)
    for i, (a, b, c) in enumerate(faces):
        body += struct.pack("<fBiiiffffi", sig[i], 3, a, b, c, dc[i, 2], dc[i, 0], dc[i, 1], opa[i], 1234)
        body += struct.pack("<9f", *rest[i])
    data = ("\n".join(head) + "\n").encode() + body
    verts = np.zeros((3, 3, 3), np.float32)
    for i, f in enumerate(faces[:2]):
        verts[i] = np.float32([V[k] for k in f])
    shs = np.zeros((3, M, 3), np.float32)
    shs[:, 0, :] = dc
    for c in range(3):
        for j in range(1, M):
            shs[:, j, c] = rest[:, c * (M - 1) + j - 1]
    return data, dict(vertices=verts.reshape(-1, 3), opacity=opa.reshape(-1, 1), sigma=sig.reshape(-1, 1), shs=shs,
                      M=M, bad_faces=1)


def two_triangle_file():
    """The canonical bytes of a 2-triangle scene with M = 2, triangle 1 stored first in the
    vertex array (cumsum = [3, 0]). Returns (bytes, inputs)."""
    verts = np.arange(18, dtype=np.float32).reshape(6, 3) * 0.5 - 1.0
    cumsum = np.int32([3, 0])
    opa = np.float32([0.25, 0.75])
    sig = np.float32([1.5, 2.5])
    shs = (np.arange(12, dtype=np.float32).reshape(2, 2, 3) - 5.0) / 8.0
    head = ["ply", "format binary_little_endian 1.0",
            "comment triangle splats: per-face opacity, sigma, SH (f_dc_c, f_rest_{c (M-1) + j - 1})",
            "element vertex 6", "property float x", "property float y", "property float z", "element face 2",
            "property list uchar int vertex_indices", "property float opacity", "property float sigma",
            "property float f_dc_0", "property float f_dc_1", "property float f_dc_2", "property float f_rest_0",
            "property float f_rest_1", "property float f_rest_2", "end_header"]
    body = b""
    for i in range(2):
        for k in range(3):
            body += struct.pack("<3f", *verts[cumsum[i] + k])
    for i in range(2):
        body += struct.pack("<B3i", 3, 3 * i, 3 * i + 1, 3 * i + 2)
        body += struct.pack("<2f", opa[i], sig[i])
        body += struct.pack("<3f", *shs[i, 0])
        body += struct.pack("<3f", shs[i, 1, 0], shs[i, 1, 1], shs[i, 1, 2])
    return ("\n".join(head) + "\n").encode() + body, dict(vertices=verts, cumsum=cumsum, opacity=opa, sigma=sig,
                                                          shs=shs)


def random_scene(P, M, seed, shuffle=True):
    rng = np.random.default_rng(seed)
    verts = rng.standard_normal((3 * P, 3)).astype(np.float32)
    cumsum = (3 * (rng.permutation(P) if shuffle else np.arange(P))).astype(np.int32)
    return dict(vertices=verts, cumsum=cumsum, opacity=rng.random((P, 1)).astype(np.float32),
                sigma=rng.random((P, 1)).astype(np.float32),
                shs=rng.standard_normal((P, M, 3)).astype(np.float32))
