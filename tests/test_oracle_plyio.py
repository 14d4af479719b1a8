"""Pins of the scene-file oracle (oracle/plyio.py) and the host-side PLY header parser,
CPU only: a file assembled field by field (shared vertices, extra and reordered
properties, an out-of-range index) decodes to the arrays it was built from; the writer's
bytes equal a hand-assembled two-triangle file; read(write(scene)) is the scene in index
order."""
import io

import numpy as np

import plyio as OP
from diff_tritracer import sceneio
from plyfix import random_scene, shared_vertex_file, two_triangle_file


def _eq(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def test_read_shared_vertex_file():
    data, want = shared_vertex_file()
    got = OP.read(data)
    for k in ("vertices", "opacity", "sigma", "shs"):
        assert _eq(got[k], want[k]), k
    assert got["M"] == 4 and got["bad_faces"] == 1
    assert got["npts"].tolist() == [3, 3, 3] and got["cumsum"].tolist() == [0, 3, 6]


def test_write_matches_hand_assembled_bytes():
    data, inp = two_triangle_file()
    assert OP.write(inp["vertices"], inp["cumsum"], inp["opacity"], inp["sigma"], inp["shs"]) == data
    # and the product's header text is the same
    assert data.startswith(sceneio.header_text(2, 2))


def test_round_trip():
    s = random_scene(500, 16, 4)
    got = OP.read(OP.write(s["vertices"], s["cumsum"], s["opacity"], s["sigma"], s["shs"]))
    want_v = s["vertices"][s["cumsum"][:, None] + np.arange(3)[None, :]].reshape(-1, 3)
    assert _eq(got["vertices"], want_v) and _eq(got["shs"], s["shs"]) and _eq(got["opacity"], s["opacity"])
    assert _eq(got["sigma"], s["sigma"]) and got["bad_faces"] == 0


def test_product_header_parser():
    data, _ = shared_vertex_file()
    f = io.BytesIO(data)
    fmt, elements = sceneio.read_header(f)
    assert fmt == "binary_little_endian"
    assert [(n, c) for n, c, _ in elements] == [("vertex", 4), ("face", 3)]
    assert elements[1][2][1] == ("vertex_indices", ("uchar", "int"))
    assert f.tell() == data.index(b"end_header\n") + len(b"end_header\n")
    offs, size = sceneio._layout(elements[1][2], "face")
    assert size == 4 + 13 + 12 + 4 + 4 + 36 and offs["opacity"] == (29, "float")
