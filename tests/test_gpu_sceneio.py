"""Scene files through the GPU (tt_scene_unpack / tt_scene_pack via diff_tritracer.sceneio)
against the oracle (oracle/plyio.py). Bars: bit-exact arrays and bit-exact file bytes.
  - a hand-assembled file (shared vertices, extra and reordered properties, unaligned
    vertex rows, an out-of-range index) loads to the oracle's arrays and reports the bad face;
  - save_ply writes the oracle's bytes (a scene whose cumsum is shuffled, M = 16 and M = 1);
  - save -> load returns the scene in index order (300K triangles), and the loaded scene
    renders like the original.
"""
import numpy as np
import pytest

import plyio as OP
from diff_tritracer import sceneio
from helpers import to_dev
from plyfix import random_scene, shared_vertex_file

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-sceneio]", *a, flush=True)


def _eq(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def test_load_hand_assembled_file(cuda, tmp_path):
    data, want = shared_vertex_file()
    p = tmp_path / "s.ply"
    p.write_bytes(data)
    _say("load hand-assembled file")
    with pytest.raises(ValueError, match="index outside"):
        sceneio.load_ply(p, cuda)
    got = sceneio.load_ply(p, cuda, allow_bad=True)
    ref = OP.read(data)
    for k in ("vertices", "opacity", "sigma", "shs"):
        assert _eq(got[k].cpu().numpy(), ref[k]), k
        assert _eq(got[k].cpu().numpy(), want[k]), k
    assert got["bad_faces"] == 1 and got["M"] == 4
    assert got["npts"].cpu().tolist() == [3, 3, 3] and got["cumsum"].cpu().tolist() == [0, 3, 6]


@pytest.mark.parametrize("P,M", [(1, 16), (129, 1), (100_003, 16)])
def test_save_writes_the_oracle_bytes(cuda, tmp_path, P, M):
    s = random_scene(P, M, 7 + P)
    p = tmp_path / "o.ply"
    _say(f"save P={P} M={M}")
    d = {k: to_dev(v, cuda) for k, v in s.items()}
    sceneio.save_ply(p, d["vertices"], d["cumsum"], d["opacity"], d["sigma"], d["shs"])
    assert p.read_bytes() == OP.write(s["vertices"], s["cumsum"], s["opacity"], s["sigma"], s["shs"])


def test_save_load_round_trip_and_render(cuda, tmp_path):
    import torch
    from diff_tritracer import scenes
    from helpers import GpuScene

    s = random_scene(300_000, 16, 11)
    d = {k: to_dev(v, cuda) for k, v in s.items()}
    p = tmp_path / "r.ply"
    _say("round trip 300K")
    sceneio.save_ply(p, d["vertices"], d["cumsum"], d["opacity"], d["sigma"], d["shs"])
    got = sceneio.load_ply(p, cuda)
    want_v = s["vertices"][s["cumsum"][:, None] + np.arange(3)[None, :]].reshape(-1, 3)
    assert _eq(got["vertices"].cpu().numpy(), want_v)
    for k in ("opacity", "sigma", "shs"):
        assert _eq(got[k].cpu().numpy(), s[k]), k
    # a scene saved and loaded renders bit-identically
    scene, ro, rd = scenes.small_config(0, 64, 64)
    d = {k: to_dev(v, cuda) for k, v in (("vertices", scene.verts), ("cumsum", scene.cumsum),
                                         ("opacity", scene.opac), ("sigma", scene.sigma), ("shs", scene.shs))}
    sceneio.save_ply(tmp_path / "c.ply", d["vertices"], d["cumsum"], d["opacity"], d["sigma"], d["shs"])
    g = sceneio.load_ply(tmp_path / "c.ply", cuda)
    back = scenes.SplatScene(g["vertices"].cpu().numpy(), g["npts"].cpu().numpy(), g["cumsum"].cpu().numpy(),
                             g["shs"].cpu().numpy(), g["opacity"].cpu().numpy(), g["sigma"].cpu().numpy())
    a = GpuScene(scene, cuda).forward(ro, rd, chunk_mode="exact")
    b = GpuScene(back, cuda).forward(ro, rd, chunk_mode="exact")
    torch.cuda.synchronize()
    assert _eq(a["color"], b["color"]) and _eq(a["others"], b["others"])
