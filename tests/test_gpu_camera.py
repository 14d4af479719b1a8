"""GPU camera-ray generation (tt_camera_rays) against the host generator
scenes.look_at_rays: bit-exact rays for several poses, sizes and fields of view."""
import numpy as np
import pytest

from diff_tritracer import capi, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("W,H,radius,seed,fov", [(1920, 1080, 4.0, 1, 60.0), (97, 61, 3.0, 2, 45.0),
                                                 (640, 480, 5.0, 3, 75.0), (1, 1, 2.0, 4, 60.0)])
def test_camera_rays_bit_exact(cuda, W, H, radius, seed, fov):
    import torch

    eye = scenes.orbit_eye(radius, seed)
    ro_ref, rd_ref = scenes.look_at_rays(W, H, eye, fov_deg=fov)
    b = scenes.camera_basis(W, H, eye, fov_deg=fov)
    ro = torch.full((H, W, 3), np.nan, device=cuda)
    rd = torch.full((H, W, 3), np.nan, device=cuda)
    capi.camera_rays(b, W, H, ro, rd)
    torch.cuda.synchronize()
    assert np.array_equal(ro.cpu().numpy().view(np.uint32), ro_ref.view(np.uint32))
    bad = np.count_nonzero(rd.cpu().numpy().view(np.uint32) != rd_ref.view(np.uint32))
    print(f"[gpu-camera] {W}x{H}: {bad} direction components differ", flush=True)
    assert bad == 0
