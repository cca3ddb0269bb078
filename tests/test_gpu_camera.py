"""Primary rays: PinholeCamera / OrthographicCamera::getRay on the device vs the
oracle (cameras/pinhole_camera.cpp:14-18, orthographic_camera.cpp:21-25),
bit-exact."""
import numpy as np
import pytest

import pyoracle as O
from swpathtracer_amd import api

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cam", [
    api.pinhole(96, 72, position=(0, 0, 3.6), look_at=(0, 0, 0), fov=40.0),
    api.pinhole(33, 17, position=(1.5, -2.0, 4.0), look_at=(0.2, 0.1, -0.3), up=(0.1, 1, 0.2), fov=70.0),
    api.orthographic(40, 30, -1.25, 1.25, -1.0, 1.5, position=(0.5, 0.5, 6)),
])
def test_camera_rays_bit_exact(device, cam):
    g = device.camera_rays(cam)
    c = np.array([O.camera_ray(cam, x + 0.5, y + 0.5) for y in range(cam.height) for x in range(cam.width)])
    assert np.array_equal(g.view(np.uint32), c.view(np.uint32))
