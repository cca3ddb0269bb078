"""Host-side input/output (CPU only): the OBJ reader's primitive order,
material mapping and normals (core/scene.cpp:32-68), the PPM P3 writer and
display transform (core/image_buffer.cpp:22-47), and the ray casters'
post-processing (integrators/*_raycaster.cpp), on hand-made inputs whose
outputs follow from those files line by line."""
import numpy as np
import pytest

from swpathtracer_amd import host_io as H

OBJ = """# two meshes
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vn 0 0 1
o floor
f 1//1 2//1 3//1 4//1
o wall
f -4 -3 -2
f 1 1 2
"""


def _write(tmp_path, text, name="m.obj"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_obj_fan_triangulation_order_and_normals(tmp_path):
    meshes = H.read_obj(_write(tmp_path, OBJ))
    assert [m[0] for m in meshes] == ["floor", "wall"]
    floor, wall = meshes[0], meshes[1]
    # quad 1 2 3 4 -> fan (1 2 3), (1 3 4)
    np.testing.assert_array_equal(floor[1], [[0, 0, 0, 1, 0, 0, 1, 1, 0], [0, 0, 0, 1, 1, 0, 0, 1, 0]])
    np.testing.assert_array_equal(floor[2], np.tile([0, 0, 1], (2, 3)))
    # negative indices relative to the end; the degenerate face (1 1 2) is dropped
    np.testing.assert_array_equal(wall[1], [[0, 0, 0, 1, 0, 0, 1, 1, 0]])
    assert not np.any(wall[2])  # no vn: flat triangle rows


def test_obj_material_mapping(tmp_path):
    path = _write(tmp_path, OBJ)
    sc = H.load_obj(path, default_material_id=7)
    assert sc.n_tris == 3 and list(sc.mat_ids) == [7, 7, 7]
    sc = H.load_obj(path, default_material_id=7, submeshes={"wall": 2})
    assert list(sc.mat_ids) == [7, 7, 2]
    sc = H.load_obj(path, default_material_id=7, render_all_submeshes=False, submeshes={"wall": 2})
    assert sc.n_tris == 1 and list(sc.mat_ids) == [2]
    assert sc.normals is None  # only flat faces left


def test_ppm_text_and_display_transform(tmp_path):
    img = np.array([[[0.0, 0.5, 1.0], [2.0, -1.0, 0.25]]], np.float32)  # 1 row, 2 pixels
    p = tmp_path / "o.ppm"
    H.write_ppm(str(p), img)
    # int(v * 255 + 0.5) after clamping: 0, 128, 255, 255, 0, 64
    assert p.read_text() == "P3\n2 1\n255\n0 128 255 255 0 64 "
    d = H.to_display(np.array([0.25, 1.5, -0.5], np.float32))
    np.testing.assert_allclose(d, [0.25 ** (1 / 2.2), 1.0, 0.0], rtol=1e-6)


def test_normal_and_depth_post_processing():
    n = np.array([[[0, 0, 0], [0, -1, 0]]], np.float32)
    np.testing.assert_array_equal(H.normal_image(n), [[[0, 0, 0], [0.5, 0.0, 0.5]]])
    d = np.array([[-1.0, 2.0, 4.0, 3.0]], np.float32)
    g = H.depth_image(d)[..., 0]
    # min over positive depths (2), max over all (4); background black
    np.testing.assert_allclose(g, [[0.0, 0.0, 1.0, 0.5]])
    g = H.depth_image(d, prescribed_min=0.0, prescribed_max=8.0)[..., 0]
    np.testing.assert_allclose(g, [[0.0, 0.25, 0.5, 0.375]])
    np.testing.assert_allclose(H.depth_image(np.array([[-1.0, 5.0]], np.float32))[..., 0], [[0.0, 1.0]])


def test_count_heatmap_endpoints():
    """t = 0 is HSL (0.65, 1, 0) = black, t = 1 is HSL (0, 1, 0.5) = pure red;
    t = 0.5 is HSL (0.325, 1, 0.25): v2 = 0.5, v1 = 0, a green-dominant colour."""
    rgb = H.count_heatmap(np.array([[10, 20, 15]], np.uint64))
    np.testing.assert_allclose(rgb[0, 0], [0, 0, 0], atol=1e-7)
    np.testing.assert_allclose(rgb[0, 1], [1, 0, 0], atol=1e-6)  # float h - 1/3 + 1 lands next to 2/3
    h = 0.325
    expect_r = 0.0 + 0.5 * (2 / 3 - (h + 1 / 3)) * 6  # 3 (h + 1/3) < 2 branch
    np.testing.assert_allclose(rgb[0, 2], [expect_r, 0.5, 0.0], atol=1e-6)
    flat = H.count_heatmap(np.array([[7, 7]], np.uint64))  # zero range, max > 0: t = 1
    np.testing.assert_allclose(flat[0, 0], [1, 0, 0], atol=1e-6)
    zero = H.count_heatmap(np.array([[0, 0]], np.uint64))
    np.testing.assert_allclose(zero, 0.0, atol=1e-7)
