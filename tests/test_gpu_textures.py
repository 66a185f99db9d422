"""Texture files in the formats the decoder reads besides PNG (BMP, TGA,
PNM): the indexed textured mesh drawn on the GPU with the noise texture
stored as each of them gives the same frame as with the PNG, and the
oracle's frame (the decoders themselves are checked byte for byte against
the reference's stb_image on the CPU, tests/test_image_formats.py)."""
import numpy as np
import pytest

from scene_runner import write_noise_png
from test_gpu_parity import assert_same, run_both

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def noise_files(tmp_path_factory):
    from PIL import Image
    d = tmp_path_factory.mktemp("texfmt")
    png = write_noise_png(str(d / "noise.png"), 256, 0x5EED0002)
    img = Image.open(png).convert("RGB")
    out = {"png": png}
    for ext, kw in (("bmp", {}), ("tga", {"compression": "tga_rle"}), ("ppm", {})):
        p = str(d / f"noise.{ext}")
        img.save(p, **kw)
        out[ext] = p
    return out


@pytest.mark.parametrize("fmt", ["bmp", "tga", "ppm"])
def test_mesh_with_texture_file_format(oracle, gpu, noise_files, fmt):
    kw = dict(width=480, height=270, p0=32, p1=64)
    ref = gpu.run("cfg/mesh", tex_path=noise_files["png"], **kw)
    a, b, sa, sb = run_both(oracle, gpu, "cfg/mesh", tex_path=noise_files[fmt], **kw)
    assert_same(a, b, sa, sb, f"cfg/mesh with a {fmt} texture")
    assert np.array_equal(b.color, ref.color)
