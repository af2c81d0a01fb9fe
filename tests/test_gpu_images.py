"""GPU parity of image input (read_pnm / read_png / read_image,
stereonet.py:262-324, :3421-3613): the lists the package returns, and the
float64 planes read_image_gpu leaves on the GPU, bit-identical to the
reference's own outputs (tests/golden/images.npz) and, for frame-sized
files, to the oracle."""
import os
import sys

import numpy as np
import pytest
import torch

import oracle
from conftest import GOLDEN, package

sys.path.insert(0, GOLDEN)
import make_images  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sv():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return package()


def _golden():
    return np.load(os.path.join(GOLDEN, "images.npz"))


def test_read_image_golden(sv, tmp_path):
    z = _golden()
    for n in z["names"]:
        p = tmp_path / str(n)
        p.write_bytes(z["file_" + n].tobytes())
        gray, rgb, W, H = sv.read_image(str(p))
        assert isinstance(gray, list) and isinstance(rgb, list)
        assert [W, H] == list(z["wh_" + n]), n
        assert np.array_equal(np.array(gray), z["gray_" + n]), n
        assert np.array_equal(np.array(rgb), z["rgb_" + n]), n
        rd = sv.read_png if str(n).startswith("png") else sv.read_pnm
        g2, r2, _, _ = rd(str(p))
        assert g2 == gray and r2 == rgb


def test_read_image_gpu_golden(sv):
    z = _golden()
    for n in z["names"]:
        gray, rgb, W, H = sv.read_image_gpu(z["file_" + n].tobytes())
        assert gray.is_cuda and gray.dtype == torch.float64 and tuple(gray.shape) == (H, W)
        assert tuple(rgb.shape) == (H, W, 3)
        assert np.array_equal(gray.cpu().numpy().ravel(), z["gray_" + n]), n
        assert np.array_equal(rgb.cpu().numpy().ravel(), z["rgb_" + n]), n


@pytest.mark.parametrize("kind", ["png_rgb8", "png_rgba16", "png_pal4", "png_gray1", "ppm16",
                                  "pgm_ascii"])
def test_frame_sized_vs_oracle(sv, kind):
    rng = np.random.default_rng(len(kind))
    W, H = 460, 370
    f = {
        "png_rgb8": lambda: make_images.make_png(rng, W, H, 2, 8, "dynamic"),
        "png_rgba16": lambda: make_images.make_png(rng, W, H, 6, 16, "fixed"),
        "png_pal4": lambda: make_images.make_png(rng, W, H, 3, 4, "dynamic", 11),
        "png_gray1": lambda: make_images.make_png(rng, W, H, 0, 1, "stored"),
        "ppm16": lambda: make_images.make_pnm(rng, W, H, "P6", 65535),
        "pgm_ascii": lambda: make_images.make_pnm(rng, W, H, "P2", 4095),
    }[kind]()
    gray, rgb, W1, H1 = sv.read_image_gpu(f)
    g0, r0, W0, H0 = oracle.read_image_bytes(f)
    assert (W1, H1) == (W0, H0) == (W, H)
    assert np.array_equal(gray.cpu().numpy().ravel(), g0)
    assert np.array_equal(rgb.cpu().numpy().ravel(), r0)


def test_batch_matches_single(sv):
    rng = np.random.default_rng(7)
    files = [make_images.make_png(rng, 97, 41, 2, 8, "dynamic") for _ in range(5)]
    files += [make_images.make_pnm(rng, 97, 41, "P5", 255) for _ in range(3)]
    gray, rgb, W, H = sv.read_images_gpu(files, threads=4)
    assert tuple(gray.shape) == (8, 41, 97) and tuple(rgb.shape) == (8, 41, 97, 3)
    for i, f in enumerate(files):
        g1, r1, _, _ = sv.read_image_gpu(f)
        assert torch.equal(gray[i], g1) and torch.equal(rgb[i], r1)
    with pytest.raises(ValueError, match="same size"):
        sv.read_images_gpu([files[0], make_images.make_pnm(rng, 5, 5, "P5", 255)])
