"""TIFF frames (Frame::from_tiff, src/frame.rs:32-190): single-channel
images of every sample type, uncompressed, PackBits, LZW or Deflate, in
strips, returned flipped up-down like the reference's convert_and_flip_into.
The decoder is checked against PIL/libtiff, an independent encoder and
decoder, on files it writes and on the reference's own ceo2_side.tif."""
import os

import numpy as np
import pytest

from dioptrin_amd.tiff import TiffError, read_tiff

Image = pytest.importorskip("PIL.Image")

HERE = os.path.dirname(os.path.abspath(__file__))
CEO2 = os.path.join(HERE, "golden", "tiff", "ceo2_side.tif")


def samples():
    rng = np.random.default_rng(3)
    base = rng.integers(0, 60000, size=(37, 53))
    base[5:9, 10:30] = 7  # runs PackBits and LZW compress
    return [(base.astype(np.uint16), None), ((base - 30000).astype(np.int32), "I"),
            ((base / 7.0).astype(np.float32), "F"), ((base % 256).astype(np.uint8), "L")]


@pytest.mark.parametrize("compression", [None, "packbits", "tiff_lzw", "tiff_deflate",
                                         "tiff_adobe_deflate"])
def test_decoder_matches_libtiff_writer(tmp_path, compression):
    for arr, mode in samples():
        path = str(tmp_path / f"im_{compression}_{arr.dtype}.tif")
        im = Image.fromarray(arr, mode=mode) if mode else Image.fromarray(arr)
        im.save(path, compression=compression)
        got = read_tiff(path)
        assert got.dtype == arr.dtype, (compression, arr.dtype)
        assert np.array_equal(got, arr[::-1]), (compression, arr.dtype)


def test_multi_strip_and_predictor(tmp_path):
    arr = samples()[0][0]
    path = str(tmp_path / "strips.tif")
    # libtiff: 5 rows per strip, LZW with horizontal differencing.
    Image.fromarray(arr).save(path, compression="tiff_lzw", tiffinfo={278: 5, 317: 2})
    assert np.array_equal(read_tiff(path), arr[::-1])
    path = str(tmp_path / "strips_deflate.tif")
    Image.fromarray(arr).save(path, compression="tiff_deflate", tiffinfo={278: 7})
    assert np.array_equal(read_tiff(path), arr[::-1])


def test_reference_ceo2_side_tiff():
    """The reference's test image (Pilatus strip, 195 x 1475 int32,
    uncompressed): our decoder gives PIL's pixels, rows flipped."""
    got = read_tiff(CEO2)
    ref = np.array(Image.open(CEO2))
    assert got.shape == (195, 1475) and got.dtype == np.int32
    assert np.array_equal(got, ref[::-1].astype(np.int32))


def test_rejects_colour_and_non_tiff(tmp_path):
    rgb = str(tmp_path / "rgb.tif")
    Image.fromarray(np.zeros((4, 5, 3), np.uint8)).save(rgb)
    with pytest.raises(TiffError, match="single-channel"):
        read_tiff(rgb)
    bad = tmp_path / "bad.tif"
    bad.write_bytes(b"not a tiff at all")
    with pytest.raises(TiffError):
        read_tiff(str(bad))
