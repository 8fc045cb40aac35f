"""Integrator input rules of the reference's batch API (dioptrin.pyi:845-904,
src/python/integrator.rs:639-790, 1867-2210), checked before any device work:
bad arrays raise for the whole call, unreadable or mismatched files report
their error in their own BatchResult, named by path."""
import numpy as np
import pytest

from dioptrin_amd import BatchResult, Integrator

Image = pytest.importorskip("PIL.Image")

PONI = dict(pixel1=1e-4, pixel2=1e-4, distance=0.05, poni1=4.5e-4, poni2=5.5e-4,
            wavelength=1e-10)


def integrator(mask_shape=None):
    ai = Integrator(dict(PONI))
    if mask_shape:
        ai.set_mask(np.zeros(mask_shape, np.uint8))
    return ai


def test_array_rules_raise_for_the_call():
    ai = integrator((9, 11))
    good = np.ones((9, 11), np.float32)
    with pytest.raises(TypeError):
        ai.batch1d("frame.tif", 10)                      # a single path
    with pytest.raises(TypeError):
        ai.batch1d([good, "frame.tif"], 10)              # paths and arrays mixed
    with pytest.raises(ValueError):
        ai.batch1d([good, np.ones((9, 12), np.float32)], 10)   # mask shape
    with pytest.raises(ValueError):
        ai.batch1d([np.ones((22, 11), np.float32)[::2]], 10)   # not C-contiguous
    with pytest.raises(TypeError):
        ai.batch1d([good.astype(np.int64)], 10)          # unsupported dtype
    with pytest.raises(TypeError):
        ai.batch1d([np.ones(11, np.float32)], 10)        # not 2D
    with pytest.raises(ValueError):
        ai.batch1d([good], 0)
    with pytest.raises(ValueError):
        integrator().batch1d([np.ones((0, 11), np.float32)], 10)  # empty
    assert ai.batch1d([], 10) == []


def test_file_errors_stay_in_their_result(tmp_path):
    ai = integrator((9, 11))
    missing = str(tmp_path / "missing.tif")
    corrupt = tmp_path / "corrupt.tif"
    corrupt.write_bytes(b"II*\x00garbage")
    wrong = str(tmp_path / "wrong.tif")
    Image.fromarray(np.ones((9, 12), np.uint16)).save(wrong)
    rgb = str(tmp_path / "rgb.tif")
    Image.fromarray(np.zeros((9, 11, 3), np.uint8)).save(rgb)
    res = ai.batch1d([missing, str(corrupt), wrong, rgb], 10)
    assert [r.filename for r in res] == [missing, str(corrupt), wrong, rgb]
    assert all(isinstance(r, BatchResult) and r.is_err() and r.result is None for r in res)
    assert "mask shape" in res[2].error and "single-channel" in res[3].error
