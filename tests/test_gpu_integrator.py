"""The Python drop-in surface on the GPU: Integrator.integrate1d /
integrate1d_with_errors / integrate1d_with_azimuthal_std / integrate2d /
batch1d / batch2d (dioptrin.pyi:597-1063, src/python/integrator.rs:
1867-2210, 2423) against GpuPlan (bit for bit) and the FP64 oracle (the
gather tolerance), BatchResult order and per-file errors, and the
reference's own CeO2 image (tests/golden/tiff/ceo2_side.tif with
tests/golden/poni/ceo2_side.poni) as real-data parity."""
import os

import numpy as np
import pytest

import oracle as O
from dioptrin_amd import GpuPlan, Integrator, load_poni
from dioptrin_amd.tiff import read_tiff

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CEO2_TIF = os.path.join(HERE, "golden", "tiff", "ceo2_side.tif")
CEO2_PONI = os.path.join(HERE, "golden", "poni", "ceo2_side.poni")
PONI = dict(pixel1=172e-6, pixel2=172e-6, distance=0.05, poni1=24 * 172e-6,
            poni2=36 * 172e-6, rot1=0.01, rot2=-0.02, wavelength=1e-10)
SHAPE = (64, 80)


def oracle_1d(poni, frame, bins, mask=None, pol=None, method="pixel_split"):
    P = O.Poni(**{k: poni[k] for k in ("pixel1", "pixel2", "distance", "poni1", "poni2",
                                        "rot1", "rot2", "rot3", "wavelength") if k in poni})
    lut = O.build_lut1d(P, *frame.shape, bins, method)
    norm = O.normalization(P, *frame.shape, pol)
    fixed = O.fixed_table(lut, mask) if mask is not None else lut
    den = O.denominators(fixed, norm)
    x = frame.astype(np.float64)
    i, s = O.gather(fixed, den, x, True)
    mag, _ = O.gather(fixed, den, np.abs(x), False)
    return lut, norm, i, s, np.abs(mag)


def close(a, b, mag, tol=2e-6):
    return np.all(np.abs(a - b) <= tol * mag)


def test_integrator_1d_entry_points():
    rng = np.random.default_rng(1)
    frame = rng.poisson(300.0, size=SHAPE).astype(np.float32)
    mask = (rng.random(SHAPE) < 0.05).astype(np.uint8)
    ai = Integrator(dict(PONI), polarization_factor=0.95)
    ai.set_mask(mask)
    plan = GpuPlan(PONI, SHAPE, n_radial=40, polarization_factor=0.95, mask=mask)
    lut, norm, i_ref, s_ref, mag = oracle_1d(PONI, frame, 40, mask, 0.95)
    r = ai.integrate1d(frame, 40)
    I, S, _ = plan.integrate_stack(frame[None], errors=True)
    assert np.array_equal(r.intensity, I[0]) and r.errors is None
    assert np.array_equal(r.radial, lut.bins())
    assert close(r.intensity, i_ref, mag)
    r = ai.integrate1d_with_errors(frame, 40)
    assert np.array_equal(r.errors, S[0]) and close(r.errors, s_ref, s_ref)
    r = ai.integrate1d_with_azimuthal_std(frame, 40, errors=True)
    i2, s2, d2 = O.azimuthal_std(lut, mask, norm, frame.astype(np.float64), errors=True)
    ok = ~np.isnan(d2)
    assert np.array_equal(np.isnan(r.azimuthal_std), ~ok)
    for a, b in ((r.intensity, i2), (r.errors, s2), (r.azimuthal_std, d2)):
        assert np.allclose(a[ok], b[ok], rtol=1e-6, atol=0)
    with pytest.raises(ValueError):
        ai.integrate1d(frame[:10], 40)   # mask shape


def test_integrator_2d_entry_point():
    frame = np.random.default_rng(2).poisson(500.0, size=SHAPE).astype(np.uint32)
    ai = Integrator(dict(PONI), method="pixel_split")
    r = ai.integrate2d(frame, 32, 36, errors=True)
    assert r.intensity.shape == (36, 32) and r.errors.shape == (36, 32)
    assert r.radial.shape == (32,) and r.azimuthal.shape == (36,)
    P = O.Poni(**PONI)
    lut, _ = O.build_lut2d(P, *SHAPE, 32, 36)
    den = O.denominators(lut, O.normalization(P, *SHAPE, None))
    i, s = O.gather(lut, den, frame.astype(np.float64))
    assert close(r.intensity.ravel(), i, np.abs(i)) and close(r.errors.ravel(), s, s)
    res = ai.batch2d([frame, frame.astype(np.float32)], 32, 36, errors=True)
    assert [x.filename for x in res] == ["array[0]", "array[1]"]
    assert np.array_equal(res[0].result.intensity, r.intensity)
    assert np.array_equal(res[1].result.intensity, r.intensity)


def test_batch1d_order_shapes_and_dtypes():
    """Frames of two shapes and three dtypes, interleaved: one result per
    frame in input order, each the bits of integrate1d on that frame."""
    rng = np.random.default_rng(3)
    ai = Integrator(dict(PONI))
    frames = []
    for k in range(7):
        shape = SHAPE if k % 2 == 0 else (40, 50)
        dt = (np.float32, np.uint16, np.uint32)[k % 3]
        frames.append(rng.poisson(100.0 + k, size=shape).astype(dt))
    res = ai.batch1d(frames, 24, errors=True, azimuthal_std=True)
    assert [r.filename for r in res] == [f"array[{k}]" for k in range(7)]
    for k, r in enumerate(res):
        one = ai.integrate1d(frames[k], 24, errors=True, azimuthal_std=True)
        assert r.is_ok() and np.array_equal(r.result.intensity, one.intensity), k
        assert np.array_equal(r.result.errors, one.errors), k
        assert np.array_equal(r.result.azimuthal_std, one.azimuthal_std, equal_nan=True), k
    stack = np.stack([f for f in frames if f.shape == SHAPE and f.dtype == np.float32])
    res3 = ai.batch1d(stack, 24)
    assert np.array_equal(res3[0].result.intensity, ai.integrate1d(frames[0], 24).intensity)


def test_batch1d_tiff_paths(tmp_path):
    """TIFF paths: each file's frame (flipped up-down on reading, as the
    reference does) gives the bits of the same array; unreadable files and
    files of another shape than the mask report errors in their own
    results without stopping the others."""
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(4)
    arrays = [rng.poisson(200.0, size=SHAPE).astype(np.uint16) for _ in range(3)]
    paths = []
    for k, a in enumerate(arrays):
        p = str(tmp_path / f"f{k}.tif")
        Image.fromarray(a).save(p, compression="tiff_lzw")
        paths.append(p)
    bad = str(tmp_path / "bad.tif")
    with open(bad, "wb") as f:
        f.write(b"MM\x00*junk")
    wrong = str(tmp_path / "wrong.tif")
    Image.fromarray(np.ones((10, 10), np.uint16)).save(wrong)
    ai = Integrator(dict(PONI))
    ai.set_mask(np.zeros(SHAPE, np.uint8))
    res = ai.batch1d([paths[0], bad, paths[1], wrong, paths[2]], 30, errors=True)
    assert [r.filename for r in res] == [paths[0], bad, paths[1], wrong, paths[2]]
    assert [r.is_ok() for r in res] == [True, False, True, False, True]
    for r, a in zip([res[0], res[2], res[4]], arrays):
        one = ai.integrate1d(np.ascontiguousarray(a[::-1]), 30, errors=True)
        assert np.array_equal(r.result.intensity, one.intensity)
        assert np.array_equal(r.result.errors, one.errors)


@pytest.mark.parametrize("method", ["pixel_split", "pyfai_full"])
def test_reference_ceo2_image_matches_oracle(method):
    """Real data: the reference's CeO2 Pilatus 300K-W image (195 x 1475
    int32, -1/-2 in the module gaps, masked) with its PONI file (tilted by
    0.55 rad): the device table equals the oracle's bit for bit, and
    intensity and sigma agree within the gather tolerance."""
    frame = read_tiff(CEO2_TIF)
    poni = load_poni(CEO2_PONI)
    mask = (frame < 0).astype(np.uint8)
    ai = Integrator(CEO2_PONI, method=method, polarization_factor=0.99,
                    precision="f32_compensated")
    ai.set_mask(mask)
    res = ai.batch1d([CEO2_TIF], 1000, errors=True)[0]
    assert res.is_ok(), res.error
    P = O.Poni(**{k: poni[k] for k in ("pixel1", "pixel2", "distance", "poni1", "poni2",
                                        "rot1", "rot2", "rot3", "wavelength")})
    lut = (O.build_lut1d(P, *frame.shape, 1000) if method == "pixel_split"
           else O.build_lut1d_pyfai(P, *frame.shape, 1000))
    plan = ai._plan(frame.shape, 1000)
    rp, col, w64, _ = plan.export_csr()
    assert np.array_equal(rp.astype(np.uint64), lut.row_ptr)
    assert np.array_equal(col.astype(np.uint64), lut.col)
    assert np.array_equal(w64.view(np.uint64), lut.w.view(np.uint64))
    norm = O.normalization(P, *frame.shape, 0.99)
    fixed = O.fixed_table(lut, mask)
    den = O.denominators(fixed, norm)
    x = frame.astype(np.float64)
    i, s = O.gather(fixed, den, x, True)
    mag, _ = O.gather(fixed, den, np.abs(x), False)
    assert close(res.result.intensity, i, np.abs(mag))
    assert close(res.result.errors, s, s)
    two = ai.integrate2d(frame, 300, 72, errors=True)
    lut2, _ = (O.build_lut2d(P, *frame.shape, 300, 72) if method == "pixel_split"
               else O.build_lut2d_pyfai(P, *frame.shape, 300, 72))
    fixed2 = O.fixed_table(lut2, mask)
    den2 = O.denominators(fixed2, norm)
    i2, s2 = O.gather(fixed2, den2, x, True)
    mag2, _ = O.gather(fixed2, den2, np.abs(x), False)
    assert close(two.intensity.ravel(), i2, np.abs(mag2))
    assert close(two.errors.ravel(), s2, s2)
