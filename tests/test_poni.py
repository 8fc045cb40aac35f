"""PONI loading (poni.rs:79-240) against the reference's own PONI files
(test_data/*.poni, copied under tests/golden/poni) and the expectations of its
unit tests (poni.rs:256-390)."""
from pathlib import Path

import pytest

from dioptrin_amd.poni import load_poni, named_detector, parse_detector_config

G = Path(__file__).parent / "golden" / "poni"


def write(tmp_path, tag, text):
    p = tmp_path / f"dioptrin_poni_{tag}.poni"
    p.write_text(text)
    return p


def test_version1_every_field():
    p = load_poni(G / "v1.poni")
    assert p == {"pixel1": 0.000172, "pixel2": 0.000172, "distance": 0.208651380603,
                 "poni1": 0.0872948482846, "poni2": 0.0799601126306,
                 "rot1": -0.0184422457059, "rot2": 0.00413760084465,
                 "rot3": -2.77645988275e-08, "wavelength": 4.066e-11, "spline_file": "None"}


def test_version2_detector_config_pixels():
    p = load_poni(G / "v2.poni")
    assert p["pixel1"] == 0.0002 and p["pixel2"] == 0.0002
    assert p["distance"] == 0.35603053579185723 and p["wavelength"] == 6.858e-11


def test_named_detector_without_pixel_size():
    p = load_poni(G / "eiger_0000.poni")
    assert p["pixel1"] == 7.5e-5 and p["pixel2"] == 7.5e-5
    assert p["rot3"] == -3.4968839362494736e-09


def test_explicit_pixel_size_kept_with_detector_name():
    p = load_poni(G / "ceo2_side.poni")
    assert p["pixel1"] == 0.00017199999999999998 and p["pixel2"] == 0.00017199999999999998
    assert p["rot1"] == -0.5523616625507498


def test_marccd_version1():
    p = load_poni(G / "LaB6_40keV_MarCCD.poni")
    assert p["pixel1"] == 7.9e-05 and p["wavelength"] == 3.1e-11


def test_detector_config_parse():
    assert parse_detector_config('{"pixel1": 0.0002, "pixel2": 0.0003, "max_shape": [605, 1070]}'
                                 ) == {"pixel1": 0.0002, "pixel2": 0.0003, "binning": None}
    assert parse_detector_config('{"orientation": 2}') == {"pixel1": None, "pixel2": None,
                                                          "binning": None}
    assert parse_detector_config('{"binning": [2, 4], "orientation": 2}')["binning"] == (2, 4)
    assert parse_detector_config('{"binning": 3}')["binning"] == (3, 3)
    assert parse_detector_config('{"pixel1": null, "binning": null}')["pixel1"] is None
    for bad in ('[1, 2]', '{"binning": 0}', '{"binning": [1, 0]}', '{"binning": 2.0}',
                '{"binning": [1, 2, 3]}', '{"binning": true}', '{"pixel1": "x"}',
                '{"pixel1": NaN}', '{"binning": -1}'):
        with pytest.raises(ValueError):
            parse_detector_config(bad)


def test_named_detector_with_binning(tmp_path):
    p = load_poni(write(tmp_path, "binned_pair", "poni_version: 2.1\nDetector: Eiger2CdTe_9M\n"
                        'Detector_config: {"binning": [2, 4], "orientation": 2}\n'
                        "Distance: 0.2\n"))
    assert p["pixel1"] == 1.5e-4 and p["pixel2"] == 3.0e-4
    p = load_poni(write(tmp_path, "binned_scalar", "poni_version: 2.1\nDetector: Eiger2CdTe_9M\n"
                        'Detector_config: {"binning": 2}\nDistance: 0.2\n'))
    assert p["pixel1"] == 1.5e-4 and p["pixel2"] == 1.5e-4


def test_detector_binning_pitches(tmp_path):
    """A detector that lists its binned pitches uses them, and rejects a
    binning it does not list."""
    p = load_poni(write(tmp_path, "mx225", "Detector: Rayonix MX225\n"
                        'Detector_config: {"binning": [2, 3]}\n'))
    assert p["pixel1"] == 7.3242e-05 and p["pixel2"] == 0.000109971
    with pytest.raises(ValueError, match="unsupported binning"):
        load_poni(write(tmp_path, "mx225_5", "Detector: MAR225\n"
                        'Detector_config: {"binning": 5}\n'))


def test_config_before_detector_name_and_v1_override(tmp_path):
    p = load_poni(write(tmp_path, "config_before_name", "poni_version: 2.1\n"
                        'Detector_config: {"orientation": 2}\nDetector: Eiger2CdTe_9M\n'
                        "Distance: 0.2\n"))
    assert p["pixel1"] == 7.5e-5 and p["pixel2"] == 7.5e-5
    p = load_poni(write(tmp_path, "v1_override", "Detector: Eiger2CdTe_9M\nPixelSize1: 0.0002\n"
                        "PixelSize2: 0.0003\nDistance: 0.2\n"))
    assert p["pixel1"] == 0.0002 and p["pixel2"] == 0.0003


def test_errors(tmp_path):
    with pytest.raises(ValueError, match="missing pixel sizes"):
        load_poni(write(tmp_path, "generic_no_pixels", "poni_version: 2.1\nDetector: Detector\n"
                        'Detector_config: {"orientation": 2}\nDistance: 0.2\n'))
    with pytest.raises(ValueError) as e:
        load_poni(write(tmp_path, "unknown_detector", "poni_version: 2.1\n"
                        "Detector: NonExistentDetector999\n"
                        'Detector_config: {"orientation": 2}\nDistance: 0.2\n'))
    assert "NonExistentDetector999" in str(e.value) and "dioptrin_poni_" in str(e.value)
    with pytest.raises(OSError):
        load_poni(G / "non_existent.poni")
    with pytest.raises(ValueError):
        load_poni(write(tmp_path, "bad_float", "PixelSize1: 1_0\nPixelSize2: 1\n"))


def test_catalogue_lookup():
    assert named_detector("eiger-9m")["max_shape"] == [3269, 3110]
    assert named_detector("Pilatus 1M")["pixel1"] == 0.000172
    with pytest.raises(KeyError):
        named_detector("NonExistentDetector999")
