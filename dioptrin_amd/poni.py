"""PONI geometry files (Poni::load, poni.rs:79-135; assign_values 137-175;
parse_detector_config 177-240) and the named-detector pitches they may rely
on (detector::named, detector.rs:671-678).

`load_poni(path)` returns the geometry as a dict with the keys GpuPlan and
Integrator take (pixel1, pixel2, distance, poni1, poni2, rot1, rot2, rot3,
wavelength) plus spline_file. Fields a file omits stay 0 (Poni::new,
poni.rs:42-55). When pixel1 or pixel2 is still 0 after the file, a named
detector supplies it: its catalogue pitch for the file's binning (the
detector's own binned pitch when it lists one, else pitch * binning).
"""
from __future__ import annotations

import json
import math
import re
from functools import lru_cache
from pathlib import Path

__all__ = ["load_poni", "named_detector", "parse_detector_config"]

_FLOAT_KEYS = {"PixelSize1": "pixel1", "PixelSize2": "pixel2", "Distance": "distance",
               "Poni1": "poni1", "Poni2": "poni2", "Rot1": "rot1", "Rot2": "rot2",
               "Rot3": "rot3", "Wavelength": "wavelength"}
# Decimal and exponent forms, inf/infinity/nan, as a float field accepts them.
_FLOAT = re.compile(r"[+-]?((\d+\.?\d*|\.\d+)([eE][+-]?\d+)?|inf|infinity|nan)", re.I)


def _parse_float(key: str, text: str) -> float:
    if not _FLOAT.fullmatch(text):
        raise ValueError(f"{key}: invalid float literal {text!r}")
    return float(text)


@lru_cache(maxsize=1)
def _catalog() -> list:
    return json.loads(Path(__file__).with_name("detectors.json").read_text())


def named_detector(name: str) -> dict:
    """Catalogue entry whose name or an alias matches `name`, ignoring case,
    spaces, '_' and '-' (detector.rs:671-678). Raises KeyError."""
    def norm(s: str) -> str:
        return s.lower().replace(" ", "").replace("_", "").replace("-", "")
    key = norm(name)
    for d in _catalog():
        if norm(d["name"]) == key or any(norm(a) == key for a in d["aliases"]):
            return d
    raise KeyError(f"Unknown detector {name}.")


def _positive_int(v, what: str) -> int:
    if isinstance(v, bool) or not isinstance(v, int) or v < 0:
        raise ValueError(f"Detector_config 'binning' {what} must be positive integers"
                         if what == "entries" else
                         "Detector_config 'binning' must be a positive integer")
    if v == 0:
        raise ValueError(f"Detector_config 'binning' {'entries ' if what == 'entries' else ''}"
                         "must be positive")
    return v


def parse_detector_config(value: str) -> dict:
    """The v2/v2.1 `Detector_config` JSON object: optional finite pixel1 and
    pixel2 and optional binning (one positive integer or a pair), as
    {"pixel1", "pixel2", "binning"} with None where absent or null."""
    def no_constants(c):
        raise ValueError(f"invalid JSON number {c}")
    obj = json.loads(value, parse_constant=no_constants)
    if not isinstance(obj, dict):
        raise ValueError("Detector_config must be a JSON object")

    def pixel(key):
        v = obj.get(key)
        if v is None:
            return None
        if isinstance(v, bool) or not isinstance(v, (int, float)) or not math.isfinite(v):
            raise ValueError(f"Detector_config '{key}' must be a finite number")
        return float(v)

    b = obj.get("binning")
    if b is None:
        binning = None
    elif isinstance(b, list) and len(b) == 2:
        binning = (_positive_int(b[0], "entries"), _positive_int(b[1], "entries"))
    elif isinstance(b, (int, float)) and not isinstance(b, bool):
        v = _positive_int(b, "scalar")
        binning = (v, v)
    else:
        raise ValueError("Detector_config 'binning' must be a positive integer or [b1, b2] array")
    return {"pixel1": pixel("pixel1"), "pixel2": pixel("pixel2"), "binning": binning}


def load_poni(path) -> dict:
    """Reads a PONI file (v1, v2, v2.1). Raises OSError for an unreadable
    file and ValueError for a malformed value, an unknown detector, an
    unsupported binning or missing pixel sizes."""
    path = str(path)
    p = {"pixel1": 0.0, "pixel2": 0.0, "distance": 0.0, "poni1": 0.0, "poni2": 0.0,
         "rot1": 0.0, "rot2": 0.0, "rot3": 0.0, "wavelength": 0.0, "spline_file": ""}
    detector = None
    binning = None
    with open(path, encoding="utf-8") as f:
        for line in f.read().splitlines():
            if line.startswith("#"):
                continue
            key, _, value = line.partition(":")
            key, value = key.strip(), value.strip()
            if key in _FLOAT_KEYS:
                p[_FLOAT_KEYS[key]] = _parse_float(key, value)
            elif key == "Detector":
                detector = value
            elif key == "Detector_config":
                cfg = parse_detector_config(value)
                if cfg["pixel1"] is not None:
                    p["pixel1"] = cfg["pixel1"]
                if cfg["pixel2"] is not None:
                    p["pixel2"] = cfg["pixel2"]
                if cfg["binning"] is not None:
                    binning = cfg["binning"]
            elif key == "SplineFile":
                p["spline_file"] = value
    if p["pixel1"] == 0.0 or p["pixel2"] == 0.0:
        if not detector or detector.lower() == "detector":
            raise ValueError(f"{path}: missing pixel sizes (pixel1/pixel2) in PONI file")
        try:
            det = named_detector(detector)
        except KeyError:
            raise ValueError(f"{path}: unknown detector '{detector}' and no pixel sizes in "
                             "Detector_config") from None
        b1, b2 = binning or (1, 1)
        pitches = {int(k): float(v) for k, v in det["binning_pitches"]}
        if pitches and (b1 not in pitches or b2 not in pitches):
            raise ValueError(f"{path}: unsupported binning [{b1}, {b2}] for detector "
                             f"'{detector}'")
        if p["pixel1"] == 0.0:
            p["pixel1"] = pitches.get(b1, det["pixel1"] * b1)
        if p["pixel2"] == 0.0:
            p["pixel2"] = pitches.get(b2, det["pixel2"] * b2)
    return p
