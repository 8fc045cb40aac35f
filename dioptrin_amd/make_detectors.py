"""Writes detectors.json, the detector pitches that PONI files may leave to a
named detector, from the reference's detector catalogue data
(src/detector/catalog.json.gz, read by src/detector.rs:659-670). Only the
fields that PONI loading uses are kept: name, aliases, pixel1, pixel2,
binning_pitches and max_shape. Run from the repository root:

    python3 dioptrin_amd/make_detectors.py /root/reference
"""
import gzip
import json
import sys
from pathlib import Path

KEEP = ("name", "aliases", "pixel1", "pixel2", "binning_pitches", "max_shape")


def main(ref: str) -> None:
    src = Path(ref) / "src" / "detector" / "catalog.json.gz"
    with gzip.open(src, "rb") as f:
        catalog = json.loads(f.read(32 * 1024 * 1024))
    rows = [{k: d.get(k) for k in KEEP} for d in catalog]
    out = Path(__file__).with_name("detectors.json")
    out.write_text(json.dumps(rows, indent=0) + "\n")
    print(f"{out}: {len(rows)} detectors")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
