"""Copies one prefiltered-map cache file the reference ships
(Content/PrecomputedMaps/NeoclassicalInterior/Prefiltered5.hdr, 128x64,
written by stb_image_write with run-length coded scanlines) as the decoder
fixture tests/golden/neoclassical_prefiltered5.hdr.  Data only."""
import shutil
from pathlib import Path

SRC = Path("/root/reference/Content/PrecomputedMaps/NeoclassicalInterior/Prefiltered5.hdr")
DST = Path(__file__).with_name("neoclassical_prefiltered5.hdr")

if __name__ == "__main__":
    shutil.copyfile(SRC, DST)
    print(DST, DST.stat().st_size)
