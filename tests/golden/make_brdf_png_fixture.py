"""Writes tests/golden/brdf_lut_png.json from the reference's
Content/PrecomputedMaps/brdf_lut.png (450x450, 256-colour palette; loaded by
Src/ImageBasedLighting.cpp:571-602 and read at vec2(NdotV, roughness) by
Shaders/PBR/PBRMaterial.glsl:110).

The PNG stores roughness 0 in its bottom row, so texel (x, y) (rows top-down)
is sampled at NdotV = (x + 0.5) / 450 and roughness = 1 - (y + 0.5) / 450.
A 16 x 16 grid of texel values (R, G) / 255 is written as (NdotV, roughness,
A, B) rows.  Run in the container that holds the reference checkout.
"""
import json
from pathlib import Path

import numpy as np
from PIL import Image

SRC = Path("/root/reference/Content/PrecomputedMaps/brdf_lut.png")
OUT = Path(__file__).resolve().parent / "brdf_lut_png.json"


def main() -> None:
    img = np.asarray(Image.open(SRC).convert("RGBA"), dtype=np.float64) / 255.0
    h, w = img.shape[:2]
    pts = []
    for gy in range(16):
        for gx in range(16):
            x = int((gx + 0.5) * w / 16)
            y = int((gy + 0.5) * h / 16)
            nv = (x + 0.5) / w
            r = 1.0 - (y + 0.5) / h
            pts.append([nv, r, float(img[y, x, 0]), float(img[y, x, 1])])
    OUT.write_text(json.dumps({"source": "Content/PrecomputedMaps/brdf_lut.png", "width": w, "height": h,
                               "orientation": "row 0 = roughness 1 (v flipped)", "points": pts}, indent=0))
    print("wrote", OUT, len(pts), "points")


if __name__ == "__main__":
    main()
