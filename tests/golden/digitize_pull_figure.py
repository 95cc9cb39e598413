"""Digitize the reference's documented tether-pull force curve.

Reads docs/src/assets/example-pull.png of the reference (the figure of the
"Tether pulling" example, docs/src/index.md:142-176: LAG, F_PULL, 19 x 19
elements, length 16, pull_speed 0.5, 32 steps of Δt = 0.5, plotted from
columns 4 (zp) and 7 (fz) of f-pull.txt) and writes the marker centres as
(zp, fz) pairs to docs_pull_force.json.

Calibration: the light-grey grid lines of the axes (zp = 2, 4, 6 and
fz = 1..6) give pixel -> data maps; markers (RGB 0.75, 0.34, 0.0) are
located by the centroid of the marker colour in a 5-pixel column strip around
each step's zp = 0.25 k, so overlapping markers of the plateau separate.
One pixel is ~0.028 in fz.

Run once in a container that holds the reference:
    python tests/golden/digitize_pull_figure.py /path/to/reference
"""
import json
import os
import sys

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    png = os.path.join(ref, "docs", "src", "assets", "example-pull.png")
    im = np.asarray(Image.open(png).convert("RGB")).astype(int)
    r, g, b = im[..., 0], im[..., 1], im[..., 2]
    marker = (abs(r - 191) < 40) & (abs(g - 87) < 40) & (b < 60)
    grey = (abs(r - g) < 6) & (abs(g - b) < 6) & (r > 200) & (r < 245)

    def lines(counts):
        idx = [i for i in range(len(counts)) if counts[i] > 100]
        groups, cur = [], [idx[0]]
        for i in idx[1:]:
            if i == cur[-1] + 1:
                cur.append(i)
            else:
                groups.append(float(np.mean(cur)))
                cur = [i]
        groups.append(float(np.mean(cur)))
        return groups

    cols = lines(grey.sum(axis=0))   # zp = 2, 4, 6
    rows = lines(grey.sum(axis=1))   # fz = 6, 5, 4, 3, 2, 1
    assert len(cols) == 3 and len(rows) == 6, (cols, rows)
    zfit = np.polyfit([2.0, 4.0, 6.0], cols, 1)   # zp -> pixel column
    ffit = np.polyfit(rows, [6.0, 5.0, 4.0, 3.0, 2.0, 1.0], 1)  # pixel row -> fz
    xs = np.arange(im.shape[1])[None, :]
    pts = []
    for k in range(1, 33):
        z = 0.25 * k
        sel = marker & (np.abs(xs - np.polyval(zfit, z)) <= 2.0)
        yy, _ = np.nonzero(sel)
        pts.append([z, round(float(np.polyval(ffit, yy.mean())), 4)])
    out = {
        "source": "docs/src/assets/example-pull.png (docs/src/index.md:142-176)",
        "example": {"motion": "LAG", "scenario": "F_PULL", "num1el": 19, "num2el": 19,
                    "length": 16.0, "pull_speed": 0.5, "dt": 0.5, "steps": 32},
        "pixel_fz": abs(float(ffit[0])),
        "zp_fz": pts,
    }
    with open(os.path.join(HERE, "docs_pull_force.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
