"""Writes tests/golden/milestone8_meta.json: facts about the reference's
docs/milestones/milestone-8*.png fixtures (Cornell preset, 800x600, default
camera) established with the oracle at 256 spp: the depth-AOV normalisation
range (min/max over the whole frame, film.rs:149-159) and the beauty bytes
that differ. Run from the repo root: python tests/golden/make_milestone8_meta.py
"""
import json
import os
import sys

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

W, H, SPP = 800, 600, 256
sc = O.Scene.cornell()
cam = O.camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, W / H)
out = sc.render(cam, W, H, SPP)
d = out[:, 9]
lo = float(d[d > 0].min())
hi = float(d[d > 0].max())
beauty = O.encode_png(out, W, H, 0)
ref = np.asarray(Image.open(os.path.join(HERE, "milestone-8.png")))
diff = np.argwhere(beauty != ref)
meta = dict(width=W, height=H, spp=SPP, depth_lo=lo, depth_hi=hi,
            beauty_diff=[[int(a) for a in x] + [int(beauty[tuple(x)]), int(ref[tuple(x)])] for x in diff])
json.dump(meta, open(os.path.join(HERE, "milestone8_meta.json"), "w"), indent=1)
print(meta)
