"""Writes the FP64 oracle reference images the statistical render tests
compare against: per-pixel mean radiance and mean of squares at 4096 spp.

    python tests/golden/make_render_references.py

Runs the oracle (oracle/oracle_render.c, the restatement of
PathTracer::Trace) with 8 threads; about a minute on 8 cores.
"""
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "rumi-renderer_amd"))

import oracle  # noqa: E402
import scenes  # noqa: E402

CASES = {
    # name: (scene factory, width, height, spp, max_depth, seed)
    "cornell64_pt_d5": (scenes.cornell_box, 64, 64, 4096, 5, 1001),
    "media64_pt_d16": (scenes.cornell_media, 64, 64, 4096, 16, 1002),
}


def main():
    for name, (factory, w, h, spp, depth, seed) in CASES.items():
        sc = factory(w, h)
        mean, m2 = oracle.render_pt(sc, w, h, spp, depth, seed=seed, nthreads=8)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, mean=mean.astype(np.float32), m2=m2.astype(np.float32),
                            spp=spp, max_depth=depth, seed=seed)
        print(name, "mean", mean.mean(axis=(0, 1)), "->", path)


if __name__ == "__main__":
    main()
