"""Writes tests/golden/cornell_ref_4096spp.npz: the Cornell-box workload
(256x256, 4 stratified wavelengths per sample, depth <= 8, Russian roulette
from generation 3, survival 0.5, domain 380-780 nm) rendered at 4096 samples
per pixel by the FP64 CPU oracle (oracle/render.c), seed 1000 (independent
of the seeds the tests render with).  Stored: weighted Y (float32) and X, Z
(float16).

Run from the repository root: python tests/golden/make_cornell_reference.py
"""
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import liar_amd as la  # noqa: E402
import oracle  # noqa: E402
from liar_amd import workloads as W  # noqa: E402


def main():
    sc, cam = W.cornell_box()
    cs = la.compile_scene(sc)
    t0 = time.time()
    xyz, w = oracle.render(cs, cam, 256, 256, 4096, wavelengths=4, max_depth=8, rr_depth=3, seed=1000,
                           threads=max(1, len(__import__("os").sched_getaffinity(0))))
    img = xyz / w
    out = Path(__file__).resolve().parent / "cornell_ref_4096spp.npz"
    np.savez_compressed(out, y=img[1].astype(np.float32), xz=img[[0, 2]].astype(np.float16), spp=4096,
                        wavelengths=4, seed=1000, seconds=time.time() - t0)
    print(f"wrote {out} in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
