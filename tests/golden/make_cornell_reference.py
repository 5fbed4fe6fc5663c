#!/usr/bin/env python3
"""Writes tests/golden/cornell_oracle_1024spp.npz: the FP64 oracle's render of
the Cornell-box workload (256x256, spectral HWSS, depth 5, Russian roulette
from depth 3) at 1024 samples per pixel, 64x the workload's 16, with render
seed 777 (no render with that seed is compared against it).

Stored per pixel (float16, a relative precision of 5e-4, far below the 1%
errors it measures): the mean XYZ, and the per-pixel variance of one
sample's Y (for the reference's own error, var / 1024).  bench.py measures the
GPU render's relative RMSE against it.

Usage: python tests/golden/make_cornell_reference.py  (about 2.5 minutes on 8 cores)
"""
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from rise_amd import pt_scene  # noqa: E402

OUT = Path(__file__).resolve().parent / "cornell_oracle_1024spp.npz"
SPP = 1024
SEED = 777


def main() -> None:
    s, r = pt_scene.cornell_box()
    r.spp = SPP
    r.seed = SEED
    t0 = time.time()
    sm, sq, cnt = oracle.render(s, r)
    mean = sm / SPP
    var_y = np.maximum(sq[..., 1] / SPP - mean[..., 1] ** 2, 0.0)
    np.savez_compressed(OUT, xyz=mean.astype(np.float16), var_y=var_y.astype(np.float16),
                        spp=np.int32(SPP), seed=np.int64(SEED))
    print(f"wrote {OUT} in {time.time() - t0:.1f} s ({cnt[0] + cnt[1]} rays)")


if __name__ == "__main__":
    main()
