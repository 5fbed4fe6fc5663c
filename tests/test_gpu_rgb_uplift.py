"""RGB-uplifted (sigmoid-polynomial) spectra on the GPU, through the C ABI.

The Cornell box with its red, green and white walls and its light given as
Rec.709 colours, uplifted through the library's Jakob-Hanika table, renders
with the same random streams as the FP64 oracle: at least 99% of pixels
within 1e-3 (relative, floored at 1% of the mean) and the image mean within
1e-3, as for step spectra.
"""
import numpy as np
import pytest

import oracle
from rise_amd import pt_scene
from rise_amd.rgb_uplift import uplift_spd
from tests.conftest import log_case
from tests.test_gpu_pt import _compare_same_streams, _render

pytestmark = pytest.mark.gpu


def test_gpu_uplifted_cornell_matches_oracle(cuda):
    log_case("cornell with uplifted RGB spectra")
    s, r = pt_scene.cornell_box()
    s.spds[0] = uplift_spd([0.75, 0.75, 0.75])
    s.spds[1] = uplift_spd([0.63, 0.06, 0.05])
    s.spds[2] = uplift_spd([0.14, 0.45, 0.09])
    light = uplift_spd([1.0, 0.85, 0.6])
    light.amp[3] = 15.0
    s.spds[3] = light
    r.width = r.height = 64
    r.spp = 16
    r.lambda_begin, r.lambda_end = 380.0, 780.0
    g, _, st = _render(s, r)
    sm, _, cnt = oracle.render(s, r)
    o = sm / r.spp
    frac, mean_rel = _compare_same_streams(g, o)
    assert frac >= 0.99, frac
    assert mean_rel < 1e-3, mean_rel
    assert st["paths"] == cnt[2] == 64 * 64 * 16
    # the red wall reflects red: its pixels' X/Y exceeds the green wall's
    assert g[32, 4, 0] / g[32, 4, 1] > g[32, 59, 0] / g[32, 59, 1]
