"""The kernels' FP32 slab test (trace_core.h node_accept_mask, run on the host
through rise_hip_debug_node_accept_host) against the exact slab test in
rational arithmetic.

The fused test computes each plane distance as q s + b with b rounded first,
which is not the decoded box's own distance rounded once; it must still never
reject a child box that the exact ray meets in [0, best / |d|].  The cases
aim rays at box corners, edges and faces (the float direction passes within
an ulp of the boundary, on either side), at distance limits a hair either
side of the box, and run axis-parallel rays (zero direction components),
whose slabs need the box padding: there a box must be taken when the origin
lies inside it along the zero axes by the padding the builder guarantees
(1e-9 of the scene extent; here 1e-9 of the node's size), and it may be
dropped when the origin lies on or outside that face.  Clear misses (the
box still missed when grown by 1e-5 of the origin's distance and the
coordinates' size) must be rejected, so a test that takes everything fails
too.
"""
import ctypes
import math
import struct
from fractions import Fraction as F

import numpy as np
import pytest

from rise_amd import _lib

INF = 1.7976931348623158e308


def _node(p, e, qlo, qhi, meta):
    """80-byte Node8: p (3 floats), exponents, imask 0, bases, meta, planes."""
    w0 = struct.pack("<3f4b", *[float(x) for x in p], *[int(x) for x in e], 0)
    w1 = struct.pack("<II8B", 0, 0, *meta)
    planes = b"".join(bytes(bytearray(int(v) for v in qlo[a])) for a in range(3))
    planes += b"".join(bytes(bytearray(int(v) for v in qhi[a])) for a in range(3))
    buf = w0 + w1 + planes
    assert len(buf) == 80
    return buf


def _box(p, e, qlo, qhi, s):
    lo = [F(float(np.float32(p[a]))) + qlo[a][s] * F(2) ** int(e[a]) for a in range(3)]
    hi = [F(float(np.float32(p[a]))) + qhi[a][s] * F(2) ** int(e[a]) for a in range(3)]
    return lo, hi


def _exact(o, d, lo, hi, best, margin=None):
    """Exact slab test of the float ray (o, d) on the box for t in [0, best / |d|].
    Returns (t_near, t_far) with t_near <= t_far, or None when empty.  With
    `margin`, zero-direction axes use the box shrunk by the margin."""
    tn, tf = F(0), None
    for a in range(3):
        oa, da = F(float(o[a])), F(float(d[a]))
        la, ha = lo[a], hi[a]
        if da == 0:
            if margin is not None:
                la, ha = la + margin[a], ha - margin[a]
            if not (la <= oa <= ha):
                return None
            continue
        t1, t2 = (la - oa) / da, (ha - oa) / da
        if t1 > t2:
            t1, t2 = t2, t1
        tn = max(tn, t1)
        tf = t2 if tf is None else min(tf, t2)
    if tf is not None and tn > tf:
        return None
    if best < INF:
        d2 = sum(F(float(x)) ** 2 for x in d)
        if tn * tn * d2 > F(best) ** 2:
            return None
    return tn, tf


def _accept(node, o, d, best):
    ray = (ctypes.c_float * 6)(*[float(x) for x in o], *[float(x) for x in d])
    out = ctypes.c_uint32()
    _lib.check(_lib.lib().rise_hip_debug_node_accept_host(node, ray, ctypes.c_double(best), ctypes.byref(out)))
    return out.value


def _random_node(rng):
    scale = 10.0 ** rng.uniform(-3, 4)
    p = np.float32(rng.uniform(-1, 1, 3) * scale)
    e = rng.integers(-22, 6, 3)
    e = np.clip(np.round(np.log2(scale)) - 8 + rng.integers(-4, 3, 3), -126, 40).astype(int)
    qlo = [[0] * 8 for _ in range(3)]
    qhi = [[0] * 8 for _ in range(3)]
    meta = [0] * 8
    for s in range(8):
        if rng.random() < 0.15:
            for a in range(3):
                qlo[a][s], qhi[a][s] = 255, 0
            continue
        meta[s] = 1 << 5
        for a in range(3):
            lo = int(rng.integers(0, 255))
            hi = int(rng.integers(lo, 256)) if rng.random() < 0.9 else lo
            qlo[a][s], qhi[a][s] = lo, min(hi, 255)
    return p, e, qlo, qhi, meta


def _boundary_point(rng, lo, hi):
    """A point on the box boundary (corner, edge or face), as floats."""
    kind = rng.integers(0, 3)
    pt = []
    for a in range(3):
        pick = rng.random()
        if kind == 0 or (kind == 1 and a < 2) or (kind == 2 and a == 0):
            v = lo[a] if pick < 0.5 else hi[a]
        else:
            v = lo[a] + (hi[a] - lo[a]) * F(int(rng.integers(0, 1000)), 1000)
        pt.append(float(v))
    return np.array(pt)


@pytest.mark.parametrize("seed", range(4))
def test_fused_slab_test_is_conservative(seed):
    rng = np.random.default_rng(1000 + seed)
    must = taken = clear_miss = wrongly_taken = 0
    for _ in range(400):
        p, e, qlo, qhi, meta = _random_node(rng)
        node = _node(p, e, qlo, qhi, meta)
        occupied = [s for s in range(8) if meta[s]]
        if not occupied:
            continue
        boxes = {s: _box(p, e, qlo, qhi, s) for s in occupied}
        for _ in range(6):
            s0 = occupied[int(rng.integers(0, len(occupied)))]
            lo, hi = boxes[s0]
            size = max(float(hi[a] - lo[a]) for a in range(3)) + abs(float(p[0])) * 1e-6 + 1e-30
            target = _boundary_point(rng, lo, hi)
            o = np.float32(target + rng.normal(0, 1, 3) * size * 10.0 ** rng.uniform(-1, 3))
            d = np.float32(target - o.astype(np.float64))
            if not np.any(d):
                continue
            r = rng.random()
            if r < 0.5:
                best = INF
            else:
                dist = float(np.linalg.norm(target - o.astype(np.float64)))
                best = dist * (1.0 + (rng.random() - 0.5) * 1e-6)
            got = _accept(node, o, d, best)
            far = float(np.max(np.abs(o.astype(np.float64) - p.astype(np.float64))))
            for s in occupied:
                ex = _exact(o, d, *boxes[s], best)
                if ex is not None:
                    must += 1
                    taken += (got >> s) & 1
                else:
                    # grown by 1e-5 of the coordinates' and the origin's distance scale
                    g = F(1, 10 ** 5) * (F(far) + max(abs(x) for x in boxes[s][0] + boxes[s][1])) + F(1, 10 ** 30)
                    tight = _exact(o, d, [x - g for x in boxes[s][0]], [x + g for x in boxes[s][1]],
                                   best * (1 + 1e-5) if best < INF else INF)
                    if tight is None:
                        clear_miss += 1
                        wrongly_taken += (got >> s) & 1
            assert not (got & ~sum(1 << s for s in occupied)), "an empty slot was taken"
    assert must > 500 and clear_miss > 500, (must, clear_miss)
    assert taken == must, f"{must - taken} of {must} boxes the exact ray meets were rejected"
    assert wrongly_taken == 0, f"{wrongly_taken} of {clear_miss} clear misses were taken"


def test_axis_parallel_rays_keep_their_slab():
    """Zero direction components: a box is taken when the origin is inside it
    by the builder's padding along those axes and the remaining axes meet
    it; a box the ray clearly misses (grown as in the test above) is
    rejected."""
    rng = np.random.default_rng(77)
    must = taken = outside = wrongly = 0
    for _ in range(400):
        p, e, qlo, qhi, meta = _random_node(rng)
        node = _node(p, e, qlo, qhi, meta)
        occupied = [s for s in range(8) if meta[s]]
        if not occupied:
            continue
        s0 = occupied[int(rng.integers(0, len(occupied)))]
        lo, hi = _box(p, e, qlo, qhi, s0)
        ext = [hi[a] - lo[a] for a in range(3)]
        zero = [bool(z) for z in rng.integers(0, 2, 3)]
        if all(zero):
            zero[int(rng.integers(0, 3))] = False
        inside = rng.random() < 0.7
        o, d = np.zeros(3, np.float32), np.zeros(3, np.float32)
        for a in range(3):
            if zero[a]:
                if inside:
                    v = lo[a] + ext[a] * F(int(rng.integers(1, 1000)), 1000)
                else:
                    v = hi[a] + max(ext[a], abs(hi[a]) * F(1, 10 ** 5) + F(1, 10 ** 30)) * F(int(rng.integers(1, 50)), 10)
                o[a] = np.float32(float(v))
            else:
                o[a] = np.float32(float(lo[a]) - float(ext[a] + 1) * rng.uniform(0.5, 50))
                d[a] = np.float32(rng.uniform(0.3, 2.0))
        got = _accept(node, o, d, INF)
        far = float(np.max(np.abs(o.astype(np.float64) - p.astype(np.float64))))
        for s in occupied:
            bl, bh = _box(p, e, qlo, qhi, s)
            margin = [(bh[a] - bl[a]) * F(1, 10 ** 9) for a in range(3)]
            g = F(1, 10 ** 5) * (F(far) + max(abs(x) for x in bl + bh)) + F(1, 10 ** 30)
            if _exact(o, d, bl, bh, INF, margin=margin) is not None:
                must += 1
                taken += (got >> s) & 1
            elif _exact(o, d, [x - g for x in bl], [x + g for x in bh], INF) is None:
                outside += 1
                wrongly += (got >> s) & 1
    assert must > 150 and outside > 150, (must, outside)
    assert taken == must
    assert wrongly == 0
