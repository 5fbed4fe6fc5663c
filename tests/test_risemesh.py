""".risemesh files (RISE's binary mesh format with its BVH2 cache block),
read and written by the library's host code (include/rise_hip_risemesh.h),
no GPU.

* Files the reference wrote (tests/golden/*.risemesh, made by
  tests/golden/make_risemesh_golden.py with its TriangleMeshGeometryIndexed::
  Serialize) read back to the exact triangles that made them; the 8-wide tree
  collapsed from their BVH2 cache gives the brute-force oracle's ids and t on
  every ray (the kernels' traversal run on host threads).
* Files written here load in the reference's Deserialize, which takes their
  BVH2 cache (its tree then has exactly the cache's node count) and answers
  the oracle's ids (skipped without oracle/_ref).
* Malformed files fail with a message (truncation at every cut point, bad
  signature or version, indices out of range, legacy tree blocks); a bad BVH2
  cache is ignored and the tree is built, as the reference rebuilds.
* risemesh_geometry in a scene file gives the same tables as the RAW2 mesh.
"""
import struct
from pathlib import Path

import numpy as np
import pytest

import oracle
from rise_amd import scenes
from rise_amd._lib import RiseHipError
from rise_amd.risemesh import RiseMesh, write_risemesh
from rise_amd.scene_file import load_scene

GOLDEN = Path(__file__).resolve().parent / "golden"


def _soup():
    return np.ascontiguousarray(scenes.soup_triangles(300, seed=11, sigma=0.08), dtype=np.float32)


def _rays(n=3000, seed=5):
    a = scenes.incoherent_rays(n, seed=seed)
    g = np.random.default_rng(seed)
    o = g.random((200, 3))
    d = np.zeros((200, 3))
    d[np.arange(200), np.arange(200) % 3] = 1.0  # axis-parallel
    return np.ascontiguousarray(np.concatenate([a, scenes.rays_from_rows(o, d)], 1), dtype=np.float32)


def test_reference_written_file_reads_back():
    m = RiseMesh(GOLDEN / "ref_soup_300.risemesh")
    info = m.info()
    assert info["version"] == 5 and info["n_tris"] == 300 and info["n_points"] == 900
    assert info["has_bvh_cache"] == 1 and info["bvh_cache_valid"] == 1 and info["n_bvh_nodes"] > 0
    assert info["face_normals"] == 1 and info["double_sided"] == 0 and info["n_coords"] == 300
    np.testing.assert_array_equal(m.triangles(), _soup())


@pytest.mark.parametrize("path", ["ref_soup_300.risemesh", "scenes/unit_cube.risemesh"])
def test_tree_from_reference_cache_matches_oracle(path):
    m = RiseMesh(GOLDEN / path)
    tris = m.triangles()
    rays = _rays()
    pid, t, st = m.trace_host(rays)
    assert st["bvh_source"] == 1  # collapsed from the file's BVH2
    bid, bt = oracle.closest(tris, rays)[:2]
    np.testing.assert_array_equal(pid, bid)
    np.testing.assert_array_equal(t, bt)
    tmax = np.where(np.isfinite(bt), bt * 0.999, 10.0).astype(np.float32)
    occ, _ = m.trace_host(rays, mode=1, tmax=tmax)
    np.testing.assert_array_equal(occ, oracle.any_hit(tris, rays, tmax))


def test_write_read_round_trip(tmp_path):
    tris = _soup()
    uv = np.zeros((300, 6), np.float32)
    uv[:, 0::2] = np.arange(300)[:, None]
    p = tmp_path / "w.risemesh"
    write_risemesh(p, tris, uv, double_sided=True)
    m = RiseMesh(p)
    info = m.info()
    assert (info["version"], info["double_sided"], info["n_tris"], info["n_coords"]) == (5, 1, 300, 300)
    assert info["bvh_cache_valid"] == 1
    np.testing.assert_array_equal(m.triangles(), tris)
    # shared corners are stored once
    cube = RiseMesh(GOLDEN / "scenes" / "unit_cube.risemesh").triangles()
    write_risemesh(tmp_path / "c.risemesh", cube, with_bvh_cache=False)
    ci = RiseMesh(tmp_path / "c.risemesh").info()
    assert ci["n_points"] == 8 and ci["has_bvh_cache"] == 0
    pid, t, st = RiseMesh(tmp_path / "c.risemesh").trace_host(_rays(200))
    assert st["bvh_source"] == 0


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref (the reference's own build) not built")
def test_reference_loads_our_file_and_its_cache(tmp_path):
    tris = _soup()
    uv = np.zeros((300, 6), np.float32)
    uv[:, 0::2] = np.arange(300)[:, None]  # the reference harness reports u as the triangle id
    p = tmp_path / "ours.risemesh"
    write_risemesh(p, tris, uv)
    ours = RiseMesh(p).info()["n_bvh_nodes"]
    ref = oracle.RefMesh(None, risemesh=p)
    assert ref.loaded_bvh2_nodes() == ours  # Deserialize took the cache instead of rebuilding
    rays = _rays()
    rid, rt, _ = ref.closest(rays)
    bid, bt = oracle.closest(tris, rays)[:2]
    np.testing.assert_array_equal(rid, bid)
    np.testing.assert_array_equal(rt, bt)


def _corrupt(tmp_path, data: bytes, name="bad.risemesh"):
    p = tmp_path / name
    p.write_bytes(data)
    return p


def test_truncated_files_fail_cleanly(tmp_path):
    data = (GOLDEN / "scenes" / "unit_cube.risemesh").read_bytes()
    # every cut before the BVH2 block is an error; cuts inside the block
    # leave a mesh without a usable cache
    bvh_at = data.index(b"BVH2")
    for cut in list(range(0, bvh_at, 7)) + [bvh_at - 1]:
        with pytest.raises(RiseHipError):
            RiseMesh(_corrupt(tmp_path, data[:cut]))
    for cut in range(bvh_at, len(data), 9):
        info = RiseMesh(_corrupt(tmp_path, data[:cut])).info()
        assert info["has_bvh_cache"] == 1 and info["bvh_cache_valid"] == 0


def test_malformed_headers_and_indices(tmp_path):
    data = bytearray((GOLDEN / "scenes" / "unit_cube.risemesh").read_bytes())
    with pytest.raises(RiseHipError, match="signature"):
        RiseMesh(_corrupt(tmp_path, b"RISEXXXX" + bytes(data[8:])))
    with pytest.raises(RiseHipError, match="version 6"):
        RiseMesh(_corrupt(tmp_path, bytes(data[:8]) + struct.pack("<I", 6) + bytes(data[12:])))
    # first triangle's first point index -> 1000 (8 points... written per corner: 36)
    npts = struct.unpack_from("<I", data, 13)[0]
    off = 13 + 4 + 24 * npts
    off += 4 + 24 * struct.unpack_from("<I", data, off)[0]   # normals
    off += 4 + 16 * struct.unpack_from("<I", data, off)[0]   # coords
    off += 4 + 24 * struct.unpack_from("<I", data, off)[0]   # colors
    bad = bytearray(data)
    struct.pack_into("<I", bad, off + 4, 1000)
    with pytest.raises(RiseHipError, match="indexes past"):
        RiseMesh(_corrupt(tmp_path, bytes(bad)))


def test_bad_cache_is_ignored(tmp_path):
    data = bytearray((GOLDEN / "ref_soup_300.risemesh").read_bytes())
    at = data.index(b"BVH2")
    n_nodes = struct.unpack_from("<I", data, at + 8)[0]
    prims_at = at + 12 + 32 * n_nodes + 4
    struct.pack_into("<I", data, prims_at, 300)  # a primitive index out of range
    m = RiseMesh(_corrupt(tmp_path, bytes(data)))
    assert m.info()["has_bvh_cache"] == 1 and m.info()["bvh_cache_valid"] == 0
    rays = _rays(500)
    pid, t, st = m.trace_host(rays)
    assert st["bvh_source"] == 0
    np.testing.assert_array_equal(pid, oracle.closest(_soup(), rays)[0])
    # a cycle (a child pointing back at the root) is refused as well
    data2 = bytearray((GOLDEN / "ref_soup_300.risemesh").read_bytes())
    struct.pack_into("<I", data2, at + 12 + 24, 0)
    assert RiseMesh(_corrupt(tmp_path, bytes(data2), "c.risemesh")).info()["bvh_cache_valid"] == 0


def _v4_bytes(tris, legacy_v3=False, legacy_tree=0):
    """A version-4 (or, legacy_v3, version-3) file built field by field."""
    n = len(tris)
    pts = tris.reshape(-1, 3).astype(np.float64)
    out = b"RISETMGI" + struct.pack("<I", 3 if legacy_v3 else 4)
    if legacy_v3:
        out += struct.pack("<I", 10) + b"\x05"
    out += b"\x01" + struct.pack("<I", len(pts)) + pts.tobytes()
    out += struct.pack("<I", 0) + struct.pack("<I", 1) + np.zeros(2).tobytes()
    idx = np.zeros((n, 3, 3), np.uint32)
    idx[:, :, 0] = np.arange(3 * n).reshape(n, 3)
    out += struct.pack("<I", n) + idx.tobytes() + b"\x00"
    if legacy_v3:
        out += b"\x00" + bytes([legacy_tree])
    return out + b"\x00"  # no BVH cache


def test_older_versions(tmp_path):
    tris = _soup()[:50]
    for v3 in (False, True):
        m = RiseMesh(_corrupt(tmp_path, _v4_bytes(tris, legacy_v3=v3)))
        assert m.info()["version"] == (3 if v3 else 4)
        np.testing.assert_array_equal(m.triangles(), tris)
    with pytest.raises(RiseHipError, match="legacy BSP/octree"):
        RiseMesh(_corrupt(tmp_path, _v4_bytes(tris, legacy_v3=True, legacy_tree=1)))


def test_scene_file_risemesh_geometry_matches_raw2():
    a, ra = load_scene(GOLDEN / "scenes" / "cornell.RISEscene")
    b, rb = load_scene(GOLDEN / "scenes" / "cornell_risemesh.RISEscene")
    assert ra == rb
    for x, y in zip(a.objects, b.objects):
        np.testing.assert_array_equal(x.tris, y.tris)
        assert (x.material, x.double_sided) == (y.material, y.double_sided)


def _cache_bytes(nodes: bytes, n_nodes: int, n_tris: int, tris: bytes = b"", magic=b"RHBVH8\0\0", node_bytes=80):
    return magic + struct.pack("<IIII", 1, node_bytes, 1, 0) + struct.pack("<QQ", n_nodes, n_tris) + nodes + tris


def _node(imask=0, child_base=0, tri_base=0, meta=(0,) * 8):
    # origin, exponents, imask, child_base, tri_base, meta, 48 quantized bounds
    return (struct.pack("<fffbbbB", 0, 0, 0, 0, 0, 0, imask) + struct.pack("<II", child_base, tri_base)
            + bytes(meta) + bytes(48))


def test_scene_cache_files_are_checked(tmp_path):
    """Malformed scene-cache files fail before anything reaches the device."""
    from rise_amd.raycast import ObjectManager

    cases = {
        "magic": _cache_bytes(_node(), 1, 0, magic=b"RHBVH9\0\0"),
        "node size": _cache_bytes(_node(), 1, 0, node_bytes=128),
        "sizes": _cache_bytes(_node(), 2, 0),
        "child range": _cache_bytes(_node(imask=1, child_base=1, meta=(1,) + (0,) * 7), 1, 0),
        "child before parent": _cache_bytes(_node(imask=1, child_base=0, meta=(1,) + (0,) * 7) + _node(), 2, 0),
        "leaf range": _cache_bytes(_node(meta=((2 << 5) | 0,) + (0,) * 7), 1, 1, tris=bytes(48)),
    }
    want = {"magic": "not a scene cache", "node size": "sizes", "sizes": "sizes", "child range": "child range",
            "child before parent": "child range", "leaf range": "leaf triangles"}
    for name, data in cases.items():
        p = tmp_path / f"{name.replace(' ', '_')}.rhbvh"
        p.write_bytes(data)
        with pytest.raises(RiseHipError, match=want[name]):
            ObjectManager.load_cache(p)
