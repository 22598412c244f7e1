"""GPU radix sort against the oracle's stable merge sort on adversarial keys:
heavy ties (stability), every tile-boundary size, device-resident counts,
and the pass counts of every configuration's key width."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
from tests import capi
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu



def _keys(n, tiles, depth_levels, rng):
    t = rng.integers(0, tiles, n).astype(np.uint64)
    d = rng.integers(0, depth_levels, n).astype(np.uint64)
    # depth bits of positive floats in [0.2, 100)
    base = np.frombuffer(np.float32(0.2).tobytes(), np.uint32)[0]
    return (t << np.uint64(32)) | (d + np.uint64(base))


def _run_sort(keys, vals, W, H, device_count=False, cap=None):
    n = len(keys)
    cap = max(n, 1) if cap is None else cap
    dev = torch.device("cuda:0")
    ku = torch.zeros(cap, dtype=torch.int64, device=dev)
    vu = torch.zeros(cap, dtype=torch.int32, device=dev)
    ku[:n] = torch.from_numpy(keys.view(np.int64))
    vu[:n] = torch.from_numpy(vals.view(np.int32))
    ks = torch.full((cap,), -1, dtype=torch.int64, device=dev)
    vs = torch.full((cap,), -1, dtype=torch.int32, device=dev)
    L = capi.lib()
    nbytes = L.ps_get_sort_buffer_size(cap)
    space = torch.zeros(nbytes, dtype=torch.uint8, device=dev)
    dn = torch.tensor([n], dtype=torch.int32, device=dev)
    capi.check(L.ps_sort_gaussian(0 if device_count else n, capi.ptr(dn) if device_count else None,
                                  cap, W, H, 16, 16, capi.ptr(space), nbytes, capi.ptr(ku),
                                  capi.ptr(vu), capi.ptr(ks), capi.ptr(vs), capi.stream()))
    torch.cuda.synchronize()
    out_k = ks.cpu().numpy().view(np.uint64)
    out_v = vs.cpu().numpy().view(np.uint32)
    # inputs preserved
    np.testing.assert_array_equal(ku[:n].cpu().numpy().view(np.uint64), keys)
    return out_k, out_v


@pytest.mark.parametrize("n", [0, 1, 2, 255, 2047, 2048, 2049, 4096 * 3 + 17, 100_000])
def test_sort_sizes_with_ties(gpu, n):
    rng = np.random.default_rng(n)
    keys = _keys(n, 40, 3, rng)  # very few distinct keys: stability is visible
    vals = rng.permutation(max(n, 1))[:n].astype(np.uint32)
    say("sort n", n)
    k, v = _run_sort(keys, vals, 160, 64)
    rk, rv = O.sort_pairs(keys, vals)
    np.testing.assert_array_equal(k[:n], rk)
    np.testing.assert_array_equal(v[:n], rv)
    if n:
        assert (k[n:] == np.uint64(2**64 - 1)).all()  # nothing written past n


@pytest.mark.parametrize("W,H", [(800, 800), (1920, 1080), (3840, 2160), (9216, 6912)])
def test_sort_key_widths(gpu, W, H):
    tiles = ((W + 15) // 16) * ((H + 15) // 16)
    rng = np.random.default_rng(W)
    n = 300_000
    keys = _keys(n, tiles, 1 << 20, rng)
    vals = np.arange(n, dtype=np.uint32)
    say("sort widths", W, H)
    k, v = _run_sort(keys, vals, W, H, device_count=True, cap=n + 5000)
    rk, rv = O.sort_pairs(keys, vals)
    np.testing.assert_array_equal(k[:n], rk)
    np.testing.assert_array_equal(v[:n], rv)


def test_sort_large_device_count_properties(gpu):
    # 12M pairs: checked by properties (sorted, permutation, stability of the
    # value order within equal keys) instead of a full oracle run.
    n = 12_000_000
    g = torch.Generator(device="cuda:0").manual_seed(5)
    dev = torch.device("cuda:0")
    tiles = 32400
    t = torch.randint(0, tiles, (n,), device=dev, generator=g, dtype=torch.int64)
    d = torch.randint(0, 1 << 12, (n,), device=dev, generator=g, dtype=torch.int64)
    ku = (t << 32) | (d + 0x3e4ccccd)
    vu = torch.arange(n, device=dev, dtype=torch.int32)
    ks = torch.empty_like(ku)
    vs = torch.empty_like(vu)
    L = capi.lib()
    nbytes = L.ps_get_sort_buffer_size(n)
    space = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    dn = torch.tensor([n], dtype=torch.int32, device=dev)
    say("sort large", n)
    capi.check(L.ps_sort_gaussian(0, capi.ptr(dn), n, 3840, 2160, 16, 16, capi.ptr(space), nbytes,
                                  capi.ptr(ku), capi.ptr(vu), capi.ptr(ks), capi.ptr(vs),
                                  capi.stream()))
    torch.cuda.synchronize()
    # sorted as unsigned 64-bit (keys are positive here, so signed compare works)
    assert bool((ks[1:] >= ks[:-1]).all())
    # stable: equal keys keep ascending original index
    same = ks[1:] == ks[:-1]
    assert bool((vs[1:][same] > vs[:-1][same]).all())
    # permutation of the input
    assert bool(torch.equal(ku[vs.long()], ks))
    assert int(torch.bincount(vs.long(), minlength=n).max()) == 1


def _keys_with_counts(counts, rng, tiles):
    """Keys whose tile t appears counts[t] times, shuffled, with few depth
    levels (ties everywhere)."""
    t = np.repeat(np.arange(len(counts), dtype=np.uint64), counts)
    rng.shuffle(t)
    d = rng.integers(0, 5, len(t)).astype(np.uint64) + np.uint64(0x3e4ccccd)
    return (t << np.uint64(32)) | d


@pytest.mark.parametrize("mode", ["segmented", "lsd"])
def test_sort_segment_sizes_around_the_lds_limit(gpu, mode, monkeypatch):
    # segments of 0, 1, 2, 63..65, 895..897, 1023..1025, 5000 and 8191..9000
    # pairs: the per-wave path and its limit, the per-workgroup LDS path and
    # its limit, and the global-memory path beyond it
    if mode == "lsd":
        monkeypatch.setenv("POLYSPLAT_SORT", "lsd")
    counts = [0, 1, 2, 63, 64, 65, 895, 896, 897, 1024, 1025, 5000, 0, 7, 8191, 8192, 8193, 9000]
    rng = np.random.default_rng(99)
    keys = _keys_with_counts(counts, rng, len(counts))
    n = len(keys)
    vals = rng.permutation(n).astype(np.uint32)
    say("sort segments", mode, n)
    k, v = _run_sort(keys, vals, 64, 64, cap=n + 3)  # 16 tiles
    rk, rv = O.sort_pairs(keys, vals)
    np.testing.assert_array_equal(k[:n], rk)
    np.testing.assert_array_equal(v[:n], rv)


@pytest.mark.parametrize("mode", ["segmented", "lsd"])
def test_sort_long_segments_with_full_depth_words(gpu, mode, monkeypatch):
    # Segments for each workgroup LDS route (up to 4096, 12288 and 32768
    # pairs) and past them, depths over all 32 bits (every digit pass
    # runs), with ties at the segment's largest depth (the padding value).
    if mode == "lsd":
        monkeypatch.setenv("POLYSPLAT_SORT", "lsd")
    counts = [1000, 4096, 4097, 5000, 8192, 8193, 12288, 12289, 16384, 16385, 20000, 32767,
              32768, 32769, 40000, 100000]
    rng = np.random.default_rng(123)
    t = np.repeat(np.arange(len(counts), dtype=np.uint64), counts)
    rng.shuffle(t)
    d = rng.integers(0, 2**32, len(t), dtype=np.uint64)
    d[rng.random(len(t)) < 0.05] = np.uint64(0xFFFFFFF0)
    d[rng.random(len(t)) < 0.05] = np.uint64(0)
    keys = (t << np.uint64(32)) | d
    n = len(keys)
    vals = rng.permutation(n).astype(np.uint32)
    say("sort long segments", mode, n)
    k, v = _run_sort(keys, vals, 64, 64, cap=n + 5)
    rk, rv = O.sort_pairs(keys, vals)
    np.testing.assert_array_equal(k[:n], rk)
    np.testing.assert_array_equal(v[:n], rv)


def test_sort_ignores_key_bits_above_end_bit(gpu):
    # The sort orders by bits [0, 32 + bit_length(tiles)) only, like the
    # reference's radix sort; higher bits ride along.  Segments whose keys
    # differ only above the sorted bits take the workgroup path.
    rng = np.random.default_rng(5)
    n = 20_000
    tiles = 40  # end bit 38
    t = rng.integers(0, 64, n).astype(np.uint64)
    junk = rng.integers(0, 4, n).astype(np.uint64) << np.uint64(6)
    d = rng.integers(0, 9, n).astype(np.uint64)
    keys = ((t | junk) << np.uint64(32)) | d
    vals = np.arange(n, dtype=np.uint32)
    k, v = _run_sort(keys, vals, 160, 64)
    assert 32 + int(tiles).bit_length() == 38
    masked = keys & np.uint64((1 << 38) - 1)
    order = np.argsort(masked, kind="stable")
    np.testing.assert_array_equal(v[:n], vals[order])
    np.testing.assert_array_equal(k[:n], keys[order])


def test_sort_lsd_and_segmented_agree_on_preprocess_keys(gpu, monkeypatch):
    from polysplat_amd import scenes
    from tests.gpu_helpers import Gpu

    # 40% large splats close to the camera: 2.3M pairs, tiles up to ~1450
    scene = scenes.make_scene(40_000, 8.0, [(0.6, -4.0, 0.5), (0.4, -1.0, 0.3)], seed=31)
    spec = scenes.make_camera(900, 500, 8.0, distance=6.0)
    g = Gpu(scene, spec, capacity=2_400_000)
    M = g.preprocess()
    say("sort agree", M)
    g.sort()
    seg_k, seg_v = g.ks[:M].clone(), g.vs[:M].clone()
    monkeypatch.setenv("POLYSPLAT_SORT", "lsd")
    g.ks.fill_(0)
    g.sort()
    assert torch.equal(g.ks[:M], seg_k) and torch.equal(g.vs[:M], seg_v)
    kt = (seg_k >> 32).cpu().numpy()
    assert np.bincount(kt).max() > 1024  # a long segment was exercised
