"""Tile orders of the reordered renderers (render.cu:4411-4866, :6959-7100)
on the CPU.

The oracle's restatement (oracle.c orc_tile_order) is pinned by what fixes
it independently: numpy's stable argsort of the negated keys (the
reference's stable descending radix sort), Morton codes with known answers
and the bit-interleaving property, the zig-zag regrouping checked group by
group against slices of the order, and the packed descriptors against the
tile ids' columns, rows and ranges.  The HIP kernels then run through the
host emulation (same sources) on zeroed and on garbage-filled buffers and
must give the oracle's outputs bit for bit, and the renderers over a tile
order the same image as the renderers without one.
"""
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from polysplat_amd import scenes
from tests import capi
from tests.emu import frames as HF


def _ranges(counts, empty_at_zero=False):
    c = np.asarray(counts, np.int64)
    ends = np.cumsum(c)
    starts = ends - c
    rg = np.stack([starts, ends], 1).astype(np.int32)
    if empty_at_zero:
        rg[c == 0] = 0
    return rg


def _morton_np(x, y):
    x = np.asarray(x, np.uint64)
    y = np.asarray(y, np.uint64)
    z = np.zeros_like(x)
    for b in range(16):
        z |= ((x >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b)
        z |= ((y >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b + 1)
    return z


def test_morton_known_answers():
    assert O.morton(0, 0) == 0
    assert O.morton(1, 0) == 1 and O.morton(0, 1) == 2 and O.morton(1, 1) == 3
    assert O.morton(2, 0) == 4 and O.morton(0, 2) == 8
    assert O.morton(0xFFFF, 0) == 0x55555555 and O.morton(0, 0xFFFF) == 0xAAAAAAAA
    assert O.morton(0xFFFF, 0xFFFF) == 0xFFFFFFFF
    # interleaving: the even bits give x back, the odd bits y
    rng = np.random.default_rng(5)
    for x, y in rng.integers(0, 1 << 16, (200, 2)):
        z = O.morton(int(x), int(y))
        xs = sum(((z >> (2 * b)) & 1) << b for b in range(16))
        ys = sum(((z >> (2 * b + 1)) & 1) << b for b in range(16))
        assert (xs, ys) == (x, y)
        # Z-order is monotone in each coordinate
        if x < 0xFFFF:
            assert O.morton(int(x) + 1, int(y)) > z


@pytest.mark.parametrize("T", [0, 1, 2, 7, 1000, 40000])
def test_oracle_order_is_the_stable_descending_sort(T):
    rng = np.random.default_rng(T)
    counts = rng.integers(0, 6, T)  # many ties
    rg = _ranges(counts, empty_at_zero=True)
    gx = max(1, int(np.sqrt(T)))
    order, keys, desc = O.tile_order(rg, gx)
    ref = np.argsort(-counts, kind="stable")
    np.testing.assert_array_equal(order, ref)
    np.testing.assert_array_equal(keys, counts[ref])
    # descriptors: column, row and range of the tile at each position
    np.testing.assert_array_equal(desc[:, 0], ref % gx)
    np.testing.assert_array_equal(desc[:, 1], ref // gx)
    np.testing.assert_array_equal(desc[:, 2:], rg[ref])
    # a permutation, heaviest first, ties by ascending tile id
    assert sorted(order.tolist()) == list(range(T))
    for a, b in zip(order[:-1], order[1:]):
        assert counts[a] > counts[b] or (counts[a] == counts[b] and a < b)


def test_oracle_order_full_width_counts():
    # counts that differ in every byte (all four digit passes run)
    rng = np.random.default_rng(9)
    counts = rng.integers(0, 1 << 30, 5000)
    counts[::7] = counts[3]  # ties at a large count
    rg = np.stack([np.zeros_like(counts), counts], 1).astype(np.int32)
    order, keys, _ = O.tile_order(rg, 50)
    ref = np.argsort(-counts, kind="stable")
    np.testing.assert_array_equal(order, ref)
    np.testing.assert_array_equal(keys, counts[ref].astype(np.uint32))


@pytest.mark.parametrize("bucket", [1, 64, 1280])
def test_oracle_morton_bucket_keys(bucket):
    rng = np.random.default_rng(bucket)
    gx, gy = 300, 70  # wider than 256 tiles: the Morton code wraps at 16 bits
    T = gx * gy
    counts = rng.integers(0, 3000, T)
    rg = _ranges(counts)
    order, keys, _ = O.tile_order(rg, gx, bucket=bucket)
    t = np.arange(T)
    b = np.minimum(counts // bucket, 0xFFFF).astype(np.uint64)
    key = (b << np.uint64(16)) | (np.uint64(0xFFFF) - (_morton_np(t % gx, t // gx)
                                                       & np.uint64(0xFFFF)))
    ref = np.argsort(-key.astype(np.int64), kind="stable")
    np.testing.assert_array_equal(order, ref)
    np.testing.assert_array_equal(keys, key[ref].astype(np.uint32))


@pytest.mark.parametrize("T,g", [(100, 7), (100, 10), (99, 100), (5, 2), (1, 3)])
def test_oracle_zigzag_groups(T, g):
    rng = np.random.default_rng(T + g)
    counts = rng.integers(0, 50, T)
    rg = _ranges(counts)
    order, _, desc = O.tile_order(rg, 10, zigzag=g)
    tiles = desc[:, 1] * 10 + desc[:, 0]
    for k, base in enumerate(range(0, T, g)):
        grp = order[base:base + g]
        want = grp[::-1] if k % 2 else grp
        np.testing.assert_array_equal(tiles[base:base + g], want)
    np.testing.assert_array_equal(desc[:, 2:], rg[tiles])


# ------------------------------------------------------------ emulated kernels
@pytest.fixture(scope="module")
def emu():
    subprocess.run([sys.executable, str(HF.HERE / "build.py")], check=True,
                   stdout=subprocess.DEVNULL)
    return HF.load(asan=False)


def _gpu_order(L, rg, gx, zig, bucket, garbage):
    T = rg.shape[0]
    alloc = HF.Buffers(garbage, 7)
    tb = L.ps_get_tile_order_sort_temp_size(T)
    temp = alloc((tb,), np.uint8)
    order, counts, ids = (alloc((max(T, 1),), np.uint32) for _ in range(3))
    desc = alloc((max(T, 1), 4), np.int32)
    rgc = np.ascontiguousarray(rg)
    capi.check_with(L, L.ps_compute_tile_order_packed(
        HF.ptr(rgc), T, gx, HF.ptr(desc), HF.ptr(order), HF.ptr(counts), HF.ptr(ids),
        HF.ptr(temp), tb, zig, bucket, None))
    return order[:T], counts[:T], ids[:T], desc[:T]


# Each case costs about 13 s in the emulation (a dozen launches of 1024
# workgroups), whatever the tile count: four cases cover ties with and
# without zig-zag, counts differing in every byte (every digit pass runs),
# Morton keys, a single tile, and garbage-filled buffers.
@pytest.mark.parametrize("counts_kind,zig,bucket,garbage", [
    ("ties", 264, 0, True), ("wide", 13, 0, False), ("ties", 792, 1280, False),
    ("single", 0, 0, True)])
def test_emulated_tile_order_matches_oracle(emu, garbage, counts_kind, zig, bucket):
    rng = np.random.default_rng({"ties": 1, "wide": 2, "uniform": 3, "single": 4}[counts_kind]
                                + zig + bucket)
    T = {"ties": 20000, "wide": 5000, "uniform": 3000, "single": 1}[counts_kind]
    if counts_kind == "wide":
        c = rng.integers(0, 1 << 30, T)
        rg = np.stack([np.zeros_like(c), c], 1).astype(np.int32)
    else:
        c = rng.integers(0, 700, T) if counts_kind == "ties" else np.full(T, 5)
        rg = _ranges(c, empty_at_zero=True)
    gx = 160 if T > 160 else T
    order, counts, ids, desc = _gpu_order(emu, rg, gx, zig, bucket, garbage)
    o_order, o_keys, o_desc = O.tile_order(rg, gx, zig, bucket)
    np.testing.assert_array_equal(order, o_order)
    np.testing.assert_array_equal(counts, o_keys)
    np.testing.assert_array_equal(ids, np.arange(T))
    np.testing.assert_array_equal(desc, o_desc)


def test_emulated_reordered_renders_equal_unordered(emu):
    L = emu
    arr = scenes.make_config_scene("c1", seed=5, P=6000)
    spec = scenes.make_camera(200, 150, 10.0, distance=9.0)
    bg = (0.2, 0.3, 0.4)
    for prec in ("fp32", "fp16"):
        f = HF.HostFrame(L, arr, spec, sh_type=prec, feat_prec=prec)
        st, n, img = f.forward(bg, capi.V_LITE)
        capi.check_with(L, st)
        gx = (spec.width + 15) // 16
        order, _, _, desc = _gpu_order(L, f.ranges.copy(), gx, 7, 0, False)
        pr = HF.PREC[prec]
        q = np.zeros(512, np.int32)
        for use_desc in (True, False):
            out = np.full_like(img, 9)
            capi.check_with(L, L.ps_render_16x16_reordered_persistent(
                n, None, spec.width, spec.height, pr, HF.ptr(f.xy), HF.ptr(f.rgbd),
                HF.ptr(f.conic), HF.ptr(f.records), HF.ptr(f.vs), HF.ptr(f.ranges),
                HF.ptr(order), HF.ptr(desc) if use_desc else None, 0, capi.floats3(bg),
                HF.ptr(out), HF.ptr(q), None))
            np.testing.assert_array_equal(out, img, err_msg=f"{prec} desc={use_desc}")
            assert not q.any()  # the queue words are left zero
        if prec == "fp16":
            # ps_forward with PS_FORWARD_TILE_ORDER: the same frame
            st, m, img2 = f.forward(bg, capi.V_LITE | capi.TILE_ORDER)
            capi.check_with(L, st)
            assert m == n
            np.testing.assert_array_equal(img2, img)
        if prec == "fp32":
            # plain renderer: row-major, over the order, over the descriptors
            base = np.full_like(img, 3)
            capi.check_with(L, L.ps_render_16x16(
                n, None, spec.width, spec.height, HF.ptr(f.xy), HF.ptr(f.rgbd),
                HF.ptr(f.conic), HF.ptr(f.ks), HF.ptr(f.vs), capi.floats3(bg), HF.ptr(base),
                None))
            out = np.full_like(img, 9)
            capi.check_with(L, L.ps_render_16x16_reordered(
                n, None, spec.width, spec.height, HF.ptr(f.xy), HF.ptr(f.rgbd),
                HF.ptr(f.conic), HF.ptr(f.ks), HF.ptr(f.vs), HF.ptr(f.ranges), HF.ptr(order),
                capi.floats3(bg), HF.ptr(out), None))
            np.testing.assert_array_equal(out, base)
            out = np.full_like(img, 9)
            capi.check_with(L, L.ps_render_16x16_reordered_desc(
                n, None, spec.width, spec.height, HF.ptr(f.xy), HF.ptr(f.rgbd),
                HF.ptr(f.conic), HF.ptr(f.vs), HF.ptr(desc), capi.floats3(bg), HF.ptr(out),
                None))
            np.testing.assert_array_equal(out, base)


def test_emulated_order_rejects_bad_arguments(emu):
    L = emu
    rg = _ranges([3, 1, 2])
    order = np.zeros(3, np.uint32)
    small = np.zeros(16, np.uint8)
    st = L.ps_compute_tile_order(HF.ptr(rg), 3, HF.ptr(order), None, None, HF.ptr(small), 16,
                                 None)
    assert st == capi.PS_ERR_INVALID
    st = L.ps_compute_tile_order(HF.ptr(rg), 3, None, None, None, None, 0, None)
    assert st == capi.PS_ERR_INVALID
    # zero tiles: nothing to do
    assert L.ps_compute_tile_order(HF.ptr(rg), 0, HF.ptr(order), None, None, None, 0,
                                   None) == capi.PS_OK
