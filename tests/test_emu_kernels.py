"""The HIP kernels themselves, run on the CPU through the host emulation
(same sources, work-items as fibers), checked against the oracle.

These cover the kernel logic without a GPU: sort semantics (ties, long
segments, key bits above end_bit), one full frame bit-exact against the
oracle, batched frames against single frames, and that no kernel reads
memory it did not write (random-filled buffers give the same results as
zeroed ones)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from polysplat_amd import scenes
from tests import capi
from tests.emu import frames as HF
from tests.gpu_helpers import oracle_frame


@pytest.fixture(scope="module")
def emu():
    subprocess.run([sys.executable, str(HF.HERE / "build.py")], check=True,
                   stdout=subprocess.DEVNULL)
    return HF.load(asan=False)


def _sort(L, keys, vals, W, H, cap=None):
    n = len(keys)
    cap = max(n, 1) if cap is None else cap
    ku = np.zeros(cap, np.uint64)
    vu = np.zeros(cap, np.uint32)
    ku[:n], vu[:n] = keys, vals
    ks = np.full(cap, 2**64 - 1, np.uint64)
    vs = np.full(cap, 2**32 - 1, np.uint32)
    nb = L.ps_get_sort_buffer_size(cap)
    space = np.zeros(nb, np.uint8)
    capi.check_with(L, L.ps_sort_gaussian(n, None, cap, W, H, 16, 16, HF.ptr(space), nb,
                                          HF.ptr(ku), HF.ptr(vu), HF.ptr(ks), HF.ptr(vs), None))
    np.testing.assert_array_equal(ku[:n], keys)  # inputs preserved
    return ks, vs


def _tie_keys(counts, rng):
    t = np.repeat(np.arange(len(counts), dtype=np.uint64), counts)
    rng.shuffle(t)
    d = rng.integers(0, 5, len(t)).astype(np.uint64) + np.uint64(0x3e4ccccd)
    return (t << np.uint64(32)) | d


@pytest.mark.parametrize("mode", ["segmented", "lsd"])
def test_sort_segments_around_the_lds_limits(emu, mode, monkeypatch):
    if mode == "lsd":
        monkeypatch.setenv("POLYSPLAT_SORT", "lsd")
    # per-wave sort (<= 896), per-workgroup sort (<= 8192), global passes
    counts = [0, 1, 2, 63, 64, 65, 895, 896, 897, 1024, 1025, 3755, 0, 7, 8191, 8192, 8193, 9000]
    rng = np.random.default_rng(99)
    keys = _tie_keys(counts, rng)
    vals = rng.permutation(len(keys)).astype(np.uint32)
    k, v = _sort(emu, keys, vals, 64, 64, cap=len(keys) + 3)
    rk, rv = O.sort_pairs(keys, vals)
    n = len(keys)
    np.testing.assert_array_equal(k[:n], rk)
    np.testing.assert_array_equal(v[:n], rv)
    assert (k[n:] == np.uint64(2**64 - 1)).all()


@pytest.mark.parametrize("threads", [1, 8])
def test_sort_long_segments_with_full_depth_words(emu, threads, monkeypatch):
    # Segments for each workgroup route (up to 4096, 12288 and 32768 pairs)
    # and past them, with depths over the full 32 bits (every digit
    # pass runs) and many ties at the segment's largest depth (the padding
    # value).
    monkeypatch.setenv("EMU_THREADS", str(threads))
    counts = [1000, 4096, 4097, 5000, 8192, 8193, 12288, 12289, 16384, 16385, 20000, 32767,
              32768, 32769, 40000]
    rng = np.random.default_rng(123)
    t = np.repeat(np.arange(len(counts), dtype=np.uint64), counts)
    rng.shuffle(t)
    d = rng.integers(0, 2**32, len(t), dtype=np.uint64)
    top = rng.random(len(t)) < 0.05
    d[top] = np.uint64(0xFFFFFFF0)  # the largest depth of each segment, repeated
    d[rng.random(len(t)) < 0.05] = np.uint64(0)
    keys = (t << np.uint64(32)) | d
    vals = rng.permutation(len(keys)).astype(np.uint32)
    k, v = _sort(emu, keys, vals, 64, 64, cap=len(keys) + 5)
    rk, rv = O.sort_pairs(keys, vals)
    n = len(keys)
    np.testing.assert_array_equal(k[:n], rk)
    np.testing.assert_array_equal(v[:n], rv)
    assert (k[n:] == np.uint64(2**64 - 1)).all()


def test_sort_orders_only_bits_below_end_bit(emu):
    rng = np.random.default_rng(5)
    n = 6000
    t = rng.integers(0, 64, n).astype(np.uint64)
    junk = rng.integers(0, 4, n).astype(np.uint64) << np.uint64(6)
    d = rng.integers(0, 9, n).astype(np.uint64)
    keys = ((t | junk) << np.uint64(32)) | d
    vals = np.arange(n, dtype=np.uint32)
    for mode in ("segmented", "lsd"):
        os.environ["POLYSPLAT_SORT"] = mode
        try:
            k, v = _sort(emu, keys, vals, 160, 64)  # 40 tiles: end bit 38
        finally:
            del os.environ["POLYSPLAT_SORT"]
        order = np.argsort(keys & np.uint64((1 << 38) - 1), kind="stable")
        np.testing.assert_array_equal(v[:n], vals[order], err_msg=mode)
        np.testing.assert_array_equal(k[:n], keys[order], err_msg=mode)


@pytest.mark.parametrize("mode", ["segmented", "lsd"])
def test_sort_with_nine_bit_digits(emu, mode, monkeypatch):
    # 9216x6912 has 248832 tiles: 18 tile bits, two 9-bit digit passes.
    if mode == "lsd":
        monkeypatch.setenv("POLYSPLAT_SORT", "lsd")
    rng = np.random.default_rng(17)
    n = 30000
    t = rng.integers(0, 248832, n).astype(np.uint64)
    t[: n // 3] = rng.integers(0, 40, n // 3)  # some long tiles as well
    d = rng.integers(0x3f000000, 0x42000000, n).astype(np.uint64)
    keys = (t << np.uint64(32)) | d
    vals = rng.permutation(n).astype(np.uint32)
    k, v = _sort(emu, keys, vals, 9216, 6912)
    rk, rv = O.sort_pairs(keys, vals)
    np.testing.assert_array_equal(k[:n], rk)
    np.testing.assert_array_equal(v[:n], rv)


def _scene():
    return scenes.make_scene(3000, 10.0, [(0.8, -4.0, 0.5), (0.2, -1.8, 0.4)], seed=3)


@pytest.mark.parametrize("garbage", [False, True])
def test_frame_bit_exact_against_the_oracle(emu, garbage):
    sc = _scene()
    spec = scenes.make_camera(200, 120, 10.0, distance=12.0)
    ref = oracle_frame(sc, spec, sh_type="fp16", feat_prec="fp16", bg=(0.1, 0.2, 0.3),
                       check_every=8, fp64=False)
    M = ref["pre"]["M"]
    f = HF.HostFrame(emu, sc, spec, capacity=M + 100, garbage=garbage, seed=7)
    st, n, img = f.forward((0.1, 0.2, 0.3))
    capi.check_with(emu, st)
    assert n == M
    np.testing.assert_array_equal(f.ks[:M], ref["keys_sorted"])
    np.testing.assert_array_equal(f.vs[:M], ref["vals_sorted"])
    assert np.abs(img.astype(int) - ref["image"].astype(int)).max() <= 1


@pytest.mark.parametrize("size,threads,quads,frames", [
    ((16, 16), 1, "0", 1), ((16, 16), 1, "1", 1), ((48, 16), 8, "0", 1), ((48, 16), 8, "1", 1),
    ((64, 32), 1, "auto", 1), ((100, 60), 8, "0", 1), ((100, 60), 8, "1", 2),
    ((400, 240), 8, "0", 2)])
def test_persistent_queue_covers_every_tile(emu, size, threads, quads, frames, monkeypatch):
    # 1, 3, 8, 28 and 375 tiles (the emulated device keeps 8 waves in
    # flight): per-XCD queues that are empty, single-tile or uneven, and
    # stealing; items that are whole tiles or (tile, quadrant) pairs, forced
    # or chosen (8 tiles: by the kernel, from the pair count); a second frame
    # on the same workspace (the queue words must be left zero).  Every pixel
    # must match the oracle.
    monkeypatch.setenv("EMU_THREADS", str(threads))
    if quads != "auto":
        monkeypatch.setenv("POLYSPLAT_RENDER_QUADS", quads)
    sc = _scene()
    spec = scenes.make_camera(size[0], size[1], 10.0, distance=12.0)
    ref = oracle_frame(sc, spec, sh_type="fp16", feat_prec="fp16", bg=(0.5, 0.25, 0.125),
                       check_every=8, fp64=False)
    f = HF.HostFrame(emu, sc, spec, capacity=ref["pre"]["M"] + 100)
    for _ in range(frames):
        st, n, img = f.forward((0.5, 0.25, 0.125))
        capi.check_with(emu, st)
        assert np.abs(img.astype(int) - ref["image"].astype(int)).max() <= 1
        assert not f.ws[:2048].any()


def test_batched_frames_equal_single_frames(emu):
    a = _scene()
    b = scenes.make_scene(1500, 6.0, [(0.7, -3.5, 0.5), (0.3, -1.5, 0.3)], seed=8)
    specs = [scenes.make_camera(200, 120, 10.0, distance=12.0),
             scenes.make_camera(97, 65, 6.0, distance=7.0)]
    bgs = [(0.0, 0.0, 0.0), (1.0, 0.5, 0.25)]
    frames = [HF.HostFrame(emu, s, sp, capacity=200_000) for s, sp in zip((a, b), specs)]
    singles = [f.forward(bg) for f, bg in zip(frames, bgs)]
    for st, _, _ in singles:
        capi.check_with(emu, st)
    for order in ([0, 1], [1, 0]):
        fr = [frames[i] for i in order]
        batch = HF.HostBatch(emu, fr, 400_000, garbage=True)
        st, counts, imgs = batch.run([bgs[i] for i in order])
        capi.check_with(emu, st)
        assert counts == [singles[i][1] for i in order]
        for img, i in zip(imgs, order):
            np.testing.assert_array_equal(img, singles[i][2])


def test_renderer_pads_partial_groups_neutrally(emu):
    # Tile lists of every length modulo the check cadence (and past one and
    # two 64-Gaussian batches): the renderer shades whole groups, the slots
    # past a list's end must change nothing.  Dense, large splats give long
    # lists; the persistent (cadence 8) and plain (cadence 2) paths both run.
    sc = scenes.make_scene(2500, 6.0, [(0.6, -3.0, 0.5), (0.4, -1.2, 0.4)], seed=31)
    spec = scenes.make_camera(96, 80, 6.0, distance=8.0)
    ref = oracle_frame(sc, spec, sh_type="fp32", feat_prec="fp32", bg=(0.3, 0.1, 0.2),
                       check_every=8, fp64=False)
    lens = np.diff(ref["ranges"], axis=1).ravel()
    assert len(set(lens % 8)) >= 6 and lens.max() > 128, lens
    f = HF.HostFrame(emu, sc, spec, sh_type="fp32", feat_prec="fp32",
                     capacity=ref["pre"]["M"] + 64, garbage=True, seed=4)
    st, n, img = f.forward((0.3, 0.1, 0.2))
    capi.check_with(emu, st)
    assert n == ref["pre"]["M"]
    assert np.abs(img.astype(int) - ref["image"].astype(int)).max() <= 1
    # plain renderer over the same sorted pairs (cadence 2)
    out = np.zeros_like(img)
    keys = f.ks[:n].copy()
    capi.check_with(emu, emu.ps_render_16x16(n, None, spec.width, spec.height, HF.ptr(f.xy),
                                             HF.ptr(f.rgbd), HF.ptr(f.conic), HF.ptr(keys),
                                             HF.ptr(f.vs), capi.floats3((0.3, 0.1, 0.2)),
                                             HF.ptr(out), None))
    ref2 = oracle_frame(sc, spec, sh_type="fp32", feat_prec="fp32", bg=(0.3, 0.1, 0.2),
                        check_every=2, fp64=False)
    assert np.abs(out.astype(int) - ref2["image"].astype(int)).max() <= 1


@pytest.mark.parametrize("quads", ["0", "1"])
def test_renderer_sparse_and_dense_groups_match_the_oracle(emu, quads, monkeypatch):
    # A city crop: sub-pixel splats (the sparse quadrant path) over a few
    # large ones (the dense path), and a mixed-size scene; images within one
    # 8-bit level of the oracle, with almost every pixel identical; with
    # whole-tile and with quadrant work items.
    monkeypatch.setenv("POLYSPLAT_RENDER_QUADS", quads)
    city = scenes.make_config_scene("c4", seed=4, P=300_000)
    cam = scenes.config_camera("c4")
    cam.width, cam.height = 320, 240
    mixed = scenes.make_scene(6000, 8.0, [(0.7, -4.5, 0.5), (0.25, -2.2, 0.4), (0.05, -0.8, 0.3)],
                              seed=13)
    for sc, spec in ((city, cam), (mixed, scenes.make_camera(200, 136, 8.0, distance=10.0))):
        ref = oracle_frame(sc, spec, sh_type="fp16", feat_prec="fp16", bg=(0.1, 0.0, 0.3),
                           check_every=8, fp64=False)
        f = HF.HostFrame(emu, sc, spec, capacity=ref["pre"]["M"] + 64, garbage=True, seed=9)
        st, n, img = f.forward((0.1, 0.0, 0.3))
        capi.check_with(emu, st)
        assert n == ref["pre"]["M"]
        d = np.abs(img.astype(int) - ref["image"].astype(int))
        assert d.max() <= 1 and (d > 0).mean() < 1e-3, (d.max(), (d > 0).mean())


# ------------------------------------------------------------------ key emission
def _emission_scene():
    # Mostly small splats plus a few that cover hundreds of tiles: most
    # Gaussians leave a hit mask for the emit pass, the large ones (rects of
    # more than 32 tiles, fragmented and massive) are projected and swept
    # again there.
    sc = scenes.make_scene(20000, 6.0, [(0.9, -3.5, 0.5), (0.097, -2.0, 0.4), (0.003, 0.2, 0.3)],
                           seed=41)
    return sc, scenes.make_camera(640, 480, 6.0, distance=7.0)


@pytest.mark.parametrize("block", [(16, 16), (8, 8), (32, 32), (16, 8), (12, 20)])
def test_preprocess_tile_sizes(emu, block):
    """Keys for tile sizes other than 16x16 (the rect bounds divide by the tile
    size; 16 multiplies by 1/16, exact for a power of two) against the oracle
    run with the same tile size."""
    sc = _scene()
    spec = scenes.make_camera(200, 120, 10.0, distance=12.0)
    bx, by = block
    cam_o = O.camera(spec.width, spec.height, spec.position.numpy(), spec.rotation.numpy(),
                     spec.focal_x, spec.focal_y, spec.zFar, spec.zNear, bx, by)
    ref = O.preprocess(sc.position.numpy(), sc.shs.numpy(), sc.opacity.numpy(),
                       sc.cov3d.numpy(), cam_o)
    M = ref["M"]
    P = sc.num_vertex
    pos = np.ascontiguousarray(sc.position.numpy())
    shs = np.ascontiguousarray(sc.shs.numpy())
    op = np.ascontiguousarray(sc.opacity.numpy())
    cov = np.ascontiguousarray(sc.cov3d.numpy())
    xy, rgbd, con = (np.zeros((P, 2), np.float32), np.zeros((P, 4), np.float32),
                     np.zeros((P, 4), np.float32))
    ku = np.zeros(M + 8, np.uint64)
    vu = np.zeros(M + 8, np.uint32)
    curr = np.zeros(1, np.int32)
    nb = emu.ps_preprocess_workspace_size(P)
    ws = np.zeros(nb, np.uint8)
    cam = capi.camera(spec, block=bx)
    cam.block_y = by
    st = emu.ps_preprocess(P, HF.ptr(pos), HF.ptr(shs), 0, HF.ptr(op), HF.ptr(cov),
                           capi.C.byref(cam), HF.ptr(xy), HF.ptr(rgbd), HF.ptr(con), None, 1,
                           HF.ptr(ku), HF.ptr(vu), M + 8, HF.ptr(curr), HF.ptr(ws), nb, None)
    capi.check_with(emu, st)
    assert int(curr[0]) == M
    np.testing.assert_array_equal(ku[:M], ref["keys"])
    np.testing.assert_array_equal(vu[:M], ref["vals"])


def _preprocess(L, sc, spec, capacity, extra=0, base=0, fill=-7, pool_taken=None):
    """ps_preprocess into key arrays `extra` entries longer than capacity.
    pool_taken (a list) receives the hit-pool words the count pass asked for
    (state words 4-5)."""
    P = sc.num_vertex
    pos = np.ascontiguousarray(sc.position.numpy())
    shs = np.ascontiguousarray(sc.shs.numpy())
    op = np.ascontiguousarray(sc.opacity.numpy())
    cov = np.ascontiguousarray(sc.cov3d.numpy())
    xy = np.zeros((P, 2), np.float32)
    rgbd = np.zeros((P, 4), np.float32)
    con = np.zeros((P, 4), np.float32)
    ku = np.full(capacity + extra, fill, np.int64).view(np.uint64)
    vu = np.full(capacity + extra, fill, np.int32).view(np.uint32)
    curr = np.array([base], np.int32)
    nb = L.ps_preprocess_workspace_size(P)
    ws = np.random.default_rng(2).integers(0, 256, nb, dtype=np.uint8)  # not cleared
    cam = capi.camera(spec)
    st = L.ps_preprocess(P, HF.ptr(pos), HF.ptr(shs), 0, HF.ptr(op), HF.ptr(cov),
                         capi.C.byref(cam), HF.ptr(xy), HF.ptr(rgbd), HF.ptr(con), None, 1,
                         HF.ptr(ku), HF.ptr(vu), capacity, HF.ptr(curr), HF.ptr(ws), nb, None)
    status = int(ws[8:12].view(np.int32)[0])  # state word 2
    count = int(ws[4:8].view(np.int32)[0])  # state word 1
    if pool_taken is not None:
        pool_taken.append(int(ws[16:24].view(np.uint64)[0]))
    return st, int(curr[0]), status, count, ku, vu, (xy, rgbd, con)


@pytest.mark.parametrize("threads", ["1", "8"])
def test_record_emission_matches_oracle_and_three_pass(emu, threads, monkeypatch):
    monkeypatch.setenv("EMU_THREADS", threads)
    sc, spec = _emission_scene()
    cam = O.camera_from_spec(spec)
    ref = O.preprocess(sc.position.numpy(), sc.shs.numpy(), sc.opacity.numpy(),
                       sc.cov3d.numpy(), cam)
    M = ref["M"]
    st, curr, status, count, ku, vu, feats = _preprocess(emu, sc, spec, M + 10, extra=5)
    capi.check_with(emu, st)
    assert (curr, status, count) == (M, 0, M)
    np.testing.assert_array_equal(ku[:M], ref["keys"])
    np.testing.assert_array_equal(vu[:M], ref["vals"])
    assert (ku[M:].view(np.int64) == -7).all()
    vis = ref["counts"] > 0
    np.testing.assert_array_equal(feats[0][vis], ref["xy"][vis])
    np.testing.assert_array_equal(feats[2][vis], ref["conic"][vis])
    np.testing.assert_array_equal(feats[1][vis], ref["rgbd"][vis])
    # three-launch form, and appending after existing pairs
    monkeypatch.setenv("POLYSPLAT_EMIT", "3pass")
    st, curr3, _, _, ku3, vu3, _ = _preprocess(emu, sc, spec, M + 10)
    capi.check_with(emu, st)
    np.testing.assert_array_equal(ku3[:M], ku[:M])
    np.testing.assert_array_equal(vu3[:M], vu[:M])
    monkeypatch.delenv("POLYSPLAT_EMIT")
    st, curr2, status, count, ku2, vu2, _ = _preprocess(emu, sc, spec, M + 100, base=77)
    capi.check_with(emu, st)
    assert (curr2, status, count) == (M + 77, 0, M + 77)
    np.testing.assert_array_equal(ku2[77:77 + M], ref["keys"])
    assert (ku2[:77].view(np.int64) == -7).all()


@pytest.mark.parametrize("threads", ["1", "8"])
def test_record_emission_overflow_writes_nothing_past_capacity(emu, threads, monkeypatch):
    monkeypatch.setenv("EMU_THREADS", threads)
    sc, spec = _emission_scene()
    cam = O.camera_from_spec(spec)
    ref = O.preprocess(sc.position.numpy(), sc.shs.numpy(), sc.opacity.numpy(),
                       sc.cov3d.numpy(), cam)
    M = ref["M"]
    for cap in (M - 1, M // 2, 0):
        st, curr, status, count, ku, vu, _ = _preprocess(emu, sc, spec, cap, extra=M)
        capi.check_with(emu, st)
        assert curr == M and status & 1 and count == 0, (cap, curr, status, count)
        assert (ku[cap:].view(np.int64) == -7).all() and (vu[cap:].view(np.int32) == -7).all()
        # what was written below capacity are whole workgroups' canonical pairs
        done = ku[:cap].view(np.int64) != -7
        np.testing.assert_array_equal(ku[:cap][done], ref["keys"][:cap][done])
    # the three-launch form writes nothing at all
    monkeypatch.setenv("POLYSPLAT_EMIT", "3pass")
    st, curr, status, count, ku, vu, _ = _preprocess(emu, sc, spec, M - 1, extra=M)
    assert curr == M and status & 1 and (ku.view(np.int64) == -7).all()
    # a negative starting offset is an overflow, with nothing written
    monkeypatch.delenv("POLYSPLAT_EMIT")
    st, curr, status, count, ku, vu, _ = _preprocess(emu, sc, spec, M + 10, base=-5)
    assert status & 1 and count == 0 and (ku.view(np.int64) == -7).all()


@pytest.mark.parametrize("threads", ["1", "8"])
def test_record_emission_with_a_full_hit_pool(emu, threads, monkeypatch):
    """Large rects keep their count-pass ballots in the hit pool; when the
    pool is full, the rest are projected and swept again.  Half of 2000
    splats cover all of an 800x600 frame (50 x 38 tiles, 30 pool words
    each): every workgroup needs more than its own 64 words, and together
    they ask the shared overflow area (8 * 16 + 16384 words for 8
    workgroups) for about twice what it holds, so both paths run."""
    monkeypatch.setenv("EMU_THREADS", threads)
    sc = scenes.make_scene(2000, 1.0, [(0.5, 0.0, 0.2), (0.3, -2.6, 0.3), (0.2, -4.0, 0.3)],
                           seed=43)
    spec = scenes.make_camera(800, 600, 1.0, distance=4.0)
    cam = O.camera_from_spec(spec)
    ref = O.preprocess(sc.position.numpy(), sc.shs.numpy(), sc.opacity.numpy(),
                       sc.cov3d.numpy(), cam)
    M = ref["M"]
    taken = []
    st, curr, status, count, ku, vu, _ = _preprocess(emu, sc, spec, M + 10, extra=5,
                                                      pool_taken=taken)
    capi.check_with(emu, st)
    assert taken[0] > 8 * 16 + 16384, taken  # more words asked for than the overflow holds
    assert (curr, status, count) == (M, 0, M)
    np.testing.assert_array_equal(ku[:M], ref["keys"])
    np.testing.assert_array_equal(vu[:M], ref["vals"])


def test_posted_count_readback(emu):
    """ps_forward(sync_count=2) posts the pair count for ps_forward_wait: the
    same count and image as a synchronous frame, overflow reported at the
    wait, every ticket waited for once (the limit of 64 outstanding tickets
    is checked on the GPU, where 65 frames take milliseconds)."""
    sc = _scene()
    spec = scenes.make_camera(200, 120, 10.0, distance=12.0)
    ref = oracle_frame(sc, spec, sh_type="fp16", feat_prec="fp16", bg=(0.1, 0.2, 0.3),
                       check_every=8, fp64=False)
    M = ref["pre"]["M"]
    f = HF.HostFrame(emu, sc, spec, capacity=M + 100)
    st, ticket, img = f.forward((0.1, 0.2, 0.3), sync=2)
    capi.check_with(emu, st)
    assert 0 <= ticket < 64 and emu.ps_forward_query(ticket) == 1
    n = capi.C.c_int(-1)
    capi.check_with(emu, emu.ps_forward_wait(ticket, capi.C.byref(n)))
    assert n.value == M
    assert np.abs(img.astype(int) - ref["image"].astype(int)).max() <= 1
    assert emu.ps_forward_wait(ticket, capi.C.byref(n)) == capi.PS_ERR_INVALID  # waited already
    assert emu.ps_forward_query(ticket) < 0
    small = HF.HostFrame(emu, sc, spec, capacity=M - 1)
    st, t2, _ = small.forward((0.1, 0.2, 0.3), sync=2)
    capi.check_with(emu, st)
    assert emu.ps_forward_wait(t2, capi.C.byref(n)) == capi.PS_ERR_OVERFLOW and n.value == M


# ------------------------------------------------------------------ depth-first pipeline
DEPTH_FIRST = 0x100
TILE_WORD = np.uint64(0xFFFFFFFF00000000)


def _flat_scene(P=600, seed=47):
    """Every Gaussian at the same depth: no digit of the depth sort runs."""
    sc = scenes.make_scene(P, 4.0, [(0.8, -3.0, 0.4), (0.2, -1.5, 0.3)], seed=seed)
    sc.position[:, 2] = 0.0
    return sc


@pytest.mark.parametrize("case,variant", [
    ("mixed", capi.V_LITE), ("mixed", capi.V_FUSED), ("mixed", capi.V_DEFAULT),
    ("pool_overflow", capi.V_LITE), ("flat", capi.V_LITE), ("empty", capi.V_LITE),
    ("routes", capi.V_LITE)])
def test_depth_first_pipeline_equals_default(emu, case, variant):
    if case == "mixed":
        sc, spec = _scene(), scenes.make_camera(200, 120, 10.0, distance=12.0)
    elif case == "pool_overflow":
        sc = scenes.make_scene(2000, 1.0, [(0.5, 0.0, 0.2), (0.3, -2.6, 0.3), (0.2, -4.0, 0.3)],
                               seed=43)
        spec = scenes.make_camera(300, 200, 1.0, distance=4.0)
    elif case == "flat":
        sc, spec = _flat_scene(), scenes.make_camera(160, 96, 4.0, distance=6.0)
    elif case == "empty":
        sc = scenes.make_scene(300, 2.0, [(1.0, -3.0, 0.3)], seed=5)
        sc.position[:, 2] -= 50.0  # behind the camera
        spec = scenes.make_camera(64, 48, 2.0)
    else:
        sc, spec = _emission_scene()
        spec = scenes.make_camera(320, 240, 6.0, distance=7.0)
    prec = "fp32" if variant == capi.V_DEFAULT else "fp16"
    bg = (0.3, 0.1, 0.6)
    ref = oracle_frame(sc, spec, sh_type=prec, feat_prec=prec, bg=bg,
                       check_every={capi.V_LITE: 8, capi.V_FUSED: 16, capi.V_DEFAULT: 2}[variant],
                       fp64=False)
    M = ref["pre"]["M"]
    std = HF.HostFrame(emu, sc, spec, sh_type=prec, feat_prec=prec, capacity=M + 40)
    st, n, img_std = std.forward(bg, variant)
    capi.check_with(emu, st)
    es = HF.HostFrame(emu, sc, spec, sh_type=prec, feat_prec=prec, capacity=M + 40,
                      garbage=True, seed=9)
    st, n2, img_es = es.forward(bg, variant | DEPTH_FIRST)
    capi.check_with(emu, st)
    assert n == n2 == M
    # sorted keys carry the tile id only (example.py:560-576, sort.cu:205-215)
    np.testing.assert_array_equal(es.ks[:M], ref["keys_sorted"] & TILE_WORD)
    np.testing.assert_array_equal(es.vs[:M], ref["vals_sorted"])
    np.testing.assert_array_equal(img_es, img_std)
    if variant == capi.V_LITE:
        np.testing.assert_array_equal(es.ranges, std.ranges)
    assert np.abs(img_es.astype(int) - ref["image"].astype(int)).max() <= 1


def test_depth_first_overflow_is_reported(emu):
    sc, spec = _scene(), scenes.make_camera(200, 120, 10.0, distance=12.0)
    ref = oracle_frame(sc, spec, sh_type="fp16", feat_prec="fp16", fp64=False)
    M = ref["pre"]["M"]
    f = HF.HostFrame(emu, sc, spec, capacity=M - 1)
    st, n, img = f.forward((0.5, 0.5, 0.5), capi.V_LITE | DEPTH_FIRST)
    assert st == capi.PS_ERR_OVERFLOW and n == M
    assert (img == 127).all()  # nothing sorted: every tile shows the background


def test_depth_first_staged_entry_points(emu):
    """ps_preprocess_depth_first then ps_es_tile_sort with the count from the
    host: the pairs in depth order, their stable tile sort, the keys' tile
    words and the ranges against the oracle."""
    sc, spec = _scene(), scenes.make_camera(200, 120, 10.0, distance=12.0)
    ref = oracle_frame(sc, spec, sh_type="fp32", feat_prec="fp32", fp64=False)
    M = ref["pre"]["M"]
    f = HF.HostFrame(emu, sc, spec, sh_type="fp32", feat_prec="fp32", capacity=M + 40,
                     garbage=True, seed=3)
    L = emu
    ws_bytes = L.ps_preprocess_workspace_size(f.P)
    ws = np.random.default_rng(4).integers(0, 256, ws_bytes, dtype=np.uint8)
    tk = np.zeros(2 * f.cap, np.uint32)
    st = L.ps_preprocess_depth_first(f.P, HF.ptr(f.pos), HF.ptr(f.shs), 0, HF.ptr(f.opac),
                                     HF.ptr(f.cov), capi.C.byref(f.cam), HF.ptr(f.xy),
                                     HF.ptr(f.rgbd), HF.ptr(f.conic), None, 0, HF.ptr(tk),
                                     HF.ptr(f.vu), f.cap, HF.ptr(f.curr), HF.ptr(ws), ws_bytes,
                                     HF.ptr(f.sort_space), f.sort_bytes, None)
    capi.check_with(L, st)
    n = int(f.curr[0])
    assert n == M
    # pairs in (depth, Gaussian id) order: ordering the oracle's canonical
    # pairs by (depth, id) and keeping each Gaussian's tiles in row-major order
    keys, vals = ref["pre"]["keys"], ref["pre"]["vals"]
    order = np.lexsort((np.arange(M), vals, keys & np.uint64(0xFFFFFFFF)))
    np.testing.assert_array_equal(tk[:M], (keys[order] >> np.uint64(32)).astype(np.uint32))
    np.testing.assert_array_equal(f.vu[:M], vals[order])
    st = L.ps_es_tile_sort(n, None, f.cap, spec.width, spec.height, 16, 16,
                           HF.ptr(f.sort_space), f.sort_bytes, HF.ptr(tk), HF.ptr(f.vu),
                           HF.ptr(tk[f.cap:]), HF.ptr(f.vs), HF.ptr(f.ks), HF.ptr(f.ranges), None)
    capi.check_with(L, st)
    np.testing.assert_array_equal(f.ks[:M], ref["keys_sorted"] & TILE_WORD)
    np.testing.assert_array_equal(f.vs[:M], ref["vals_sorted"])
    np.testing.assert_array_equal(f.ranges, O.tile_ranges(ref["keys_sorted"], f.tiles, 1))
    # aliasing outputs and inputs is refused
    assert L.ps_es_tile_sort(n, None, f.cap, spec.width, spec.height, 16, 16,
                             HF.ptr(f.sort_space), f.sort_bytes, HF.ptr(tk), HF.ptr(f.vu),
                             HF.ptr(tk), HF.ptr(f.vs), HF.ptr(f.ks), None, None) == capi.PS_ERR_INVALID
