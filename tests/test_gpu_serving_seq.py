"""Serving loop on the GPU with the single-stream executors (frames waited
for one by one, and frames enqueued back to back with their pair counts
read back later): every request of a mixed stream (three resident scenes,
three resolutions, orbit poses, requests grouped by the scheduler, key
buffers grown on overflow) is byte-identical to the same request rendered
alone through the single-frame entry point with buffers of its own."""
import math

import numpy as np
import pytest
import torch

from polysplat_amd import _C
from polysplat_amd import serving as SV
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu

RES = ((320, 180), (640, 360), (800, 600))
SIZES = (30_000, 80_000, 200_000)
BOX = 10.0


def _alone(sc, r, radius):
    dev = torch.device("cuda:0")
    cam = SV.orbit_camera(r.width, r.height, r.azimuth, radius)
    cap = r.pairs + 16
    z = lambda n, dt: torch.zeros(n, dtype=dt, device=dev)  # noqa: E731
    ku, vu, ks, vs = z(cap, torch.int64), z(cap, torch.int32), z(cap, torch.int64), z(cap, torch.int32)
    space = z(_C.ops.get_sort_buffer_size(cap), torch.int8)
    curr = z(1, torch.int32)
    P = sc["P"]
    e = torch.empty(0, dtype=torch.float32, device=dev)
    rec = z(24 * P, torch.uint8)
    tiles = ((r.width + 15) // 16) * ((r.height + 15) // 16)
    ranges = torch.zeros((tiles, 2), dtype=torch.int32, device=dev)
    out = torch.zeros((r.height, r.width, 3), dtype=torch.uint8, device=dev)
    ws = z(_C.ops.forward_workspace_size(P), torch.uint8)
    n = _C.ops.forward_frame(sc["pos"], sc["shs"], sc["opac"], sc["cov"], r.width, r.height,
                             cam.position, cam.rotation, cam.focal_x, cam.focal_y, cam.zFar,
                             cam.zNear, 1, 1, curr, space, ku, vu, ks, vs, e, e, e, rec, ranges,
                             torch.tensor(r.bg, dtype=torch.float32), out, ws, True)
    return n, out.cpu().numpy()


@pytest.mark.parametrize("executor", ["SequentialExecutor", "PipelinedExecutor",
                                      "PipelinedExecutor+graphs",
                                      "PipelinedExecutor+graphs+tile_order"])
def test_serving_equals_frames_rendered_alone(gpu, executor):
    radius = 1.5 * BOX
    scenes = SV.make_resident_scenes(SIZES, device="cuda:0", box=BOX,
                                     mix=[(0.9, -3.5, 0.5), (0.1, -2.0, 0.4)])
    if "+graphs" in executor:
        # "+tile_order": every frame heaviest-first (thresholds at one pair)
        order = 1 if executor.endswith("+tile_order") else 0
        ex = SV.PipelinedExecutor(scenes, radius=radius, graphs=True, tile_order_pairs=order,
                                  tile_order_per_tile=order)
    else:
        ex = getattr(SV, executor)(scenes, radius=radius)
    est = SV.Estimates()
    say("sequential serving calibrate")
    SV.calibrate(ex, est, len(SIZES), resolutions=RES, azimuths=(0.0,), repeats=1)
    sched = SV.Scheduler(SIZES, 4 << 30, est, max_frames=6, max_batch_ms=50.0)
    reqs = SV.poisson_requests(36, 4000.0, len(SIZES), resolutions=RES, seed=11)
    for i, r in enumerate(reqs):
        r.bg = ((i % 3) / 2.0, 0.25, 1.0 - (i % 2))
    loop = SV.ServingLoop(sched, ex, keep_outputs=True)
    say("sequential serving run", len(reqs))
    done = loop.run(reqs)
    torch.cuda.synchronize()
    assert len(done) == len(reqs) and not sched.rejected
    summary = SV.summarize(done, sched.rejected, wall_s=ex.now_s(), batches=loop.batches)
    say("summary", summary, "grows", ex.grows)
    assert all(r.latency_ms > 0 and math.isfinite(r.latency_ms) for r in done)
    if "+graphs" in executor:
        # one capture per (scene, resolution), again after each key-buffer growth
        assert 0 < ex.captures <= len(SIZES) * len(RES) * (1 + ex.grows)
    if executor.endswith("+tile_order"):
        assert ex.ordered_frames > 0
    for r in done:
        n, img = _alone(scenes[r.scene], r, radius)
        assert n == r.pairs, (r.rid, n, r.pairs)
        mix = loop.outputs[r.rid].cpu().numpy()
        np.testing.assert_array_equal(mix, img, err_msg=f"request {r.rid} differs in the stream")
