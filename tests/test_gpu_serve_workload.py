"""The serving workload's own configuration on the device: its three resident synthetic scenes (the
10,242-triangle field, the 262,144-triangle sphere soup and the 1,048,578-triangle field) and a seeded
subset of its job stream (bench.draw_jobs, the stream the benchmark serves) that covers every
resolution (256^2 to 1920x1080), sample count (4 to 256) and depth (4, 8, 16) and every scene.

With the server's default configuration (2M path slots, the express lane for small jobs, SRPT), every
job must be bit-identical alone and in the mix. A job the server runs unsplit (at least min_items
pixels in its lane) must also be bit-identical to the single-job renderer; a job split into sample
sub-streams is within the radiance tolerance of it (99.9 % of pixels within 0.1 %).
"""
import os
import sys

import numpy as np
import pytest

from photon_amd import gpu, host
from photon_amd.types import RenderSettings

from gpu_util import bitwise_fraction, radiance_agreement, uploaded

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from bench import DEPTHS, RESOLUTIONS, SPPS, draw_jobs  # noqa: E402

pytestmark = pytest.mark.gpu

SLOTS = 1 << 21
MAX_SAMPLES = 80_000_000  # keeps the three renders of each job to a few seconds


def pick_jobs(limit=14):
    """The first jobs of the benchmark's stream that add a resolution, sample count, depth, scene,
    resolution-and-sample-count pair or scene-and-depth pair not yet covered (at most `limit`, skipping
    jobs above MAX_SAMPLES); plus the first 1080p job at 16 spp or more."""
    stream = draw_jobs(2000, seed=2024)
    seen = {"res": set(), "spp": set(), "depth": set(), "scene": set(), "res_spp": set(), "scene_depth": set()}
    out = []
    for j in stream:
        if j["w"] * j["h"] * j["spp"] > MAX_SAMPLES or len(out) >= limit:
            continue
        keys = {"res": (j["w"], j["h"]), "spp": j["spp"], "depth": j["depth"], "scene": j["scene"],
                "res_spp": (j["w"], j["h"], j["spp"]), "scene_depth": (j["scene"], j["depth"])}
        if any(v not in seen[k] for k, v in keys.items()):
            out.append(j)
            for k, v in keys.items():
                seen[k].add(v)
    big = next(j for j in stream if (j["w"], j["h"]) == (1920, 1080) and j["spp"] >= 16 and j["w"] * j["h"] * j["spp"] <= MAX_SAMPLES)
    if big not in out:
        out.append(big)
    assert seen["res"] == set(RESOLUTIONS) and seen["spp"] == set(SPPS) and seen["depth"] == set(DEPTHS) and seen["scene"] == {0, 1, 2}
    return out


@pytest.fixture(scope="module")
def scenes():
    return [host.Scene.field(10, seed=7, env_width=2048).build(), host.Scene.spheres(256, box=100.0, alpha_fraction=0.1, seed=1).build(),
            host.Scene.field(1024, seed=1, env_width=2048).build()]


def _submit(srv, ids, scenes, j):
    return srv.submit(ids[j["scene"]], j["w"], j["h"], j["spp"], scenes[j["scene"]].camera_view(j["w"] / j["h"]), max_depth=j["depth"],
                      rr_depth=3, seed=j["seed"])


def test_serving_workload_jobs_alone_and_in_the_mix(scenes):
    jobs = pick_jobs()
    assert [s.num_tris for s in scenes] == [10242, 262144, 1048578]
    print(f"{len(jobs)} jobs: " + ", ".join(f"{j['w']}x{j['h']}x{j['spp']} d{j['depth']} s{j['scene']}" for j in jobs))
    # Alone: one job at a time on an idle server.
    srv = gpu.Server(path_slots=SLOTS)
    ids = [srv.add_scene(s) for s in scenes]
    alone = []
    for j in jobs:
        k = _submit(srv, ids, scenes, j)
        srv.run_until_idle()
        alone.append(srv.result(k))
    srv.close()
    # In the mix: all submitted at once, in another order, and stepped together.
    srv = gpu.Server(path_slots=SLOTS)
    ids = [srv.add_scene(s) for s in scenes]
    order = list(range(len(jobs)))[::-1]
    handles = {i: _submit(srv, ids, scenes, jobs[i]) for i in order}
    steps = 0
    while srv.pending:
        srv.step()
        steps += 1
        assert steps < 200000
    mixed = {i: srv.result(h) for i, h in handles.items()}
    srv.close()
    # Express lane: jobs of at most 16 x (slots / 2) samples, split below slots / 2 pixels; main lane: split
    # below slots / 8 pixels.
    express_slots = SLOTS // 2
    for i, j in enumerate(jobs):
        assert mixed[i].tobytes() == alone[i].tobytes(), f"job {i} differs in the mix"
        pixels, samples = j["w"] * j["h"], j["w"] * j["h"] * j["spp"]
        unsplit = pixels >= (express_slots if samples <= 16 * express_slots else SLOTS // 8)
        r = uploaded(scenes[j["scene"]], j["w"], j["h"], engine=gpu.ENGINE_WAVEFRONT)
        rs = RenderSettings.make(j["w"], j["h"], max_depth=j["depth"], rr_depth=3, seed=j["seed"])
        r.render(scenes[j["scene"]].camera_view(j["w"] / j["h"]), rs, j["spp"])
        ref = r.download_image()
        r.close()
        agree, bits = radiance_agreement(alone[i], ref), bitwise_fraction(alone[i], ref)
        print(f"job {i} {j['w']}x{j['h']}x{j['spp']} d{j['depth']} scene {j['scene']}: {'unsplit' if unsplit else 'split'}, "
              f"{100 * agree:.3f}% within 0.1% of the renderer, {100 * bits:.2f}% bitwise")
        assert agree >= 0.999, i
        if unsplit:
            assert alone[i].tobytes() == ref.tobytes(), f"unsplit job {i} differs from the renderer"
