"""The serving loop: jobs over resident scenes, admitted by frame-buffer footprint, batched together.

A job's image depends only on the job and the server's min_items (how small jobs are split into sample
sub-streams), never on what else shares the wavefronts: it must be bit-identical to the same job served
alone, whatever the arrival order, slot count, schedule and admission budget. Without splitting it is
also bit-identical to the single-job renderer (which follows the reference's GpuRenderer::render); with
splitting it is within the workload's radiance tolerance of it.
"""
import numpy as np
import pytest

from photon_amd import gpu, host
from photon_amd.types import RenderSettings

from gpu_util import radiance_agreement, uploaded

pytestmark = pytest.mark.gpu

# (scene name, width, height, spp, max_depth, seed): mixed resolutions, sample counts and depths.
JOBS = [("cornell-c1", 64, 64, 8, 8, 1), ("materials", 96, 54, 4, 6, 2), ("glass", 40, 30, 12, 10, 3),
        ("cornell-c1", 33, 17, 5, 3, 4), ("furnace", 48, 48, 2, 8, 5), ("materials", 20, 90, 6, 12, 6)]


@pytest.fixture(scope="module")
def scenes():
    return {n: host.Scene.builtin(n).build() for n in sorted({j[0] for j in JOBS})}


def _alone(scene, w, h, spp, depth, seed):
    r = uploaded(scene, w, h, engine=gpu.ENGINE_WAVEFRONT)
    r.render(scene.camera_view(w / h), RenderSettings.make(w, h, max_depth=depth, rr_depth=3, seed=seed), spp)
    img = r.download_image()
    r.close()
    return img


def _serve(scenes, jobs, order, **cfg):
    srv = gpu.Server(**cfg)
    ids = {n: srv.add_scene(s) for n, s in scenes.items()}
    jid = {}
    for k in order:
        n, w, h, spp, depth, seed = jobs[k]
        jid[k] = srv.submit(ids[n], w, h, spp, scenes[n].camera_view(w / h), max_depth=depth, seed=seed)
    srv.run_until_idle()
    out = {}
    for k in order:
        st = srv.status(jid[k])
        assert st.state == gpu.JOB_DONE and st.samples_done == st.samples == st.pixels * jobs[k][3]
        assert st.submit_ms <= st.admit_ms <= st.done_ms
        out[k] = srv.result(jid[k])
    srv.close()
    return out


def test_each_job_alone_equals_the_renderer(scenes):
    for k, (n, w, h, spp, depth, seed) in enumerate(JOBS):
        ref = _alone(scenes[n], w, h, spp, depth, seed)
        got = _serve(scenes, JOBS, [k], path_slots=1 << 14, min_items=1)[k]
        assert got.tobytes() == ref.tobytes(), f"job {k} ({n} {w}x{h})"


def test_split_jobs_match_the_renderer_and_themselves(scenes):
    """Every job split into sample sub-streams: within tolerance of the renderer, bitwise alone vs mix."""
    cfg = dict(path_slots=1 << 14, min_items=50000)
    alone = {k: _serve(scenes, JOBS, [k], **cfg)[k] for k in range(len(JOBS))}
    mixed = _serve(scenes, JOBS, [5, 2, 0, 4, 1, 3], path_slots=3000, min_items=50000, pass_samples=7000)
    for k, (n, w, h, spp, depth, seed) in enumerate(JOBS):
        ref = _alone(scenes[n], w, h, spp, depth, seed)
        agree = radiance_agreement(alone[k], ref)
        print(f"job {k}: split agreement {100 * agree:.3f}%")
        assert agree >= 0.999
        assert mixed[k].tobytes() == alone[k].tobytes(), f"job {k}"


@pytest.mark.parametrize("cfg", [dict(path_slots=1 << 14), dict(path_slots=3000, iterations_per_step=2),
                                 dict(path_slots=1 << 16, max_jobs=2), dict(path_slots=5000, memory_budget=64 * 64 * 16),
                                 dict(path_slots=4096, pass_samples=3000), dict(path_slots=2000, pass_samples=1),
                                 dict(path_slots=3000, pass_samples=5000, schedule=gpu.SCHEDULE_FIFO),
                                 dict(path_slots=3000, pass_samples=4000, schedule=gpu.SCHEDULE_HRRN),
                                 dict(path_slots=3000, pass_samples=5000, schedule=gpu.SCHEDULE_SHARE),
                                 dict(path_slots=1 << 14, schedule=gpu.SCHEDULE_SHARE, iterations_per_step=1),
                                 dict(path_slots=3000, pass_samples=4000, schedule=gpu.SCHEDULE_DEADLINE),
                                 dict(path_slots=1 << 14, schedule=gpu.SCHEDULE_DEADLINE, deadline_stretch=0.5, iterations_per_step=1),
                                 dict(path_slots=1 << 14, express_slots=-1),
                                 dict(path_slots=4096, express_slots=512, express_samples=40000, express_iterations=1),
                                 dict(path_slots=2048, express_slots=2048, express_samples=12000, iterations_per_step=1)])
def test_jobs_in_a_mix_equal_jobs_alone(scenes, cfg):
    # The same splitting of small jobs in both runs (it is part of what defines a job's image).
    alone = {k: _serve(scenes, JOBS, [k], path_slots=1 << 14, min_items=2048)[k] for k in range(len(JOBS))}
    order = [3, 0, 5, 1, 4, 2]
    mixed = _serve(scenes, JOBS, order, min_items=2048, **cfg)
    for k in order:
        assert mixed[k].tobytes() == alone[k].tobytes(), f"job {k} differs in the mix ({cfg})"


def test_jobs_submitted_while_others_run(scenes):
    """Admission between steps: new work items start after those already handed out."""
    srv = gpu.Server(path_slots=2048, iterations_per_step=2, min_items=1)
    ids = {n: srv.add_scene(s) for n, s in scenes.items()}
    jid = {}
    pending = list(range(len(JOBS)))
    steps = 0
    while pending or srv.pending:
        if pending and steps % 3 == 0:
            k = pending.pop(0)
            n, w, h, spp, depth, seed = JOBS[k]
            jid[k] = srv.submit(ids[n], w, h, spp, scenes[n].camera_view(w / h), max_depth=depth, seed=seed)
        srv.step()
        steps += 1
        assert steps < 100000
    for k, j in jid.items():
        n, w, h, spp, depth, seed = JOBS[k]
        ref = _alone(scenes[n], w, h, spp, depth, seed)
        assert srv.result(j).tobytes() == ref.tobytes(), f"job {k}"
    srv.close()


def test_express_lane_runs_beside_the_main_lane(scenes):
    """Small jobs in the express lane finish while a large job holds the main lane."""
    big = ("materials", 320, 240, 64, 8, 7)
    jobs = JOBS + [big]
    srv = gpu.Server(path_slots=1 << 14, express_slots=4096, express_samples=40000, min_items=1)
    ids = {n: srv.add_scene(s) for n, s in scenes.items()}
    n, w, h, spp, depth, seed = big
    jb = srv.submit(ids[n], w, h, spp, scenes[n].camera_view(w / h), max_depth=depth, seed=seed)
    js = {}
    for k, (n, w, h, spp, depth, seed) in enumerate(JOBS):
        js[k] = srv.submit(ids[n], w, h, spp, scenes[n].camera_view(w / h), max_depth=depth, seed=seed)
    srv.run_until_idle()
    small_done = max(srv.status(j).done_ms for j in js.values())
    assert small_done < srv.status(jb).done_ms
    for k, j in js.items():
        n, w, h, spp, depth, seed = jobs[k]
        assert srv.result(j).tobytes() == _alone(scenes[n], w, h, spp, depth, seed).tobytes(), f"job {k}"
    n, w, h, spp, depth, seed = big
    assert srv.result(jb).tobytes() == _alone(scenes[n], w, h, spp, depth, seed).tobytes()
    srv.close()


def test_errors():
    srv = gpu.Server(path_slots=1024)
    with pytest.raises(RuntimeError):
        srv.submit(0, 8, 8, 1, host.Scene.builtin("cornell").build().camera_view(1.0))
    sid = srv.add_scene(host.Scene.builtin("cornell").build())
    j = srv.submit(sid, 8, 8, 1, host.Scene.builtin("cornell").build().camera_view(1.0))
    with pytest.raises(RuntimeError, match="not done"):
        srv.result(j)
    srv.run_until_idle()
    srv.result(j)
    with pytest.raises(RuntimeError, match="already"):
        srv.result(j)
    srv.close()
