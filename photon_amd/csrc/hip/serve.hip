// The serving loop (photon_hip.h, ph_server_*): resident scenes, FIFO admission of render jobs by HBM
// footprint, and continuous batching of all admitted jobs' paths in one wavefront engine.
//
// A step admits the queued jobs that fit (frame buffer bytes within the budget, a free job-table entry),
// then advances every path in flight by a fixed number of bounces. A job runs as passes of at most
// `pass_samples` pixel samples; its next pass is queued once the previous one has finished. Unstarted
// work is served shortest-remaining-job first (or passes first come, first served), so a short job is
// not stuck behind a long one. Jobs finish independently, and a finished job's image can be read while
// the others keep running.
//
// A job with fewer than `min_items` pixels runs each pixel's samples as S sub-streams (S divides spp)
// in separate partial images, which the result adds in sub-stream order: its paths can then fill more
// slots than it has pixels. Within a sub-stream the samples are taken in order, pass after pass. An
// image therefore depends only on the job and the server's min_items, never on what else shared the
// wavefronts; with S = 1 it is bit-identical to the renderer's.
#include <chrono>
#include <deque>
#include <memory>
#include <vector>

#include "device_scene.h"
#include "internal.h"
#include "wavefront.h"

using namespace phd;
using namespace phk;

namespace {

struct ServedScene {
    DeviceScene scene;
    DevBuf<int> mat_class;
};

struct Job {
    ph_job_desc desc{};
    ph_job_status status{};
    int entry = -1;
    size_t bytes = 0;
    int split = 1;          // sub-streams per pixel
    int64_t items = 0;      // work items of a pass: pixels x split
    int pass_spp = 1;       // samples per pass
    int next_k = 0;         // first sample of the next pass to queue
    int passes_queued = 0;  // passes queued so far
    int pass_k0 = 0, pass_k1 = 0;
    bool pass_pending = false;  // the next pass could not be queued yet
    DevBuf<float4> accum;
};

}  // namespace

struct ph_server {
    ph_server_config cfg{};
    hipStream_t stream = nullptr;
    WavefrontEngine engine;
    std::vector<std::unique_ptr<ServedScene>> scenes;
    std::vector<std::unique_ptr<Job>> jobs;
    std::deque<int> queue;   // submitted, not admitted (FIFO)
    std::vector<int> running;
    int64_t mem_used = 0;
    int finished = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now_ms() const { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }

    ph_server(const ph_server_config& c) : cfg(c) {
        if (cfg.memory_budget <= 0) cfg.memory_budget = int64_t(16) << 30;
        if (cfg.path_slots <= 0) cfg.path_slots = 1 << 21;
        if (cfg.iterations_per_step <= 0) cfg.iterations_per_step = 8;
        if (cfg.max_jobs <= 0 || cfg.max_jobs > WavefrontEngine::kMaxJobs) cfg.max_jobs = WavefrontEngine::kMaxJobs;
        if (cfg.pass_samples <= 0) cfg.pass_samples = int64_t(4) * cfg.path_slots;
        if (cfg.schedule != PH_SCHEDULE_SRPT && cfg.schedule != PH_SCHEDULE_FIFO) throw std::invalid_argument("unknown schedule");
        if (cfg.min_items == 0) cfg.min_items = std::max(1, cfg.path_slots / 8);
        if (cfg.min_items < 0) cfg.min_items = 1;
        PH_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        engine.serve_begin(stream, cfg.path_slots);
    }
    ~ph_server() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }

    void admit() {
        while (!queue.empty() && int(running.size()) < cfg.max_jobs) {
            Job& j = *jobs[size_t(queue.front())];
            // A job larger than the whole budget is admitted alone rather than never.
            if (mem_used + int64_t(j.bytes) > cfg.memory_budget && !running.empty()) break;
            j.accum.allocate(size_t(j.items));
            j.accum.zero(stream);
            const ServedScene& sc = *scenes[size_t(j.desc.scene)];
            WavefrontJob w{};
            w.scene = sc.scene.view();
            w.camera = j.desc.camera;
            w.width = j.desc.width;
            w.height = j.desc.height;
            w.spp = j.desc.spp;
            w.frame0 = j.desc.frame0;
            w.seed = j.desc.seed;
            w.opt = PathOpts{j.desc.path.max_depth, j.desc.path.rr_depth, j.desc.path.clamp, j.desc.path.nee};
            w.accum = j.accum.ptr;
            w.albedo = nullptr;
            w.nd = nullptr;
            w.mat_class = sc.mat_class.ptr;
            w.split = j.split;
            const int entry = engine.serve_admit(stream, w);
            if (entry < 0) {
                j.accum.release();
                break;
            }
            j.entry = entry;
            j.status.state = PH_JOB_RUNNING;
            j.status.admit_ms = now_ms();
            const int64_t per_pass = cfg.pass_samples / std::max<int64_t>(1, j.items);
            j.pass_spp = int(std::min<int64_t>(std::max<int64_t>(per_pass, 1), j.desc.spp / j.split));
            queue_next_pass(j);
            mem_used += int64_t(j.bytes);
            running.push_back(queue.front());
            queue.pop_front();
        }
    }

    void queue_next_pass(Job& j) {
        const int span = j.desc.spp / j.split;
        const int k0 = j.next_k, k1 = std::min(span, k0 + j.pass_spp);
        // Shortest remaining work first, or passes in the order they are queued.
        const double priority = cfg.schedule == PH_SCHEDULE_FIFO ? 0.0 : double(j.items) * double(span - k0);
        if (!engine.serve_enqueue(j.entry, k0, k1, priority)) {
            j.pass_pending = true;
            return;
        }
        j.pass_pending = false;
        j.pass_k0 = k0;
        j.pass_k1 = k1;
        j.next_k = k1;
        j.passes_queued++;
    }

    void step() {
        admit();
        if (running.empty()) return;
        engine.serve_run(stream, cfg.iterations_per_step);
        const double t = now_ms();
        std::vector<int> still;
        for (int id : running) {
            Job& j = *jobs[size_t(id)];
            const int64_t px = j.items;
            const int64_t done = engine.serve_pixels_done(j.entry);  // finished items, all passes so far
            const int64_t in_pass = done - px * (j.passes_queued - 1);
            j.status.samples_done = px * j.pass_k0 + in_pass * (j.pass_k1 - j.pass_k0);
            if (j.pass_pending) {
                queue_next_pass(j);
            } else if (done == px * j.passes_queued) {
                if (j.next_k < j.desc.spp / j.split) {
                    queue_next_pass(j);
                } else {
                    j.status.state = PH_JOB_DONE;
                    j.status.done_ms = t;
                    engine.serve_release(j.entry);
                    j.entry = -1;
                    finished++;
                    continue;
                }
            }
            still.push_back(id);
        }
        running.swap(still);
    }
};

extern "C" {

int ph_server_create(const ph_server_config* cfg, ph_server** out) {
    if (!out) return set_error("null argument");
    return guarded([&] {
        ph_server_config c{};
        if (cfg) c = *cfg;
        *out = new ph_server(c);
        return 0;
    });
}

void ph_server_destroy(ph_server* s) { delete s; }

int ph_server_set_wavefront_config(ph_server* s, const ph_wavefront_config* cfg) {
    if (!s || !cfg) return set_error("null argument");
    return guarded([&] {
        s->engine.set_config(*cfg);
        return 0;
    });
}

int ph_server_add_scene(ph_server* s, const ph_scene_desc* d) {
    if (!s) return set_error("null server");
    return guarded([&] {
        DeviceScene::validate(d);
        auto sc = std::make_unique<ServedScene>();
        sc->scene.upload(*d, s->stream);
        std::vector<int> cls = material_classes(sc->scene.view(), s->stream);
        sc->mat_class.upload(cls.data(), cls.size(), s->stream);
        PH_CHECK(hipStreamSynchronize(s->stream));
        s->scenes.push_back(std::move(sc));
        return int(s->scenes.size()) - 1;
    });
}

int ph_server_submit(ph_server* s, const ph_job_desc* d) {
    if (!s || !d) return set_error("null argument");
    if (d->scene < 0 || d->scene >= int(s->scenes.size())) return set_error("unknown scene id");
    if (d->width <= 0 || d->height <= 0 || d->spp <= 0) return set_error("a job needs a positive size and sample count");
    if (int64_t(d->width) * d->height * d->spp >= int64_t(INT32_MAX)) return set_error("job too large: pixels x spp must stay below 2^31");
    if (d->width > 65536 || d->height > 65536) return set_error("job image too large");
    if (d->path.max_depth <= 0 || d->path.max_depth > 0xFFFF) return set_error("max_depth out of range");
    return guarded([&] {
        auto j = std::make_unique<Job>();
        j->desc = *d;
        j->status.state = PH_JOB_QUEUED;
        j->status.pixels = d->width * d->height;
        j->status.samples = int64_t(j->status.pixels) * d->spp;
        j->status.submit_ms = s->now_ms();
        // Sub-streams: the largest divisor of spp that keeps pixels x split within min_items (or 1).
        const int64_t want = (int64_t(s->cfg.min_items) + j->status.pixels - 1) / j->status.pixels;
        for (int k = 1; k <= d->spp && k <= want; k++)
            if (d->spp % k == 0) j->split = k;
        j->items = int64_t(j->status.pixels) * j->split;
        j->bytes = size_t(j->items) * sizeof(float4);
        s->jobs.push_back(std::move(j));
        const int id = int(s->jobs.size()) - 1;
        s->queue.push_back(id);
        return id;
    });
}

int ph_server_step(ph_server* s) {
    if (!s) return set_error("null server");
    return guarded([&] {
        s->step();
        return s->finished;
    });
}

int ph_server_status(const ph_server* s, int job, ph_job_status* out) {
    if (!s || !out) return set_error("null argument");
    if (job < 0 || job >= int(s->jobs.size())) return set_error("unknown job id");
    *out = s->jobs[size_t(job)]->status;
    return 0;
}

int ph_server_result(ph_server* s, int job, float* rgb) {
    if (!s || !rgb) return set_error("null argument");
    if (job < 0 || job >= int(s->jobs.size())) return set_error("unknown job id");
    Job& j = *s->jobs[size_t(job)];
    if (j.status.state != PH_JOB_DONE) return set_error(j.status.state == PH_JOB_RELEASED ? "the job's result was already read" : "the job is not done");
    return guarded([&] {
        const size_t n = size_t(j.status.pixels);
        std::vector<float4> raw(size_t(j.items));
        PH_CHECK(hipMemcpyAsync(raw.data(), j.accum.ptr, raw.size() * sizeof(float4), hipMemcpyDeviceToHost, s->stream));
        PH_CHECK(hipStreamSynchronize(s->stream));
        const float inv = 1.0f / float(j.desc.spp);
        for (size_t i = 0; i < n; i++) {
            // The partial images in sub-stream order (one plane: the accumulation itself).
            float r = raw[i].x, g = raw[i].y, b = raw[i].z;
            for (int p = 1; p < j.split; p++) {
                const float4& v = raw[size_t(p) * n + i];
                r += v.x;
                g += v.y;
                b += v.z;
            }
            rgb[3 * i] = r * inv;
            rgb[3 * i + 1] = g * inv;
            rgb[3 * i + 2] = b * inv;
        }
        j.accum.release();
        s->mem_used -= int64_t(j.bytes);
        j.status.state = PH_JOB_RELEASED;
        return 0;
    });
}

int ph_server_release(ph_server* s, int job) {
    if (!s) return set_error("null server");
    if (job < 0 || job >= int(s->jobs.size())) return set_error("unknown job id");
    Job& j = *s->jobs[size_t(job)];
    if (j.status.state != PH_JOB_DONE) return set_error(j.status.state == PH_JOB_RELEASED ? "the job's result was already read" : "the job is not done");
    j.accum.release();
    s->mem_used -= int64_t(j.bytes);
    j.status.state = PH_JOB_RELEASED;
    return 0;
}

int ph_server_pending(const ph_server* s) { return s ? int(s->queue.size() + s->running.size()) : set_error("null server"); }

double ph_server_now_ms(const ph_server* s) { return s ? s->now_ms() : 0.0; }

}  // extern "C"
