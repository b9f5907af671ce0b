// The serving loop (photon_hip.h, ph_server_*): resident scenes, FIFO admission of render jobs by HBM
// footprint, and continuous batching of all admitted jobs' paths in shared wavefronts.
//
// A step admits the queued jobs that fit (frame buffer bytes within the budget, a free job-table entry),
// then advances every path in flight by a fixed number of bounces. A job runs as passes of at most
// `pass_samples` pixel samples; its next pass is queued once the previous one has finished. Unstarted
// work is served shortest-remaining-job first (or passes first come, first served), so a short job is
// not stuck behind a long one. Jobs finish independently, and a finished job's image can be read while
// the others keep running.
//
// Jobs of at most `express_samples` samples run in an express lane: a second engine with a small slot
// pool on a stream of its own, which runs concurrently with the main lane and as often as it has work
// during the main lane's run. Its iterations cover few paths and are short, so a short job's latency is
// set by its own depth rather than by iterations over millions of other paths.
//
// A job with fewer than `min_items` pixels runs each pixel's samples as S sub-streams (S divides spp)
// in separate partial images, which the result adds in sub-stream order: its paths can then fill more
// slots than it has pixels. Within a sub-stream the samples are taken in order, pass after pass. An
// image therefore depends only on the job and the server's configuration, never on what else shared
// the wavefronts; with S = 1 it is bit-identical to the renderer's.
#include <atomic>
#include <chrono>
#include <deque>
#include <exception>
#include <memory>
#include <thread>
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
    double deadline = 0.0;      // PH_SCHEDULE_DEADLINE: fixed when the job is admitted
    DevBuf<float4> accum;
};

}  // namespace

// One engine with its own stream and slots. The main lane runs every job by default; the express lane
// (when configured) runs the jobs of at most `express_samples` samples on a small slot pool of its own,
// concurrently with the main lane, so that a short job's bounces are not each as long as an iteration
// over millions of paths.
struct Lane {
    WavefrontEngine engine;
    hipStream_t stream = nullptr;
    int slots = 0;           // 0: lane off
    int min_items = 1;       // split threshold of this lane's jobs
    int iterations = 1;      // bounces per engine run
    int64_t pass_samples = 0;
    std::deque<int> queue;   // submitted, not admitted (FIFO)
    std::vector<int> running;
    double rate = 0.0;       // samples per ms of the lane's recent runs (response-ratio estimates)
};

struct ph_server {
    ph_server_config cfg{};
    int device = 0;
    Lane lanes[2];  // 0: main, 1: express
    std::vector<std::unique_ptr<ServedScene>> scenes;
    std::vector<std::unique_ptr<Job>> jobs;
    int64_t mem_used = 0;
    int finished = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now_ms() const { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
    hipStream_t stream() const { return lanes[0].stream; }

    ph_server(const ph_server_config& c) : cfg(c) {
        if (cfg.memory_budget <= 0) cfg.memory_budget = int64_t(16) << 30;
        if (cfg.path_slots <= 0) cfg.path_slots = 1 << 21;
        if (cfg.iterations_per_step <= 0) cfg.iterations_per_step = 4;
        if (cfg.max_jobs <= 0 || cfg.max_jobs > WavefrontEngine::kMaxJobs) cfg.max_jobs = WavefrontEngine::kMaxJobs;
        if (cfg.schedule != PH_SCHEDULE_SRPT && cfg.schedule != PH_SCHEDULE_FIFO && cfg.schedule != PH_SCHEDULE_HRRN && cfg.schedule != PH_SCHEDULE_SHARE &&
            cfg.schedule != PH_SCHEDULE_DEADLINE)
            throw std::invalid_argument("unknown schedule");
        if (!(cfg.deadline_stretch >= 0.0)) throw std::invalid_argument("deadline_stretch must be >= 0");
        if (cfg.deadline_stretch == 0.0) cfg.deadline_stretch = 4.0;
        if (cfg.express_slots == 0) cfg.express_slots = std::max(1, cfg.path_slots / 2);
        if (cfg.express_slots < 0) cfg.express_slots = 0;
        if (cfg.express_samples <= 0) cfg.express_samples = int64_t(16) * cfg.express_slots;
        if (cfg.express_iterations <= 0) cfg.express_iterations = cfg.iterations_per_step;
        PH_CHECK(hipGetDevice(&device));
        const int slots[2] = {cfg.path_slots, cfg.express_slots};
        for (int k = 0; k < 2; k++) {
            Lane& L = lanes[k];
            L.slots = slots[k];
            if (L.slots == 0) continue;
            L.iterations = k == 0 ? cfg.iterations_per_step : cfg.express_iterations;
            L.pass_samples = cfg.pass_samples > 0 ? cfg.pass_samples : int64_t(4) * L.slots;
            // Small jobs are split until they fill an eighth of the main lane, or all of the express lane.
            L.min_items = cfg.min_items > 0 ? cfg.min_items : cfg.min_items < 0 ? 1 : (k == 0 ? std::max(1, L.slots / 8) : L.slots);
            PH_CHECK(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
            L.engine.serve_begin(L.stream, L.slots);
            // Processor sharing hands out a few waves' worth of items per job per round.
            if (cfg.schedule == PH_SCHEDULE_SHARE) L.engine.serve_share(std::max(64, L.slots / 2048));
        }
    }
    ~ph_server() {
        for (Lane& L : lanes)
            if (L.stream) {
                (void)hipStreamSynchronize(L.stream);
                (void)hipStreamDestroy(L.stream);
            }
    }

    int lane_of(const Job& j) const { return lanes[1].slots > 0 && j.status.samples <= cfg.express_samples ? 1 : 0; }

    bool idle() const { return lanes[0].running.empty() && lanes[1].running.empty(); }

    void admit(Lane& L) {
        while (!L.queue.empty() && int(L.running.size()) < cfg.max_jobs) {
            Job& j = *jobs[size_t(L.queue.front())];
            // A job larger than the whole budget is admitted alone rather than never.
            if (mem_used + int64_t(j.bytes) > cfg.memory_budget && !idle()) break;
            j.accum.allocate(size_t(j.items));
            j.accum.zero(L.stream);
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
            const int entry = L.engine.serve_admit(L.stream, w);
            if (entry < 0) {
                j.accum.release();
                break;
            }
            j.entry = entry;
            j.status.state = PH_JOB_RUNNING;
            j.status.admit_ms = now_ms();
            // The service estimate: the job's samples at a fixed nominal rate of 1e6 samples/ms. Only the
            // order of deadlines matters, and a fixed rate keeps the order of jobs admitted at different
            // times consistent (the lanes' measured rates vary with the mix).
            j.deadline = j.status.submit_ms + cfg.deadline_stretch * double(j.status.samples) * 1e-6;
            const int64_t per_pass = L.pass_samples / std::max<int64_t>(1, j.items);
            j.pass_spp = int(std::min<int64_t>(std::max<int64_t>(per_pass, 1), j.desc.spp / j.split));
            queue_next_pass(L, j);
            mem_used += int64_t(j.bytes);
            L.running.push_back(L.queue.front());
            L.queue.pop_front();
        }
    }

    // Lower runs first. SRPT: the job's remaining samples; FIFO: queue order; HRRN: minus the response
    // ratio (waiting time + estimated remaining service) / estimated remaining service; DEADLINE: the
    // job's deadline.
    double priority_of(const Lane& L, const Job& j, int k0) const {
        const double remaining = double(j.items) * double(j.desc.spp / j.split - k0);
        if (cfg.schedule == PH_SCHEDULE_FIFO || cfg.schedule == PH_SCHEDULE_SHARE) return 0.0;
        if (cfg.schedule == PH_SCHEDULE_SRPT) return remaining;
        if (cfg.schedule == PH_SCHEDULE_DEADLINE) return j.deadline;
        const double service = remaining / std::max(L.rate, 1e-3);
        const double waited = now_ms() - j.status.submit_ms;
        return -(waited + service) / std::max(service, 1e-6);
    }

    void reprioritize(Lane& L) {
        if (cfg.schedule != PH_SCHEDULE_HRRN) return;
        for (int id : L.running) {
            const Job& j = *jobs[size_t(id)];
            if (!j.pass_pending) L.engine.serve_reprioritize(j.entry, priority_of(L, j, j.pass_k0));
        }
    }

    void queue_next_pass(Lane& L, Job& j) {
        const int span = j.desc.spp / j.split;
        const int k0 = j.next_k, k1 = std::min(span, k0 + j.pass_spp);
        const double priority = priority_of(L, j, k0);
        if (!L.engine.serve_enqueue(j.entry, k0, k1, priority)) {
            j.pass_pending = true;
            return;
        }
        j.pass_pending = false;
        j.pass_k0 = k0;
        j.pass_k1 = k1;
        j.next_k = k1;
        j.passes_queued++;
    }

    // After a run of the lane's engine: progress of its jobs, next passes, completions at time t.
    void advance(Lane& L, double t, double run_ms) {
        std::vector<int> still;
        int64_t progress = 0;
        for (int id : L.running) {
            Job& j = *jobs[size_t(id)];
            const int64_t px = j.items;
            const int64_t done = L.engine.serve_pixels_done(j.entry);  // finished items, all passes so far
            const int64_t in_pass = done - px * (j.passes_queued - 1);
            const int64_t before = j.status.samples_done;
            j.status.samples_done = px * j.pass_k0 + in_pass * (j.pass_k1 - j.pass_k0);
            progress += j.status.samples_done - before;
            if (j.pass_pending) {
                queue_next_pass(L, j);
            } else if (done == px * j.passes_queued) {
                if (j.next_k < j.desc.spp / j.split) {
                    queue_next_pass(L, j);
                } else {
                    j.status.state = PH_JOB_DONE;
                    j.status.done_ms = t;
                    L.engine.serve_release(j.entry);
                    j.entry = -1;
                    finished++;
                    continue;
                }
            }
            still.push_back(id);
        }
        L.running.swap(still);
        if (run_ms > 0 && progress > 0) L.rate = L.rate > 0 ? 0.8 * L.rate + 0.2 * (double(progress) / run_ms) : double(progress) / run_ms;
    }

    void step() {
        Lane& M = lanes[0];
        Lane& X = lanes[1];
        admit(M);
        if (X.slots > 0) admit(X);
        const bool run_m = !M.running.empty(), run_x = !X.running.empty();
        if (run_m && run_x) {
            // The main lane's run on a thread of its own; meanwhile the express lane runs as many times
            // as it has work, finishing its jobs as it goes. The lanes share nothing but the device.
            std::atomic<bool> m_done{false};
            double m_end = 0.0;
            std::exception_ptr m_err;
            reprioritize(M);
            const double m_start = now_ms();
            std::thread th([&] {
                try {
                    PH_CHECK(hipSetDevice(device));
                    M.engine.serve_run(M.stream, M.iterations);
                } catch (...) {
                    m_err = std::current_exception();
                }
                m_end = now_ms();
                m_done.store(true, std::memory_order_release);
            });
            try {
                do {
                    reprioritize(X);
                    const double t0 = now_ms();
                    X.engine.serve_run(X.stream, X.iterations);
                    const double t1 = now_ms();
                    advance(X, t1, t1 - t0);
                } while (!m_done.load(std::memory_order_acquire) && !X.running.empty());
            } catch (...) {
                th.join();
                throw;
            }
            th.join();
            if (m_err) std::rethrow_exception(m_err);
            advance(M, m_end, m_end - m_start);
        } else if (run_m || run_x) {
            Lane& L = run_m ? M : X;
            reprioritize(L);
            const double t0 = now_ms();
            L.engine.serve_run(L.stream, L.iterations);
            const double t1 = now_ms();
            advance(L, t1, t1 - t0);
        }
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
        for (Lane& L : s->lanes)
            if (L.slots > 0) L.engine.set_config(*cfg);
        return 0;
    });
}

int ph_server_add_scene(ph_server* s, const ph_scene_desc* d) {
    if (!s) return set_error("null server");
    return guarded([&] {
        DeviceScene::validate(d);
        auto sc = std::make_unique<ServedScene>();
        sc->scene.upload(*d, s->stream());
        std::vector<int> cls = material_classes(sc->scene.view(), s->stream());
        sc->mat_class.upload(cls.data(), cls.size(), s->stream());
        PH_CHECK(hipStreamSynchronize(s->stream()));
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
        Lane& L = s->lanes[s->lane_of(*j)];
        // Sub-streams: the largest divisor of spp that keeps pixels x split within the lane's min_items (or 1).
        const int64_t want = (int64_t(L.min_items) + j->status.pixels - 1) / j->status.pixels;
        for (int k = 1; k <= d->spp && k <= want; k++)
            if (d->spp % k == 0) j->split = k;
        j->items = int64_t(j->status.pixels) * j->split;
        j->bytes = size_t(j->items) * sizeof(float4);
        s->jobs.push_back(std::move(j));
        const int id = int(s->jobs.size()) - 1;
        L.queue.push_back(id);
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
        PH_CHECK(hipMemcpyAsync(raw.data(), j.accum.ptr, raw.size() * sizeof(float4), hipMemcpyDeviceToHost, s->stream()));
        PH_CHECK(hipStreamSynchronize(s->stream()));
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

int ph_server_pending(const ph_server* s) {
    if (!s) return set_error("null server");
    size_t n = 0;
    for (const Lane& L : s->lanes) n += L.queue.size() + L.running.size();
    return int(n);
}

double ph_server_now_ms(const ph_server* s) { return s ? s->now_ms() : 0.0; }

}  // extern "C"
