// server.hip - continuous-batching serving loop (no counterpart in the reference).
//
// Requests (a scene plus a step count) queue up; each poll admits queued
// requests while their arena fits the HBM budget and a slot is free, advances
// every resident scene by the same number of ticks, and retires the finished
// ones. In batched mode one launch per kernel covers every resident scene (the
// kernels walk a device table of scene descriptors), and one tick of the whole
// batch is captured once as a HIP graph and replayed until the batch changes.
// In plain mode each scene's tick is launched on its own, one scene per launch:
// the same kernels, the comparison the batched path is measured against.
//
// Rounds are pipelined: a poll launches its round and returns without waiting;
// the next poll first admits new requests (uploads and set-up on a second
// stream, overlapping the running round), then waits for the round, accounts
// and retires, admits into the freed slots and launches the next round. Scene
// preparation in submit (host FP64 work) thus overlaps the GPU too.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>

#include "runtime.hpp"

using namespace emb;

namespace {

enum ReqState { QUEUED, RUNNING, DONE, CANCELLED };

struct Request {
    int64_t id = 0;
    HostScene hs;
    int64_t steps = 0, done = 0;
    size_t footprint = 0;
    double time0 = 0.0;
    double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;   // host clock (s): request latency and queue wait
    ReqState state = QUEUED;
    Resident r;
    std::vector<double> pos;                   // kept results
    std::vector<float> vel, ex, ey, bz;
    bool has_result = false;
    std::string fault;                         // a device-side numeric flag raised during the run
};

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct EmbServer {
    EmbDevice *dev = nullptr;
    EmbSettings s{};
    EmbServerConfig cfg{};
    hipStream_t stream = nullptr;
    std::map<int64_t, std::unique_ptr<Request>> all;
    std::deque<Request *> queue;
    std::vector<Request *> active;
    int64_t next_id = 1;
    size_t used_bytes = 0;
    // batch
    DevScene *d_scenes = nullptr;
    int32_t *d_prefix = nullptr;
    Batch batch{};
    int fdtd_max = 1;
    bool dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipStream_t up_stream = nullptr;           // admissions: uploads and set-up kernels
    hipEvent_t up_event = nullptr;
    std::vector<Request *> inflight;           // the launched round's scenes
    int64_t inflight_k = 0;                    // and its tick count
    std::vector<DevScene> host_scenes;         // sources of the scene-table copies
    std::vector<int32_t> host_prefix;
    // stats
    std::vector<float> latencies_ms;   // one sample per scene-step: the batch tick it took part in
    std::vector<float> request_ms;     // per completed request: submit -> done
    std::vector<float> wait_ms;        // per admitted request: submit -> admitted
    int64_t cell_steps = 0, completed = 0;
    double t_first = -1.0, t_last = 0.0;
    int peak = 0;
};

namespace {

int drop_graph(EmbServer *sv) {
    if (sv->graph_exec) (void)hipGraphExecDestroy(sv->graph_exec);
    if (sv->graph) (void)hipGraphDestroy(sv->graph);
    sv->graph_exec = nullptr;
    sv->graph = nullptr;
    return EMB_OK;
}

// Scene table and block prefixes of the resident scenes.
int rebuild_batch(EmbServer *sv) {
    const int n = (int)sv->active.size();
    std::vector<DevScene> &scenes = sv->host_scenes;
    std::vector<int32_t> &prefix = sv->host_prefix;
    scenes.clear();
    prefix.assign((size_t)KC_COUNT * (n + 1), 0);
    sv->fdtd_max = 1;
    sv->batch.even_fdtd = 0;
    sv->batch.top_small = 1;
    for (int k = 0; k < KC_COUNT; k++) sv->batch.total[k] = 0;
    for (int i = 0; i < n; i++) {
        const DevScene &d = sv->active[(size_t)i]->r.ds;
        scenes.push_back(d);
        int blocks[KC_COUNT];
        scene_blocks(d, blocks);
        for (int k = 0; k < KC_COUNT; k++) prefix[(size_t)k * (n + 1) + i + 1] = prefix[(size_t)k * (n + 1) + i] + blocks[k];
        sv->fdtd_max = std::max(sv->fdtd_max, d.fdtd_substeps);
        if (d.fdtd_substeps % 2 == 0) sv->batch.even_fdtd = 1;
        if ((d.cols * d.rows + SCAN_TILE - 1) / SCAN_TILE > 256) sv->batch.top_small = 0;
    }
    for (int k = 0; k < KC_COUNT; k++) sv->batch.total[k] = prefix[(size_t)k * (n + 1) + n];
    if (n > 0) {
        // Stream-ordered after every earlier round; the sources live in the server
        // until the next rebuild, which comes after this round has been waited for.
        EMB_HIP(hipMemcpyAsync(sv->d_scenes, scenes.data(), scenes.size() * sizeof(DevScene), hipMemcpyHostToDevice, sv->stream));
        EMB_HIP(hipMemcpyAsync(sv->d_prefix, prefix.data(), prefix.size() * 4, hipMemcpyHostToDevice, sv->stream));
    }
    sv->batch.scenes = sv->d_scenes;
    sv->batch.prefix = sv->d_prefix;
    sv->batch.n_scenes = n;
    drop_graph(sv);
    sv->dirty = false;
    return EMB_OK;
}

int capture_graph(EmbServer *sv) {
    EMB_HIP(hipStreamBeginCapture(sv->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = tick_kernels(sv->batch, sv->fdtd_max, sv->stream);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(sv->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(EMB_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    sv->graph = g;
    EMB_HIP(hipGraphInstantiate(&sv->graph_exec, g, nullptr, nullptr, 0));
    return EMB_OK;
}

int retire(EmbServer *sv, Request *q) {
    Resident &r = q->r;
    if (q->state == RUNNING) {   // a numeric flag (k_rest's wait limit, charge range) fails the request loudly
        int32_t flags[2] = {0, 0};
        EMB_HIP(hipMemcpy(flags, &r.ds.ctr->rest_error, sizeof(flags), hipMemcpyDeviceToHost));
        q->fault = numeric_fault(flags[0], flags[1]);
    }
    if (q->state == RUNNING && sv->cfg.keep_results && q->fault.empty()) {
        const DevScene &d = r.ds;
        const size_t n = (size_t)d.n, nx = d.nx, ny = d.ny;
        q->pos.resize(2 * n);
        q->vel.resize(2 * n);
        q->ex.resize((nx + 1) * ny);
        q->ey.resize(nx * (ny + 1));
        q->bz.resize((nx + 1) * (ny + 1));
        std::vector<double> p(2 * n);
        int rc = read_positions(d, p.data());
        if (rc != EMB_OK) return rc;
        for (size_t i = 0; i < 2 * n; i++) q->pos[i] = p[i];
        std::vector<float> a(n), b(n);
        EMB_HIP(hipMemcpy(a.data(), d.vx, n * 4, hipMemcpyDeviceToHost));
        EMB_HIP(hipMemcpy(b.data(), d.vy, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) q->vel[2 * i] = a[i], q->vel[2 * i + 1] = b[i];
        EMB_HIP(hipMemcpy(q->ex.data(), d.ex, q->ex.size() * 4, hipMemcpyDeviceToHost));
        EMB_HIP(hipMemcpy(q->ey.data(), d.ey, q->ey.size() * 4, hipMemcpyDeviceToHost));
        EMB_HIP(hipMemcpy(q->bz.data(), d.bz, q->bz.size() * 4, hipMemcpyDeviceToHost));
        q->has_result = true;
    }
    free_resident(r);
    sv->used_bytes -= q->footprint;
    return EMB_OK;
}

}  // namespace

extern "C" {

int emb_server_create(EmbDevice *dev, const EmbSettings *settings, const EmbServerConfig *config, EmbServer **out) {
    try {
        if (!dev || !settings || !config || !out) return fail(EMB_E_INVALID, "null argument");
        if (config->max_slots < 1 || config->max_slots > 4096) return fail(EMB_E_INVALID, "max_slots must be 1..4096");
        if (config->ticks_per_round < 1 || config->ticks_per_round > TICK_TABLE)
            return fail(EMB_E_INVALID, "ticks_per_round must be 1..64");
        if (config->hbm_budget_bytes <= 0) return fail(EMB_E_INVALID, "hbm_budget_bytes must be positive");
        EMB_HIP(hipSetDevice(dev->ordinal));
        auto sv = std::make_unique<EmbServer>();
        sv->dev = dev;
        sv->s = *settings;
        sv->cfg = *config;
        EMB_HIP(hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking));
        EMB_HIP(hipStreamCreateWithFlags(&sv->up_stream, hipStreamNonBlocking));
        EMB_HIP(hipEventCreateWithFlags(&sv->up_event, hipEventDisableTiming));
        EMB_HIP(hipMalloc(&sv->d_scenes, sizeof(DevScene) * config->max_slots));
        EMB_HIP(hipMalloc(&sv->d_prefix, sizeof(int32_t) * KC_COUNT * (config->max_slots + 1)));
        EMB_HIP(hipEventCreate(&sv->ev0));
        EMB_HIP(hipEventCreate(&sv->ev1));
        *out = sv.release();
        return EMB_OK;
    } catch (const std::exception &e) {
        return fail(EMB_E_INVALID, e.what());
    }
}

int emb_server_submit(EmbServer *sv, const EmbScene *scene, int64_t steps, int64_t *id) {
    try {
        if (!sv || !scene || !id || steps < 0) return fail(EMB_E_INVALID, "bad argument");
        int trc = ensure_tables(sv->dev, sv->s.electron_softening, sv->s.smoothing_passes, sv->s.electron_cutoff);
        if (trc != EMB_OK) return trc;
        auto q = std::make_unique<Request>();
        const std::string why = prepare_scene(sv->s, *scene, q->hs);
        if (!why.empty()) return fail(EMB_E_INVALID, why);
        q->footprint = scene_bytes(sv->s, q->hs.cols, q->hs.rows, q->hs.n, (int)q->hs.stacks.size(),
                                   (int)q->hs.stack_cells.size(), (int)q->hs.bats.size(), (int)q->hs.ac_freq.size());
        // Its arena and its DST-I matrices must fit the budget on their own.
        const size_t sines = sine_set_bytes(q->hs.nx) + (q->hs.ny != q->hs.nx ? sine_set_bytes(q->hs.ny) : 0);
        if ((int64_t)(q->footprint + sines) > sv->cfg.hbm_budget_bytes)
            return fail(EMB_E_NOMEM, "scene footprint exceeds the HBM budget");
        q->steps = steps;
        q->time0 = scene->time;
        q->id = sv->next_id++;
        q->t_submit = now_s();
        *id = q->id;
        sv->queue.push_back(q.get());
        sv->all[q->id] = std::move(q);
        if (sv->t_first < 0.0) sv->t_first = now_s();
        return EMB_OK;
    } catch (const std::exception &e) {
        return fail(EMB_E_INVALID, e.what());
    }
}

}  // extern "C"

namespace {

// Admit queued requests, in arrival order, while a slot and the budget allow:
// upload and set up on the admission stream; the main stream waits for it.
int admit(EmbServer *sv) {
    bool any = false;
    while (!sv->queue.empty() && (int)sv->active.size() < sv->cfg.max_slots) {
        Request *q = sv->queue.front();
        // The budget covers the resident arenas and every sine set on the device;
        // idle sets (no resident scene) are freed to make room.
        const int64_t budget = sv->cfg.hbm_budget_bytes;
        const size_t need = q->footprint + sine_missing_bytes(sv->dev, q->hs.nx, q->hs.ny);
        if ((int64_t)(sv->used_bytes + sine_held_bytes(sv->dev) + need) > budget &&
            (int64_t)(sv->used_bytes + need) <= budget)
            trim_sine(sv->dev, (size_t)(budget - (int64_t)(sv->used_bytes + need)));
        if ((int64_t)(sv->used_bytes + sine_held_bytes(sv->dev) + need) > budget) break;
        sv->queue.pop_front();
        int rc = make_resident(sv->dev, sv->s, std::move(q->hs), sv->up_stream, q->r);
        if (rc != EMB_OK) {
            free_resident(q->r);
            q->state = CANCELLED;
            return rc;
        }
        q->r.time = q->time0;
        EMB_HIP(setup_kernels(q->r.batch, sv->up_stream));
        EMB_HIP(finish_sync_kernel(q->r.batch, sv->up_stream));
        sv->used_bytes += q->footprint;
        q->state = RUNNING;
        q->t_admit = now_s();
        sv->wait_ms.push_back((float)((q->t_admit - q->t_submit) * 1e3));
        any = true;
        if (q->steps == 0) {   // nothing to run: finished on admission
            EMB_HIP(hipStreamSynchronize(sv->up_stream));
            int rc2 = retire(sv, q);
            q->state = DONE;
            q->t_done = now_s();
            sv->request_ms.push_back((float)((q->t_done - q->t_submit) * 1e3));
            sv->completed++;
            if (rc2 != EMB_OK) return rc2;
            continue;
        }
        sv->active.push_back(q);
        sv->dirty = true;
    }
    if (any) {
        EMB_HIP(hipEventRecord(sv->up_event, sv->up_stream));
        EMB_HIP(hipStreamWaitEvent(sv->stream, sv->up_event, 0));
    }
    sv->peak = std::max(sv->peak, (int)sv->active.size());
    return EMB_OK;
}

// Wait for the launched round, account its steps and latency, retire what finished.
int finish_round(EmbServer *sv) {
    if (sv->inflight.empty()) return EMB_OK;
    EMB_HIP(hipEventSynchronize(sv->ev1));
    float ms = 0.0f;
    EMB_HIP(hipEventElapsedTime(&ms, sv->ev0, sv->ev1));
    const int64_t k = sv->inflight_k;
    const float per_step = ms / (float)k;
    for (Request *q : sv->inflight) {
        q->done += k;
        sv->cell_steps += (int64_t)q->r.ds.nx * q->r.ds.ny * k;
        for (int64_t t = 0; t < k; t++) sv->latencies_ms.push_back(per_step);
    }
    sv->inflight.clear();
    std::vector<Request *> keep;
    const double t_round = now_s();   // the round is complete on the device by now
    for (Request *q : sv->active) {
        if (q->done >= q->steps) {
            q->t_done = t_round;
            sv->request_ms.push_back((float)((q->t_done - q->t_submit) * 1e3));
            int rc = retire(sv, q);
            q->state = DONE;
            sv->completed++;
            sv->dirty = true;
            if (rc != EMB_OK) return rc;
        } else {
            keep.push_back(q);
        }
    }
    sv->active.swap(keep);
    sv->t_last = now_s();
    return EMB_OK;
}

// Launch k ticks of every resident scene, without waiting.
int launch_round(EmbServer *sv) {
    if (sv->active.empty()) return EMB_OK;
    if (sv->dirty) {
        int rc = rebuild_batch(sv);
        if (rc != EMB_OK) return rc;
    }
    int64_t k = sv->cfg.ticks_per_round;
    for (Request *q : sv->active) k = std::min(k, q->steps - q->done);
    for (Request *q : sv->active) {
        int rc = enqueue_tick_tables(q->r, (int)k, sv->stream);
        if (rc != EMB_OK) return rc;
    }
    const bool graphs = sv->cfg.batched && sv->cfg.use_graph;
    if (graphs && !sv->graph_exec) {
        int rc = capture_graph(sv);
        if (rc != EMB_OK) return rc;
    }
    EMB_HIP(hipEventRecord(sv->ev0, sv->stream));
    for (int64_t t = 0; t < k; t++) {
        if (graphs) {
            EMB_HIP(hipGraphLaunch(sv->graph_exec, sv->stream));
        } else if (sv->cfg.batched) {
            EMB_HIP(tick_kernels(sv->batch, sv->fdtd_max, sv->stream));
        } else {
            for (Request *q : sv->active) EMB_HIP(tick_kernels(q->r.batch, q->r.ds.fdtd_substeps, sv->stream));
        }
    }
    EMB_HIP(hipEventRecord(sv->ev1, sv->stream));
    sv->inflight = sv->active;
    sv->inflight_k = k;
    return EMB_OK;
}

}  // namespace

extern "C" {

int emb_server_cancel(EmbServer *sv, int64_t id) {
    try {
        if (!sv) return fail(EMB_E_INVALID, "null server");
        auto it = sv->all.find(id);
        if (it == sv->all.end()) return fail(EMB_E_INVALID, "no such request");
        Request *q = it->second.get();
        if (q->state == QUEUED) {
            sv->queue.erase(std::find(sv->queue.begin(), sv->queue.end(), q));
        } else if (q->state == RUNNING) {
            EMB_HIP(hipSetDevice(sv->dev->ordinal));
            int rc = finish_round(sv);   // settle the running round first
            if (rc != EMB_OK) return rc;
            if (q->state == RUNNING) {
                EMB_HIP(hipStreamSynchronize(sv->up_stream));
                sv->active.erase(std::find(sv->active.begin(), sv->active.end(), q));
                q->state = CANCELLED;
                rc = retire(sv, q);
                if (rc != EMB_OK) return rc;
                sv->dirty = true;
            }
        }
        if (q->state != DONE) q->state = CANCELLED;
        return EMB_OK;
    } catch (const std::exception &e) {
        return fail(EMB_E_INVALID, e.what());
    }
}

int emb_server_poll(EmbServer *sv, int32_t *active_out) {
    try {
        if (!sv) return fail(EMB_E_INVALID, "null server");
        EMB_HIP(hipSetDevice(sv->dev->ordinal));
        int rc = admit(sv);                  // overlaps the running round
        if (rc == EMB_OK) rc = finish_round(sv);
        if (rc == EMB_OK) rc = admit(sv);    // into the slots just freed
        if (rc == EMB_OK) rc = launch_round(sv);
        if (rc != EMB_OK) return rc;
        if (active_out) *active_out = (int32_t)(sv->active.size() + sv->queue.size());
        return EMB_OK;
    } catch (const std::exception &e) {
        return fail(EMB_E_INVALID, e.what());
    }
}

int emb_server_done(EmbServer *sv, int64_t id) {
    if (!sv) return fail(EMB_E_INVALID, "null server");
    auto it = sv->all.find(id);
    if (it == sv->all.end()) return fail(EMB_E_INVALID, "no such request");
    if (it->second->state == CANCELLED) return fail(EMB_E_STATE, "request was cancelled");
    return it->second->state == DONE ? 1 : 0;
}

int emb_server_result(EmbServer *sv, int64_t id, double *pos, double *vel, double *ex, double *ey, double *bz) {
    if (!sv) return fail(EMB_E_INVALID, "null server");
    auto it = sv->all.find(id);
    if (it == sv->all.end()) return fail(EMB_E_INVALID, "no such request");
    Request *q = it->second.get();
    if (q->state != DONE) return fail(EMB_E_STATE, "request not finished");
    if (!q->fault.empty()) return fail(EMB_E_NUMERIC, q->fault);
    if (!q->has_result) return fail(EMB_E_STATE, "result not kept (keep_results = 0) or already read");
    auto out = [](double *dst, const auto &src) {
        if (dst)
            for (size_t i = 0; i < src.size(); i++) dst[i] = src[i];
    };
    out(pos, q->pos);
    out(vel, q->vel);
    out(ex, q->ex);
    out(ey, q->ey);
    out(bz, q->bz);
    q->pos = {};
    q->vel = q->ex = q->ey = q->bz = {};
    q->has_result = false;
    return EMB_OK;
}

int emb_server_stats(EmbServer *sv, EmbServerStats *out) {
    if (!sv || !out) return fail(EMB_E_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->completed = sv->completed;
    out->cell_steps = sv->cell_steps;
    out->elapsed_s = sv->t_first < 0.0 ? 0.0 : std::max(0.0, sv->t_last - sv->t_first);
    out->cell_steps_per_s = out->elapsed_s > 0.0 ? (double)sv->cell_steps / out->elapsed_s : 0.0;
    out->step_samples = (int64_t)sv->latencies_ms.size();
    // Nearest-rank percentile of a sample (0 when empty).
    auto pct = [](std::vector<float> v, double p) {
        if (v.empty()) return 0.0;
        const size_t k = std::min(v.size() - 1, (size_t)std::floor(p / 100.0 * (double)(v.size() - 1) + 0.5));
        std::nth_element(v.begin(), v.begin() + (long)k, v.end());
        return (double)v[k];
    };
    out->step_latency_p50_ms = pct(sv->latencies_ms, 50.0);
    out->step_latency_p99_ms = pct(sv->latencies_ms, 99.0);
    out->request_latency_p50_ms = pct(sv->request_ms, 50.0);
    out->request_latency_p99_ms = pct(sv->request_ms, 99.0);
    out->queue_wait_p50_ms = pct(sv->wait_ms, 50.0);
    out->queue_wait_p99_ms = pct(sv->wait_ms, 99.0);
    out->peak_resident = sv->peak;
    return EMB_OK;
}

int emb_server_destroy(EmbServer *sv) {
    if (!sv) return EMB_OK;
    (void)hipSetDevice(sv->dev->ordinal);
    (void)hipStreamSynchronize(sv->stream);
    (void)hipStreamSynchronize(sv->up_stream);
    for (Request *q : sv->active) free_resident(q->r);
    drop_graph(sv);
    if (sv->d_scenes) (void)hipFree(sv->d_scenes);
    if (sv->d_prefix) (void)hipFree(sv->d_prefix);
    if (sv->ev0) (void)hipEventDestroy(sv->ev0);
    if (sv->ev1) (void)hipEventDestroy(sv->ev1);
    if (sv->up_event) (void)hipEventDestroy(sv->up_event);
    (void)hipStreamDestroy(sv->up_stream);
    (void)hipStreamDestroy(sv->stream);
    delete sv;
    return EMB_OK;
}

}  // extern "C"
