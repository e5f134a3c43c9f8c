// Serving loop: scene requests (initial state, parameters, step count) arrive at any time,
// are admitted by device-memory footprint, packed into one batched substep chain, stepped
// frame by frame while resident in HBM, and retired with their result copied back.
//
// Continuous batching: a worker thread owns the device work.  Between frames it admits
// queued requests that fit the budget and retires finished ones; when the set of resident
// scenes changes it repacks them into contiguous slot / table / brick ranges inside buffers
// allocated once at the server's capacity, rebuilds the scene table and re-captures the
// per-substep hipGraph.  Every scene's arithmetic depends only on its own state (see
// sph_batch.hip), so a result is bit-identical to stepping that scene alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sph_world.h"
#include "sph_host.hpp"
#include "sph_kernels.hpp"

using namespace sph;

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

enum ReqState { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3 };

struct Request {
    uint64_t id = 0;
    sph_params params{};
    uint32_t n = 0, T = 0;
    uint64_t bricks = 0;
    uint64_t footprint = 0;
    std::vector<float> pos, vel;  // input state, then the result
    uint32_t steps_left = 0;
    int state = kQueued;
    std::string error;
    Clock::time_point t_submit, t_admit, t_done;
    uint32_t p_begin = 0;  // slot range while resident
    bool fresh = true;     // state not yet on the device
};

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t lo = (size_t)pos;
    const size_t hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

}  // namespace

struct sph_server {
    sph_server_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    Chain chain;
    SceneConst *d_scenes = nullptr;
    uint16_t *d_scene_of = nullptr;
    GraphCache graph;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::vector<Request *> resident;
    std::map<uint64_t, std::unique_ptr<Request>> requests;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    std::string fatal;  // set when a device error stops the worker
    bool cfg_bricks = false;  // brick kernels for the resident scenes (else one lane each)

    uint64_t resident_bytes = 0, resident_particles = 0, resident_cells = 0, resident_bricks = 0;
    uint64_t submitted = 0, completed = 0, failed = 0, frames = 0;
    double particle_steps = 0.0, busy_s = 0.0;
    std::vector<double> step_ms, scene_ms, wait_ms;
};

namespace {

// Force-list entries per slot when the configuration leaves them at 0: room for the
// ~100 neighbours of a breaking dam or a settled column at the target density; a denser
// particle walks its cells instead (the same result).
constexpr uint32_t kServerListEntries = 128;

// Device bytes a resident scene occupies (the chain's per-slot buffers with force lists of
// `list_entries` per slot, and its table and halo grid).
uint64_t scene_footprint(uint32_t n, uint32_t T, uint32_t grid_cells, uint32_t list_entries) {
    return (uint64_t)n * (136u + 4u * list_entries) + (uint64_t)T * 24u + (uint64_t)grid_cells * 8u;
}

bool fits(const sph_server *s, const Request *r) {
    return s->resident.size() < s->cfg.max_scenes &&
           s->resident_bytes + r->footprint <= s->cfg.hbm_budget_bytes &&
           s->resident_particles + r->n <= s->cfg.max_particles &&
           s->resident_cells + r->T <= s->cfg.max_table_cells &&
           (!s->cfg_bricks || s->resident_bricks + r->bricks <= s->chain.b_cap);
}

// Lays the resident scenes out contiguously, moves continuing scenes' state to their new
// slots, uploads the state of newly admitted scenes and rebuilds the scene tables.
int repack(sph_server *s, const std::vector<uint32_t> &old_begin) {
    const size_t count = s->resident.size();
    std::vector<SceneConst> table(count);
    uint32_t pb = 0, tb = 0, bb = 0, ub = 0, gb = 0;
    bool moved = false;
    for (size_t i = 0; i < count; ++i) {
        Request *r = s->resident[i];
        SceneConst c = make_scene_const(r->params, r->T, tb, pb, pb + r->n, bb, ub, gb);
        // <= 2n + 1024 buckets per scene: within the capacity; a halo grid that does not
        // fit what is left of the grid capacity is dropped (hash-table lookups instead)
        ub += scene_buckets(c);
        if ((uint64_t)gb + c.grid_cells > s->chain.g_cap) c.grid_cells = 0u;
        gb += c.grid_cells;
        if (s->cfg_bricks && r->bricks) bb += (uint32_t)r->bricks;
        else c.use_bricks = 0u;
        table[i] = c;
        if (!r->fresh && old_begin[i] != pb) moved = true;
        r->p_begin = pb;
        pb += r->n;
        tb += r->T;
    }
    hipError_t e = hipSuccess;
    if (moved) {
        // Continuing scenes whose slot range changed go through the (idle between frames)
        // sorted-state buffers, so that a scene moving into a range another one vacates
        // never overlaps itself; scenes that keep their slots (a long-lived scene admitted
        // first keeps slot 0) are not touched.
        for (size_t i = 0; i < count && e == hipSuccess; ++i) {
            Request *r = s->resident[i];
            if (r->fresh || old_begin[i] == r->p_begin) continue;
            const size_t bytes = (size_t)r->n * 16;
            e = hipMemcpyAsync(s->chain.b.pos_sorted + r->p_begin, s->chain.b.pos + old_begin[i], bytes,
                               hipMemcpyDeviceToDevice, s->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(s->chain.b.vel_sorted + r->p_begin, s->chain.b.vel + old_begin[i], bytes,
                                   hipMemcpyDeviceToDevice, s->stream);
        }
        for (size_t i = 0; i < count && e == hipSuccess; ++i) {
            Request *r = s->resident[i];
            if (r->fresh || old_begin[i] == r->p_begin) continue;
            const size_t bytes = (size_t)r->n * 16;
            e = hipMemcpyAsync(s->chain.b.pos + r->p_begin, s->chain.b.pos_sorted + r->p_begin, bytes,
                               hipMemcpyDeviceToDevice, s->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(s->chain.b.vel + r->p_begin, s->chain.b.vel_sorted + r->p_begin, bytes,
                                   hipMemcpyDeviceToDevice, s->stream);
        }
    }
    for (size_t i = 0; i < count && e == hipSuccess; ++i) {
        Request *r = s->resident[i];
        if (!r->fresh) continue;
        const size_t bytes = (size_t)r->n * 16;
        e = hipMemcpyAsync(s->chain.b.pos + r->p_begin, r->pos.data(), bytes, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(s->chain.b.vel + r->p_begin, r->vel.data(), bytes, hipMemcpyHostToDevice, s->stream);
        r->fresh = false;
    }
    if (e == hipSuccess && count)
        e = hipMemcpyAsync(s->d_scenes, table.data(), count * sizeof(SceneConst), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess && count) e = launch_fill_scene_of(s->stream, s->d_scenes, (uint32_t)count, s->d_scene_of, pb);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return hip_fail(e, "repack");
    s->chain.configure(pb, tb, bb ? bb : 1u, ub ? ub : 1u, gb);
    s->chain.bricks = s->cfg_bricks && bb > 0;
    s->graph.drop();
    return SPH_OK;
}

void worker_main(sph_server *s) {
    hipSetDevice(s->device);
    std::vector<uint32_t> old_begin;
    bool changed = false;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->stop) return;  // resident and queued scenes are abandoned
            // Admit in arrival order while the head of the queue fits.
            while (!s->pending.empty() && fits(s, s->pending.front())) {
                Request *r = s->pending.front();
                s->pending.pop_front();
                r->state = kRunning;
                r->t_admit = Clock::now();
                s->wait_ms.push_back(std::chrono::duration<double, std::milli>(r->t_admit - r->t_submit).count());
                s->resident.push_back(r);
                s->resident_bytes += r->footprint;
                s->resident_particles += r->n;
                s->resident_cells += r->T;
                s->resident_bricks += r->bricks;
                changed = true;
            }
            if (s->resident.empty()) {
                if (s->stop) return;
                s->cv_work.wait(lk, [s] { return s->stop || !s->pending.empty(); });
                continue;
            }
        }
        // Only this thread modifies `resident`, so it can be read without the lock here.
        if (changed) {
            old_begin.resize(s->resident.size());
            for (size_t i = 0; i < s->resident.size(); ++i) old_begin[i] = s->resident[i]->p_begin;
            int rc = repack(s, old_begin);
            if (rc) {
                std::lock_guard<std::mutex> lk(s->mu);
                s->fatal = last_error();
                s->stop = true;
                s->cv_done.notify_all();
                return;
            }
            changed = false;
        }
        const ManyScenes sc{s->d_scenes, s->d_scene_of, (uint32_t)s->resident.size()};
        const Clock::time_point t0 = Clock::now();
        hipError_t e = hipSuccess;
        if (!s->graph.exec) {
            int rc = s->graph.ensure_begin(s->stream);
            if (rc == SPH_OK) {
                hipError_t le = s->chain.enqueue(s->stream, sc, nullptr);
                rc = s->graph.end(s->stream, le);
            }
            if (rc) e = hipErrorUnknown;
        }
        for (uint32_t k = 0; k < s->cfg.substeps && e == hipSuccess; ++k) e = hipGraphLaunch(s->graph.exec, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        const double dt = seconds_since(t0);
        if (e != hipSuccess) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->fatal = std::string("device error in a batched frame: ") + hipGetErrorString(e);
            s->stop = true;
            s->cv_done.notify_all();
            return;
        }
        std::lock_guard<std::mutex> lk(s->mu);
        s->frames++;
        s->busy_s += dt;
        s->step_ms.push_back(dt * 1e3);
        s->particle_steps += (double)s->resident_particles;
        std::vector<Request *> keep;
        keep.reserve(s->resident.size());
        for (Request *r : s->resident) {
            if (--r->steps_left > 0) {
                keep.push_back(r);
                continue;
            }
            const size_t bytes = (size_t)r->n * 16;
            hipError_t de = hipMemcpyAsync(r->pos.data(), s->chain.b.pos + r->p_begin, bytes, hipMemcpyDeviceToHost, s->stream);
            if (de == hipSuccess)
                de = hipMemcpyAsync(r->vel.data(), s->chain.b.vel + r->p_begin, bytes, hipMemcpyDeviceToHost, s->stream);
            if (de == hipSuccess) de = hipStreamSynchronize(s->stream);
            r->t_done = Clock::now();
            if (de != hipSuccess) {
                r->state = kFailed;
                r->error = hipGetErrorString(de);
                s->failed++;
            } else {
                r->state = kDone;
                s->completed++;
                s->scene_ms.push_back(std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count());
            }
            s->resident_bytes -= r->footprint;
            s->resident_particles -= r->n;
            s->resident_cells -= r->T;
            s->resident_bricks -= r->bricks;
            changed = true;
        }
        s->resident.swap(keep);
        if (changed) s->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

int sph_server_create(const sph_server_config *cfg, sph_server **out) {
    if (!cfg || !out) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    *out = nullptr;
    if (cfg->max_scenes == 0 || cfg->max_scenes > 65535u)
        return fail(SPH_ERR_INVALID_ARG, "max_scenes must be in [1, 65535]");
    if (cfg->max_particles == 0 || cfg->max_particles >= (1u << 30))
        return fail(SPH_ERR_INVALID_ARG, "max_particles must be in [1, 2^30)");
    if (cfg->max_table_cells == 0 || cfg->max_table_cells >= (1u << 31))
        return fail(SPH_ERR_INVALID_ARG, "max_table_cells must be in [1, 2^31)");
    if (cfg->substeps == 0) return fail(SPH_ERR_INVALID_ARG, "substeps must be >= 1");
    if (cfg->list_entries > (uint32_t)kList)
        return fail(SPH_ERR_INVALID_ARG, "list_entries must be at most %d", kList);
    sph_server *s = new (std::nothrow) sph_server();
    if (!s) return fail(SPH_ERR_OUT_OF_MEMORY, "host allocation");
    s->cfg = *cfg;
    if (s->cfg.max_bricks == 0) s->cfg.max_bricks = 1u << 22;
    if (s->cfg.list_entries == 0) s->cfg.list_entries = kServerListEntries;
    s->cfg_bricks = std::getenv("SPH_SERVER_BRICKS") && std::strcmp(std::getenv("SPH_SERVER_BRICKS"), "1") == 0;
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(e, "stream creation");
    }
    const uint64_t bucket_cap = 2ull * cfg->max_particles + 1024ull * cfg->max_scenes;
    const uint64_t grid_want = 16ull * cfg->max_particles + 4096ull * cfg->max_scenes;
    const uint64_t grid_cap = grid_want < (1ull << 30) ? grid_want : (1ull << 30);
    if (bucket_cap >= (1ull << 32)) {
        hipStreamDestroy(s->stream);
        delete s;
        return fail(SPH_ERR_CAPACITY, "max_particles / max_scenes too large");
    }
    int rc = s->chain.allocate(cfg->max_particles, cfg->max_table_cells, s->cfg.max_bricks,
                               (uint32_t)bucket_cap, (uint32_t)grid_cap, s->cfg.list_entries);
    if (rc) {
        hipStreamDestroy(s->stream);
        delete s;
        return rc;
    }
    e = hipMalloc((void **)&s->d_scenes, (size_t)cfg->max_scenes * sizeof(SceneConst));
    if (e == hipSuccess) e = hipMalloc((void **)&s->d_scene_of, (size_t)cfg->max_particles * 2);
    if (e != hipSuccess) {
        s->chain.release();
        if (s->d_scenes) hipFree(s->d_scenes);
        hipStreamDestroy(s->stream);
        delete s;
        return fail(SPH_ERR_OUT_OF_MEMORY, "scene tables: %s", hipGetErrorString(e));
    }
    s->chain.configure(0, 1, 1, 1, 0);
    s->worker = std::thread(worker_main, s);
    *out = s;
    return SPH_OK;
}

int sph_server_submit(sph_server *s, const sph_params *params, uint32_t n, const float *pos4,
                      const float *vel4, uint32_t steps, uint64_t *request_id) {
    if (!s || !request_id) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    int rc = validate_params(params);
    if (rc) return rc;
    if (n == 0) return fail(SPH_ERR_INVALID_ARG, "particle count must be > 0");
    if (steps == 0) return fail(SPH_ERR_INVALID_ARG, "steps must be >= 1");
    if (params->substeps != s->cfg.substeps)
        return fail(SPH_ERR_INVALID_ARG, "request substeps %u differ from the server's %u", params->substeps,
                    s->cfg.substeps);
    rc = validate_state(*params, n, pos4, vel4);
    if (rc) return rc;
    auto r = std::make_unique<Request>();
    r->params = *params;
    r->n = n;
    r->T = params->table_cells ? params->table_cells : n;
    r->params.table_cells = r->T;
    const SceneConst c = make_scene_const(r->params, r->T, 0, 0, n, 0, 0, 0);
    r->bricks = scene_bricks(c);
    r->footprint = scene_footprint(n, r->T, c.grid_cells, s->cfg.list_entries);
    if (r->footprint > s->cfg.hbm_budget_bytes || n > s->cfg.max_particles || r->T > s->cfg.max_table_cells ||
        (s->cfg_bricks && r->bricks > s->chain.b_cap))
        return fail(SPH_ERR_CAPACITY, "request exceeds the server's capacity");
    r->pos.assign(pos4, pos4 + (size_t)n * 4);
    r->vel.assign(vel4, vel4 + (size_t)n * 4);
    r->steps_left = steps;
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fail(SPH_ERR_STATE, "server stopped: %s", s->fatal.c_str());
    r->id = s->next_id++;
    *request_id = r->id;
    s->pending.push_back(r.get());
    s->requests[r->id] = std::move(r);
    s->submitted++;
    s->cv_work.notify_one();
    return SPH_OK;
}

int sph_server_poll(sph_server *s, uint64_t id) {
    if (!s) return fail(SPH_ERR_INVALID_ARG, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) return fail(SPH_ERR_INVALID_ARG, "unknown request %llu", (unsigned long long)id);
    return it->second->state;
}

int sph_server_wait(sph_server *s, uint64_t id, double timeout_s) {
    if (!s) return fail(SPH_ERR_INVALID_ARG, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) return fail(SPH_ERR_INVALID_ARG, "unknown request %llu", (unsigned long long)id);
    Request *r = it->second.get();
    auto ready = [&] { return r->state >= kDone || s->stop; };
    if (timeout_s < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
        return fail(SPH_ERR_STATE, "request %llu not finished within the timeout", (unsigned long long)id);
    if (r->state == kFailed) return fail(SPH_ERR_DEVICE, "request %llu failed: %s", (unsigned long long)id, r->error.c_str());
    if (r->state != kDone) return fail(SPH_ERR_STATE, "server stopped: %s", s->fatal.c_str());
    return SPH_OK;
}

int sph_server_result(sph_server *s, uint64_t id, float *pos4, float *vel4) {
    if (!s) return fail(SPH_ERR_INVALID_ARG, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) return fail(SPH_ERR_INVALID_ARG, "unknown request %llu", (unsigned long long)id);
    Request *r = it->second.get();
    if (r->state != kDone) return fail(SPH_ERR_STATE, "request %llu is not done", (unsigned long long)id);
    if (pos4) std::memcpy(pos4, r->pos.data(), r->pos.size() * sizeof(float));
    if (vel4) std::memcpy(vel4, r->vel.data(), r->vel.size() * sizeof(float));
    s->requests.erase(it);
    return SPH_OK;
}

int sph_server_get_stats(sph_server *s, sph_server_stats *out) {
    if (!s || !out) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    sph_server_stats st{};
    st.submitted = s->submitted;
    st.completed = s->completed;
    st.failed = s->failed;
    st.frames = s->frames;
    st.resident_scenes = (uint32_t)s->resident.size();
    st.queued = (uint32_t)s->pending.size();
    st.resident_particles = s->resident_particles;
    st.resident_bytes = s->resident_bytes;
    st.step_p50_ms = percentile(s->step_ms, 0.50);
    st.step_p99_ms = percentile(s->step_ms, 0.99);
    st.scene_p50_ms = percentile(s->scene_ms, 0.50);
    st.scene_p99_ms = percentile(s->scene_ms, 0.99);
    st.wait_p50_ms = percentile(s->wait_ms, 0.50);
    st.wait_p99_ms = percentile(s->wait_ms, 0.99);
    st.particle_steps_per_s = s->busy_s > 0 ? s->particle_steps / s->busy_s : 0.0;
    *out = st;
    return SPH_OK;
}

void sph_server_destroy(sph_server *s) {
    if (!s) return;
    {
        // The worker finishes the frame in flight and exits; queued and resident
        // requests are dropped.
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    if (s->worker.joinable()) s->worker.join();
    hipStreamSynchronize(s->stream);
    s->graph.drop();
    s->chain.release();
    if (s->d_scenes) hipFree(s->d_scenes);
    if (s->d_scene_of) hipFree(s->d_scene_of);
    hipStreamDestroy(s->stream);
    delete s;
}

}  // extern "C"
