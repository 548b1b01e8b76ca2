// serve.hip -- continuous-batching server over one batched libastro context.
//
// Requests (independent systems with their own step counts) wait in a FIFO queue.  At every
// tick the server settles what the device has run (when a request has just finished, and
// every kSettleEvery ticks), retires finished systems, evicts failed ones, admits queued
// requests while the batch stays within its HBM budget and its particle, system and
// tree-node capacities, lays the batch out again when it changed (only systems whose place
// changes move), and enqueues one batched step: eagerly on the first tick of a layout, then
// as one captured graph replayed while the batch stays the same.  Each system's arithmetic
// depends only on its own particles (own cube, keys, tree and root; no cross-system
// reduction), so a result is the same alone and in any mix.
//
// Load shedding: with max_wait_s set, a request whose estimated wait before its last step
// (the work ahead of it and its own, over the measured throughput) exceeds it is refused at
// submit.  Failure isolation: the integrator flags every system whose state turns
// non-finite or leaves max_extent; such a system is evicted with status FAILED and the
// others continue bit-identically.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "ctx.hpp"

using namespace astro;

namespace {

constexpr uint32_t kSettleEvery = 16;  // ticks the host may run ahead of the device

struct Request {
    uint64_t id = 0;
    uint32_t n = 0, steps = 0, done = 0;
    uint32_t offset = 0;     // position in the batch while resident
    bool resident = false;   // state lives in the batch (host copies released)
    uint32_t status = ASTRO_REQUEST_DONE;
    std::vector<float4> pos, vel;
    std::vector<uint32_t> type;
    double t_submit = 0, t_admit = -1, t_done = -1;
    uint64_t work() const { return (uint64_t)n * (steps - done); }
};

uint64_t nodes_for(const std::vector<uint32_t>& sizes, uint32_t bits) {
    uint64_t t = 0;
    for (uint32_t l = 0; l <= bits; l++) {
        uint32_t b = 3 * bits - 3 * l;
        uint64_t possible = b >= 63 ? (1ull << 63) : (1ull << b);
        for (uint32_t n : sizes) t += std::min<uint64_t>(n, possible);
    }
    return t;
}

}  // namespace

struct astro_server {
    astro_ctx* ctx = nullptr;
    astro_server_config cfg{};
    std::string err;
    std::deque<std::unique_ptr<Request>> queue;
    std::vector<std::unique_ptr<Request>> active;
    std::vector<astro_completion> completions;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> results;
    float4 *stage_pos = nullptr, *stage_vel = nullptr;
    uint32_t* stage_type = nullptr;
    uint32_t* guard = nullptr;       // [max_systems] failure flags raised by the integrator
    uint64_t bytes_per_particle = 0;
    bool layout_dirty = false;
    bool unsettled = false;          // ticks enqueued since the last settle
    uint32_t stable_ticks = 0;       // ticks since the last layout change
    uint32_t since_settle = 0;
    uint64_t queued_work = 0;        // particle-steps waiting in the queue
    uint64_t ticks = 0, particle_steps = 0;
    // throughput estimate for load shedding: the highest particle-steps per second of busy
    // time between two settles
    double rate = 0.0, settle_t = 0.0;
    uint64_t settle_ps = 0;
    std::chrono::steady_clock::time_point epoch = std::chrono::steady_clock::now();
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - epoch).count();
    }
};

static int sfail(astro_server* s, int code, const std::string& m) {
    if (s) s->err = m;
    return code;
}

extern "C" const char* astro_server_last_error(const astro_server* s) {
    return s ? s->err.c_str() : "null server";
}

extern "C" int astro_server_destroy(astro_server* s) {
    if (!s) return ASTRO_OK;
    if (s->ctx) (void)hipSetDevice(s->ctx->device);
    if (s->stage_pos) (void)hipFree(s->stage_pos);
    if (s->stage_vel) (void)hipFree(s->stage_vel);
    if (s->stage_type) (void)hipFree(s->stage_type);
    if (s->guard) (void)hipFree(s->guard);
    astro_destroy(s->ctx);
    delete s;
    return ASTRO_OK;
}

extern "C" int astro_server_create(const astro_server_config* cfg, const astro_params* p,
                                   astro_server** out) {
    if (!cfg || !p || !out || cfg->max_particles == 0 || cfg->max_systems == 0 ||
        !(cfg->max_extent >= 0.0f) || !(cfg->max_wait_s >= 0.0))
        return ASTRO_ERR_ARG;
    if (p->lighting_stars) return ASTRO_ERR_ARG;  // requests carry no colours through the batch
    *out = nullptr;
    auto* s = new astro_server();
    s->cfg = *cfg;
    astro_config cc{};
    cc.max_particles = cfg->max_particles;
    cc.max_tree = cfg->max_particles;
    cc.bits_per_axis = p->bits_per_axis;
    cc.max_systems = cfg->max_systems;
    cc.emit_octree_nodes = ASTRO_EMIT_NEVER;
    cc.device = cfg->device;
    int rc = astro_create(&cc, &s->ctx);
    if (rc) {
        delete s;
        return rc;
    }
    if ((rc = astro_set_params(s->ctx, p))) {
        astro_server_destroy(s);
        return rc;
    }
    astro_mass_bounds(s->ctx, nullptr, 0, 4, true);  // no systems yet; admissions fold in
    const uint64_t NP = cfg->max_particles;
    if (hipMalloc(&s->stage_pos, NP * sizeof(float4)) != hipSuccess ||
        hipMalloc(&s->stage_vel, NP * sizeof(float4)) != hipSuccess ||
        hipMalloc(&s->stage_type, NP * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&s->guard, cfg->max_systems * sizeof(uint32_t)) != hipSuccess) {
        astro_server_destroy(s);
        return ASTRO_ERR_NOMEM;
    }
    launch_fill_u32(s->guard, 0, cfg->max_systems, s->ctx->stream);
    s->ctx->guard = s->guard;
    s->ctx->guard_extent = cfg->max_extent;
    // device bytes held per particle of capacity: context buffers + tree nodes + staging
    s->bytes_per_particle = (s->ctx->device_bytes + 36 * NP + NP - 1) / NP;
    *out = s;
    return ASTRO_OK;
}

extern "C" int astro_server_device_bytes(const astro_server* s, uint64_t* bytes,
                                         uint64_t* per_particle) {
    if (!s) return ASTRO_ERR_ARG;
    if (bytes) *bytes = s->ctx->device_bytes + 36ull * s->cfg.max_particles;
    if (per_particle) *per_particle = s->bytes_per_particle;
    return ASTRO_OK;
}

extern "C" int astro_server_submit(astro_server* s, uint64_t id, const PositionMass* pos,
                                   const VelocityRadius* vel, const ColorType* col, uint32_t n,
                                   uint32_t steps) {
    if (!s || !pos || n == 0) return ASTRO_ERR_ARG;
    const uint64_t budget = s->cfg.hbm_budget ? s->cfg.hbm_budget : UINT64_MAX;
    if (n > s->cfg.max_particles || (uint64_t)n * s->bytes_per_particle > budget ||
        nodes_for({n}, s->ctx->params.bits_per_axis) > s->ctx->node_capacity)
        return sfail(s, ASTRO_ERR_CAPACITY, "request can never fit the server");
    if (s->cfg.max_wait_s > 0.0 && s->rate > 0.0 && steps) {
        // the later of two bounds on its last step: the work ahead of it and its own, and its
        // steps at the tick length of the batch it joins (ticks grow with the batch)
        uint64_t ahead = s->queued_work, batch = n;
        for (auto& r : s->active) {
            ahead += r->work();
            batch += r->n;
        }
        for (auto& q : s->queue) batch += q->n;
        batch = std::min<uint64_t>(batch, s->cfg.max_particles);
        const double work = (double)std::max(ahead + (uint64_t)n * steps, batch * steps);
        if (work / s->rate > s->cfg.max_wait_s)
            return sfail(s, ASTRO_ERR_BUSY, "load shed: the estimated wait exceeds max_wait_s");
    }
    for (uint32_t i = 0; i < n; i++) {
        bool ok = std::isfinite(pos[i].position[0]) && std::isfinite(pos[i].position[1]) &&
                  std::isfinite(pos[i].position[2]) && std::isfinite(pos[i].mass);
        if (vel)
            ok = ok && std::isfinite(vel[i].velocity[0]) && std::isfinite(vel[i].velocity[1]) &&
                 std::isfinite(vel[i].velocity[2]);
        if (!ok) return sfail(s, ASTRO_ERR_ARG, "request holds a non-finite position, mass or velocity");
    }
    auto r = std::make_unique<Request>();
    r->id = id;
    r->n = n;
    r->steps = steps;
    r->t_submit = s->now();
    r->pos.resize(n);
    r->vel.resize(n);
    r->type.assign(n, 0u);
    for (uint32_t i = 0; i < n; i++) {
        r->pos[i] = make_float4(pos[i].position[0], pos[i].position[1], pos[i].position[2],
                                pos[i].mass);
        r->vel[i] = vel ? make_float4(vel[i].velocity[0], vel[i].velocity[1], vel[i].velocity[2],
                                      vel[i].radius)
                        : make_float4(0, 0, 0, 0);
        if (col) r->type[i] = col[i].type;
    }
    if (steps == 0) {  // nothing to run: complete at once
        r->t_admit = r->t_done = r->t_submit;
        s->completions.push_back({id, n, 0, r->t_submit, r->t_admit, r->t_done,
                                  ASTRO_REQUEST_DONE, 0});
        s->results[id] = std::move(r);
        return ASTRO_OK;
    }
    s->queued_work += (uint64_t)n * steps;
    s->queue.push_back(std::move(r));
    return ASTRO_OK;
}

// Moves a resident request out of the batch with its current state (the stream is idle).
static int finish(astro_server* s, std::unique_ptr<Request>& r, uint32_t status, double t) {
    astro_ctx* c = s->ctx;
    r->pos.resize(r->n);
    r->vel.resize(r->n);
    HIPCHK(c, hipMemcpyAsync(r->pos.data(), c->pos + r->offset, r->n * sizeof(float4),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(r->vel.data(), c->vel + r->offset, r->n * sizeof(float4),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    r->status = status;
    r->t_done = t;
    r->resident = false;
    s->completions.push_back({r->id, r->n, r->steps, r->t_submit, r->t_admit, r->t_done, status,
                              r->done});
    s->layout_dirty = true;
    s->results[r->id] = std::move(r);
    return ASTRO_OK;
}

// Waits for the enqueued ticks, then retires finished requests and evicts failed ones (every
// active one if the device raised its error word).  Returns ASTRO_ERR_DEVICE in that case.
static int settle(astro_server* s, astro_tick_info* ti) {
    astro_ctx* c = s->ctx;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double t = s->now();
    if (s->particle_steps > s->settle_ps && t > s->settle_t) {
        // the highest rate seen: the device's rate with a full batch (a batch of small
        // systems alone runs far slower and would shed requests the device can take)
        const double r = (double)(s->particle_steps - s->settle_ps) / (t - s->settle_t);
        s->rate = std::max(s->rate, r);
    }
    s->settle_t = t;
    s->settle_ps = s->particle_steps;
    s->unsettled = false;
    s->since_settle = 0;
    const int dev = astro_check(c);  // reads and clears the error word
    if (dev && dev != ASTRO_ERR_DEVICE) return dev;
    const uint32_t ns = (uint32_t)s->active.size();
    std::vector<uint32_t> flags(ns, 0u);
    if (ns) {
        HIPCHK(c, hipMemcpyAsync(flags.data(), s->guard, ns * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    std::vector<std::unique_ptr<Request>> keep;
    for (uint32_t k = 0; k < ns; k++) {
        auto& r = s->active[k];
        const bool failed = dev != 0 || flags[k] != 0u;
        if (!failed && r->done < r->steps) {
            keep.push_back(std::move(r));
            continue;
        }
        if (int rc = finish(s, r, failed ? ASTRO_REQUEST_FAILED : ASTRO_REQUEST_DONE, t)) return rc;
        if (ti) {
            ti->completed += failed ? 0u : 1u;
            ti->failed += failed ? 1u : 0u;
        }
    }
    s->active = std::move(keep);
    if (dev) {
        std::string m = c->err;
        return sfail(s, ASTRO_ERR_DEVICE, m + " (every active request failed)");
    }
    return ASTRO_OK;
}

// Whether the front of the queue fits next to the active systems.
static bool front_fits(const astro_server* s) {
    if (s->queue.empty()) return false;
    const Request* r = s->queue.front().get();
    std::vector<uint32_t> sizes;
    uint64_t parts = r->n;
    for (auto& a : s->active) {
        sizes.push_back(a->n);
        parts += a->n;
    }
    sizes.push_back(r->n);
    const uint64_t budget = s->cfg.hbm_budget ? s->cfg.hbm_budget : UINT64_MAX;
    return sizes.size() <= s->cfg.max_systems && parts <= s->cfg.max_particles &&
           parts * s->bytes_per_particle <= budget &&
           nodes_for(sizes, s->ctx->params.bits_per_axis) <= s->ctx->node_capacity;
}

// Admit queued requests (FIFO, no overtaking) while the batch fits.
static uint32_t admit(astro_server* s) {
    uint32_t admitted = 0;
    std::vector<uint32_t> sizes;
    uint64_t parts = 0;
    for (auto& r : s->active) {
        sizes.push_back(r->n);
        parts += r->n;
    }
    const uint64_t budget = s->cfg.hbm_budget ? s->cfg.hbm_budget : UINT64_MAX;
    const double t = s->now();
    while (!s->queue.empty()) {
        Request* r = s->queue.front().get();
        if (sizes.size() + 1 > s->cfg.max_systems || parts + r->n > s->cfg.max_particles ||
            (parts + r->n) * s->bytes_per_particle > budget)
            break;
        sizes.push_back(r->n);
        if (nodes_for(sizes, s->ctx->params.bits_per_axis) > s->ctx->node_capacity) {
            sizes.pop_back();
            break;
        }
        parts += r->n;
        astro_mass_bounds(s->ctx, &r->pos[0].w, r->n, 4, false);  // kept conservative
        r->t_admit = t;
        s->queued_work -= (uint64_t)r->n * r->steps;
        s->active.push_back(std::move(s->queue.front()));
        s->queue.pop_front();
        admitted++;
        s->layout_dirty = true;
    }
    return admitted;
}

// Lays the active systems out contiguously in their order: resident systems whose place
// changes move on the device (through the staging buffers, as the ranges may overlap), the
// others stay; newcomers are copied from the host into their places.
static int relayout(astro_server* s) {
    astro_ctx* c = s->ctx;
    hipStream_t st = c->stream;
    uint32_t off = 0;
    std::vector<uint32_t> offsets;
    std::vector<Request*> moved;
    for (auto& r : s->active) {
        offsets.push_back(off);
        if (r->resident && r->offset != off) {
            HIPCHK(c, hipMemcpyAsync(s->stage_pos + off, c->pos + r->offset, r->n * sizeof(float4),
                                     hipMemcpyDeviceToDevice, st));
            HIPCHK(c, hipMemcpyAsync(s->stage_vel + off, c->vel + r->offset, r->n * sizeof(float4),
                                     hipMemcpyDeviceToDevice, st));
            HIPCHK(c, hipMemcpyAsync(s->stage_type + off, c->type + r->offset,
                                     r->n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
            moved.push_back(r.get());
        }
        off += r->n;
    }
    for (size_t k = 0; k < s->active.size(); k++) {
        Request* r = s->active[k].get();
        const uint32_t o = offsets[k];
        if (r->resident && r->offset != o) {
            HIPCHK(c, hipMemcpyAsync(c->pos + o, s->stage_pos + o, r->n * sizeof(float4),
                                     hipMemcpyDeviceToDevice, st));
            HIPCHK(c, hipMemcpyAsync(c->vel + o, s->stage_vel + o, r->n * sizeof(float4),
                                     hipMemcpyDeviceToDevice, st));
            HIPCHK(c, hipMemcpyAsync(c->type + o, s->stage_type + o, r->n * sizeof(uint32_t),
                                     hipMemcpyDeviceToDevice, st));
        } else if (!r->resident) {  // newly admitted: from the host, then release the copies
            HIPCHK(c, hipMemcpyAsync(c->pos + o, r->pos.data(), r->n * sizeof(float4),
                                     hipMemcpyHostToDevice, st));
            HIPCHK(c, hipMemcpyAsync(c->vel + o, r->vel.data(), r->n * sizeof(float4),
                                     hipMemcpyHostToDevice, st));
            HIPCHK(c, hipMemcpyAsync(c->type + o, r->type.data(), r->n * sizeof(uint32_t),
                                     hipMemcpyHostToDevice, st));
        }
    }
    const uint32_t n = off;
    offsets.push_back(n);
    c->nsys = (uint32_t)s->active.size();
    c->sys_off = offsets;
    if (c->nsys) {
        HIPCHK(c, hipMemcpyAsync(c->d_sys_off, c->sys_off.data(), (c->nsys + 1) * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, st));
        launch_fill_sys_of(c->sys_of, c->d_sys_off, c->nsys, n, st);
        launch_fill_u32(s->guard, 0, c->nsys, st);  // flags follow the systems' order
    }
    if (int rc = astro_set_counts(c, n, n)) return rc;  // synchronises the stream
    for (size_t k = 0; k < s->active.size(); k++) {
        auto& r = s->active[k];
        r->offset = offsets[k];
        if (!r->resident) {
            r->resident = true;
            std::vector<float4>().swap(r->pos);
            std::vector<float4>().swap(r->vel);
            std::vector<uint32_t>().swap(r->type);
        }
    }
    // the traversal block: 1,024 threads, each block with the top layers of its own system's
    // tree in LDS (measured on 256 systems of 1,024 to 65,536 particles: 4.33 ms per step
    // against 5.92 with blocks of 256 and no LDS top)
    c->force_block = 1024u;
    if (int rc = astro_refresh_layers(c)) return rc;  // also drops the captured step graph
    s->layout_dirty = false;
    s->stable_ticks = 0;
    return ASTRO_OK;
}

static bool any_finished(const astro_server* s) {
    for (auto& r : s->active)
        if (r->done >= r->steps) return true;
    return false;
}

extern "C" int astro_server_sync(astro_server* s, astro_tick_info* info) {
    if (!s) return ASTRO_ERR_ARG;
    astro_tick_info ti{};
    int rc = settle(s, &ti);
    ti.active_after = (uint32_t)s->active.size();
    ti.queued = (uint32_t)s->queue.size();
    ti.synced = 1;
    if (info) *info = ti;
    return rc;
}

extern "C" int astro_server_tick(astro_server* s, float dt, astro_tick_info* info) {
    if (!s) return ASTRO_ERR_ARG;
    astro_ctx* c = s->ctx;
    astro_tick_info ti{};
    const double t0 = s->now();
    int rc = ASTRO_OK;
    auto out = [&](int code) {
        ti.active_after = (uint32_t)s->active.size();
        ti.step_ms = (s->now() - t0) * 1e3;
        if (info) *info = ti;
        return code;
    };
    // settle before a layout change too: the failure flags follow the systems' order
    if (s->unsettled &&
        (any_finished(s) || s->since_settle >= kSettleEvery || s->layout_dirty || front_fits(s))) {
        ti.synced = 1;
        if ((rc = settle(s, &ti))) return out(rc);
    }
    ti.admitted = admit(s);
    if (s->layout_dirty && (rc = relayout(s))) return out(sfail(s, rc, c->err));
    ti.active = (uint32_t)s->active.size();
    ti.queued = (uint32_t)s->queue.size();
    ti.particles = c->ns;
    if (!s->active.empty()) {
        // the first tick of a layout runs eagerly; from the second the step is captured once
        // and replayed while the batch stays the same
        c->use_graph = s->stable_ticks >= 1;
        if (!s->unsettled) {  // the throughput is measured over busy time only
            s->settle_t = s->now();
            s->settle_ps = s->particle_steps;
        }
        if ((rc = astro_step(c, dt, 1))) return out(sfail(s, rc, c->err));
        for (auto& r : s->active) r->done++;  // counted as enqueued
        s->particle_steps += c->ns;
        s->ticks++;
        s->stable_ticks++;
        s->since_settle++;
        s->unsettled = true;
        // settle at once when a request has run its last step, so that its completion time
        // includes no idle tick
        if (any_finished(s)) {
            ti.synced = 1;
            if ((rc = settle(s, &ti))) return out(rc);
        }
    }
    return out(ASTRO_OK);
}

extern "C" int astro_server_poll(astro_server* s, astro_completion* out, uint32_t max,
                                 uint32_t* n_out) {
    if (!s || !n_out || (!out && max)) return ASTRO_ERR_ARG;
    const uint32_t n = (uint32_t)std::min<size_t>(max, s->completions.size());
    for (uint32_t k = 0; k < n; k++) out[k] = s->completions[k];
    s->completions.erase(s->completions.begin(), s->completions.begin() + n);
    *n_out = n;
    return ASTRO_OK;
}

extern "C" int astro_server_fetch(astro_server* s, uint64_t id, PositionMass* pos,
                                  VelocityRadius* vel) {
    if (!s) return ASTRO_ERR_ARG;
    auto it = s->results.find(id);
    if (it == s->results.end()) return sfail(s, ASTRO_ERR_STATE, "no finished request with that id");
    Request& r = *it->second;
    for (uint32_t i = 0; i < r.n; i++) {
        if (pos) {
            std::memset(&pos[i], 0, sizeof(PositionMass));
            pos[i].position[0] = r.pos[i].x;
            pos[i].position[1] = r.pos[i].y;
            pos[i].position[2] = r.pos[i].z;
            pos[i].mass = r.pos[i].w;
        }
        if (vel) {
            std::memset(&vel[i], 0, sizeof(VelocityRadius));
            vel[i].velocity[0] = r.vel[i].x;
            vel[i].velocity[1] = r.vel[i].y;
            vel[i].velocity[2] = r.vel[i].z;
            vel[i].radius = r.vel[i].w;
        }
    }
    s->results.erase(it);
    return ASTRO_OK;
}

extern "C" int astro_server_counters(const astro_server* s, uint64_t* ticks,
                                     uint64_t* particle_steps, uint32_t* queued,
                                     uint32_t* active) {
    if (!s) return ASTRO_ERR_ARG;
    if (ticks) *ticks = s->ticks;
    if (particle_steps) *particle_steps = s->particle_steps;
    if (queued) *queued = (uint32_t)s->queue.size();
    if (active) *active = (uint32_t)s->active.size();
    return ASTRO_OK;
}
