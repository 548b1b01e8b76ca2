// serve.hip -- continuous-batching server over one batched libastro context.
//
// Requests (independent systems with their own step counts) wait in a FIFO queue.  At every
// tick boundary the server retires systems that have finished, admits queued requests while
// the batch stays within its HBM budget, particle, system and tree-node capacities, compacts
// the surviving state on the device, and runs one batched step for every active system.
// Each system's arithmetic depends only on its own particles (own cube, keys, tree and
// root; no cross-system reduction), so a result is the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "ctx.hpp"

using namespace astro;

namespace {

struct Request {
    uint64_t id = 0;
    uint32_t n = 0, steps = 0, done = 0;
    uint32_t offset = 0;  // position in the batch while active
    std::vector<float4> pos, vel;
    std::vector<uint32_t> type;
    double t_submit = 0, t_admit = -1, t_done = -1;
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
    uint64_t bytes_per_particle = 0;
    bool layout_dirty = false;
    uint64_t ticks = 0, particle_steps = 0;
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
    if (s->stage_pos) (void)hipFree(s->stage_pos);
    if (s->stage_vel) (void)hipFree(s->stage_vel);
    if (s->stage_type) (void)hipFree(s->stage_type);
    astro_destroy(s->ctx);
    delete s;
    return ASTRO_OK;
}

extern "C" int astro_server_create(const astro_server_config* cfg, const astro_params* p,
                                   astro_server** out) {
    if (!cfg || !p || !out || cfg->max_particles == 0 || cfg->max_systems == 0)
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
    cc.emit_octree_nodes = 0;
    cc.device = cfg->device;
    int rc = astro_create(&cc, &s->ctx);
    if (rc) {
        delete s;
        return rc;
    }
    if ((rc = astro_set_params(s->ctx, p)) || (rc = astro_set_use_graph(s->ctx, 0))) {
        astro_server_destroy(s);
        return rc;
    }
    astro_mass_bounds(s->ctx, nullptr, 0, 4, true);  // no systems yet; admissions fold in
    const uint64_t NP = cfg->max_particles;
    if (hipMalloc(&s->stage_pos, NP * sizeof(float4)) != hipSuccess ||
        hipMalloc(&s->stage_vel, NP * sizeof(float4)) != hipSuccess ||
        hipMalloc(&s->stage_type, NP * sizeof(uint32_t)) != hipSuccess) {
        astro_server_destroy(s);
        return ASTRO_ERR_NOMEM;
    }
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
        s->completions.push_back({id, n, 0, r->t_submit, r->t_admit, r->t_done});
        s->results[id] = std::move(r);
        return ASTRO_OK;
    }
    s->queue.push_back(std::move(r));
    return ASTRO_OK;
}

// Move finished systems out of the batch: download their state, record completion.
static int retire(astro_server* s, uint32_t* completed) {
    astro_ctx* c = s->ctx;
    std::vector<std::unique_ptr<Request>> keep;
    const double t = s->now();
    for (auto& r : s->active) {
        if (r->done < r->steps) {
            keep.push_back(std::move(r));
            continue;
        }
        r->pos.resize(r->n);
        r->vel.resize(r->n);
        HIPCHK(c, hipMemcpyAsync(r->pos.data(), c->pos + r->offset, r->n * sizeof(float4),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(r->vel.data(), c->vel + r->offset, r->n * sizeof(float4),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        r->t_done = t;
        s->completions.push_back({r->id, r->n, r->steps, r->t_submit, r->t_admit, r->t_done});
        (*completed)++;
        s->layout_dirty = true;
        s->results[r->id] = std::move(r);
    }
    s->active = std::move(keep);
    return ASTRO_OK;
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
        s->active.push_back(std::move(s->queue.front()));
        s->queue.pop_front();
        admitted++;
        s->layout_dirty = true;
    }
    return admitted;
}

// Lay the active systems out contiguously: survivors move on the device (through the staging
// buffers), newcomers are copied from the host.
static int relayout(astro_server* s) {
    astro_ctx* c = s->ctx;
    hipStream_t st = c->stream;
    uint32_t off = 0;
    std::vector<uint32_t> offsets;
    for (auto& r : s->active) {
        offsets.push_back(off);
        if (r->t_admit >= 0 && r->done > 0) {  // resident: device copy
            HIPCHK(c, hipMemcpyAsync(s->stage_pos + off, c->pos + r->offset, r->n * sizeof(float4),
                                     hipMemcpyDeviceToDevice, st));
            HIPCHK(c, hipMemcpyAsync(s->stage_vel + off, c->vel + r->offset, r->n * sizeof(float4),
                                     hipMemcpyDeviceToDevice, st));
            HIPCHK(c, hipMemcpyAsync(s->stage_type + off, c->type + r->offset,
                                     r->n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
        } else {  // newly admitted: host copy
            HIPCHK(c, hipMemcpyAsync(s->stage_pos + off, r->pos.data(), r->n * sizeof(float4),
                                     hipMemcpyHostToDevice, st));
            HIPCHK(c, hipMemcpyAsync(s->stage_vel + off, r->vel.data(), r->n * sizeof(float4),
                                     hipMemcpyHostToDevice, st));
            HIPCHK(c, hipMemcpyAsync(s->stage_type + off, r->type.data(), r->n * sizeof(uint32_t),
                                     hipMemcpyHostToDevice, st));
        }
        off += r->n;
    }
    const uint32_t n = off;
    if (n) {
        HIPCHK(c, hipMemcpyAsync(c->pos, s->stage_pos, n * sizeof(float4), hipMemcpyDeviceToDevice, st));
        HIPCHK(c, hipMemcpyAsync(c->vel, s->stage_vel, n * sizeof(float4), hipMemcpyDeviceToDevice, st));
        HIPCHK(c, hipMemcpyAsync(c->type, s->stage_type, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    }
    for (size_t k = 0; k < s->active.size(); k++) {
        auto& r = s->active[k];
        r->offset = offsets[k];
        if (r->done == 0) {  // host copies are no longer needed once resident
            std::vector<float4>().swap(r->pos);
            std::vector<float4>().swap(r->vel);
            std::vector<uint32_t>().swap(r->type);
        }
    }
    offsets.push_back(n);
    c->nsys = (uint32_t)s->active.size();
    c->sys_off = offsets;
    if (c->nsys) {
        HIPCHK(c, hipMemcpyAsync(c->d_sys_off, c->sys_off.data(), (c->nsys + 1) * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, st));
        launch_fill_sys_of(c->sys_of, c->d_sys_off, c->nsys, n, st);
    }
    HIPCHK(c, hipStreamSynchronize(st));
    if (int rc = astro_set_counts(c, n, n)) return rc;
    if (int rc = astro_refresh_layers(c)) return rc;
    s->layout_dirty = false;
    return ASTRO_OK;
}

extern "C" int astro_server_tick(astro_server* s, float dt, astro_tick_info* info) {
    if (!s) return ASTRO_ERR_ARG;
    astro_ctx* c = s->ctx;
    astro_tick_info ti{};
    int rc = retire(s, &ti.completed);
    if (rc) return sfail(s, rc, c->err);
    ti.admitted = admit(s);
    if (s->layout_dirty && (rc = relayout(s))) return sfail(s, rc, c->err);
    ti.active = (uint32_t)s->active.size();
    ti.queued = (uint32_t)s->queue.size();
    ti.particles = c->ns;
    if (!s->active.empty()) {
        const double t0 = s->now();
        if ((rc = astro_enqueue_step(c, dt))) return sfail(s, rc, c->err);
        if ((rc = astro_check(c))) return sfail(s, rc, c->err);  // syncs
        ti.step_ms = (s->now() - t0) * 1e3;
        for (auto& r : s->active) r->done++;
        s->particle_steps += c->ns;
        s->ticks++;
        // retire right away so completion times include no idle tick
        rc = retire(s, &ti.completed);
        if (rc) return sfail(s, rc, c->err);
    }
    ti.active_after = (uint32_t)s->active.size();
    if (info) *info = ti;
    return ASTRO_OK;
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
