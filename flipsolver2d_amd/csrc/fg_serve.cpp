// fg_serve.cpp -- continuous batching of scene simulations (fg_server_*).
//
// The reference runs one FlipSolver per scene and calls stepFrame() in its
// application loop (AutoBench/benchmarkrunnerapplication.cpp:88-92).  Here a
// server keeps many scenes resident on one device and advances all of them
// by one frame per tick through fg_batch_step_frame, admitting queued
// requests in arrival order as HBM budget and slots free up.  Host-only code
// on top of the C-ABI.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <string>
#include <unordered_map>
#include <vector>

#include "flipgpu.h"

namespace fg {
int report_error(int code, const std::string& msg);
}

namespace {

int bad(const char* msg) { return fg::report_error(FG_E_INVALID, msg); }

struct Req {
    int64_t id = 0;
    fg_scene_params params{};
    std::vector<int32_t> snv, fnv;
    std::vector<float> sv, fv, friction, fvisc;
    int32_t n_solids = 0, n_fluids = 0;
    uint32_t seed = 0;
    int32_t frames = 0;
    bool keep = false;
    uint64_t bytes = 0;
    fg_scene* scene = nullptr;
    fg_request_status status{};
};

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t lo = (size_t)std::floor(pos), hi = (size_t)std::ceil(pos);
    const double f = pos - (double)lo;
    return v[lo] * (1.0 - f) + v[hi] * f;
}

}  // namespace

struct fg_server {
    fg_ctx* ctx = nullptr;
    uint64_t budget = 0;
    int32_t max_resident = 0;
    std::chrono::steady_clock::time_point t0;
    int64_t next_id = 1;
    std::unordered_map<int64_t, Req> reqs;
    std::deque<int64_t> queue;      // arrival order
    std::vector<int64_t> resident;  // admitted, still stepping
    uint64_t resident_bytes = 0;    // scenes held (stepping or kept)
    // summary
    std::vector<double> frame_ms, admit_wait_ms;
    int64_t frames = 0, psub = 0, submitted = 0, completed = 0, rejected = 0, failed = 0, cancelled = 0;
    int64_t budget_waits = 0, cap_waits = 0;
    double busy_ms = 0;
    uint64_t peak_bytes = 0;
    int32_t peak_resident = 0;

    double now_ms() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

void drop_scene(fg_server* srv, Req& r)
{
    if (!r.scene) return;
    fg_scene_destroy(r.scene);
    r.scene = nullptr;
    srv->resident_bytes -= r.bytes;
}

// The head request cannot be admitted, for good: it leaves the queue as
// REJECTED with the code that refused it, and the queue moves on.
void reject_head(fg_server* srv, Req& r, int code)
{
    r.status.state = FG_REQ_REJECTED;
    r.status.error = code;
    r.status.finish_ms = srv->now_ms();
    srv->rejected++;
    srv->queue.pop_front();
}

int admit(fg_server* srv, int* admitted)
{
    *admitted = 0;
    while (!srv->queue.empty()) {
        Req& r = srv->reqs[srv->queue.front()];
        if ((int32_t)srv->resident.size() >= srv->max_resident) {
            srv->cap_waits++;
            break;
        }
        if (srv->resident_bytes + r.bytes > srv->budget) {   // head waits: no overtaking
            srv->budget_waits++;
            break;
        }
        fg_scene* sc = nullptr;
        int rc = fg_scene_create(srv->ctx, &r.params, &sc);
        if (rc == FG_E_BUDGET || rc == FG_E_DEVICE) {
            // the context's budget or the device is full: wait while this
            // server holds scenes that will free memory, else it never fits
            if (srv->resident_bytes > 0) break;
            reject_head(srv, r, rc);
            continue;
        }
        if (rc != FG_OK) {
            reject_head(srv, r, rc);
            continue;
        }
        {
            fg_scene_objects o{};
            o.solids = fg_polygons{r.n_solids, r.snv.data(), r.sv.data()};
            o.friction = r.friction.empty() ? nullptr : r.friction.data();
            o.fluids = fg_polygons{r.n_fluids, r.fnv.data(), r.fv.data()};
            o.fluid_viscosity = r.fvisc.empty() ? nullptr : r.fvisc.data();
            o.seed = r.seed;
            rc = fg_scene_init_objects(sc, &o);
        }
        if (rc != FG_OK) {
            // e.g. FG_E_CAPACITY: its fluid seeds more particles than max_particles
            fg_scene_destroy(sc);
            if (rc == FG_E_DEVICE) return rc;
            reject_head(srv, r, rc);
            continue;
        }
        r.scene = sc;
        srv->resident_bytes += r.bytes;
        r.status.state = FG_REQ_RESIDENT;
        r.status.admit_ms = srv->now_ms();
        fg_scene_particle_count(sc, &r.status.particles);
        srv->admit_wait_ms.push_back(r.status.admit_ms - r.status.submit_ms);
        srv->resident.push_back(r.id);
        srv->queue.pop_front();
        (*admitted)++;
    }
    return FG_OK;
}

}  // namespace

extern "C" {

int fg_server_create(fg_ctx* ctx, uint64_t hbm_budget_bytes, int32_t max_resident, fg_server** out)
{
    if (!ctx || !out) return bad("NULL argument");
    if (max_resident <= 0 || hbm_budget_bytes == 0) return bad("max_resident and hbm_budget_bytes must be > 0");
    auto* s = new fg_server();
    s->ctx = ctx;
    s->budget = hbm_budget_bytes;
    s->max_resident = max_resident;
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return FG_OK;
}

int fg_server_destroy(fg_server* srv)
{
    if (!srv) return FG_OK;
    for (auto& kv : srv->reqs) drop_scene(srv, kv.second);
    delete srv;
    return FG_OK;
}

int fg_server_submit(fg_server* srv, const fg_request* q, int64_t* id_out)
{
    if (!srv || !q || !id_out) return bad("NULL argument");
    if (q->frames < 1) return bad("frames must be >= 1");
    if ((q->n_solids > 0 && (!q->solid_nverts || !q->solid_verts)) ||
        (q->n_fluids > 0 && (!q->fluid_nverts || !q->fluid_verts)) || q->n_solids < 0 || q->n_fluids < 0)
        return bad("bad geometry arrays");
    uint64_t bytes = 0;
    const int rc = fg_scene_bytes(&q->params, &bytes);
    if (rc != FG_OK) return rc;
    Req r;
    r.id = srv->next_id++;
    r.params = q->params;
    r.n_solids = q->n_solids;
    r.n_fluids = q->n_fluids;
    int64_t ns = 0, nf = 0;
    for (int32_t k = 0; k < q->n_solids; k++) {
        if (q->solid_nverts[k] < 3) return bad("a polygon needs 3 vertices");
        r.snv.push_back(q->solid_nverts[k]);
        ns += q->solid_nverts[k];
    }
    for (int32_t k = 0; k < q->n_fluids; k++) {
        if (q->fluid_nverts[k] < 3) return bad("a polygon needs 3 vertices");
        r.fnv.push_back(q->fluid_nverts[k]);
        nf += q->fluid_nverts[k];
    }
    if (ns) r.sv.assign(q->solid_verts, q->solid_verts + 2 * ns);
    if (nf) r.fv.assign(q->fluid_verts, q->fluid_verts + 2 * nf);
    if (q->friction && q->n_solids > 0) r.friction.assign(q->friction, q->friction + q->n_solids);
    if (q->fluid_viscosity && q->n_fluids > 0) r.fvisc.assign(q->fluid_viscosity, q->fluid_viscosity + q->n_fluids);
    if (r.snv.empty()) r.snv.push_back(0);
    if (r.fnv.empty()) r.fnv.push_back(0);
    if (r.sv.empty()) r.sv.assign(2, 0.f);
    if (r.fv.empty()) r.fv.assign(2, 0.f);
    r.seed = q->seed;
    r.frames = q->frames;
    r.keep = q->keep != 0;
    r.bytes = bytes;
    r.status.submit_ms = srv->now_ms();
    srv->submitted++;
    if (bytes > srv->budget) {
        r.status.state = FG_REQ_REJECTED;
        r.status.error = FG_E_BUDGET;
        srv->rejected++;
    } else {
        r.status.state = FG_REQ_QUEUED;
        srv->queue.push_back(r.id);
    }
    *id_out = r.id;
    srv->reqs.emplace(r.id, std::move(r));
    return FG_OK;
}

int fg_server_tick(fg_server* srv, fg_tick_stats* st)
{
    if (!srv) return bad("server is NULL");
    const double t_start = srv->now_ms();
    fg_tick_stats out{};
    int admitted = 0;
    int rc = admit(srv, &admitted);
    if (rc != FG_OK) return rc;
    const double t_admitted = srv->now_ms();
    out.admitted = admitted;
    const size_t n = srv->resident.size();
    if (n > 0) {
        std::vector<fg_scene*> scenes(n);
        for (size_t k = 0; k < n; k++) scenes[k] = srv->reqs[srv->resident[k]].scene;
        std::vector<fg_stats> stats(n);
        rc = fg_batch_step_frame(srv->ctx, scenes.data(), (int32_t)n, stats.data());
        if (rc != FG_OK && rc != FG_E_NUMERIC) return rc;   // FG_E_NUMERIC: see stats[k].status
        const double t_done = srv->now_ms();
        std::vector<int64_t> still;
        for (size_t k = 0; k < n; k++) {
            Req& r = srv->reqs[srv->resident[k]];
            const fg_stats& s = stats[k];
            if (s.status != FG_OK) {   // quarantined: this request ends, the others go on
                r.status.state = FG_REQ_FAILED;
                r.status.error = s.status;
                r.status.finish_ms = t_done;
                srv->failed++;
                out.retired++;
                if (!r.keep) drop_scene(srv, r);
                continue;
            }
            r.status.frames_done++;
            r.status.particles = s.particles;
            r.status.particle_substeps += s.particle_substeps;
            r.status.max_frame_ms = std::max(r.status.max_frame_ms, s.frame_ms);
            srv->frame_ms.push_back(s.frame_ms);
            out.frame_ms_max = std::max(out.frame_ms_max, s.frame_ms);
            out.particle_substeps += s.particle_substeps;
            if (r.status.frames_done >= r.frames) {
                r.status.state = FG_REQ_DONE;
                r.status.finish_ms = t_done;
                srv->completed++;
                out.retired++;
                if (!r.keep) drop_scene(srv, r);
            } else {
                still.push_back(r.id);
            }
        }
        out.stepped = (int32_t)n;
        srv->resident.swap(still);
        srv->frames += (int64_t)n;
        srv->psub += out.particle_substeps;
    }
    out.queued = (int32_t)srv->queue.size();
    out.resident = (int32_t)srv->resident.size();
    out.resident_bytes = srv->resident_bytes;
    srv->peak_bytes = std::max(srv->peak_bytes, srv->resident_bytes);
    srv->peak_resident = std::max<int32_t>(srv->peak_resident, (int32_t)n);
    out.admit_ms = t_admitted - t_start;
    out.tick_ms = srv->now_ms() - t_start;
    srv->busy_ms += out.tick_ms;
    if (st) *st = out;
    return FG_OK;
}

int fg_server_status(fg_server* srv, int64_t id, fg_request_status* out)
{
    if (!srv || !out) return bad("NULL argument");
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return bad("unknown request id");
    *out = it->second.status;
    return FG_OK;
}

int fg_server_scene(fg_server* srv, int64_t id, fg_scene** out)
{
    if (!srv || !out) return bad("NULL argument");
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return bad("unknown request id");
    if (!it->second.scene) return fg::report_error(FG_E_STATE, "request holds no scene (queued, released or rejected)");
    *out = it->second.scene;
    return FG_OK;
}

int fg_server_release(fg_server* srv, int64_t id)
{
    if (!srv) return bad("server is NULL");
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return bad("unknown request id");
    Req& r = it->second;
    if (r.status.state != FG_REQ_DONE && r.status.state != FG_REQ_FAILED)
        return fg::report_error(FG_E_STATE, "request has not finished (fg_server_cancel stops it)");
    drop_scene(srv, r);
    if (r.status.state == FG_REQ_DONE) r.status.state = FG_REQ_RELEASED;
    return FG_OK;
}

int fg_server_cancel(fg_server* srv, int64_t id)
{
    if (!srv) return bad("server is NULL");
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return bad("unknown request id");
    Req& r = it->second;
    switch (r.status.state) {
    case FG_REQ_QUEUED:
        srv->queue.erase(std::find(srv->queue.begin(), srv->queue.end(), id));
        break;
    case FG_REQ_RESIDENT:
        srv->resident.erase(std::find(srv->resident.begin(), srv->resident.end(), id));
        drop_scene(srv, r);
        break;
    case FG_REQ_DONE:
    case FG_REQ_FAILED:
        return fg_server_release(srv, id);
    default:
        return fg::report_error(FG_E_STATE, "request already ended");
    }
    r.status.state = FG_REQ_CANCELLED;
    r.status.error = FG_E_CANCELLED;
    r.status.finish_ms = srv->now_ms();
    srv->cancelled++;
    return FG_OK;
}

int fg_server_summary_get(fg_server* srv, fg_server_summary* out)
{
    if (!srv || !out) return bad("NULL argument");
    fg_server_summary s{};
    s.frames = srv->frames;
    s.particle_substeps = srv->psub;
    s.busy_ms = srv->busy_ms;
    s.p50_frame_ms = percentile(srv->frame_ms, 0.50);
    s.p99_frame_ms = percentile(srv->frame_ms, 0.99);
    s.max_frame_ms = srv->frame_ms.empty() ? 0.0 : *std::max_element(srv->frame_ms.begin(), srv->frame_ms.end());
    s.p50_admit_wait_ms = percentile(srv->admit_wait_ms, 0.50);
    s.p99_admit_wait_ms = percentile(srv->admit_wait_ms, 0.99);
    s.submitted = srv->submitted;
    s.completed = srv->completed;
    s.rejected = srv->rejected;
    s.peak_resident_bytes = srv->peak_bytes;
    s.peak_resident = srv->peak_resident;
    s.failed = srv->failed;
    s.cancelled = srv->cancelled;
    s.budget_waits = srv->budget_waits;
    s.cap_waits = srv->cap_waits;
    *out = s;
    return FG_OK;
}

}  // extern "C"
