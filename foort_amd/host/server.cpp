// server.cpp -- render server (no counterpart in the reference): requests of
// any metric, camera and resolution, each an adaptive mesh of a few
// iterations, run concurrently on one GPU by continuous batching.
//
// Requests whose run descriptions differ only in the metric parameters, the
// camera and the screen form a group; each group has one device queue
// (foort_gpu_queue_*): a ring of pixels that one persistent kernel drains, its
// lanes refilled from whichever request pushed the next pixel.  A group thread
// owns the queue: it pushes the iteration batches that are ready, runs
// time-sliced launches on the queue's stream, and collects the batches that are
// done.  Mesh work -- building a request's mesh, feeding it an iteration's
// values, queueing the next iteration -- runs on a pool of host threads, so
// the GPU keeps integrating other requests' pixels meanwhile.  Each request has
// at most one batch on the device: its mesh needs those values before it can
// choose the next pixels.
#include "host.hpp"

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace foort_host {
namespace {

using Clock = std::chrono::steady_clock;

// device bytes per ring pixel: screen index (16), request slot (4), termination,
// quadrant, passes (4 each), closest radius (8)
constexpr uint64_t kBytesPerPixel = 16 + 4 + 4 + 4 + 4 + 8;
constexpr uint32_t kSlotsPerQueue = 1024;

// The fields a queue's group shares (the queue checks the same on the device).
std::string group_key(const foort_run_desc &d)
{
    std::string k;
    auto add = [&k](const auto &v) { k.append(reinterpret_cast<const char *>(&v), sizeof v); };
    add(d.metric), add(d.rlog_scale), add(d.integrator), add(d.christoffel), add(d.step_size);
    add(d.smallest_step), add(d.derivative_h), add(d.verlet_tolerance), add(d.term_flags), add(d.precision);
    add(d.horizon_epsilon), add(d.horizon_update_n), add(d.sphere_radius), add(d.sphere_update_n);
    add(d.timeout_max_steps), add(d.timeout_update_n), add(d.theta_epsilon), add(d.theta_update_n);
    add(d.diag_flags), add(d.ep_threshold), add(d.ep_update.update_n_steps), add(d.ep_update.update_start);
    add(d.ep_update.update_finish), add(d.cr_update.update_n_steps), add(d.cr_update.update_start);
    add(d.cr_update.update_finish), add(d.step_cap), add(d.lanes_per_cu), add(d.refill_lanes);
    add(d.refill_interval), add(d.mixed_switch_radius), add(d.mixed_pole_margin);
    return k;
}

struct Group;

// The output handler the program would build for this run description and mesh with
// Output.FilePrefix = prefix, FileExtension = ext, TimeStamp = false, FirstLineInfo = true
// (FOORT/Config.cpp:1079-1155): one file per diagnostic in the order of CreateDiagnosticVector.
std::unique_ptr<OutputHandler> make_output(const foort_run_desc &d, const foort_mesh_desc &m, const Mesh &mesh,
                                           const std::string &prefix, const std::string &ext)
{
    RunConfig rc;
    rc.desc = d;
    rc.mesh = m;
    rc.integrator_name = d.integrator == FOORT_INTEGRATOR_VERLET ? "Verlet" : "RK4";
    const std::vector<uint32_t> order = diagnostic_order(d.diag_flags, d.value_diag);
    std::vector<std::string> names;
    std::vector<char> kinds;
    for (uint32_t f : order) {
        names.push_back(diag_name(f));
        kinds.push_back(f == FOORT_DIAGFLAG_CLOSESTRADIUS ? OutputHandler::kFixed : OutputHandler::kInt);
    }
    auto out = std::make_unique<OutputHandler>(prefix, "", ext, names, ~0UL - 1, ~0UL,
                                               first_line_info(rc, mesh.getFullDescriptionStr()));
    out->setValueKinds(kinds);
    return out;
}

struct Req {
    foort_request req{};
    foort_request_result res{};
    Group *group = nullptr;
    std::unique_ptr<Mesh> mesh;
    uint32_t vd = 0;
    bool finished = false;         // the mesh is done (or the request failed)
    uint32_t slot = UINT32_MAX;
    bool direct = false;           // GeodesicPosition: run whole by a pool thread on a context of its own
    std::vector<uint64_t> input_px; // InputCertainPixelsMesh: the request's (row, column) list, copied
    std::vector<uint64_t> idx;     // the batch to push: (row, column) pairs
    uint64_t n = 0, first = 0;     // its pixels and, once pushed, its ring position
    uint64_t pushed = 0;           // pixels pushed for the slot since it was added
    uint64_t steps_seen = 0;       // the slot's step count at the last look
    std::vector<int32_t> term, quad, passes;
    std::vector<double> closest;
    // output files (when the request names a prefix): the handler and the diagnostics' order
    std::string out_prefix, out_ext;
    std::unique_ptr<OutputHandler> out;
    std::vector<uint32_t> order;
};

struct Group {
    std::string key;
    foort_run_desc desc{};
    foort_gpu_queue *q = nullptr;
    uint64_t capacity = 0;
    std::thread thr;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Req *> events;      // requests with a batch ready, or finished
    bool stop = false;
};

} // namespace
} // namespace foort_host

using namespace foort_host;

struct foort_server {
    foort_server_config cfg{};
    Clock::time_point t0 = Clock::now();
    std::mutex mu;
    std::condition_variable cv_task, cv_done;
    std::deque<std::pair<Req *, bool>> tasks; // (request, true: build its mesh; false: take its values)
    std::deque<foort_request_result> done;
    uint64_t submitted = 0, finished = 0;
    bool stopping = false;  // no new requests
    bool pool_stop = false; // every request finished: the pool may end
    std::vector<std::thread> pool;
    std::map<std::string, std::unique_ptr<Group>> groups;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    void post_task(Req *r, bool build)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            tasks.emplace_back(r, build);
        }
        cv_task.notify_one();
    }

    static void post_event(Req *r)
    {
        Group *g = r->group;
        {
            std::lock_guard<std::mutex> lk(g->mu);
            g->events.push_back(r);
        }
        g->cv.notify_one();
    }

    void complete(Req *r)
    {
        r->res.end_s = now();
        {
            std::lock_guard<std::mutex> lk(mu);
            done.push_back(r->res);
            ++finished;
        }
        cv_done.notify_all();
        delete r;
    }

    // ---- host pool: mesh work ----
    // Take the last batch's values (unless `build`), then move the mesh on to its next
    // non-empty iteration and hand the batch to the group thread (or finish).
    void mesh_step(Req *r, bool build)
    {
        try {
            auto value = [r](uint32_t diag, uint64_t i) -> double {
                return diag == FOORT_DIAGFLAG_FOURCOLORSCREEN    ? r->quad[i]
                       : diag == FOORT_DIAGFLAG_EQUATORIALPASSES ? r->passes[i]
                                                                 : r->closest[i];
            };
            if (build) {
                r->res.start_s = now();
                r->mesh = make_mesh(r->req.mesh);
                r->order = diagnostic_order(r->req.desc.diag_flags, r->req.desc.value_diag);
                r->vd = r->order.front();
                if (!r->out_prefix.empty())
                    r->out = make_output(r->req.desc, r->req.mesh, *r->mesh, r->out_prefix, r->out_ext);
            } else {
                if (r->out)
                    r->out->PrepareForOutput(r->n);
                for (uint64_t i = 0; i < r->n; ++i) {
                    r->mesh->GeodesicFinished(i, MeshValue{value(r->vd, i), 0});
                    if (r->out) {
                        double vals[4];
                        for (size_t k = 0; k < r->order.size(); ++k)
                            vals[k] = value(r->order[k], i);
                        r->out->NewGeodesicValues(i, r->idx[2 * i], r->idx[2 * i + 1], vals);
                    }
                }
                r->mesh->EndCurrentLoop();
                r->res.pixels += r->n;
                ++r->res.iterations;
            }
            for (;;) {
                if (r->mesh->IsFinished()) {
                    if (r->out)
                        r->out->OutputFinished();
                    r->finished = true;
                    break;
                }
                r->n = r->mesh->getCurNrGeodesics();
                if (r->n) {
                    r->idx.resize(2 * r->n);
                    r->mesh->screenIndices(r->idx.data());
                    r->res.footprint_bytes = std::max(r->res.footprint_bytes, r->n * kBytesPerPixel);
                    break;
                }
                r->mesh->EndCurrentLoop(); // an empty iteration
                ++r->res.iterations;
            }
        } catch (const std::exception &e) {
            set_error(e.what());
            r->res.status = FOORT_EINVAL;
            r->finished = true;
        }
        if (r->finished && r->slot == UINT32_MAX)
            complete(r); // the group thread holds nothing of it
        else
            post_event(r);
    }

    // A request with GeodesicPosition, whose saved points the queue kernel does not keep: the
    // pool thread runs it whole, as the program does (FOORT/Main.cpp:313-419), on a context of
    // its own -- each iteration one foort_gpu_trace_positions batch, its values to the mesh and
    // its texts to the output files.
    void run_direct(Req *r)
    {
        r->res.start_s = now();
        foort_gpu_ctx *ctx = nullptr;
        int rc = foort_gpu_open(&r->req.desc, &ctx);
        if (rc) {
            set_error(std::string("render server: ") + foort_gpu_last_error(nullptr));
            r->res.status = rc;
            complete(r);
            return;
        }
        try {
            r->mesh = make_mesh(r->req.mesh);
            r->order = diagnostic_order(r->req.desc.diag_flags, r->req.desc.value_diag);
            r->vd = r->order.front();
            if (!r->out_prefix.empty())
                r->out = make_output(r->req.desc, r->req.mesh, *r->mesh, r->out_prefix, r->out_ext);
            const bool grid = r->req.mesh.type == FOORT_MESH_SIMPLESQUARE;
            while (!r->mesh->IsFinished() && !rc) {
                const uint64_t n = r->mesh->getCurNrGeodesics();
                if (r->out)
                    r->out->PrepareForOutput(n);
                if (n) {
                    r->idx.resize(2 * n);
                    r->mesh->screenIndices(r->idx.data());
                    r->term.assign(n, 0), r->quad.assign(n, 0), r->passes.assign(n, 0), r->closest.assign(n, 0);
                    foort_geodesic_results view{r->term.data(), r->quad.data(), r->passes.data(), r->closest.data(),
                                                nullptr,        nullptr,        nullptr,          nullptr};
                    foort_positions pos{};
                    foort_stats st{};
                    rc = foort_gpu_trace_positions(ctx, r->mesh->rowColumnSize(), n, grid ? nullptr : r->idx.data(),
                                                   &view, &pos, &st);
                    if (rc) {
                        set_error(std::string("render server: ") + foort_gpu_last_error(ctx));
                        break;
                    }
                    Positions p;
                    p.offsets.assign(pos.offsets, pos.offsets + n + 1);
                    p.points.assign(pos.points, pos.points + 4 * pos.total);
                    foort_positions_free(&pos);
                    r->res.steps += st.steps;
                    r->res.kernel_ms += st.device_ms;
                    r->res.pixels += n;
                    r->res.footprint_bytes = std::max(r->res.footprint_bytes, n * kBytesPerPixel + 32 * pos.total);
                    for (uint64_t i = 0; i < n; ++i) {
                        r->mesh->GeodesicFinished(i, r->vd == FOORT_DIAGFLAG_GEODESICPOSITION
                                                         ? p.last_angles(i)
                                                         : MeshValue{r->vd == FOORT_DIAGFLAG_FOURCOLORSCREEN ? r->quad[i]
                                                                     : r->vd == FOORT_DIAGFLAG_EQUATORIALPASSES
                                                                         ? r->passes[i]
                                                                         : r->closest[i],
                                                                     0});
                        if (r->out) {
                            std::vector<std::string> texts;
                            for (uint32_t d : r->order)
                                texts.push_back(d == FOORT_DIAGFLAG_GEODESICPOSITION
                                                    ? p.text(i)
                                                    : diag_value_text(d, r->quad[i], r->passes[i], r->closest[i]));
                            r->out->NewGeodesicOutput(i, r->idx[2 * i], r->idx[2 * i + 1], std::move(texts));
                        }
                    }
                }
                r->mesh->EndCurrentLoop();
                ++r->res.iterations;
            }
            if (!rc && r->out)
                r->out->OutputFinished();
        } catch (const std::exception &e) {
            set_error(e.what());
            rc = FOORT_EINVAL;
        }
        foort_gpu_close(ctx);
        r->res.status = rc;
        complete(r);
    }

    void pool_thread()
    {
        for (;;) {
            std::pair<Req *, bool> t;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_task.wait(lk, [&] { return pool_stop || !tasks.empty(); });
                if (tasks.empty())
                    return;
                t = tasks.front();
                tasks.pop_front();
            }
            if (t.first->direct)
                run_direct(t.first);
            else
                mesh_step(t.first, t.second);
        }
    }

    // ---- group thread: the device queue ----
    void fail_request(Group &g, Req *r, int status)
    {
        r->res.status = status;
        if (r->slot != UINT32_MAX && g.q)
            (void)foort_gpu_queue_remove(g.q, r->slot);
        r->slot = UINT32_MAX;
        complete(r);
    }

    void group_thread(Group &g)
    {
        int broken = foort_gpu_queue_open(&g.desc, g.capacity, kSlotsPerQueue, &g.q);
        if (broken)
            set_error(std::string("render server: ") + foort_gpu_queue_last_error(nullptr));
        std::deque<Req *> waiting;          // batches not yet pushed, in arrival order
        std::vector<Req *> on_device;       // one batch each on the device
        std::vector<uint64_t> done_steps(2 * kSlotsPerQueue);
        for (;;) {
            std::deque<Req *> ev;
            {
                std::unique_lock<std::mutex> lk(g.mu);
                if (on_device.empty())
                    g.cv.wait(lk, [&] { return g.stop || !g.events.empty(); });
                ev.swap(g.events);
                if (g.stop && ev.empty() && waiting.empty() && on_device.empty())
                    break;
            }
            for (Req *r : ev) {
                if (r->finished) {
                    if (r->slot != UINT32_MAX)
                        (void)foort_gpu_queue_remove(g.q, r->slot);
                    r->slot = UINT32_MAX;
                    complete(r);
                } else {
                    waiting.push_back(r);
                }
            }
            if (broken) {
                for (Req *r : waiting)
                    fail_request(g, r, broken);
                waiting.clear();
                for (Req *r : on_device)
                    fail_request(g, r, broken);
                on_device.clear();
                continue;
            }
            // push ready batches in arrival order while slots and ring room last
            while (!waiting.empty()) {
                Req *r = waiting.front();
                if (r->n > g.capacity) {
                    waiting.pop_front();
                    set_error("render server: an iteration batch exceeds the ring");
                    fail_request(g, r, FOORT_ENOMEM);
                    continue;
                }
                if (r->slot == UINT32_MAX) {
                    uint32_t s = 0;
                    const int rc = foort_gpu_queue_add(g.q, &r->req.desc, r->mesh->rowColumnSize(), &s);
                    if (rc == FOORT_ENOMEM)
                        break; // every slot taken: wait for a request to finish
                    if (rc) {
                        waiting.pop_front();
                        set_error(std::string("render server: ") + foort_gpu_queue_last_error(g.q));
                        fail_request(g, r, rc);
                        continue;
                    }
                    r->slot = s;
                    r->pushed = r->steps_seen = 0;
                }
                const int rc = foort_gpu_queue_push(g.q, r->slot, r->n, r->idx.data(), &r->first);
                if (rc == FOORT_ENOMEM)
                    break; // the ring is full: wait for batches to finish
                waiting.pop_front();
                if (rc) {
                    fail_request(g, r, rc);
                    continue;
                }
                r->pushed += r->n;
                on_device.push_back(r);
            }
            if (on_device.empty())
                continue;
            foort_stats st{};
            uint64_t pushed = 0, fin = 0;
            int rc = foort_gpu_queue_run(g.q, cfg.launch_us, &st);
            if (!rc)
                rc = foort_gpu_queue_progress(g.q, done_steps.data(), &pushed, &fin);
            if (rc) {
                set_error(std::string("render server: ") + foort_gpu_queue_last_error(g.q));
                broken = rc;
                continue;
            }
            // the launch's time, shared out by the steps each request finished in it
            uint64_t total = 0;
            for (Req *r : on_device)
                total += done_steps[2 * r->slot + 1] - r->steps_seen;
            for (size_t i = 0; i < on_device.size();) {
                Req *r = on_device[i];
                const uint64_t s = done_steps[2 * r->slot + 1];
                if (total)
                    r->res.kernel_ms += st.device_ms * double(s - r->steps_seen) / double(total);
                r->res.steps += s - r->steps_seen;
                r->steps_seen = s;
                if (done_steps[2 * r->slot] < r->pushed) {
                    ++i;
                    continue;
                }
                r->term.resize(r->n), r->quad.resize(r->n), r->passes.resize(r->n), r->closest.resize(r->n);
                foort_geodesic_results out{r->term.data(), r->quad.data(), r->passes.data(), r->closest.data(),
                                           nullptr,        nullptr,        nullptr,          nullptr};
                rc = foort_gpu_queue_results(g.q, r->first, r->n, &out);
                on_device[i] = on_device.back();
                on_device.pop_back();
                if (rc) {
                    fail_request(g, r, rc);
                    continue;
                }
                post_task(r, false);
            }
        }
        foort_gpu_queue_close(g.q);
    }
};

extern "C" {

void foort_server_config_init(foort_server_config *cfg)
{
    if (!cfg)
        return;
    cfg->workers = 4;
    cfg->launch_us = 2000;
    cfg->hbm_budget_bytes = uint64_t(4) << 30;
}

int foort_server_create(const foort_server_config *cfg, foort_server **srv)
{
    if (!srv) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    auto *s = new foort_server();
    foort_server_config_init(&s->cfg);
    if (cfg) {
        if (cfg->workers)
            s->cfg.workers = cfg->workers;
        if (cfg->launch_us)
            s->cfg.launch_us = cfg->launch_us;
        if (cfg->hbm_budget_bytes)
            s->cfg.hbm_budget_bytes = cfg->hbm_budget_bytes;
    }
    if (s->cfg.workers > 256 || s->cfg.hbm_budget_bytes < 256 * kBytesPerPixel) {
        delete s;
        set_error("at most 256 workers and a budget of at least 256 pixels");
        return FOORT_EINVAL;
    }
    for (uint32_t w = 0; w < s->cfg.workers; ++w)
        s->pool.emplace_back([s] { s->pool_thread(); });
    *srv = s;
    return FOORT_OK;
}

int foort_server_submit(foort_server *srv, const foort_request *req)
{
    if (!srv || !req) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    const bool listed = req->mesh.type == FOORT_MESH_INPUTCERTAINPIXELS;
    if (listed) {
        // the list is checked here as the console dialogue checks each entry (FOORT/Mesh.cpp:81-137)
        const uint64_t side = static_cast<uint64_t>(std::sqrt(static_cast<double>(req->mesh.total_pixels)));
        if (req->mesh.pixel_count && !req->mesh.pixels) {
            set_error("InputCertainPixelsMesh: NULL pixel list");
            return FOORT_EINVAL;
        }
        for (uint64_t i = 0; i < 2 * req->mesh.pixel_count; ++i)
            if (req->mesh.pixels[i] >= side) {
                set_error("InputCertainPixelsMesh: pixel " + std::to_string(i / 2) + " lies outside the " +
                          std::to_string(side) + " x " + std::to_string(side) + " screen");
                return FOORT_EINVAL;
            }
    }
    auto *r = new Req();
    r->req = *req;
    r->direct = (req->desc.diag_flags & FOORT_DIAGFLAG_GEODESICPOSITION) != 0;
    if (listed)
        r->input_px.assign(req->mesh.pixels, req->mesh.pixels + 2 * req->mesh.pixel_count);
    // the request keeps its own copy of a pixel list (the caller's may go away)
    r->req.mesh.pixels = listed ? r->input_px.data() : nullptr;
    r->req.mesh.pixel_count = listed ? req->mesh.pixel_count : 0;
    if (req->output_prefix && *req->output_prefix) {
        r->out_prefix = req->output_prefix;
        r->out_ext = req->output_extension ? req->output_extension : "";
    }
    r->req.output_prefix = r->req.output_extension = nullptr;
    r->res.tag = req->tag;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stopping) {
            delete r;
            set_error("server is shutting down");
            return FOORT_EINVAL;
        }
        r->res.submit_s = srv->now();
        if (r->direct) {
            ++srv->submitted;
            srv->tasks.emplace_back(r, true);
            srv->cv_task.notify_one();
            return FOORT_OK;
        }
        const std::string key = group_key(req->desc);
        std::unique_ptr<Group> &g = srv->groups[key];
        if (!g) {
            g = std::make_unique<Group>();
            g->key = key;
            g->desc = req->desc;
            uint64_t cap = 256;
            while (cap * 2 * kBytesPerPixel <= srv->cfg.hbm_budget_bytes && cap < (uint64_t(1) << 32))
                cap *= 2;
            g->capacity = cap;
            Group *gp = g.get();
            g->thr = std::thread([srv, gp] { srv->group_thread(*gp); });
        }
        r->group = g.get();
        ++srv->submitted;
        srv->tasks.emplace_back(r, true);
    }
    srv->cv_task.notify_one();
    return FOORT_OK;
}

int foort_server_results(foort_server *srv, foort_request_result *out, int cap, double timeout_s)
{
    if (!srv || (cap > 0 && !out)) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    std::unique_lock<std::mutex> lk(srv->mu);
    if (timeout_s > 0)
        srv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return !srv->done.empty(); });
    int n = 0;
    while (n < cap && !srv->done.empty()) {
        out[n++] = srv->done.front();
        srv->done.pop_front();
    }
    return n;
}

int foort_server_drain(foort_server *srv)
{
    if (!srv) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    std::unique_lock<std::mutex> lk(srv->mu);
    srv->cv_done.wait(lk, [&] { return srv->finished == srv->submitted; });
    return FOORT_OK;
}

void foort_server_destroy(foort_server *srv)
{
    if (!srv)
        return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stopping = true; // no new requests; those submitted still finish
    }
    foort_server_drain(srv);
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->pool_stop = true;
    }
    srv->cv_task.notify_all();
    for (auto &t : srv->pool)
        t.join();
    for (auto &kv : srv->groups) {
        Group &g = *kv.second;
        {
            std::lock_guard<std::mutex> lk(g.mu);
            g.stop = true;
        }
        g.cv.notify_all();
        g.thr.join();
    }
    delete srv;
}

} // extern "C"
