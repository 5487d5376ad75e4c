// server.cpp -- render server (no counterpart in the reference): requests of
// any metric, camera and resolution, each an adaptive mesh of a few
// iterations, run concurrently on one GPU.
//
// A fixed set of workers each owns a GPU context (its stream, device staging
// buffers and persistent-lane state stay allocated from request to request;
// foort_gpu_configure switches the run description) and runs one request at a
// time, iteration by iteration: the mesh hands out its pixels, the GPU
// integrates them on the worker's stream, the value diagnostic goes back to the
// mesh.  Requests are admitted in arrival order once the device memory they
// may need -- their largest possible batch of pixel indices and results --
// fits in the budget next to the requests already running.
#include "host.hpp"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

namespace foort_host {
namespace {

using Clock = std::chrono::steady_clock;

// device bytes per pixel of a batch: screen index (16) + termination, quadrant,
// passes (4 each) + closest radius (8)
constexpr uint64_t kBytesPerPixel = 16 + 4 + 4 + 4 + 8;

uint64_t max_batch_pixels(const foort_mesh_desc &m)
{
    switch (m.type) {
    case FOORT_MESH_SIMPLESQUARE: return m.total_pixels;
    case FOORT_MESH_INPUTCERTAINPIXELS: return m.pixel_count;
    default: {
        // the initial grid, or an iteration's queue: at most 5 new pixels per subdivided
        // square, and never more than MaxPixels in all (0 = no limit)
        const uint64_t per_iter = m.iteration_pixels > (uint64_t(1) << 40) ? (uint64_t(1) << 40) : 5 * m.iteration_pixels;
        uint64_t most = std::max<uint64_t>(m.initial_pixels, per_iter);
        if (m.max_pixels)
            most = std::min<uint64_t>(most, std::max(m.max_pixels, m.initial_pixels));
        return most;
    }
    }
}

struct Job {
    foort_request req;
    foort_request_result res{};
    uint64_t need = 0;
};

} // namespace
} // namespace foort_host

using namespace foort_host;

struct foort_server {
    foort_server_config cfg{};
    Clock::time_point t0 = Clock::now();
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job> pending;             // arrival order
    std::deque<foort_request_result> done;
    uint64_t in_flight = 0;              // admitted bytes
    uint64_t running = 0, submitted = 0, finished = 0;
    bool stopping = false;
    std::vector<std::thread> workers;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    void run_job(Job &j, foort_gpu_ctx *&ctx)
    {
        foort_request_result &r = j.res;
        r.start_s = now();
        int rc = FOORT_OK;
        if (!ctx)
            rc = foort_gpu_open(&j.req.desc, &ctx);
        else
            rc = foort_gpu_configure(ctx, &j.req.desc);
        if (rc != FOORT_OK) {
            r.status = rc;
            return;
        }
        try {
            std::unique_ptr<Mesh> mesh = make_mesh(j.req.mesh);
            const uint32_t vd = diagnostic_order(j.req.desc.diag_flags, j.req.desc.value_diag).front();
            std::vector<uint64_t> idx;
            std::vector<int32_t> term, quad, passes;
            std::vector<double> closest, value;
            while (!mesh->IsFinished()) {
                const uint64_t n = mesh->getCurNrGeodesics();
                idx.resize(2 * n);
                term.resize(n), quad.resize(n), passes.resize(n), closest.resize(n), value.resize(n);
                mesh->screenIndices(idx.data());
                foort_geodesic_results out{term.data(), quad.data(), passes.data(), closest.data(),
                                           nullptr,     nullptr,     nullptr,       nullptr};
                foort_stats st{};
                if (n) {
                    rc = foort_gpu_trace_pixels(ctx, mesh->rowColumnSize(), n, idx.data(), &out, &st);
                    if (rc != FOORT_OK)
                        break;
                }
                for (uint64_t i = 0; i < n; ++i)
                    mesh->GeodesicFinished(i, vd == FOORT_DIAGFLAG_FOURCOLORSCREEN   ? quad[i]
                                              : vd == FOORT_DIAGFLAG_EQUATORIALPASSES ? passes[i]
                                                                                      : closest[i]);
                mesh->EndCurrentLoop();
                r.pixels += n;
                r.steps += st.steps;
                r.kernel_ms += st.device_ms;
                ++r.iterations;
            }
        } catch (const std::exception &e) {
            set_error(e.what());
            rc = FOORT_EINVAL;
        }
        r.status = rc;
    }

    void worker()
    {
        foort_gpu_ctx *ctx = nullptr;
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lk(mu);
                // admit the oldest request once its memory fits (always when nothing runs)
                cv_work.wait(lk, [&] {
                    return stopping || (!pending.empty() &&
                                        (running == 0 || in_flight + pending.front().need <= cfg.hbm_budget_bytes));
                });
                if (pending.empty() && stopping)
                    break;
                if (pending.empty())
                    continue;
                j = pending.front();
                pending.pop_front();
                in_flight += j.need;
                ++running;
            }
            cv_work.notify_all();
            run_job(j, ctx);
            j.res.end_s = now();
            {
                std::lock_guard<std::mutex> lk(mu);
                in_flight -= j.need;
                --running;
                ++finished;
                done.push_back(j.res);
            }
            cv_work.notify_all();
            cv_done.notify_all();
        }
        foort_gpu_close(ctx);
    }
};

extern "C" {

void foort_server_config_init(foort_server_config *cfg)
{
    if (!cfg)
        return;
    cfg->workers = 4;
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
        if (cfg->hbm_budget_bytes)
            s->cfg.hbm_budget_bytes = cfg->hbm_budget_bytes;
    }
    if (s->cfg.workers > 64) {
        delete s;
        set_error("at most 64 workers");
        return FOORT_EINVAL;
    }
    for (uint32_t w = 0; w < s->cfg.workers; ++w)
        s->workers.emplace_back([s] { s->worker(); });
    *srv = s;
    return FOORT_OK;
}

int foort_server_submit(foort_server *srv, const foort_request *req)
{
    if (!srv || !req) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    Job j;
    j.req = *req;
    j.req.mesh.pixels = nullptr; // a request's mesh is adaptive or a grid; pixel lists are not kept
    j.req.mesh.pixel_count = 0;
    if (j.req.mesh.type == FOORT_MESH_INPUTCERTAINPIXELS) {
        set_error("the server takes grid and subdivision meshes");
        return FOORT_EINVAL;
    }
    j.need = max_batch_pixels(j.req.mesh) * kBytesPerPixel;
    j.res.tag = req->tag;
    j.res.footprint_bytes = j.need;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stopping) {
            set_error("server is shutting down");
            return FOORT_EINVAL;
        }
        j.res.submit_s = srv->now();
        srv->pending.push_back(j);
        ++srv->submitted;
    }
    srv->cv_work.notify_all();
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
        srv->stopping = true;
    }
    srv->cv_work.notify_all();
    for (auto &t : srv->workers)
        t.join();
    delete srv;
}

} // extern "C"
