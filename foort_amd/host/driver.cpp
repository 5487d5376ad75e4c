// driver.cpp -- the program of FOORT/Main.cpp:158-433 around the GPU batch
// call, and the C ABI of foort_host.h.
//
// Per mesh iteration the reference runs, for every pixel index on an OpenMP
// thread, SetNewInitialConditions -> Reset -> Update until terminated ->
// GeodesicFinished + NewGeodesicOutput (FOORT/Main.cpp:320-414).  Here the
// iteration's screen indices go to the GPU in one call
// (foort_gpu_trace_pixels / foort_gpu_trace_grid); the returned diagnostic
// values then feed the mesh and the output handler in index order, exactly
// the values and texts Geodesic::getDiagnosticFinalValue and getAllOutputStr
// form (FOORT/Geodesic.cpp:111-149).
#include "host.hpp"

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace foort_host {

namespace {
thread_local std::string g_error;

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Add a batch's statistics to the run's, and print the warnings the reference prints per
// pixel in ViewScreen::SetNewInitialConditions (FOORT/ViewScreen.cpp:112-116), once per
// pixel as it does (after the batch rather than during it).
void add_stats(foort_stats &total, const foort_stats &st)
{
    total.geodesics += st.geodesics;
    total.steps += st.steps;
    total.launches += st.launches;
    total.device_ms += st.device_ms;
    total.future_pointing += st.future_pointing;
    total.outward_pointing += st.outward_pointing;
    total.nonfinite += st.nonfinite;
    total.graph_launches += st.graph_launches;
    for (uint64_t k = 0; k < st.future_pointing; ++k)
        screen_output("Initial velocity of geodesic is future-pointing (should be past-pointing)!", Level::Warning);
    for (uint64_t k = 0; k < st.outward_pointing; ++k)
        screen_output("Initial velocity of geodesic is outward-pointing (should be inward-pointing)!", Level::Warning);
}

// the value and text of one diagnostic for geodesic i
struct BatchResults {
    std::vector<int32_t> term, quadrant, passes;
    std::vector<double> closest;
    explicit BatchResults(uint64_t n) : term(n), quadrant(n), passes(n), closest(n) {}
    foort_geodesic_results view()
    {
        return foort_geodesic_results{term.data(), quadrant.data(), passes.data(), closest.data(),
                                      nullptr,     nullptr,         nullptr,       nullptr};
    }
    double value(uint32_t diag, uint64_t i) const
    {
        switch (diag) {
        case FOORT_DIAGFLAG_FOURCOLORSCREEN: return static_cast<double>(quadrant[i]);
        case FOORT_DIAGFLAG_EQUATORIALPASSES: return static_cast<double>(passes[i]);
        default: return closest[i];
        }
    }
    std::string text(uint32_t diag, uint64_t i) const
    {
        return diag_value_text(diag, quadrant[i], passes[i], closest[i]);
    }
};

int copy_out(const std::string &s, char *buf, size_t cap)
{
    if (buf && cap) {
        size_t n = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return static_cast<int>(s.size());
}

struct ConfigError {
    int code;
};

// Autotuner results are kept in a small text file (FOORT_TUNE_CACHE, else
// ~/.cache/foort_tune.txt), one line per configuration: the key (what the kernels
// depend on, and the tuning mode) and the chosen schedule, block size, refill
// interval and refill threshold.  A later run of the same configuration reuses them.
std::string tune_cache_path()
{
    if (const char *p = std::getenv("FOORT_TUNE_CACHE"))
        return p;
    if (const char *h = std::getenv("HOME"))
        return std::string(h) + "/.cache/foort_tune.txt";
    return "";
}

std::string tune_key(const foort_run_desc &d, int mode)
{
    std::ostringstream k;
    k << "metric=" << d.metric << " rlog=" << d.rlog_scale << " integ=" << d.integrator << " chris=" << d.christoffel
      << " prec=" << d.precision << " terms=" << d.term_flags << " diags=" << d.diag_flags << " mode=" << mode
      << " sched=" << (mode == 2 ? -1 : d.schedule);
    return k.str();
}

bool tune_lookup(const std::string &key, foort_run_desc &d)
{
    std::ifstream f(tune_cache_path());
    std::string line;
    while (std::getline(f, line)) {
        const size_t bar = line.find(" | ");
        if (bar == std::string::npos || line.compare(0, bar, key) != 0 || bar != key.size())
            continue;
        std::istringstream v(line.substr(bar + 3));
        int sched, block;
        unsigned interval, refill;
        if (v >> sched >> block >> interval >> refill) {
            d.schedule = sched;
            d.block_size = block;
            d.refill_interval = interval;
            d.refill_lanes = refill;
            return true;
        }
    }
    return false;
}

void tune_store(const std::string &key, const foort_run_desc &d)
{
    const std::string path = tune_cache_path();
    if (path.empty())
        return;
    std::error_code ec;
    std::filesystem::create_directories(std::filesystem::path(path).parent_path(), ec);
    std::ofstream f(path, std::ios::app);
    if (f)
        f << key << " | " << d.schedule << " " << d.block_size << " " << d.refill_interval << " " << d.refill_lanes
          << "\n";
}

// the program body; throws ConfigError for the reference's early exits
int run(const char *cfg_path, const foort_program_options &o)
{
    set_output_level(4);
    screen_output("FOORT (gfx950 GPU path) in configuration mode.", Level::Proc);
    if (!cfg_path) {
        screen_output("No configuration file given. Exiting...\n", Level::Warning);
        throw ConfigError{FOORT_EINVAL};
    }
    std::unique_ptr<Setting> root;
    try {
        root = parse_config_file(cfg_path);
    } catch (const ParseError &e) {
        if (e.line == 0)
            std::cerr << "Config file not found at " << cfg_path << ".\n Exiting...\n";
        else
            std::cerr << "Unable to parse config file correctly: parse error at " << cfg_path << ":" << e.line << " - "
                      << e.message << ".\n"
                      << "Remember that all numbers must be given as values, e.g. \"3.14/2.0\" is now allowed.\n"
                         "Exiting...\n";
        set_error(e.line == 0 ? "configuration file not found" : "configuration parse error: " + e.message);
        throw ConfigError{FOORT_EINVAL};
    }
    screen_output("Initializing all object using configuration file...", Level::Proc);
    RunConfig rc;
    read_run_config(*root, rc, true);
    if (o.christoffel >= 0)
        rc.desc.christoffel = o.christoffel;
    if (o.precision >= 0)
        rc.desc.precision = o.precision;
    if (o.schedule >= 0)
        rc.desc.schedule = o.schedule;
    if (o.screen_output_level >= 0)
        set_output_level(o.screen_output_level);
    if (o.output_prefix) {
        rc.prefix = o.output_prefix;
        rc.to_files = !rc.prefix.empty();
    }

    std::unique_ptr<Mesh> mesh = make_mesh(rc.mesh);
    const std::string first_line = first_line_info(rc, mesh->getFullDescriptionStr());
    const std::vector<uint32_t> order = diagnostic_order(rc.desc.diag_flags, rc.desc.value_diag);
    std::vector<std::string> names;
    for (uint32_t f : order)
        names.push_back(diag_name(f));
    OutputHandler out(rc.to_files ? rc.prefix : "", rc.timestamp, rc.extension, names,
                      rc.to_files ? rc.to_cache : ~0UL - 1, rc.to_files ? rc.per_file : ~0UL,
                      rc.to_files && rc.first_line_on ? first_line : "");
    {
        // the diagnostics' texts: std::to_string of the quadrant and of the passes (ints) and of the
        // closest radius (a double), FOORT/Diagnostics.cpp:204-208, :383-387, :449-453
        std::vector<char> kinds;
        for (uint32_t f : order)
            kinds.push_back(f == FOORT_DIAGFLAG_CLOSESTRADIUS ? OutputHandler::kFixed : OutputHandler::kInt);
        out.setValueKinds(kinds);
    }
    screen_output("Done loading options from configuration file.", Level::Proc);

    // the objects, as FOORT/Main.cpp:262-299 lists them
    const Level L = Level::Subproc;
    screen_output("\n--------------------------------", L);
    screen_output("LIST OF ALL INITIALIZED OBJECTS:", L);
    screen_output("Metric: " + metric_description(rc.desc) + ".", L);
    screen_output("Geodesic source: No source.", L);
    screen_output("Diagnostics turned on: ", L);
    screen_output("<begin list>", L);
    for (uint32_t f : order)
        screen_output(diag_description(f, rc) + ".", L);
    screen_output("<end list>", L);
    screen_output("Terminations turned on:", L);
    screen_output("<begin list>", L);
    std::string terms = termination_descriptions(rc);
    for (size_t p = 0; (p = terms.find(", ")) != std::string::npos;) {
        screen_output(terms.substr(0, p) + ".", L);
        terms.erase(0, p + 2);
    }
    screen_output("<end list>", L);
    screen_output(viewscreen_description(rc, mesh->getFullDescriptionStr()) + ".", L);
    screen_output(integrator_description(rc), L);
    screen_output(out.getFullDescriptionStr(), L);
    screen_output("--------------------------------\n", L);

    const bool gp = (rc.desc.diag_flags & FOORT_DIAGFLAG_GEODESICPOSITION) != 0;
    if (gp && o.tracer && !o.positions_tracer) {
        set_error("a run with GeodesicPosition needs a positions tracer next to the batch tracer");
        screen_output(g_error, Level::Warning);
        return FOORT_EINVAL;
    }

    foort_gpu_ctx *gpu = nullptr;
    if (!o.tracer) {
        int rcode = foort_gpu_open(&rc.desc, &gpu);
        if (rcode != FOORT_OK) {
            set_error(std::string("GPU path: ") + foort_gpu_last_error(nullptr));
            screen_output(g_error, Level::Warning);
            return rcode;
        }
    }
    struct Closer {
        foort_gpu_ctx *c;
        ~Closer() { foort_gpu_close(c); }
    } closer{gpu};

    const bool grid = rc.mesh.type == FOORT_MESH_SIMPLESQUARE;
    int tuned_class = -1; // the largest batch size (log2) tuned for so far
    uint64_t iteration = 0;
    foort_stats total{};
    // wall time of the phases of the loop: screen indices, integration, results into the mesh and
    // the output cache, mesh refinement (EndCurrentLoop), writing the files
    double t_index = 0, t_trace = 0, t_results = 0, t_refine = 0, t_write = 0;
    auto lap = [](std::chrono::steady_clock::time_point &t) {
        auto now = std::chrono::steady_clock::now();
        double d = std::chrono::duration<double>(now - t).count();
        t = now;
        return d;
    };
    auto t_total = std::chrono::steady_clock::now();
    while (!mesh->IsFinished()) {
        screen_output("Starting new integration loop.", Level::Proc);
        auto t_iter = std::chrono::steady_clock::now();
        auto t = t_iter;
        const uint64_t n = mesh->getCurNrGeodesics();
        screen_output("Integrating " + std::to_string(static_cast<long long>(n)) + " geodesics on the GPU...",
                      Level::Proc);
        out.PrepareForOutput(n);
        t_write += lap(t);
        std::vector<uint64_t> idx(2 * n);
        mesh->screenIndices(idx.data());
        BatchResults res(n);
        t_index += lap(t);
        foort_geodesic_results view = res.view();
        const uint64_t side = mesh->rowColumnSize();
        int rcode = FOORT_OK;
        Positions positions;
        if (n > 0 && gp) {
            if (o.positions_tracer) {
                positions.offsets.assign(n + 1, 0);
                rcode = o.positions_tracer(o.tracer_user, &rc.desc, side, n, idx.data(), &view,
                                           positions.offsets.data(), positions.points.data(),
                                           positions.points.size() / 4);
                if (!rcode && positions.offsets[n] > positions.points.size() / 4) {
                    positions.points.resize(4 * positions.offsets[n]);
                    rcode = o.positions_tracer(o.tracer_user, &rc.desc, side, n, idx.data(), &view,
                                               positions.offsets.data(), positions.points.data(),
                                               positions.points.size() / 4);
                }
                if (rcode)
                    set_error("positions tracer failed with code " + std::to_string(rcode));
            } else {
                foort_positions pos{};
                foort_stats st{};
                rcode = foort_gpu_trace_positions(gpu, side, n, grid ? nullptr : idx.data(), &view, &pos, &st);
                if (rcode) {
                    set_error(std::string("GPU path: ") + foort_gpu_last_error(gpu));
                } else {
                    positions.offsets.assign(pos.offsets, pos.offsets + n + 1);
                    positions.points.assign(pos.points, pos.points + 4 * pos.total);
                }
                foort_positions_free(&pos);
                add_stats(total, st);
            }
        } else if (n > 0) {
            if (o.tracer) {
                rcode = o.tracer(o.tracer_user, &rc.desc, side, n, idx.data(), &view);
                if (rcode)
                    set_error("batch tracer failed with code " + std::to_string(rcode));
            } else {
                // Tuning by batch size: geodesic lengths grow more uneven as an adaptive mesh
                // refines (the later, larger iterations hug the image's edges), so the options
                // are chosen again whenever a batch of at least 65536 pixels reaches a larger
                // power of two than any before (the uniform initial grid of an adaptive mesh
                // excepted).  The choice per size is kept in the tune cache; a new one is timed
                // on the whole batch (at most its middle 2^21 pixels): a part of a batch does not
                // show its tail of long geodesics, which decides between the schedules.
                const bool adaptive = rc.mesh.type == FOORT_MESH_SQUARESUBDIVISION ||
                                      rc.mesh.type == FOORT_MESH_SQUARESUBDIVISIONV2;
                const int size_class = n >= 65536 ? 63 - __builtin_clzll(n) : -1;
                if (rc.autotune && size_class > tuned_class && (!adaptive || iteration > 0)) {
                    tuned_class = size_class;
                    const std::string key =
                        tune_key(rc.desc, rc.autotune) + " size=2^" + std::to_string(size_class);
                    foort_run_desc chosen = rc.desc;
                    if (tune_lookup(key, chosen)) {
                        if (foort_gpu_configure(gpu, &chosen) == FOORT_OK)
                            screen_output("Autotuner: options of an earlier run for batches of 2^" +
                                              std::to_string(size_class) + " pixels (" + tune_cache_path() + ").",
                                          Level::Subproc);
                    } else {
                        const uint64_t w = std::min<uint64_t>(n, uint64_t(1) << 21), a = (n - w) / 2;
                        foort_stats tst{};
                        if (foort_gpu_autotune(gpu, side, w, idx.data() + 2 * a,
                                               rc.autotune == 2 ? FOORT_TUNE_SCHEDULE : 0u, &chosen, &tst) ==
                            FOORT_OK) {
                            tune_store(key, chosen);
                            screen_output(
                                "Autotuner: " +
                                    std::string(chosen.schedule == FOORT_SCHEDULE_ONESHOT
                                                    ? "one-shot schedule, block size " +
                                                          std::to_string(chosen.block_size)
                                                    : "persistent schedule, refill every " +
                                                          std::to_string(chosen.refill_interval) + " steps once " +
                                                          std::to_string(chosen.refill_lanes) + " lanes are idle") +
                                    " (" + std::to_string(tst.device_ms) + " ms on " + std::to_string(w) +
                                    " pixels of a batch of " + std::to_string(n) + ").",
                                Level::Subproc);
                        } else
                            screen_output(std::string("Autotuner failed: ") + foort_gpu_last_error(gpu),
                                          Level::Warning);
                    }
                }
                // With Output.LoopMessageFrequency set, the iteration is integrated in chunks of
                // about that many geodesics (at least 65,536, so the GPU stays full) and the
                // reference's progress line follows each chunk (FOORT/Main.cpp:354-371); each
                // geodesic's result does not depend on the chunks.
                const unsigned long freq = loop_message_frequency();
                const uint64_t chunk =
                    freq == ~0UL ? n : std::max<uint64_t>(uint64_t(1) << 16, (freq + 255) / 256 * 256);
                auto t_loop = std::chrono::steady_clock::now();
                for (uint64_t a = 0; a < n && rcode == FOORT_OK; a += chunk) {
                    const uint64_t c = std::min(chunk, n - a);
                    foort_geodesic_results part = view;
                    part.term += a, part.quadrant += a, part.equat_passes += a, part.closest_radius += a;
                    foort_stats st{};
                    rcode = grid ? foort_gpu_trace_grid(gpu, side, a, c, &part, &st)
                                 : foort_gpu_trace_pixels(gpu, side, c, idx.data() + 2 * a, &part, &st);
                    if (rcode)
                        set_error(std::string("GPU path: ") + foort_gpu_last_error(gpu));
                    add_stats(total, st);
                    if (!rcode && a + c < n) {
                        const double el = seconds_since(t_loop), speed = double(a + c) / el;
                        screen_output("Approx. at geodesic " + std::to_string(static_cast<long long>(a + c)) + " (" +
                                          std::to_string(el) + "s elapsed; speed: " +
                                          std::to_string(static_cast<long>(speed)) +
                                          " geod/s; est. loop time remaining: " +
                                          std::to_string(double(n - a - c) / speed) + "s)...",
                                      Level::Subproc);
                    }
                }
            }
        }
        if (rcode != FOORT_OK) {
            screen_output(g_error, Level::Warning);
            return rcode;
        }
        t_trace += lap(t);
        const uint32_t vd = order.front();
        // GeodesicFinished and NewGeodesicOutput write the slot of their own index only
        // (as in the reference's parallel loop), so the geodesics are handed over in parallel
        if (!gp) {
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < static_cast<long long>(n); ++i) {
                const uint64_t u = static_cast<uint64_t>(i);
                mesh->GeodesicFinished(u, MeshValue{res.value(vd, u), 0});
                double vals[4];
                for (size_t k = 0; k < order.size(); ++k)
                    vals[k] = res.value(order[k], u);
                out.NewGeodesicValues(u, idx[2 * i], idx[2 * i + 1], vals);
            }
        } else {
            // the texts (long ones for GeodesicPosition) are formed in parallel and handed over in order
            std::vector<std::vector<std::string>> texts(n);
#pragma omp parallel for schedule(dynamic, 256)
            for (long long i = 0; i < static_cast<long long>(n); ++i) {
                const uint64_t u = static_cast<uint64_t>(i);
                mesh->GeodesicFinished(u, vd == FOORT_DIAGFLAG_GEODESICPOSITION ? positions.last_angles(u)
                                                                                : MeshValue{res.value(vd, u), 0});
                for (uint32_t d : order)
                    texts[u].push_back(d == FOORT_DIAGFLAG_GEODESICPOSITION ? positions.text(u) : res.text(d, u));
            }
            for (uint64_t u = 0; u < n; ++u)
                out.NewGeodesicOutput(u, idx[2 * u], idx[2 * u + 1], std::move(texts[u]));
        }
        t_results += lap(t);
        const double it = seconds_since(t_iter), tt = seconds_since(t_total);
        screen_output("Integration loop done. Time taken for integration loop: " + std::to_string(it) + "s (" +
                          std::to_string(it / 60) + "m); total time elapsed: " + std::to_string(tt) + "s (" +
                          std::to_string(tt / 60) + "m).",
                      Level::Proc);
        t = std::chrono::steady_clock::now();
        mesh->EndCurrentLoop();
        t_refine += lap(t);
        if (!mesh->IsFinished())
            out.FormatAhead(); // this iteration's lines, formatted in the background while the next one runs
        ++iteration;
    }
    const double tt = seconds_since(t_total);
    screen_output("All integration finished! Total time elapsed: " + std::to_string(tt) + "s (" +
                      std::to_string(tt / 60) + "m).",
                  Level::Proc);
    auto t = std::chrono::steady_clock::now();
    out.OutputFinished();
    t_write += lap(t);
    screen_output("Time by phase: screen indices " + std::to_string(t_index) + "s, integration " +
                      std::to_string(t_trace) + "s (kernels " + std::to_string(total.device_ms / 1e3) +
                      "s), results to mesh and output " + std::to_string(t_results) + "s, mesh refinement " +
                      std::to_string(t_refine) + "s, writing files " + std::to_string(t_write) + "s.",
                  Level::Subproc);
    screen_output("FOORT finished. Goodbye!", Level::Proc);
    std::cout.flush();
    if (o.stats)
        *o.stats = total;
    return FOORT_OK;
}

template <class F>
int guarded(F &&f)
{
    try {
        return f();
    } catch (const ConfigError &e) {
        return e.code;
    } catch (const ParseError &e) {
        set_error("configuration parse error at line " + std::to_string(e.line) + ": " + e.message);
        return FOORT_EINVAL;
    } catch (const std::bad_alloc &) {
        set_error("out of host memory");
        return FOORT_ENOMEM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return FOORT_EINVAL;
    }
}
} // namespace

std::string set_error(const std::string &msg)
{
    g_error = msg;
    return g_error;
}

} // namespace foort_host

using namespace foort_host;

struct foort_mesh {
    std::unique_ptr<Mesh> m;
};

struct foort_config {
    RunConfig rc;
};

struct foort_output {
    std::unique_ptr<OutputHandler> h;
};

extern "C" {

const char *foort_host_last_error(void) { return g_error.c_str(); }

int foort_mesh_create(const foort_mesh_desc *desc, foort_mesh **mesh)
{
    if (!desc || !mesh) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    return guarded([&] {
        if (desc->type == FOORT_MESH_SQUARESUBDIVISION || desc->type == FOORT_MESH_SQUARESUBDIVISIONV2) {
            if (desc->max_subdivide < 1 || desc->max_subdivide > 40 || desc->initial_pixels < 4)
                throw std::runtime_error("subdivision mesh needs max_subdivide in 1..40 and at least 2x2 pixels");
        } else if (desc->total_pixels < 4) {
            throw std::runtime_error("a mesh needs at least 2x2 pixels");
        }
        auto *m = new foort_mesh{make_mesh(*desc)};
        *mesh = m;
        return FOORT_OK;
    });
}

void foort_mesh_destroy(foort_mesh *mesh) { delete mesh; }

uint64_t foort_mesh_count(const foort_mesh *mesh) { return mesh ? mesh->m->getCurNrGeodesics() : 0; }

uint64_t foort_mesh_row_column_size(const foort_mesh *mesh) { return mesh ? mesh->m->rowColumnSize() : 0; }

int foort_mesh_screen_indices(const foort_mesh *mesh, uint64_t *screen_index)
{
    if (!mesh || !screen_index) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    mesh->m->screenIndices(screen_index);
    return FOORT_OK;
}

int foort_mesh_finished(foort_mesh *mesh, const double *values)
{
    if (!mesh || (!values && mesh->m->getCurNrGeodesics())) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    const uint64_t n = mesh->m->getCurNrGeodesics();
    for (uint64_t i = 0; i < n; ++i)
        mesh->m->GeodesicFinished(i, MeshValue{values[i], 0});
    return FOORT_OK;
}

int foort_mesh_finished_pairs(foort_mesh *mesh, const double *pairs)
{
    if (!mesh || (!pairs && mesh->m->getCurNrGeodesics())) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    const uint64_t n = mesh->m->getCurNrGeodesics();
    for (uint64_t i = 0; i < n; ++i)
        mesh->m->GeodesicFinished(i, MeshValue{pairs[2 * i], pairs[2 * i + 1]});
    return FOORT_OK;
}

int foort_mesh_end_loop(foort_mesh *mesh)
{
    if (!mesh) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    return guarded([&] {
        mesh->m->EndCurrentLoop();
        return FOORT_OK;
    });
}

int foort_mesh_is_finished(const foort_mesh *mesh) { return mesh ? mesh->m->IsFinished() : 1; }

int foort_mesh_description(const foort_mesh *mesh, char *buf, size_t cap)
{
    if (!mesh) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    return copy_out(mesh->m->getFullDescriptionStr(), buf, cap);
}

int foort_config_read(const char *path, const char *text, foort_config **cfg)
{
    if (!cfg || (!path && !text)) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    return guarded([&] {
        std::unique_ptr<Setting> root = path ? parse_config_file(path) : parse_config(text);
        auto *c = new foort_config{};
        read_run_config(*root, c->rc, false);
        *cfg = c;
        return FOORT_OK;
    });
}

void foort_config_free(foort_config *cfg) { delete cfg; }

int foort_config_run_desc(const foort_config *cfg, foort_run_desc *desc)
{
    if (!cfg || !desc) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    *desc = cfg->rc.desc;
    return FOORT_OK;
}

int foort_config_mesh_desc(const foort_config *cfg, foort_mesh_desc *desc)
{
    if (!cfg || !desc) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    *desc = cfg->rc.mesh;
    return FOORT_OK;
}

int foort_config_first_line(const foort_config *cfg, char *buf, size_t cap)
{
    if (!cfg) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    return guarded([&] {
        std::unique_ptr<Mesh> m = make_mesh(cfg->rc.mesh);
        return copy_out(first_line_info(cfg->rc, m->getFullDescriptionStr()), buf, cap);
    });
}

int foort_config_diag_names(const foort_config *cfg, char *buf, size_t cap)
{
    if (!cfg) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    std::string s;
    for (uint32_t f : diagnostic_order(cfg->rc.desc.diag_flags, cfg->rc.desc.value_diag))
        s += (s.empty() ? "" : " ") + diag_name(f);
    return copy_out(s, buf, cap);
}

int foort_output_create(const char *prefix, const char *timestamp, const char *extension,
                        const char *const *diag_names, int n_diags, uint64_t to_cache, uint64_t per_file,
                        const char *first_line, foort_output **out)
{
    if (!out || n_diags < 0 || (n_diags && !diag_names)) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    return guarded([&] {
        std::vector<std::string> names(diag_names, diag_names + n_diags);
        *out = new foort_output{std::make_unique<OutputHandler>(prefix ? prefix : "", timestamp ? timestamp : "",
                                                                extension ? extension : "", names, to_cache,
                                                                per_file == 0 ? ~0UL : per_file,
                                                                first_line ? first_line : "")};
        return FOORT_OK;
    });
}

void foort_output_destroy(foort_output *out) { delete out; }

int foort_output_prepare(foort_output *out, uint64_t count)
{
    if (!out) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    return guarded([&] {
        out->h->PrepareForOutput(count);
        return FOORT_OK;
    });
}

int foort_output_geodesic(foort_output *out, uint64_t index, uint64_t row, uint64_t column, const char *const *texts)
{
    if (!out) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    return guarded([&] {
        std::vector<std::string> t;
        for (int k = 0; texts && texts[k]; ++k)
            t.emplace_back(texts[k]);
        out->h->NewGeodesicOutput(index, row, column, std::move(t));
        return FOORT_OK;
    });
}

int foort_output_finished(foort_output *out)
{
    if (!out) {
        set_error("NULL argument");
        return FOORT_EINVAL;
    }
    return guarded([&] {
        out->h->OutputFinished();
        return FOORT_OK;
    });
}

void foort_program_options_init(foort_program_options *opts)
{
    if (!opts)
        return;
    std::memset(opts, 0, sizeof *opts);
    opts->christoffel = -1;
    opts->precision = -1;
    opts->schedule = -1;
    opts->screen_output_level = -1;
}

int foort_program_run(const char *cfg_path, const foort_program_options *opts)
{
    foort_program_options o;
    foort_program_options_init(&o);
    if (opts)
        o = *opts;
    return guarded([&] { return run(cfg_path, o); });
}

} // extern "C"
