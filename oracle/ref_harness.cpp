// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.  Driver around the reference's
// own sources (FOORT/Metric.cpp, Integrators.cpp, Geodesic.cpp,
// Terminations.cpp, Diagnostics.cpp, ViewScreen.cpp, Mesh.cpp, InputOutput.cpp,
// Utilities.cpp, compiled where they lie under the reference tree by
// oracle/Makefile).  It plays the part of FOORT/Main.cpp + FOORT/Config.cpp:
// it builds the reference objects from a foort_run_desc the way
// Config::GetMetric / InitializeDiagnostics / InitializeTerminations /
// GetViewScreen / GetGeodesicIntegrator do (FOORT/Config.cpp:123-1071) and
// runs the loop of FOORT/Main.cpp:373-400 for a list of pixels.  Config.cpp
// itself needs libconfig++, which is not installed, so the static option
// members it defines (FOORT/Config.cpp:11-28) are defined here.
#include "Geodesic.h"
#include "InputOutput.h"
#include "Integrators.h"
#include "Metric.h"
#include "Mesh.h"
#include "Utilities.h"
#include "ViewScreen.h"

#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "foort_gpu.h"

std::unique_ptr<GeodesicPositionOptions> GeodesicPositionDiagnostic::DiagOptions;
std::unique_ptr<EquatorialPassesOptions> EquatorialPassesDiagnostic::DiagOptions;
std::unique_ptr<ClosestRadiusOptions> ClosestRadiusDiagnostic::DiagOptions;
std::unique_ptr<HorizonTermOptions> HorizonTermination::TermOptions;
std::unique_ptr<BoundarySphereTermOptions> BoundarySphereTermination::TermOptions;
std::unique_ptr<TimeOutTermOptions> TimeOutTermination::TermOptions;
std::unique_ptr<ThetaSingularityTermOptions> ThetaSingularityTermination::TermOptions;

namespace {

std::unique_ptr<Metric> make_metric(const foort_run_desc *d)
{
    const double *P = d->metric_params;
    bool rl = d->rlog_scale != 0;
    switch (d->metric) {
    case FOORT_METRIC_KERR: return std::unique_ptr<Metric>(new KerrMetric(P[0], rl));
    case FOORT_METRIC_FLATSPACE: return std::unique_ptr<Metric>(new FlatSpaceMetric());
    case FOORT_METRIC_RASHEEDLARSEN: return std::unique_ptr<Metric>(new RasheedLarsenMetric(P[0], P[1], P[2], P[3], rl));
    case FOORT_METRIC_JOHANNSEN: return std::unique_ptr<Metric>(new JohannsenMetric(P[0], P[1], P[2], P[3], P[4], rl));
    default: return std::unique_ptr<Metric>(new MankoNovikovMetric(P[0], P[1], rl));
    }
}

// A Mesh that hands out a fixed list of screen indices, scaled like the
// subdivision meshes (FOORT/Mesh.cpp:199-206).
class ListMesh final : public Mesh
{
public:
    ListMesh(pixelcoord side, std::vector<ScreenIndex> px, DiagBitflag valdiag)
        : Mesh(valdiag), m_side{side}, m_px{std::move(px)} {}
    largecounter getCurNrGeodesics() const final { return m_px.size(); }
    void getNewInitConds(largecounter index, ScreenPoint &u, ScreenIndex &s) const final
    {
        s = m_px[index];
        u = ScreenPoint{s[0] * 1.0 / static_cast<real>(m_side - 1), s[1] * 1.0 / static_cast<real>(m_side - 1)};
    }
    void GeodesicFinished(largecounter, std::vector<real>) final {}
    void EndCurrentLoop() final {}
    bool IsFinished() const final { return false; }

private:
    pixelcoord m_side;
    std::vector<ScreenIndex> m_px;
};

struct Run {
    std::unique_ptr<Metric> metric;
    std::unique_ptr<Source> source;
    std::unique_ptr<ViewScreen> view;
    DiagBitflag diags{}, valdiag{};
    TermBitflag terms{};
    GeodesicIntegratorFunc integ{};
};

void setup(Run &R, const foort_run_desc *d, std::unique_ptr<Mesh> mesh)
{
    SetOutputLevel(OutputLevel::Level_0_WARNING);
    R.metric = make_metric(d);
    R.source = std::unique_ptr<Source>(new NoSource(R.metric.get()));
    const auto *sph = dynamic_cast<const SphericalHorizonMetric *>(R.metric.get());
    bool rl = sph && sph->getrLogScale();
    R.diags = static_cast<DiagBitflag>(d->diag_flags & (Diag_FourColorScreen | Diag_EquatorialPasses | Diag_ClosestRadius |
                                                        Diag_GeodesicPosition));
    R.valdiag = static_cast<DiagBitflag>(d->value_diag);
    GeodesicPositionDiagnostic::DiagOptions = std::make_unique<GeodesicPositionOptions>(
        d->gp_output_steps,
        UpdateFrequency{d->gp_update.update_n_steps, d->gp_update.update_start != 0, d->gp_update.update_finish != 0});
    EquatorialPassesDiagnostic::DiagOptions = std::make_unique<EquatorialPassesOptions>(
        d->ep_threshold, UpdateFrequency{d->ep_update.update_n_steps, d->ep_update.update_start != 0, d->ep_update.update_finish != 0});
    ClosestRadiusDiagnostic::DiagOptions = std::make_unique<ClosestRadiusOptions>(
        rl, UpdateFrequency{d->cr_update.update_n_steps, d->cr_update.update_start != 0, d->cr_update.update_finish != 0});
    R.terms = static_cast<TermBitflag>(d->term_flags);
    if (!sph)
        R.terms &= ~Term_Horizon;
    if (sph)
        HorizonTermination::TermOptions = std::make_unique<HorizonTermOptions>(
            sph->getHorizonRadius(), sph->getrLogScale(), d->horizon_epsilon, d->horizon_update_n);
    BoundarySphereTermination::TermOptions = std::make_unique<BoundarySphereTermOptions>(d->sphere_radius, rl, d->sphere_update_n);
    TimeOutTermination::TermOptions = std::make_unique<TimeOutTermOptions>(d->timeout_max_steps, d->timeout_update_n);
    ThetaSingularityTermination::TermOptions = std::make_unique<ThetaSingularityTermOptions>(d->theta_epsilon, d->theta_update_n);
    Integrators::epsilon = d->step_size;
    Integrators::Derivative_hval = d->derivative_h;
    Integrators::SmallestPossibleStepsize = d->smallest_step;
    Integrators::VerletVelocityTolerance = d->verlet_tolerance;
    R.integ = d->integrator == FOORT_INTEGRATOR_VERLET ? Integrators::IntegrateGeodesicStep_Verlet
                                                       : Integrators::IntegrateGeodesicStep_RK4;
    Point pos{d->camera_position[0], d->camera_position[1], d->camera_position[2], d->camera_position[3]};
    R.view = std::unique_ptr<ViewScreen>(new ViewScreen(pos, OneIndex{0, -1, 0, 0},
        ScreenPoint{d->screen_size[0], d->screen_size[1]}, ScreenPoint{d->screen_center[0], d->screen_center[1]},
        std::move(mesh), R.metric.get()));
}

// The reference's own meshes, by FOORT_MESH_* code (InputCertainPixelsMesh runs its
// console dialogue on this process's standard input and output).
std::unique_ptr<Mesh> make_ref_mesh(int type, uint64_t pixels, uint64_t max_pixels, uint64_t iteration, int max_sub,
                                    int to_final, DiagBitflag val)
{
    switch (type) {
    case 0: return std::unique_ptr<Mesh>(new SimpleSquareMesh(pixels, val));
    case 1: return std::unique_ptr<Mesh>(new InputCertainPixelsMesh(pixels, val));
    case 2:
        return std::unique_ptr<Mesh>(new SquareSubdivisionMesh(max_pixels, pixels, max_sub, iteration, to_final != 0, val));
    case 3:
        return std::unique_ptr<Mesh>(new SquareSubdivisionMeshV2(max_pixels, pixels, max_sub, iteration, to_final != 0, val));
    default: return nullptr;
    }
}

void copy_text(const std::string &s, char *buf, size_t cap)
{
    if (!buf || !cap)
        return;
    size_t n = std::min(cap - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
}

} // namespace

extern "C" {

// Drive a reference mesh with values read from a grid: every iteration's queue of
// screen indices is appended to out_idx as (row, col) pairs and each geodesic
// reports value_grid[row * grid_side + col] to Mesh::GeodesicFinished.  Returns
// the number of pixels handed out, -1 if out_idx or iter_counts is too small,
// -2 if a screen index falls outside the grid.
long long ref_mesh_run(int type, uint64_t pixels, uint64_t max_pixels, uint64_t iteration, int max_sub, int to_final,
                       uint32_t valdiag, const double *value_grid, uint64_t grid_side, uint64_t *out_idx, uint64_t cap,
                       uint64_t *iter_counts, int iter_cap, int *n_iters, char *desc, size_t desc_cap)
{
    SetOutputLevel(OutputLevel::Level_0_WARNING);
    EquatorialPassesDiagnostic::DiagOptions = std::make_unique<EquatorialPassesOptions>(0.01, UpdateFrequency{1, true, true});
    ClosestRadiusDiagnostic::DiagOptions = std::make_unique<ClosestRadiusOptions>(false, UpdateFrequency{1, true, true});
    std::unique_ptr<Mesh> mesh =
        make_ref_mesh(type, pixels, max_pixels, iteration, max_sub, to_final, static_cast<DiagBitflag>(valdiag));
    if (!mesh)
        return -3;
    copy_text(mesh->getFullDescriptionStr(), desc, desc_cap);
    long long total = 0;
    int it = 0;
    while (!mesh->IsFinished()) {
        const uint64_t n = mesh->getCurNrGeodesics();
        if (it >= iter_cap || total + (long long)n > (long long)cap)
            return -1;
        iter_counts[it++] = n;
        for (uint64_t i = 0; i < n; ++i) {
            ScreenPoint u;
            ScreenIndex s;
            mesh->getNewInitConds(i, u, s);
            if (s[0] >= grid_side || s[1] >= grid_side)
                return -2;
            out_idx[2 * (total + i)] = s[0];
            out_idx[2 * (total + i) + 1] = s[1];
            mesh->GeodesicFinished(i, std::vector<real>{value_grid[s[0] * grid_side + s[1]]});
        }
        total += n;
        mesh->EndCurrentLoop();
    }
    *n_iters = it;
    return total;
}

// Utilities::GetFirstLineInfoString for a run description and one of the reference meshes.
void ref_first_line(const foort_run_desc *d, int mesh_type, uint64_t pixels, uint64_t max_pixels, uint64_t iteration,
                    int max_sub, int to_final, char *buf, size_t cap)
{
    Run R;
    setup(R, d, make_ref_mesh(mesh_type, pixels, max_pixels, iteration, max_sub, to_final,
                              static_cast<DiagBitflag>(d->value_diag)));
    Integrators::IntegratorDescription = d->integrator == FOORT_INTEGRATOR_VERLET ? "Verlet" : "RK4";
    copy_text(Utilities::GetFirstLineInfoString(R.metric.get(), R.source.get(), R.diags, R.valdiag, R.terms, R.view.get()),
              buf, cap);
}

// Drive the reference GeodesicOutputHandler: batch b holds batch_sizes[b] geodesics; geodesic i of
// batch b has the screen index text "b i " and, for diagnostic k, the text "v<b>_<i>_<k>".
void ref_output_run(const char *prefix, const char *timestamp, const char *ext, const char *const *names, int ndiag,
                    uint64_t to_cache, uint64_t per_file, const char *first_line, const uint64_t *batch_sizes, int nbatch)
{
    SetOutputLevel(OutputLevel::Level_0_WARNING);
    GeodesicOutputHandler h(prefix, timestamp, ext, std::vector<std::string>(names, names + ndiag), to_cache, per_file,
                            first_line);
    for (int b = 0; b < nbatch; ++b) {
        h.PrepareForOutput(batch_sizes[b]);
        for (uint64_t i = 0; i < batch_sizes[b]; ++i) {
            std::vector<std::string> v{std::to_string(b) + " " + std::to_string(i) + " "};
            for (int k = 0; k < ndiag; ++k)
                v.push_back("v" + std::to_string(b) + "_" + std::to_string(i) + "_" + std::to_string(k));
            h.NewGeodesicOutput(i, std::move(v));
        }
    }
    h.OutputFinished();
}

// The reference program (FOORT/Main.cpp:205-427) on a run description, one of the
// reference meshes and a file output handler: every mesh iteration integrates its
// pixels (in parallel), then hands the values to the mesh and the texts to the
// output handler in index order.  Returns the number of geodesics integrated.
long long ref_program_run(const foort_run_desc *d, int mesh_type, uint64_t pixels, uint64_t max_pixels,
                          uint64_t iteration, int max_sub, int to_final, const char *prefix, const char *ext,
                          uint64_t to_cache, uint64_t per_file, int first_line_on, int nthreads)
{
    Run R;
    setup(R, d, make_ref_mesh(mesh_type, pixels, max_pixels, iteration, max_sub, to_final,
                              static_cast<DiagBitflag>(d->value_diag)));
    Integrators::IntegratorDescription = d->integrator == FOORT_INTEGRATOR_VERLET ? "Verlet" : "RK4";
    std::string first = first_line_on ? Utilities::GetFirstLineInfoString(R.metric.get(), R.source.get(), R.diags,
                                                                           R.valdiag, R.terms, R.view.get())
                                      : std::string();
    GeodesicOutputHandler out(prefix, "", ext, Utilities::GetDiagNameStrings(R.diags, R.valdiag), to_cache, per_file,
                              first);
    long long total = 0;
    while (!R.view->IsFinished()) {
        const long long n = static_cast<long long>(R.view->getCurNrGeodesics());
        out.PrepareForOutput(static_cast<largecounter>(n));
        std::vector<std::vector<real>> values(n);
        std::vector<std::vector<std::string>> texts(n);
#pragma omp parallel num_threads(nthreads)
        {
            Geodesic geo(R.metric.get(), R.source.get(), R.diags, R.valdiag, R.terms, R.integ);
#pragma omp for schedule(dynamic, 1)
            for (long long i = 0; i < n; ++i) {
                Point p0;
                OneIndex v0;
                ScreenIndex s;
                R.view->SetNewInitialConditions(static_cast<largecounter>(i), p0, v0, s);
                geo.Reset(s, p0, v0);
                while (geo.getTermCondition() == Term::Continue)
                    geo.Update();
                values[i] = geo.getDiagnosticFinalValue();
                texts[i] = geo.getAllOutputStr();
            }
        }
        // the mesh's done-flags are a std::vector<bool>, so its callbacks run on one thread here
        for (long long i = 0; i < n; ++i) {
            R.view->GeodesicFinished(static_cast<largecounter>(i), values[i]);
            out.NewGeodesicOutput(static_cast<largecounter>(i), texts[i]);
        }
        total += n;
        R.view->EndCurrentLoop();
    }
    out.OutputFinished();
    return total;
}


// Integrate pixels with the reference code: screen_index (row, col) pairs, or
// grid indices first..first+count-1 of a SimpleSquareMesh(size*size).
int ref_trace(const foort_run_desc *d, uint64_t size, uint64_t first, uint64_t count, const uint64_t *screen_index,
              int32_t *term, int32_t *quadrant, int32_t *passes, double *closest, double *fpos, double *fvel,
              double *lambda, uint64_t *steps, int nthreads)
{
    Run R;
    std::unique_ptr<Mesh> mesh;
    DiagBitflag val = static_cast<DiagBitflag>(d->value_diag);
    if (screen_index) {
        std::vector<ScreenIndex> px(count);
        for (uint64_t i = 0; i < count; ++i)
            px[i] = ScreenIndex{screen_index[2 * i], screen_index[2 * i + 1]};
        mesh = std::unique_ptr<Mesh>(new ListMesh(size, std::move(px), val));
    } else {
        mesh = std::unique_ptr<Mesh>(new SimpleSquareMesh(size * size, val));
    }
    setup(R, d, std::move(mesh));
    // diagnostics 'FourColorScreen' and 'EquatPasses' come out exactly through their strings;
    // ClosestRadius at full precision through a second Geodesic whose only diagnostic it is.
    bool want_cr = (R.diags & Diag_ClosestRadius) != 0;
#pragma omp parallel num_threads(nthreads)
    {
        Geodesic geo(R.metric.get(), R.source.get(), R.diags, R.valdiag, R.terms, R.integ);
        Geodesic geo_cr(R.metric.get(), R.source.get(), Diag_ClosestRadius, Diag_ClosestRadius, R.terms, R.integ);
#pragma omp for schedule(dynamic, 1)
        for (long long ii = 0; ii < (long long)count; ++ii) {
            Point p0;
            OneIndex v0;
            ScreenIndex s;
            R.view->SetNewInitialConditions(first + ii, p0, v0, s);
            geo.Reset(s, p0, v0);
            uint64_t n = 0;
            while (geo.getTermCondition() == Term::Continue) {
                geo.Update();
                ++n;
            }
            term[ii] = static_cast<int32_t>(geo.getTermCondition());
            steps[ii] = n;
            lambda[ii] = geo.getCurrentLambda();
            Point fp = geo.getCurrentPos();
            OneIndex fv = geo.getCurrentVel();
            std::memcpy(&fpos[4 * ii], fp.data(), 32);
            std::memcpy(&fvel[4 * ii], fv.data(), 32);
            quadrant[ii] = 0;
            passes[ii] = 0;
            closest[ii] = -1;
            // getAllOutputStr: [screen index, one string per diagnostic in vector order]
            std::vector<std::string> out = geo.getAllOutputStr();
            DiagnosticUniqueVector order = CreateDiagnosticVector(R.diags, R.valdiag, nullptr);
            for (size_t k = 0; k < order.size(); ++k) {
                std::string nm = order[k]->getNameStr();
                if (nm == "FourColorScreen")
                    quadrant[ii] = std::stoi(out[k + 1]);
                else if (nm == "EquatPasses")
                    passes[ii] = std::stoi(out[k + 1]);
            }
            if (want_cr) {
                geo_cr.Reset(s, p0, v0);
                while (geo_cr.getTermCondition() == Term::Continue)
                    geo_cr.Update();
                closest[ii] = geo_cr.getDiagnosticFinalValue()[0];
            }
        }
    }
    return 0;
}

// The per-pixel work of FOORT/Main.cpp:373-400 exactly once per pixel (initial
// conditions, Reset, Update to termination, the final value for the mesh and
// the output strings), for timing the reference; returns the total step count.
unsigned long long ref_time(const foort_run_desc *d, uint64_t size, uint64_t count, const uint64_t *screen_index,
                            int nthreads)
{
    Run R;
    std::vector<ScreenIndex> px(count);
    for (uint64_t i = 0; i < count; ++i)
        px[i] = ScreenIndex{screen_index[2 * i], screen_index[2 * i + 1]};
    setup(R, d, std::unique_ptr<Mesh>(new ListMesh(size, std::move(px), static_cast<DiagBitflag>(d->value_diag))));
    unsigned long long total = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : total)
    {
        Geodesic geo(R.metric.get(), R.source.get(), R.diags, R.valdiag, R.terms, R.integ);
#pragma omp for
        for (long long ii = 0; ii < (long long)count; ++ii) {
            Point p0;
            OneIndex v0;
            ScreenIndex s;
            R.view->SetNewInitialConditions(ii, p0, v0, s);
            geo.Reset(s, p0, v0);
            while (geo.getTermCondition() == Term::Continue) {
                geo.Update();
                ++total;
            }
            volatile size_t sink = geo.getDiagnosticFinalValue().size() + geo.getAllOutputStr().size();
            (void)sink;
        }
    }
    return total;
}

void ref_metric(const foort_run_desc *d, const double *p, double *gdd, double *guu)
{
    auto M = make_metric(d);
    Point P{p[0], p[1], p[2], p[3]};
    TwoIndex a = M->getMetric_dd(P), b = M->getMetric_uu(P);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            gdd[4 * i + j] = a[i][j];
            guu[4 * i + j] = b[i][j];
        }
}

void ref_christoffel(const foort_run_desc *d, const double *p, double *G)
{
    auto M = make_metric(d);
    Integrators::Derivative_hval = d->derivative_h;
    ThreeIndex c = M->getChristoffel_udd(Point{p[0], p[1], p[2], p[3]});
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k)
                G[16 * i + 4 * j + k] = c[i][j][k];
}

void ref_step(const foort_run_desc *d, const double *x, const double *v, double *nx, double *nv, double *h)
{
    Run R;
    setup(R, d, std::unique_ptr<Mesh>(new SimpleSquareMesh(4, static_cast<DiagBitflag>(d->value_diag))));
    Point np{}, vv{};
    real step{};
    R.integ(Point{x[0], x[1], x[2], x[3]}, OneIndex{v[0], v[1], v[2], v[3]}, np, vv, step, R.metric.get(), R.source.get());
    std::memcpy(nx, np.data(), 32);
    std::memcpy(nv, vv.data(), 32);
    *h = step;
}

void ref_initial(const foort_run_desc *d, uint64_t size, uint64_t row, uint64_t col, double *pos, double *vel)
{
    Run R;
    std::vector<ScreenIndex> px{ScreenIndex{row, col}};
    setup(R, d, std::unique_ptr<Mesh>(new ListMesh(size, px, static_cast<DiagBitflag>(d->value_diag))));
    Point p;
    OneIndex v;
    ScreenIndex s;
    R.view->SetNewInitialConditions(0, p, v, s);
    std::memcpy(pos, p.data(), 32);
    std::memcpy(vel, v.data(), 32);
}

} // extern "C"
