// config.cpp -- what the reference's Config namespace does with a parsed
// configuration (FOORT/Config.cpp:94-1155), the screen messages of
// FOORT/InputOutput.cpp:12-56 and the description strings that make up the
// first line of every output file (FOORT/Utilities.cpp:52-85).  Option names,
// defaults, types and fall-backs are the reference's; the result is a
// foort_run_desc for the GPU path plus the mesh and output settings.
#include "host.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <ctime>
#include <iostream>
#include <limits>
#include <sstream>

namespace foort_host {

namespace {
int g_level = 1;                       // FOORT/InputOutput.cpp:14
unsigned long g_loop_freq = ~0UL;      // FOORT/InputOutput.cpp:17
constexpr double kPi = 3.1415926535;   // FOORT/Geometry.h:44

struct SettingError {
    std::string what;
};

std::string lower(std::string s)
{
    for (auto &c : s)
        c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

// the setting group `name` under g, or nullptr
const Setting *group(const Setting &g, const char *name)
{
    const Setting *c = g.child(name);
    return c && c->type == Setting::Group ? c : nullptr;
}
} // namespace

void set_output_level(int level) { g_level = level; }

void screen_output(const std::string &text, Level lvl, bool newline)
{
    if (static_cast<int>(lvl) > g_level)
        return;
    if (lvl == Level::Warning)
        std::cout << "WARNING: ";
    if (lvl == Level::Debug)
        std::cout << "DEBUG MSG: ";
    std::cout << text;
    if (newline)
        std::cout << '\n';
}

void set_loop_message_frequency(unsigned long f) { g_loop_freq = f; }
unsigned long loop_message_frequency() { return g_loop_freq; }

std::string to_str(double v)
{
    char b[512];
    std::snprintf(b, sizeof b, "%f", v);
    return b;
}

std::string to_str_sci(double v)
{
    std::ostringstream s;
    s << std::scientific << v;
    return s.str();
}

std::string point_str(const double *p, int n)
{
    std::string s = "(";
    for (int i = 0; i < n - 1; ++i)
        s += to_str(p[i]) + ", ";
    return s + to_str(p[n - 1]) + ")";
}

// ---------------------------------------------------------------------------
// descriptions
// ---------------------------------------------------------------------------

std::vector<uint32_t> diagnostic_order(uint32_t all, uint32_t val)
{
    static const uint32_t kFlags[4] = {FOORT_DIAGFLAG_FOURCOLORSCREEN, FOORT_DIAGFLAG_GEODESICPOSITION,
                                       FOORT_DIAGFLAG_EQUATORIALPASSES, FOORT_DIAGFLAG_CLOSESTRADIUS};
    std::vector<uint32_t> v;
    for (uint32_t f : kFlags) {
        if (!(all & f))
            continue;
        v.push_back(f);
        if (val & f)
            std::rotate(v.begin(), v.begin() + 1, v.end());
    }
    return v;
}

std::string diag_name(uint32_t flag)
{
    switch (flag) {
    case FOORT_DIAGFLAG_FOURCOLORSCREEN: return "FourColorScreen";
    case FOORT_DIAGFLAG_GEODESICPOSITION: return "GeodesicPosition";
    case FOORT_DIAGFLAG_EQUATORIALPASSES: return "EquatPasses";
    case FOORT_DIAGFLAG_CLOSESTRADIUS: return "ClosestRadius";
    default: return "";
    }
}

std::string diag_description(uint32_t flag, const RunConfig &rc)
{
    switch (flag) {
    case FOORT_DIAGFLAG_FOURCOLORSCREEN: return "Four-color screen";
    case FOORT_DIAGFLAG_GEODESICPOSITION:
        return "Geodesic position (output " + std::to_string(rc.desc.gp_output_steps) + " steps, updates every " +
               std::to_string(static_cast<unsigned long>(rc.desc.gp_update.update_n_steps)) + " steps)";
    case FOORT_DIAGFLAG_EQUATORIALPASSES:
        return "Equatorial passes (threshold = " + to_str(rc.desc.ep_threshold) + ")";
    case FOORT_DIAGFLAG_CLOSESTRADIUS: return "Closest radius";
    default: return "";
    }
}

double value_distance(uint32_t valdiag, const MeshValue &a, const MeshValue &b)
{
    switch (valdiag) {
    case FOORT_DIAGFLAG_FOURCOLORSCREEN: return std::fabs(a.v0 - b.v0) < 1.0 ? 0.0 : 1.0;
    case FOORT_DIAGFLAG_GEODESICPOSITION: // arc length between (theta, phi), FOORT/Diagnostics.cpp:310-321
        return std::acos(std::cos(a.v0) * std::cos(b.v0) + std::sin(a.v0) * std::sin(b.v0) * std::cos(a.v1 - b.v1));
    default: return std::fabs(a.v0 - b.v0); // EquatorialPasses, ClosestRadius
    }
}

std::string metric_description(const foort_run_desc &d)
{
    const double *P = d.metric_params;
    const std::string coord = d.rlog_scale ? "using logarithmic r coord" : "using normal r coord";
    switch (d.metric) {
    case FOORT_METRIC_KERR: return "Kerr (a = " + to_str(P[0]) + ", " + coord + ")";
    case FOORT_METRIC_FLATSPACE: return "Flat space";
    case FOORT_METRIC_RASHEEDLARSEN: {
        // the constructor stores the parameters rescaled to unit mass (FOORT/Metric.cpp:244-250)
        double s = (P[2] + P[3]) / 4.0;
        return "Rasheed-Larsen (m = " + to_str(P[0] / s) + ", a = " + to_str(P[1] / s) + ", p = " + to_str(P[2] / s) +
               ", q = " + to_str(P[3] / s) + ", " + coord + ")";
    }
    case FOORT_METRIC_JOHANNSEN:
        return "Johannsen (a = " + to_str(P[0]) + ", alpha13 = " + to_str(P[1]) + ", alpha22 = " + to_str(P[2]) +
               ", alpha52 = " + to_str(P[3]) + ", epsilon3 = " + to_str(P[4]) + ", " + coord + ")";
    default:
        return "Manko-Novikov (a = " + to_str(P[0]) + ", alpha3 = " + to_str(P[1]) + ", " + coord + ")";
    }
}

std::string termination_descriptions(const RunConfig &rc)
{
    const foort_run_desc &d = rc.desc;
    std::string s;
    if (d.term_flags & FOORT_TERMFLAG_HORIZON)
        s += "Horizon (stop at " + to_str(1 + d.horizon_epsilon) + "x(horizon radius)), ";
    if (d.term_flags & FOORT_TERMFLAG_BOUNDARYSPHERE)
        s += "Boundary sphere (R = " + to_str(d.sphere_radius) + "), ";
    if (d.term_flags & FOORT_TERMFLAG_TIMEOUT)
        s += "Time out (max integration steps: " + std::to_string(static_cast<unsigned long>(d.timeout_max_steps)) + "), ";
    if (d.term_flags & FOORT_TERMFLAG_THETASINGULARITY)
        s += "Theta singularity (epsilon: " + to_str(d.theta_epsilon) + "), ";
    return s;
}

std::string integrator_description(const RunConfig &rc)
{
    std::string s = "Integrator: " + rc.integrator_name;
    if (rc.integrator_name == "Verlet")
        s += " (velocity tolerance: " + to_str_sci(rc.desc.verlet_tolerance) + ")";
    return s + ", basic step size: " + to_str_sci(rc.desc.step_size) + ", min. step size: " +
           to_str_sci(rc.desc.smallest_step) + ", derivative h: " + to_str_sci(rc.desc.derivative_h);
}

std::string viewscreen_description(const RunConfig &rc, const std::string &mesh_description)
{
    return "ViewScreen position: " + point_str(rc.desc.camera_position, 4) +
           ", screen size: " + point_str(rc.desc.screen_size, 2) + ", " + mesh_description;
}

std::string first_line_info(const RunConfig &rc, const std::string &mesh_description)
{
    std::string diags = "Diagnostics: ";
    for (uint32_t f : diagnostic_order(rc.desc.diag_flags, rc.desc.value_diag))
        diags += diag_description(f, rc) + ", ";
    return "Metric: " + metric_description(rc.desc) + "; " + "Source: No source; " + diags + "; " +
           "Terminations: " + termination_descriptions(rc) + "; " + viewscreen_description(rc, mesh_description) +
           "; " + integrator_description(rc);
}

// ---------------------------------------------------------------------------
// Config::* (FOORT/Config.cpp:94-1155)
// ---------------------------------------------------------------------------

namespace {

void init_screen_output(const Setting &root, RunConfig &rc)
{
    int lvl = 4;
    set_output_level(4);
    if (const Setting *o = group(root, "Output")) {
        o->lookup("ScreenOutputLevel", lvl);
        lvl = std::min(std::max(lvl, 0), 5);
        set_output_level(lvl);
        unsigned long f = loop_message_frequency();
        o->lookup_counter("LoopMessageFrequency", f);
        set_loop_message_frequency(f);
    }
    rc.screen_output_level = lvl;
}

void get_metric(const Setting &root, RunConfig &rc)
{
    foort_run_desc &d = rc.desc;
    d.metric = FOORT_METRIC_KERR;
    d.metric_params[0] = 0.5;
    d.rlog_scale = 0;
    rc.horizon_metric = true;
    auto param = [](const Setting &m, const char *metric, const char *name, double &v) {
        if (!m.lookup(name, v))
            screen_output(std::string(metric) + ": no value for " + name + " given. Using default: " + to_str(v) + ".",
                          kOtherDefault);
    };
    try {
        const Setting *m = group(root, "Metric");
        if (!m)
            throw SettingError{"No metric settings found."};
        std::string name;
        if (!m->lookup("Name", name))
            throw SettingError{"No metric settings found."};
        name = lower(name);
        bool rlog = false;
        double *P = d.metric_params;
        if (name == "kerr") {
            double a = 0.5;
            param(*m, "Kerr", "a", a);
            m->lookup("RLogScale", rlog);
            d.metric = FOORT_METRIC_KERR;
            P[0] = a;
        } else if (name == "flatspace") {
            d.metric = FOORT_METRIC_FLATSPACE;
            rc.horizon_metric = false;
        } else if (name == "rasheedlarsen" || name == "rasheed-larsen") {
            double mm = 1.0, a = 0.5, p = 2.0, q = 2.0;
            param(*m, "Rasheed-Larsen", "m", mm);
            param(*m, "Rasheed-Larsen", "a", a);
            param(*m, "Rasheed-Larsen", "p", p);
            param(*m, "Rasheed-Larsen", "q", q);
            m->lookup("RLogScale", rlog);
            d.metric = FOORT_METRIC_RASHEEDLARSEN;
            P[0] = mm, P[1] = a, P[2] = p, P[3] = q;
        } else if (name == "johannsen") {
            double a = 0.7, a13 = 2.0, a22 = 0.0, a52 = 0.0, e3 = 0.0;
            param(*m, "Johannsen", "a", a);
            param(*m, "Johannsen", "alpha13", a13);
            param(*m, "Johannsen", "alpha22", a22);
            param(*m, "Johannsen", "alpha52", a52);
            param(*m, "Johannsen", "epsilon3", e3);
            m->lookup("RLogScale", rlog);
            d.metric = FOORT_METRIC_JOHANNSEN;
            P[0] = a, P[1] = a13, P[2] = a22, P[3] = a52, P[4] = e3;
        } else if (name == "mankonovikov" || name == "manko-novikov") {
            double a = 0.0, al3 = 5.0;
            param(*m, "Manko-Novikov", "a", a);
            param(*m, "Manko-Novikov", "alpha3", al3);
            m->lookup("RLogScale", rlog);
            d.metric = FOORT_METRIC_MANKONOVIKOV;
            P[0] = a, P[1] = al3;
        } else {
            throw SettingError{"No metric settings found."};
        }
        d.rlog_scale = rc.horizon_metric && rlog;
    } catch (const SettingError &e) {
        screen_output(e.what + " Using default metric (Kerr with a = 0.5).", kImportantDefault);
        d.metric = FOORT_METRIC_KERR;
        d.metric_params[0] = 0.5;
        d.rlog_scale = 0;
        rc.horizon_metric = true;
    }
}

void get_source(const Setting &root)
{
    try {
        const Setting *s = group(root, "Source");
        if (!s)
            throw SettingError{"No geodesic source settings found."};
        std::string name;
        if (!s->lookup("Name", name) || lower(name) != "nosource")
            throw SettingError{"No source settings found."};
    } catch (const SettingError &e) {
        screen_output(e.what + " Using default source (No source.).", kOtherDefault);
    }
}

void read_update(const Setting &g, unsigned long &n, int32_t &start, int32_t &finish)
{
    g.lookup_counter("UpdateFrequency", n);
    if (n == 0) {
        bool s = start != 0, f = finish != 0;
        g.lookup("UpdateStart", s);
        g.lookup("UpdateFinish", f);
        start = s, finish = f;
    }
}

void init_diagnostics(const Setting &root, RunConfig &rc)
{
    foort_run_desc &d = rc.desc;
    uint32_t all = 0, val = 0;
    try {
        const Setting *ds = group(root, "Diagnostics");
        if (!ds)
            throw SettingError{"No diagnostic settings found."};
        auto on = [ds](const char *name) -> const Setting * {
            const Setting *g = group(*ds, name);
            bool v = false;
            return g && g->exists("On") && g->lookup("On", v) && v ? g : nullptr;
        };
        auto use_for_mesh = [&val](const Setting *g, uint32_t flag) {
            bool v = false;
            if (val == 0 && g->lookup("UseForMesh", v) && v)
                val = flag;
        };
        if (const Setting *g = on("FourColorScreen")) {
            all |= FOORT_DIAGFLAG_FOURCOLORSCREEN;
            use_for_mesh(g, FOORT_DIAGFLAG_FOURCOLORSCREEN);
        }
        if (const Setting *g = on("GeodesicPosition")) {
            all |= FOORT_DIAGFLAG_GEODESICPOSITION;
            unsigned long n = 1;
            int32_t s = 0, f = 1;
            read_update(*g, n, s, f);
            unsigned long steps = 0;
            g->lookup_counter("OutputSteps", steps);
            rc.desc.gp_output_steps = steps;
            rc.desc.gp_update = {n, s, f};
            use_for_mesh(g, FOORT_DIAGFLAG_GEODESICPOSITION);
        }
        if (const Setting *g = on("EquatorialPasses")) {
            all |= FOORT_DIAGFLAG_EQUATORIALPASSES;
            unsigned long n = 1;
            int32_t s = 1, f = 1;
            read_update(*g, n, s, f);
            double thr = 0.01;
            g->lookup("Threshold", thr);
            d.ep_threshold = thr;
            d.ep_update = {n, s, f};
            use_for_mesh(g, FOORT_DIAGFLAG_EQUATORIALPASSES);
        }
        if (const Setting *g = on("ClosestRadius")) {
            all |= FOORT_DIAGFLAG_CLOSESTRADIUS;
            unsigned long n = 1;
            int32_t s = 1, f = 1;
            read_update(*g, n, s, f);
            d.cr_update = {n, s, f};
            use_for_mesh(g, FOORT_DIAGFLAG_CLOSESTRADIUS);
        }
        if (all == 0)
            throw SettingError{"No diagnostics turned on."};
        if (val == 0) {
            if (all & FOORT_DIAGFLAG_FOURCOLORSCREEN) {
                screen_output("No mesh diagnostic set; using FourColorScreen.", kOtherDefault);
                val = FOORT_DIAGFLAG_FOURCOLORSCREEN;
            } else {
                throw SettingError{
                    "Diagnostics turned on but no Mesh diagnostic selected, and FourColorScreen not turned on."};
            }
        }
    } catch (const SettingError &e) {
        screen_output(e.what + " Using default diagnostic(s) (FourColorScreen).", kImportantDefault);
        all = val = FOORT_DIAGFLAG_FOURCOLORSCREEN;
    }
    d.diag_flags = all;
    d.value_diag = val;
}

void init_terminations(const Setting &root, RunConfig &rc)
{
    foort_run_desc &d = rc.desc;
    uint32_t all = 0;
    try {
        const Setting *ts = group(root, "Terminations");
        if (!ts)
            throw SettingError{"No termination settings found."};
        auto on = [ts](const char *name) -> const Setting * {
            const Setting *g = group(*ts, name);
            bool v = false;
            return g && g->exists("On") && g->lookup("On", v) && v ? g : nullptr;
        };
        if (const Setting *g = on("Horizon")) {
            if (!rc.horizon_metric) {
                screen_output("Horizon Termination turned on but metric does not have horizon! Turning off Horizon "
                              "Termination.",
                              kImportantDefault);
            } else {
                all |= FOORT_TERMFLAG_HORIZON;
                double eps = 0.01;
                g->lookup("Epsilon_Horizon", eps);
                unsigned long n = 1;
                g->lookup_counter("UpdateFrequency", n);
                d.horizon_epsilon = eps;
                d.horizon_update_n = n;
            }
        }
        if (const Setting *g = on("BoundarySphere")) {
            all |= FOORT_TERMFLAG_BOUNDARYSPHERE;
            double R = 1000;
            g->lookup("SphereRadius", R);
            unsigned long n = 1;
            g->lookup_counter("UpdateFrequency", n);
            d.sphere_radius = R;
            d.sphere_update_n = n;
        }
        if (const Setting *g = on("TimeOut")) {
            all |= FOORT_TERMFLAG_TIMEOUT;
            unsigned long steps = 10000, n = 1;
            g->lookup_counter("MaxSteps", steps);
            g->lookup_counter("UpdateFrequency", n);
            d.timeout_max_steps = steps;
            d.timeout_update_n = n;
        }
        if (const Setting *g = on("ThetaSingularity")) {
            all |= FOORT_TERMFLAG_THETASINGULARITY;
            double eps = 1e-5;
            g->lookup("Epsilon", eps);
            unsigned long n = 1;
            g->lookup_counter("UpdateFrequency", n);
            d.theta_epsilon = eps;
            d.theta_update_n = n;
        }
        if (all == 0)
            throw SettingError{"No terminations turned on."};
    } catch (const SettingError &e) {
        screen_output(e.what + " Using default termination(s) (BoundarySphere and TimeOut).", kImportantDefault);
        all = FOORT_TERMFLAG_BOUNDARYSPHERE | FOORT_TERMFLAG_TIMEOUT;
        d.sphere_radius = 1000;
        d.sphere_update_n = 1;
        d.timeout_max_steps = 10000;
        d.timeout_update_n = 1;
    }
    d.term_flags = all;
}

// InputCertainPixelsMesh's console dialogue (FOORT/Mesh.cpp:81-141)
void read_input_pixels(unsigned long side, std::vector<uint64_t> &px)
{
    const std::string prefix = "InputCertainPixelsMesh message: ";
    screen_output(prefix + "Screen is a square with width/height = " + std::to_string(side) + ".", Level::Warning);
    for (;;) {
        long long x = 0, y = 0;
        screen_output(prefix + "Please enter coordinates for a pixel (x and y, separated by space); coordinates must "
                               "lie between 0 and " +
                          std::to_string(side - 1) + " (enter negative number to stop): ",
                      Level::Warning);
        std::cin >> x;
        if (x < 0 || !std::cin)
            break;
        std::cin >> y;
        if (y < 0 || !std::cin)
            break;
        if (static_cast<unsigned long>(x) > side - 1 || static_cast<unsigned long>(y) > side - 1) {
            screen_output(prefix + "Invalid coordinates. Please try again.", Level::Warning);
        } else {
            px.push_back(static_cast<uint64_t>(x));
            px.push_back(static_cast<uint64_t>(y));
            screen_output(prefix + "Pixel (" + std::to_string(x) + ", " + std::to_string(y) + ") added.",
                          Level::Warning);
        }
        std::cin.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
        std::cin.clear();
    }
    if (px.empty())
        screen_output("No pixels added to integration list!", Level::Warning);
}

void get_mesh(const Setting &root, RunConfig &rc, bool read_stdin)
{
    foort_mesh_desc &m = rc.mesh;
    m = foort_mesh_desc{};
    m.value_diag = rc.desc.value_diag;
    try {
        const Setting *vs = group(root, "ViewScreen");
        const Setting *ms = vs ? group(*vs, "Mesh") : nullptr;
        if (!ms)
            throw SettingError{"No Mesh settings found."};
        std::string type;
        if (!ms->lookup("Type", type))
            throw SettingError{"No Mesh Type specified."};
        if (type == "SimpleSquareMesh" || type == "InputCertainPixelsMesh") {
            unsigned long total = 100 * 100;
            ms->lookup_counter("TotalPixels", total);
            m.total_pixels = total;
            if (type == "SimpleSquareMesh") {
                m.type = FOORT_MESH_SIMPLESQUARE;
            } else {
                m.type = FOORT_MESH_INPUTCERTAINPIXELS;
                if (read_stdin)
                    read_input_pixels(static_cast<unsigned long>(std::sqrt(static_cast<double>(total))),
                                      rc.input_pixels);
            }
        } else if (type == "SquareSubdivisionMesh" || type == "SquareSubdivisionMeshV2") {
            unsigned long initial = 100, maxp = 100, iter = 100;
            int maxsub = 1;
            bool tofinal = false;
            ms->lookup_counter("InitialPixels", initial);
            ms->lookup_counter("MaxPixels", maxp);
            ms->lookup_counter("IterationPixels", iter);
            ms->lookup("MaxSubdivide", maxsub);
            if (maxp < initial)
                maxp = initial;
            if (maxsub < 1) {
                screen_output("Invalid MaxSubdivide level given. Using MaxSubdivide = 1.", kOtherDefault);
                maxsub = 1;
            }
            ms->lookup("InitialSubdivisionToFinal", tofinal);
            m.type = type == "SquareSubdivisionMesh" ? FOORT_MESH_SQUARESUBDIVISION : FOORT_MESH_SQUARESUBDIVISIONV2;
            m.initial_pixels = initial;
            m.max_pixels = maxp;
            m.iteration_pixels = iter;
            m.max_subdivide = maxsub;
            m.initial_subdivision_to_final = tofinal;
        } else {
            throw SettingError{"Incorrect Mesh Type specified."};
        }
    } catch (const SettingError &e) {
        screen_output(e.what + " Using default (SimpleSquareMesh with 100x100 pixels).", kImportantDefault);
        m = foort_mesh_desc{};
        m.type = FOORT_MESH_SIMPLESQUARE;
        m.total_pixels = 100 * 100;
        m.value_diag = rc.desc.value_diag;
    }
    m.pixels = rc.input_pixels.empty() ? nullptr : rc.input_pixels.data();
    m.pixel_count = rc.input_pixels.size() / 2;
}

void get_viewscreen(const Setting &root, RunConfig &rc, bool read_stdin)
{
    foort_run_desc &d = rc.desc;
    double pos[4] = {0, 1000, kPi / 2.0, 0};
    double dir[4] = {0, -1, 0, 0};
    double size[2] = {10.0, 10.0}, center[2] = {0.0, 0.0};
    try {
        const Setting *vs = group(root, "ViewScreen");
        if (!vs)
            throw SettingError{"No view screen settings found."};
        static const char *kAxes[4] = {"t", "r", "theta", "phi"};
        if (const Setting *p = group(*vs, "Position"))
            for (int i = 0; i < 4; ++i)
                p->lookup(kAxes[i], pos[i]);
        if (const Setting *p = group(*vs, "Direction"))
            for (int i = 0; i < 4; ++i)
                p->lookup(kAxes[i], dir[i]);
        if (const Setting *p = group(*vs, "ScreenSize"))
            p->lookup("x", size[0]), p->lookup("y", size[1]);
        if (const Setting *p = group(*vs, "ScreenCenter"))
            p->lookup("x", center[0]), p->lookup("y", center[1]);
    } catch (const SettingError &e) {
        screen_output(e.what + " Using default ViewScreen Settings.", kImportantDefault);
    }
    get_mesh(root, rc, read_stdin);
    for (int i = 0; i < 4; ++i) {
        d.camera_position[i] = pos[i];
        rc.direction[i] = dir[i];
    }
    d.screen_size[0] = size[0], d.screen_size[1] = size[1];
    d.screen_center[0] = center[0], d.screen_center[1] = center[1];
    // FOORT/ViewScreen.h:51-55
    if (!(dir[0] == 0 && dir[1] == -1 && dir[2] == 0 && dir[3] == 0))
        screen_output("ViewScreen is only supported pointing inwards at the moment; Direction = {0, -1, 0, 0} will "
                      "be used",
                      Level::Warning);
}

void get_integrator(const Setting &root, RunConfig &rc)
{
    foort_run_desc &d = rc.desc;
    d.integrator = FOORT_INTEGRATOR_RK4;
    rc.integrator_name = "RK4";
    try {
        const Setting *is = group(root, "Integrator");
        if (!is)
            throw SettingError{"No integrator settings found."};
        double step = d.step_size;
        if (!is->lookup("StepSize", step))
            screen_output("Using default integrator stepsize: " + to_str(d.step_size) + ".", kOtherDefault);
        else
            d.step_size = step;
        is->lookup("DerivativeH", d.derivative_h);
        is->lookup("SmallestPossibleStepsize", d.smallest_step);
        std::string type;
        if (!is->lookup("Type", type))
            throw SettingError{"No integrator settings found."};
        type = lower(type);
        if (type == "rk4") {
            d.integrator = FOORT_INTEGRATOR_RK4;
            rc.integrator_name = "RK4";
        } else if (type == "verlet") {
            d.integrator = FOORT_INTEGRATOR_VERLET;
            rc.integrator_name = "Verlet";
            is->lookup("VerletVelocityTolerance", d.verlet_tolerance);
        } else {
            throw SettingError{"No integrator settings found."};
        }
    } catch (const SettingError &e) {
        screen_output(e.what + " Using default integrator (RK4 integrator).", kImportantDefault);
    }
}

// Execution options of the GPU path: an optional group the reference does not read.
void get_gpu_options(const Setting &root, RunConfig &rc)
{
    foort_run_desc &d = rc.desc;
    const Setting *g = group(root, "GPU");
    if (!g)
        return;
    std::string v;
    if (g->lookup("Christoffel", v)) {
        v = lower(v);
        if (v == "fd" || v == "finitedifference")
            d.christoffel = FOORT_CHRISTOFFEL_FD;
        else if (v == "analytic")
            d.christoffel = FOORT_CHRISTOFFEL_ANALYTIC;
        else
            screen_output("GPU.Christoffel must be \"FD\" or \"Analytic\"; using FD.", Level::Warning);
    }
    if (g->lookup("Precision", v)) {
        v = lower(v);
        if (v == "fp64")
            d.precision = FOORT_PRECISION_FP64;
        else if (v == "mixed")
            d.precision = FOORT_PRECISION_MIXED;
        else
            screen_output("GPU.Precision must be \"FP64\" or \"Mixed\"; using FP64.", Level::Warning);
    }
    if (g->lookup("Schedule", v)) {
        v = lower(v);
        if (v == "oneshot")
            d.schedule = FOORT_SCHEDULE_ONESHOT;
        else if (v == "persistent")
            d.schedule = FOORT_SCHEDULE_PERSISTENT;
        else
            screen_output("GPU.Schedule must be \"OneShot\" or \"Persistent\"; using OneShot.", Level::Warning);
    }
    unsigned long n = d.step_cap;
    if (g->lookup_counter("StepCap", n))
        d.step_cap = n;
    int i = 0;
    if (g->lookup("BlockSize", i))
        d.block_size = i;
    n = d.steps_per_launch;
    if (g->lookup_counter("StepsPerLaunch", n))
        d.steps_per_launch = static_cast<uint32_t>(n);
    n = d.lanes_per_cu;
    if (g->lookup_counter("LanesPerCU", n))
        d.lanes_per_cu = static_cast<uint32_t>(n);
    n = d.refill_lanes;
    if (g->lookup_counter("RefillLanes", n))
        d.refill_lanes = static_cast<uint32_t>(n);
    n = d.refill_interval;
    if (g->lookup_counter("RefillInterval", n))
        d.refill_interval = static_cast<uint32_t>(n);
    g->lookup("MixedSwitchRadius", d.mixed_switch_radius);
    g->lookup("MixedPoleMargin", d.mixed_pole_margin);
    n = d.launch_us;
    if (g->lookup_counter("LaunchMicroseconds", n))
        d.launch_us = static_cast<uint32_t>(n);
    n = d.tail_geodesics;
    if (g->lookup_counter("TailGeodesics", n))
        d.tail_geodesics = static_cast<uint32_t>(std::min<unsigned long>(n, UINT32_MAX));
    bool b = d.use_graph != 0;
    if (g->lookup("UseGraph", b))
        d.use_graph = b;
    if (g->lookup("Autotune", v)) {
        v = lower(v);
        if (v == "none" || v == "off")
            rc.autotune = 0;
        else if (v == "options")
            rc.autotune = 1;
        else if (v == "schedule")
            rc.autotune = 2;
        else
            screen_output("GPU.Autotune must be \"None\", \"Options\" or \"Schedule\"; not tuning.", Level::Warning);
    }
}

std::string timestamp_now()
{
    std::time_t t = std::time(nullptr);
    std::tm tm = *std::localtime(&t);
    char b[64];
    std::strftime(b, sizeof b, "%y%m%d-%H%M%S", &tm);
    return b;
}

void get_output(const Setting &root, RunConfig &rc)
{
    rc.to_files = false;
    rc.prefix = rc.timestamp = rc.extension = "";
    rc.to_cache = ~0UL - 1;
    rc.per_file = ~0UL;
    rc.first_line_on = true;
    try {
        const Setting *o = group(root, "Output");
        if (!o)
            throw SettingError{"No output handler settings found."};
        std::string prefix;
        if (!o->lookup("FilePrefix", prefix))
            throw SettingError{"No output file name prefix found."};
        std::string ext;
        o->lookup("FileExtension", ext);
        bool ts = true;
        o->lookup("TimeStamp", ts);
        unsigned long cache = ~0UL - 1, per = ~0UL;
        o->lookup_counter("GeodesicsToCache", cache);
        o->lookup_counter("GeodesicsPerFile", per);
        if (per == 0)
            per = ~0UL;
        bool first = true;
        o->lookup("FirstLineInfo", first);
        rc.to_files = true;
        rc.prefix = prefix;
        rc.extension = ext;
        rc.timestamp = ts ? timestamp_now() : "";
        rc.to_cache = cache;
        rc.per_file = per;
        rc.first_line_on = first;
    } catch (const SettingError &e) {
        screen_output(e.what + " Will do all output to console.", kImportantDefault);
    }
}

} // namespace

void read_run_config(const Setting &root, RunConfig &rc, bool read_pixels_from_stdin)
{
    foort_desc_init(&rc.desc);
    init_screen_output(root, rc);
    get_metric(root, rc);
    get_source(root);
    init_diagnostics(root, rc);
    init_terminations(root, rc);
    get_viewscreen(root, rc, read_pixels_from_stdin);
    get_integrator(root, rc);
    get_gpu_options(root, rc);
    get_output(root, rc);
}

} // namespace foort_host
