// host.hpp -- host side of the drop-in: screen messages, the configuration
// semantics of FOORT/Config.cpp, the description strings that form the first
// line of every output file, the meshes of FOORT/Mesh.cpp and the output
// handler of FOORT/InputOutput.cpp.  The GPU does the integration
// (foort_gpu.h); everything here runs once per pixel at most.
#pragma once

#include <cstdint>
#include <future>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "cfg.hpp"
#include "foort_gpu.h"
#include "foort_host.h"

namespace foort_host {

// ---- screen output: FOORT/InputOutput.cpp:12-56 ----
enum class Level { Warning = 0, Proc = 1, Subproc = 2, AllDetail = 3, Debug = 4, Max = 5 };
void set_output_level(int level);
void screen_output(const std::string &text, Level lvl = Level::AllDetail, bool newline = true);
void set_loop_message_frequency(unsigned long f);
unsigned long loop_message_frequency();
// levels of the reference's "default used" messages (FOORT/Config.h:43-46)
constexpr Level kOtherDefault = Level::Proc;
constexpr Level kImportantDefault = Level::Warning;

// ---- number texts as std::to_string and std::scientific write them ----
std::string to_str(double v);       // "%f"
std::string to_str_sci(double v);   // "%e"
std::string point_str(const double *p, int n); // toString(std::array<real, n>), FOORT/Geometry.h:98-112

// ---- the objects Config::* builds, as plain values ----
struct RunConfig {
    foort_run_desc desc{};
    bool horizon_metric = true;   // the metric is a SphericalHorizonMetric (all but FlatSpace)
    double direction[4] = {0, -1, 0, 0};
    foort_mesh_desc mesh{};
    std::vector<uint64_t> input_pixels; // InputCertainPixelsMesh list (read from stdin)
    std::string integrator_name = "RK4"; // Integrators::IntegratorDescription
    // output handler (Config::GetOutputHandler)
    bool to_files = false;
    std::string prefix, timestamp, extension;
    unsigned long to_cache = ~0UL - 1;
    unsigned long per_file = ~0UL;
    bool first_line_on = true;
    int screen_output_level = 4;
    // GPU group: tune the execution options on the first mesh iteration (0 none, 1 the
    // schedule's own options, 2 also the schedule)
    int autotune = 0;
};

// Config::InitializeScreenOutput ... Config::GetOutputHandler in the order of
// FOORT/Main.cpp:202-235.  Reads the console for InputCertainPixelsMesh when
// read_pixels_from_stdin is true.
void read_run_config(const Setting &root, RunConfig &rc, bool read_pixels_from_stdin);

// CreateDiagnosticVector order (FOORT/Diagnostics.cpp:14-99): the diagnostics in
// flag order, each followed by std::rotate(begin, begin + 1, end) when it is the
// value diagnostic.  The first entry supplies the mesh value
// (Geodesic::getDiagnosticFinalValue, FOORT/Geodesic.cpp:141-149).
std::vector<uint32_t> diagnostic_order(uint32_t all, uint32_t val);
std::string diag_name(uint32_t flag);
std::string diag_description(uint32_t flag, const RunConfig &rc);
std::string metric_description(const foort_run_desc &d);
std::string termination_descriptions(const RunConfig &rc);
std::string integrator_description(const RunConfig &rc);
std::string viewscreen_description(const RunConfig &rc, const std::string &mesh_description);
std::string first_line_info(const RunConfig &rc, const std::string &mesh_description);

// FinalDataValDistance of the value diagnostic (FOORT/Diagnostics.cpp:214-221, :395-401, :460-464)
// A geodesic's value for the mesh (getDiagnosticFinalValue): one number, or (theta, phi) for
// GeodesicPosition (FOORT/Diagnostics.cpp:303-308)
struct MeshValue {
    double v0 = 0, v1 = 0;
};
double value_distance(uint32_t valdiag, const MeshValue &a, const MeshValue &b);

// ---- meshes: FOORT/Mesh.h:24-61 ----
class Mesh {
public:
    explicit Mesh(uint32_t valdiag) : m_valdiag(valdiag) {}
    virtual ~Mesh() = default;
    virtual uint64_t getCurNrGeodesics() const = 0;
    virtual void getNewInitConds(uint64_t index, uint64_t &row, uint64_t &col) const = 0;
    virtual void GeodesicFinished(uint64_t index, MeshValue value) = 0;
    virtual void EndCurrentLoop() = 0;
    virtual bool IsFinished() const = 0;
    virtual std::string getFullDescriptionStr() const = 0;
    virtual uint64_t rowColumnSize() const = 0;
    // all screen indices of the current iteration as (row, column) pairs
    void screenIndices(uint64_t *out) const;

protected:
    double distance(const MeshValue &a, const MeshValue &b) const { return value_distance(m_valdiag, a, b); }
    uint32_t m_valdiag;
};

std::unique_ptr<Mesh> make_mesh(const foort_mesh_desc &d);

// getFullDataStr of FourColorScreen, EquatorialPasses (std::to_string of the int) and
// ClosestRadius (std::to_string of the double), FOORT/Diagnostics.cpp:202-206, :382-386, :448-452
inline std::string diag_value_text(uint32_t diag, int32_t quadrant, int32_t passes, double closest)
{
    switch (diag) {
    case FOORT_DIAGFLAG_FOURCOLORSCREEN: return std::to_string(quadrant);
    case FOORT_DIAGFLAG_EQUATORIALPASSES: return std::to_string(passes);
    default: return std::to_string(closest);
    }
}

// The kept positions of a batch (GeodesicPosition): pixel i has points offsets[i] .. offsets[i + 1].
struct Positions {
    std::vector<uint64_t> offsets;
    std::vector<double> points;
    // GeodesicPositionDiagnostic::getFullDataStr, FOORT/Diagnostics.cpp:286-301:
    // "<points> ;; " and then "x " for every coordinate of every point
    std::string text(uint64_t i) const
    {
        const uint64_t a = offsets[i], b = offsets[i + 1];
        std::string s = std::to_string(static_cast<unsigned long long>(b - a)) + " ;; ";
        for (uint64_t k = 4 * a; k < 4 * b; ++k)
            s += std::to_string(points[k]) + " ";
        return s;
    }
    // getFinalDataVal (:303-308): (theta, phi) of the last kept point.  The reference reads
    // the last element of an empty list when nothing was saved; here that gives (0, 0).
    MeshValue last_angles(uint64_t i) const
    {
        const uint64_t b = offsets[i + 1];
        if (b == offsets[i])
            return MeshValue{};
        return MeshValue{points[4 * (b - 1) + 2], points[4 * (b - 1) + 3]};
    }
};

// ---- GeodesicOutputHandler: FOORT/InputOutput.h:62-133 ----
class OutputHandler {
public:
    OutputHandler(std::string prefix, std::string timestamp, std::string extension,
                  std::vector<std::string> diag_names, unsigned long to_cache = ~0UL - 1,
                  unsigned long per_file = ~0UL, std::string first_line = "");
    void PrepareForOutput(uint64_t count);
    // screen index text (row, column) and one text per diagnostic
    void NewGeodesicOutput(uint64_t index, uint64_t row, uint64_t col, std::vector<std::string> texts);
    // the same with the diagnostics' values, written as texts when the files are written:
    // kind[d] = kInt (std::to_string of an int) or kFixed (std::to_string of a double, "%f")
    enum Kind : char { kInt = 1, kFixed = 2 };
    void setValueKinds(std::vector<char> kinds) { m_kinds = std::move(kinds); }
    void NewGeodesicValues(uint64_t index, uint64_t row, uint64_t col, const double *values);
    // Start formatting the lines of the entries given since the last call, in the background
    // (the entries are copied); WriteCached writes what is ready.  Values entries only.
    void FormatAhead();
    void OutputFinished();
    std::string getFullDescriptionStr() const;
    bool writesToConsole() const { return m_console; }

private:
    static constexpr int kMaxValues = 4;
    static constexpr uint64_t kNoText = ~uint64_t(0);
    struct Entry {              // plain data, so the cache grows by copying bytes
        uint64_t row, col;
        double values[kMaxValues];
        uint64_t text;          // index into m_texts, or kNoText: the entry holds values
    };
    void AppendText(std::string &b, const Entry &e, int d) const;
    void WriteCached();
    void OpenForFirstTime(const std::string &name);
    std::string FileName(int diag, unsigned short filenr) const;

    const std::string m_prefix, m_timestamp, m_extension;
    const std::vector<std::string> m_names;
    const unsigned long m_to_cache;
    const unsigned long m_per_file;
    const bool m_first_line_on;
    const std::string m_first_line;
    bool m_console = false;
    unsigned long m_prev_cached = 0;
    unsigned long m_in_file = 0;
    unsigned short m_full_files = 0;
    std::vector<Entry> m_cache;
    std::vector<std::vector<std::string>> m_texts; // texts given through NewGeodesicOutput
    std::vector<char> m_kinds;
    // lines formatted ahead: cache entries [lo, hi), one text per diagnostic
    struct Ahead {
        uint64_t lo, hi;
        std::future<std::vector<std::string>> text;
    };
    std::vector<Ahead> m_ahead;
    uint64_t m_ahead_end = 0;
};

std::string set_error(const std::string &msg); // records the thread's last host error

} // namespace foort_host
