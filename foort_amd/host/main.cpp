// main.cpp -- `foort <file.cfg> [options]`: the reference's command line
// (FOORT/Main.cpp:158-200, one argument, the configuration file) running the
// integration on the GPU.  Like the reference it exits with status 0 when the
// configuration file is missing or does not parse; a failure of the GPU path
// exits with status 1.
//
// Options (none has a counterpart in the reference; each overrides the
// configuration's optional GPU group):
//   --christoffel fd|analytic   finite-difference (default) or exact derivatives
//   --precision fp64|mixed      binary64 throughout (default) or binary32 far field
//   --schedule oneshot|persistent
//   --output-prefix PREFIX      replaces Output.FilePrefix
//   --screen-output-level N     replaces Output.ScreenOutputLevel
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "foort_host.h"

static int pick(const char *v, const char *a, const char *b)
{
    if (!std::strcmp(v, a))
        return 0;
    if (!std::strcmp(v, b))
        return 1;
    std::fprintf(stderr, "unknown option value '%s' (expected %s or %s)\n", v, a, b);
    std::exit(2);
}

int main(int argc, char **argv)
{
    foort_program_options o;
    foort_program_options_init(&o);
    const char *cfg = nullptr;
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        const char *v = i + 1 < argc ? argv[i + 1] : nullptr;
        if (!std::strcmp(a, "--christoffel") && v)
            o.christoffel = pick(v, "fd", "analytic"), ++i;
        else if (!std::strcmp(a, "--precision") && v)
            o.precision = pick(v, "fp64", "mixed"), ++i;
        else if (!std::strcmp(a, "--schedule") && v)
            o.schedule = pick(v, "oneshot", "persistent"), ++i;
        else if (!std::strcmp(a, "--output-prefix") && v)
            o.output_prefix = v, ++i;
        else if (!std::strcmp(a, "--screen-output-level") && v)
            o.screen_output_level = std::atoi(v), ++i;
        else if (!cfg && a[0] != '-')
            cfg = a;
        else {
            std::fprintf(stderr, "unknown argument '%s'\n", a);
            return 2;
        }
    }
    foort_stats st{};
    o.stats = &st;
    int rc = foort_program_run(cfg, &o);
    if (rc == FOORT_OK) {
        if (st.geodesics)
            std::printf("GPU: %llu geodesics, %llu integration steps, %.3f s of kernel time\n",
                        static_cast<unsigned long long>(st.geodesics), static_cast<unsigned long long>(st.steps),
                        st.device_ms / 1000.0);
        return 0;
    }
    // the reference's early exits (no file, unreadable, syntax error) end with status 0
    const std::string err = foort_host_last_error();
    if (!cfg || err.rfind("configuration", 0) == 0)
        return 0;
    std::fprintf(stderr, "foort: %s\n", err.c_str());
    return 1;
}
