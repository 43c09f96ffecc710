// Handler lifecycle driver for the sanitizer builds of the host library
// (test infrastructure; linked with lbm_cpu_stand_in.cpp, no GPU).
//
// Replays, through the C API of include/portlbm_c.h, the sequences the
// Python tests and callers run -- the snapshot restart of
// tests/test_gpu_snapshot.py (a reference run with fields and populations
// read, a first run saved at step 123, a second run loaded from it and
// finished), pause/resume from another thread, re-initialisation, debug
// mode, CSV export -- for every algorithm x layout x storage, several
// rounds each, and checks what the host library promises (restart bit for
// bit, iteration counts).  Exit code 0: every check passed; AddressSanitizer,
// UBSan or ThreadSanitizer report and fail the run on any memory or thread
// error.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <string>
#include <thread>
#include <vector>

#include <unistd.h>

#include "portlbm_c.h"

namespace
{
int g_failures = 0;

#define CHECK(expr)                                                                            \
    do                                                                                         \
    {                                                                                          \
        if (!(expr))                                                                           \
        {                                                                                      \
            ++g_failures;                                                                      \
            std::fprintf(stderr, "FAILED line %d: %s (%s)\n", __LINE__, #expr, portlbm_last_error()); \
        }                                                                                      \
    } while (0)

std::string doc(const char *alg, const char *layout, const char *storage, const char *rounding, unsigned steps,
                unsigned fui, unsigned w = 96, unsigned h = 40, bool debug = false)
{
    const unsigned wg = (std::strcmp(alg, "npol") == 0 || std::strcmp(alg, "nsol") == 0) ? 64 : 128;
    char buf[1024];
    std::snprintf(buf, sizeof buf,
                  "{\"algorithmic\": {\"algorithm\": \"%s\", \"dataLayout\": \"%s\", \"debugMode\": %s, "
                  "\"frameUpdateInterval\": %u, \"timeSteps\": %u, \"workGroupSize\": %u}, "
                  "\"domain\": {\"horizontalNodes\": %u, \"verticalNodes\": %u, \"scenario\": \"circle\"}, "
                  "\"physical\": {\"inletDensity\": 1.05, \"inletVelocity\": {\"x\": 0.0, \"y\": 0.0}, "
                  "\"outletDensity\": 1.0, \"outletVelocity\": {\"x\": 0.0, \"y\": 0.0}, \"relaxationTime\": 0.7}, "
                  "\"mi355x\": {\"storage\": \"%s\", \"rounding\": \"%s\"}}",
                  alg, layout, debug ? "true" : "false", fui, steps, wg, w, h, storage, rounding);
    return buf;
}

struct Handler
{
    portlbm_handler *h = nullptr;
    explicit Handler(const std::string &json) { CHECK(portlbm_create_from_json(json.c_str(), &h) == PORTLBM_OK); }
    ~Handler() { portlbm_destroy(h); }
    uint32_t extent(int k)
    {
        uint32_t e[9] = {};
        CHECK(portlbm_get_extents(h, e) == PORTLBM_OK);
        return e[k];
    }
    size_t inner() { return size_t(extent(0) - 2) * (extent(1) - 2); }
    std::vector<double> field(int f)
    {
        std::vector<double> v(inner());
        CHECK(portlbm_get_field(h, f, v.data(), v.size()) == PORTLBM_OK);
        return v;
    }
    std::vector<double> populations()
    {
        std::vector<double> v(9ull * extent(8));
        CHECK(portlbm_get_populations(h, v.data(), v.size()) == PORTLBM_OK);
        return v;
    }
    uint32_t iteration()
    {
        uint32_t it = 0;
        CHECK(portlbm_get_iteration(h, &it) == PORTLBM_OK);
        return it;
    }
    void run()
    {
        CHECK(portlbm_start(h) == PORTLBM_OK);
        CHECK(portlbm_block_until_finished(h) == PORTLBM_OK);
    }
};

void restart_case(const char *alg, const char *layout, const char *storage, const char *rounding,
                  const std::string &snap)
{
    std::vector<std::vector<double>> expect;
    std::vector<double> expect_pop;
    {
        Handler ref(doc(alg, layout, storage, rounding, 300, 50));
        ref.run();
        for (int f = 0; f < 4; ++f)
            expect.push_back(ref.field(f));
        expect_pop = ref.populations();
    }
    std::vector<double> at_save;
    {
        Handler first(doc(alg, layout, storage, rounding, 123, 50));
        first.run();
        CHECK(portlbm_save_snapshot(first.h, snap.c_str()) == PORTLBM_OK);
        at_save = first.field(PORTLBM_DENSITY);
    }
    {
        Handler second(doc(alg, layout, storage, rounding, 300, 50));
        CHECK(portlbm_load_snapshot(second.h, snap.c_str()) == PORTLBM_OK);
        CHECK(second.iteration() == 123);
        CHECK(second.field(PORTLBM_DENSITY) == at_save);
        second.run();
        CHECK(second.iteration() == 300);
        for (int f = 0; f < 4; ++f)
            CHECK(second.field(f) == expect[f]);
        CHECK(second.populations() == expect_pop);
    }
}

void pause_resume_case(const char *alg, const char *layout, const char *storage)
{
    Handler h(doc(alg, layout, storage, "nearest", 1500, 7, 48, 20));
    CHECK(portlbm_start(h.h) == PORTLBM_OK);
    std::thread reader([&] {
        double progress = 0;
        for (int i = 0; i < 50; ++i)
            CHECK(portlbm_get_progress(h.h, &progress) == PORTLBM_OK);
    });
    std::this_thread::sleep_for(std::chrono::microseconds(200));
    CHECK(portlbm_pause(h.h) == PORTLBM_OK);
    reader.join();
    const uint32_t at_pause = h.iteration();
    CHECK(at_pause <= 1500);
    (void)h.field(PORTLBM_X_VELOCITY);
    CHECK(portlbm_start(h.h) == PORTLBM_OK);
    CHECK(portlbm_block_until_finished(h.h) == PORTLBM_OK);
    CHECK(h.iteration() == 1500);
    CHECK(portlbm_initialize(h.h) == PORTLBM_OK);
    CHECK(h.iteration() == 0);
    // destroyed while running: the destructor stops and joins the worker
    CHECK(portlbm_start(h.h) == PORTLBM_OK);
}

void debug_case(const char *alg, const char *layout, const char *storage)
{
    Handler h(doc(alg, layout, storage, "nearest", 3, 1, 8, 4, true));
    h.run();
    size_t needed = 0;
    CHECK(portlbm_debug_history(h.h, PORTLBM_DENSITY, nullptr, 0, &needed) == PORTLBM_OK);
    std::vector<double> hist(needed);
    CHECK(portlbm_debug_history(h.h, PORTLBM_DENSITY, hist.data(), hist.size(), &needed) == PORTLBM_OK);
    CHECK(needed == 3 * h.inner());
}
}  // namespace

// usage: handler_lifecycle [rounds] [layouts]; layouts "all" (default) or
// "collision" (the collision layout only, every algorithm and storage)
int main(int argc, char **argv)
{
    const int rounds = argc > 1 ? std::atoi(argv[1]) : 2;
    const bool all_layouts = argc <= 2 || std::strcmp(argv[2], "all") == 0;
    const std::filesystem::path dir = std::filesystem::temp_directory_path() /
                                      ("portlbm_lifecycle_" + std::to_string(::getpid()));
    std::filesystem::create_directories(dir);
    const std::string snap = (dir / "s.plbm").string();
    const char *algs[] = {"nptl", "lptl", "npol", "nsol"};
    const char *layouts[] = {"collision", "stream", "bundle"};
    const char *storages[] = {"f64", "f32", "f16-shifted", "bf16-shifted"};
    int cases = 0;
    for (int round = 0; round < rounds; ++round)
        for (const char *alg : algs)
            for (const char *layout : layouts)
                for (const char *storage : storages)
                {
                    if (!all_layouts && std::strcmp(layout, "collision") != 0)
                        continue;
                    const char *rounding = (round & 1) && storage[1] != '3' && storage[1] != '6' ? "stochastic"
                                                                                                 : "nearest";
                    restart_case(alg, layout, storage, rounding, snap);
                    pause_resume_case(alg, layout, storage);
                    ++cases;
                }
    for (const char *alg : algs)
        debug_case(alg, "stream", "f64");
    {
        Handler h(doc("lptl", "bundle", "f64", "nearest", 20, 5));
        h.run();
        CHECK(portlbm_export_csv(h.h, (dir / "r.csv").string().c_str()) == PORTLBM_OK);
        Handler other(doc("lptl", "bundle", "f64", "nearest", 10, 5));
        CHECK(portlbm_save_snapshot(h.h, snap.c_str()) == PORTLBM_OK);
        CHECK(portlbm_load_snapshot(other.h, snap.c_str()) != PORTLBM_OK);  // past its timeSteps
        CHECK(portlbm_load_snapshot(other.h, (dir / "missing").string().c_str()) != PORTLBM_OK);
    }
    std::filesystem::remove_all(dir);
    std::printf("lifecycle: %d cases x (restart, pause/resume) + debug + csv: %s\n", cases,
                g_failures ? "FAILED" : "ok");
    return g_failures ? 1 : 0;
}
