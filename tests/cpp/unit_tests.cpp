// Host-side unit tests of the C++ surface, re-expressing the reference's
// Catch2 suites tests/test_access.cpp, tests/test_domain.cpp and
// tests/test_properties.cpp with a minimal harness.  No GPU is touched.
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "lbm/benchmark/benchmark.hpp"
#include "lbm/core/access.hpp"
#include "lbm/core/domain.hpp"
#include "lbm/core/properties.hpp"
#include "lbm/exceptions/exceptions.hpp"
#include "lbm/file_interaction/file_interaction.hpp"

namespace
{
int g_failures = 0;
int g_checks = 0;
std::string g_case;

void record(bool ok, const char *expr, int line)
{
    ++g_checks;
    if (!ok)
    {
        ++g_failures;
        std::cerr << "FAILED [" << g_case << "] line " << line << ": " << expr << "\n";
    }
}

#define CHECK(expr) record(static_cast<bool>(expr), #expr, __LINE__)
#define CHECK_FALSE(expr) record(!static_cast<bool>(expr), "!(" #expr ")", __LINE__)
#define CHECK_THROWS_AS(expr, type)                                                                      \
    do                                                                                                   \
    {                                                                                                    \
        bool thrown_ = false;                                                                            \
        try                                                                                              \
        {                                                                                                \
            expr;                                                                                        \
        }                                                                                                \
        catch (const type &)                                                                             \
        {                                                                                                \
            thrown_ = true;                                                                              \
        }                                                                                                \
        catch (...)                                                                                      \
        {                                                                                                \
        }                                                                                                \
        record(thrown_, #expr " throws " #type, __LINE__);                                               \
    } while (0)
#define CHECK_NOTHROW(expr)                                                                              \
    do                                                                                                   \
    {                                                                                                    \
        bool ok_ = true;                                                                                 \
        try                                                                                              \
        {                                                                                                \
            expr;                                                                                        \
        }                                                                                                \
        catch (...)                                                                                      \
        {                                                                                                \
            ok_ = false;                                                                                 \
        }                                                                                                \
        record(ok_, #expr " does not throw", __LINE__);                                                  \
    } while (0)

std::vector<std::pair<std::string, std::function<void()>>> &registry()
{
    static std::vector<std::pair<std::string, std::function<void()>>> r;
    return r;
}

struct Register
{
    Register(const char *name, std::function<void()> fn) { registry().emplace_back(name, std::move(fn)); }
};
#define TEST_CASE(name, fn) static Register reg_##fn(name, fn)

using namespace lbm::core::access;
using namespace lbm::core::power_functions;
using lbm::core::Domain;
using lbm::core::Properties;

// ---- access (reference tests/test_access.cpp) ------------------------------
void collision_accessor()
{
    constexpr unsigned int N = 999;
    CHECK(CollisionAccessor::at(0, 0, N) == 0);
    CHECK(CollisionAccessor::at(0, 8, N) == 8);
    CHECK(CollisionAccessor::at(1, 0, N) == 9);
    CHECK(CollisionAccessor::at(1, 8, N) == 17);
    CHECK(CollisionAccessor::at(5, 3, N) == 48);
    for (unsigned int node = 0; node < 4; ++node)
        for (unsigned int dir = 0; dir < 9; ++dir)
        {
            const unsigned int idx = CollisionAccessor::at(node, dir, N);
            CHECK(idx / 9 == node);
            CHECK(idx % 9 == dir);
        }
}
TEST_CASE("CollisionAccessor: AoS layout index", collision_accessor);

void stream_accessor()
{
    constexpr unsigned int total = 100;
    CHECK(StreamAccessor::at(0, 0, total) == 0);
    CHECK(StreamAccessor::at(1, 0, total) == 1);
    CHECK(StreamAccessor::at(0, 1, total) == 100);
    CHECK(StreamAccessor::at(5, 3, total) == 305);
    for (unsigned int node = 0; node < 5; ++node)
        for (unsigned int dir = 0; dir < 9; ++dir)
        {
            const unsigned int idx = StreamAccessor::at(node, dir, total);
            CHECK(idx / total == dir);
            CHECK(idx % total == node);
        }
}
TEST_CASE("StreamAccessor: SoA layout index", stream_accessor);

void bundle_accessor()
{
    constexpr unsigned int total = 100;
    const unsigned int expect0[9] = {0, 1, 2, 300, 301, 302, 600, 601, 602};
    for (unsigned int d = 0; d < 9; ++d)
        CHECK(BundleAccessor::at(0, d, total) == expect0[d]);
    CHECK(BundleAccessor::at(1, 0, total) == 3);
    CHECK(BundleAccessor::at(1, 3, total) == 303);
    CHECK(BundleAccessor::at(1, 6, total) == 603);
    for (unsigned int node = 0; node < 3; ++node)
        for (unsigned int dir = 0; dir < 9; ++dir)
            CHECK(BundleAccessor::at(node, dir, total) == 3u * (dir / 3u) * total + (dir % 3u) + 3u * node);
}
TEST_CASE("BundleAccessor: bundled-SoA layout index", bundle_accessor);

void node_index()
{
    constexpr unsigned int H = 10;
    CHECK(get_node_index(0, 0, H) == 0);
    CHECK(get_node_index(1, 0, H) == 1);
    CHECK(get_node_index(9, 0, H) == 9);
    CHECK(get_node_index(0, 1, H) == 10);
    CHECK(get_node_index(3, 2, H) == 23);
    for (unsigned int y = 0; y < 4; ++y)
        for (unsigned int x = 0; x < H; ++x)
            CHECK(get_node_index(x, y, H) == x + y * H);
}
TEST_CASE("get_node_index: row-major (x,y) to linear", node_index);

void neighbor()
{
    constexpr unsigned int H = 10, N = 55;
    CHECK(get_neighbor(N, 4, H) == N);
    CHECK(get_neighbor(N, 5, H) == N + 1);
    CHECK(get_neighbor(N, 3, H) == N - 1);
    CHECK(get_neighbor(N, 7, H) == N + H);
    CHECK(get_neighbor(N, 1, H) == N - H);
    CHECK(get_neighbor(N, 8, H) == N + H + 1);
    CHECK(get_neighbor(N, 6, H) == N + H - 1);
    CHECK(get_neighbor(N, 2, H) == N - H + 1);
    CHECK(get_neighbor(N, 0, H) == N - H - 1);
    CHECK(get_neighbor(55, 8, 10) == 66);
}
TEST_CASE("get_neighbor: D2Q9 direction offsets", neighbor);

void result_index()
{
    constexpr unsigned int H = 10, domain_count = 64;
    CHECK(get_result_index(1u, 1u, H) == 0u);
    CHECK(get_result_index(2u, 1u, H) == 1u);
    CHECK(get_result_index(8u, 1u, H) == 7u);
    CHECK(get_result_index(1u, 2u, H) == H - 2u);
    CHECK(get_result_index(2u, 2u, H) == H - 2u + 1u);
    CHECK(get_result_index(1u, 1u, H, domain_count, 0u) == 0u);
    CHECK(get_result_index(1u, 1u, H, domain_count, 1u) == domain_count);
    CHECK(get_result_index(2u, 1u, H, domain_count, 2u) == 1u + 2u * domain_count);
    CHECK(get_result_index(H + 1u, H) == 0u);
    CHECK(get_result_index(H + 2u, H) == 1u);
    CHECK(get_result_index(2u * H + 1u, H) == H - 2u);
}
TEST_CASE("get_result_index: three overloads strip the ghost layer", result_index);

void buffered_node_access()
{
    // tiles of 4 x 3 real nodes: a buffer column after x = 4, 8, ...; a buffer
    // row after y = 3, 6, ...
    CHECK(decomposed::BufferedNodeAccess::get_index(1, 1, 3, 4, 100) == 1 + 100);
    CHECK(decomposed::BufferedNodeAccess::get_index(4, 3, 3, 4, 100) == 4 + 3 * 100);
    CHECK(decomposed::BufferedNodeAccess::get_index(5, 3, 3, 4, 100) == 6 + 3 * 100);
    CHECK(decomposed::BufferedNodeAccess::get_index(5, 4, 3, 4, 100) == 6 + 5 * 100);
    CHECK(decomposed::NonBufferedNodeAccess::get_index(5, 4, 3, 4, 100) == 5 + 4 * 100);
}
TEST_CASE("BufferedNodeAccess skips buffer rows and columns", buffered_node_access);

void powers()
{
    CHECK_FALSE(is_power_of_4(1));
    CHECK(is_power_of_4(4));
    CHECK(is_power_of_4(16));
    CHECK(is_power_of_4(64));
    CHECK(is_power_of_4(256));
    CHECK_FALSE(is_power_of_4(2));
    CHECK_FALSE(is_power_of_4(8));
    CHECK_FALSE(is_power_of_4(32));
    CHECK_FALSE(is_power_of_4(6));
    const std::size_t p4[][2] = {{1, 0}, {4, 1}, {16, 2}, {64, 3}, {256, 4}, {2, 0}, {8, 0}, {32, 0}};
    for (auto &c : p4)
        CHECK(which_power_of_4(c[0]) == c[1]);
    const std::size_t p2[][2] = {{1, 0},  {2, 1}, {4, 2}, {8, 3},  {16, 4}, {32, 5},
                                 {64, 6}, {3, 0}, {6, 0}, {12, 0}, {24, 0}};
    for (auto &c : p2)
        CHECK(which_power_of_2(c[0]) == c[1]);
}
TEST_CASE("power functions", powers);

// ---- domain (reference tests/test_domain.cpp) ------------------------------
Properties make_props(const std::string &algorithm, unsigned int inner_v, unsigned int inner_h, unsigned int wg)
{
    return Properties(algorithm, "stream", false, wg, 10, 10, "Hagen-Poiseuille", inner_v, inner_h, 0.0, 0.0, 1.2,
                      0.0, 0.0, 1.0, 0.6);
}

void domain_lptl()
{
    auto props = make_props("lptl", 10, 6, 64);
    Domain dom(props);
    CHECK(dom.vertical_nodes == 12u);
    CHECK(dom.horizontal_nodes == 8u);
    CHECK(dom.total_node_count == 12u * 8u);
    CHECK(dom.subdomain_count_vertical == 1u);
    CHECK(dom.subdomain_count_horizontal == 1u);
    CHECK(dom.subdomain_vertical_nodes == 12u);
    CHECK(dom.subdomain_horizontal_nodes == 8u);
}
TEST_CASE("Domain: lptl preserves ghost-extended extents", domain_lptl);

void domain_nptl()
{
    {
        auto props = make_props("nptl", 10, 10, 4);
        Domain dom(props);
        CHECK(dom.total_node_count == dom.horizontal_nodes * dom.vertical_nodes);
        CHECK(dom.horizontal_nodes >= 12u);
        CHECK(dom.vertical_nodes >= 12u);
        CHECK(dom.subdomain_horizontal_nodes == 2u);
        CHECK(dom.subdomain_vertical_nodes == 2u);
    }
    {
        auto props = make_props("nptl", 10, 10, 8);
        Domain dom(props);
        CHECK(dom.total_node_count == dom.horizontal_nodes * dom.vertical_nodes);
        CHECK(dom.subdomain_horizontal_nodes * dom.subdomain_vertical_nodes == 8u);
    }
    {
        auto props = make_props("nptl", 10, 10, 7);
        Domain dom(props);
        CHECK(dom.subdomain_vertical_nodes == 1u);
        CHECK(dom.subdomain_horizontal_nodes == 7u);
    }
    {
        // 256 x 64 inner, wg 256: 16 x 16 tiles, no padding
        auto props = make_props("nptl", 64, 256, 256);
        Domain dom(props);
        CHECK(dom.horizontal_nodes == 258u);
        CHECK(dom.vertical_nodes == 66u);
        CHECK(dom.subdomain_count_horizontal == 16u);
        CHECK(dom.subdomain_count_vertical == 4u);
    }
}
TEST_CASE("Domain: nptl structural invariants", domain_nptl);

void domain_npol()
{
    {
        auto props = make_props("npol", 10, 10, 4);
        Domain dom(props);
        CHECK(dom.total_node_count == dom.horizontal_nodes * dom.vertical_nodes);
        CHECK(dom.subdomain_horizontal_nodes * dom.subdomain_vertical_nodes == 4u);
    }
    {
        auto props = make_props("npol", 20, 20, 16);
        Domain dom(props);
        CHECK(dom.subdomain_horizontal_nodes == 4u);
        CHECK(dom.subdomain_vertical_nodes == 4u);
    }
    {
        auto props = make_props("npol", 64, 256, 256);
        Domain dom(props);
        CHECK(dom.horizontal_nodes == 273u);
        CHECK(dom.vertical_nodes == 69u);
        CHECK(dom.total_node_count == 18837u);
    }
    {
        auto props = make_props("npol", 2048, 8192, 1024);
        Domain dom(props);
        CHECK(dom.horizontal_nodes == 8449u);
        CHECK(dom.vertical_nodes == 2113u);
    }
}
TEST_CASE("Domain: npol structural invariants", domain_npol);

void domain_nsol()
{
    {
        auto props = make_props("nsol", 10, 10, 8);
        Domain dom(props);
        CHECK(dom.total_node_count == dom.horizontal_nodes * dom.vertical_nodes);
    }
    {
        auto props = make_props("nsol", 64, 256, 256);
        Domain dom(props);
        CHECK(dom.subdomain_vertical_nodes == 15u);
        CHECK(dom.subdomain_horizontal_nodes == 14u);
        CHECK(dom.horizontal_nodes == 286u);
        CHECK(dom.vertical_nodes == 81u);
    }
    {
        auto props = make_props("nsol", 2048, 8192, 1024);
        Domain dom(props);
        CHECK(dom.horizontal_nodes == 8495u);
        CHECK(dom.vertical_nodes == 2145u);
        CHECK(dom.total_node_count == 18221775u);
    }
    {
        auto props = make_props("nsol", 10, 10, 5);  // below the minimum of 6, no settings file
        CHECK_THROWS_AS(Domain dom(props), lbm::exceptions::json::PropertyArgumentException);
    }
}
TEST_CASE("Domain: nsol structural invariants", domain_nsol);

// ---- properties (reference tests/test_properties.cpp) -----------------------
Properties valid_props()
{
    return Properties("lptl", "stream", false, 64, 100, 100, "Hagen-Poiseuille", 30, 126, 0.0, 0.0, 1.3, 0.0, 0.0,
                      1.0, 0.6);
}

void properties_ctor()
{
    auto p = valid_props();
    CHECK(p.algorithm == "lptl");
    CHECK(p.data_layout == "stream");
    CHECK(p.scenario == "Hagen-Poiseuille");
    CHECK(p.debug_mode == false);
    CHECK(p.vertical_nodes == 32u);
    CHECK(p.horizontal_nodes == 128u);
    CHECK(p.domain_node_count == 30u * 126u);
    CHECK(p.total_unexpanded_node_count == 32u * 128u);
    CHECK(p.inlet_density == 1.3);
    CHECK(p.outlet_density == 1.0);
    CHECK(p.relaxation_time == 0.6);
    CHECK(p.work_group_size == 64u);
    CHECK(p.time_steps == 100u);
    CHECK(p.settings_path.empty());
}
TEST_CASE("Properties constructor: field storage and ghost-layer accounting", properties_ctor);

void properties_valid()
{
    for (const char *alg : {"nptl", "lptl", "npol", "nsol"})
    {
        auto p = valid_props();
        p.algorithm = alg;
        CHECK_NOTHROW(p.validate());
    }
    for (const char *layout : {"stream", "collision", "bundle"})
    {
        auto p = valid_props();
        p.data_layout = layout;
        CHECK_NOTHROW(p.validate());
    }
    for (const char *sc : {"Hagen-Poiseuille", "walls", "circle", "square", "plate", "skyscraper", "wing", "porous",
                           "porous-circles", "cylinder"})
    {
        auto p = valid_props();
        p.scenario = sc;
        CHECK_NOTHROW(p.validate());
    }
    Properties p("lptl", "stream", false, 1, 1, 1, "Hagen-Poiseuille", 2, 2, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 0.5);
    CHECK_NOTHROW(p.validate());
}
TEST_CASE("Properties::validate(): valid inputs do not throw", properties_valid);

void properties_invalid()
{
    using Exc = lbm::exceptions::json::PropertyArgumentException;
    {
        Properties p("lptl", "stream", false, 64, 100, 100, "Hagen-Poiseuille", 1, 30, 0.0, 0.0, 1.3, 0.0, 0.0, 1.0,
                     0.6);
        CHECK_THROWS_AS(p.validate(), Exc);
    }
    {
        Properties p("lptl", "stream", false, 64, 100, 100, "Hagen-Poiseuille", 30, 1, 0.0, 0.0, 1.3, 0.0, 0.0, 1.0,
                     0.6);
        CHECK_THROWS_AS(p.validate(), Exc);
    }
    std::vector<std::function<void(Properties &)>> breakers = {
        [](Properties &p) { p.time_steps = 0; },
        [](Properties &p) { p.frame_update_interval = 0; },
        [](Properties &p) { p.work_group_size = 0; },
        [](Properties &p) { p.inlet_density = 0.0; },
        [](Properties &p) { p.inlet_density = -1.0; },
        [](Properties &p) { p.outlet_density = 0.0; },
        [](Properties &p) { p.relaxation_time = 0.0; },
        [](Properties &p) { p.algorithm = "not-an-algorithm"; },
        [](Properties &p) { p.data_layout = "not-a-layout"; },
        [](Properties &p) { p.scenario = "not-a-scenario"; },
        [](Properties &p) { p.extensions.storage = "f8"; },
    };
    for (auto &b : breakers)
    {
        auto p = valid_props();
        b(p);
        CHECK_THROWS_AS(p.validate(), Exc);
    }
}
TEST_CASE("Properties::validate(): invalid inputs throw PropertyArgumentException", properties_invalid);

std::string g_settings_path;

void json_bundled()
{
    if (g_settings_path.empty())
        return;
    CHECK(std::filesystem::exists(g_settings_path));
    Properties p = lbm::file_interaction::json_to_properties(g_settings_path);
    CHECK(p.algorithm == "lptl");
    CHECK(p.data_layout == "bundle");
    CHECK(p.scenario == "circle");
    CHECK(p.debug_mode == false);
    CHECK(p.time_steps == 1000000u);
    CHECK(p.work_group_size == 1024u);
    CHECK(p.inlet_density > 0.0);
    CHECK(p.outlet_density > 0.0);
    CHECK(p.relaxation_time > 0.0);
    CHECK(p.settings_path == g_settings_path);
    CHECK_NOTHROW(p.validate());
}
TEST_CASE("json_to_properties: parses bundled settings/settings.json", json_bundled);

void json_roundtrip()
{
    auto orig = valid_props();
    const auto tmp = std::filesystem::temp_directory_path() / "portlbm_test_roundtrip.json";
    orig.settings_path = tmp.string();
    lbm::file_interaction::properties_to_json(orig);
    CHECK(std::filesystem::exists(tmp));
    Properties loaded = lbm::file_interaction::json_to_properties(tmp.string());
    CHECK(loaded.algorithm == orig.algorithm);
    CHECK(loaded.data_layout == orig.data_layout);
    CHECK(loaded.scenario == orig.scenario);
    CHECK(loaded.debug_mode == orig.debug_mode);
    CHECK(loaded.time_steps == orig.time_steps);
    CHECK(loaded.work_group_size == orig.work_group_size);
    CHECK(loaded.inlet_density == orig.inlet_density);
    CHECK(loaded.outlet_density == orig.outlet_density);
    CHECK(loaded.relaxation_time == orig.relaxation_time);
    // stored node counts are written as stored: reading back adds ghosts again
    CHECK(loaded.horizontal_nodes == orig.horizontal_nodes + 2);
    // the -2 offset used for self-correction restores the original extents
    Properties corrected = lbm::file_interaction::json_to_properties(tmp.string(), -2);
    CHECK(corrected.horizontal_nodes == orig.horizontal_nodes);
    CHECK(corrected.vertical_nodes == orig.vertical_nodes);
    std::filesystem::remove(tmp);
}
TEST_CASE("json round-trip: properties_to_json -> json_to_properties", json_roundtrip);

void json_errors()
{
    using lbm::file_interaction::JsonValue;
    CHECK_THROWS_AS(JsonValue::parse("{\"a\": }"), lbm::exceptions::json::UnknownPropertyException);
    CHECK_THROWS_AS(lbm::file_interaction::json_value_to_properties(JsonValue::parse("{\"algorithmic\": {}}")),
                    lbm::exceptions::json::MissingPropertyException);
    JsonValue v = JsonValue::parse("{\"s\": \"a\\u00e9\\n\", \"n\": 18446744073709551615, \"f\": -1.5e-3}");
    CHECK(v.at("s").as_string("s") == "a\xc3\xa9\n");
    CHECK(v.at("n").as_unsigned("n") == 18446744073709551615ull);
    CHECK(v.at("f").as_double("f") == -1.5e-3);
    CHECK(JsonValue::parse(v.dump()).at("f").as_double("f") == -1.5e-3);
}
TEST_CASE("JSON reader: syntax errors, escapes, exact integers", json_errors);

// benchmark.cpp:30-62 for a device maximum of 1024 (gfx950): per quarter, the
// powers of two plus the first odd and the first even other size.
void benchmark_work_group_sizes()
{
    using lbm::benchmark::get_work_group_sizes;
    const std::vector<std::size_t> expect = {1024, 1023, 1022, 768, 767, 512, 511, 510, 256, 255, 254};
    CHECK(get_work_group_sizes(1024) == expect);
    const std::vector<std::size_t> small = {16, 15, 14, 12, 11, 8, 7, 6, 4, 3, 2};
    CHECK(get_work_group_sizes(16) == small);
}
TEST_CASE("Benchmark: work-group sizes per device maximum", benchmark_work_group_sizes);

void benchmark_progress_and_time_format()
{
    lbm::benchmark::Progress p(4);
    p.advance(2.0);
    CHECK(p.current_test_rows == 1 && p.progress == 25.0 && p.current_average_time == 2.0);
    CHECK(p.remaining_time_estimation == 6.0);
    p.advance(4.0);
    CHECK(p.current_average_time == 3.0 && p.remaining_time_estimation == 6.0 && p.progress == 50.0);
    CHECK(lbm::benchmark::seconds_to_time_format(90061) == "1d, 1h, 1m, 1s");
    CHECK(lbm::benchmark::seconds_to_time_format(59) == "0d, 0h, 0m, 59s");
}
TEST_CASE("Benchmark: progress and remaining-time format", benchmark_progress_and_time_format);

void benchmark_settings_file()
{
    const std::filesystem::path path = std::filesystem::path(g_settings_path).parent_path() / "benchmark.json";
    const auto s = lbm::benchmark::BenchmarkSettings::from_file(path.string());
    CHECK(s.test_runs == 10 && s.test_linear_two_lattice);
    CHECK((s.data_layouts == std::vector<std::string>{"collision", "stream", "bundle"}));
    CHECK((s.algorithms == std::vector<std::string>{"nptl", "npol", "nsol"}));
    CHECK((s.phase_2_scenarios == std::vector<std::string>{"plate", "wing", "porous"}));
    bool threw = false;
    try
    {
        lbm::benchmark::BenchmarkSettings::from_file("/nonexistent/benchmark.json");
    }
    catch (const lbm::exceptions::Exception &)
    {
        threw = true;
    }
    CHECK(threw);
}
TEST_CASE("Benchmark: settings/benchmark.json", benchmark_settings_file);

// ---- golden known answers ----------------------------------------------------
// tests/golden/reference_kats.json: the access and power-function known answers
// of the reference's tests/test_access.cpp as data (inputs, expected outputs,
// reference file:line), checked against the C++ surface.
void golden_known_answers()
{
    if (g_settings_path.empty())
        return;
    const std::filesystem::path file = std::filesystem::path(g_settings_path).parent_path().parent_path() /
                                       "tests" / "golden" / "reference_kats.json";
    CHECK(std::filesystem::exists(file));
    std::ifstream in(file);
    std::stringstream text;
    text << in.rdbuf();
    using lbm::file_interaction::JsonValue;
    const JsonValue doc = JsonValue::parse(text.str());
    const auto u = [](const JsonValue &c, const char *key) { return static_cast<unsigned int>(c.at(key).as_unsigned(key)); };
    int cases = 0;
    for (const JsonValue &c : doc.at("accessor").items())
    {
        const std::string &layout = c.at("layout").as_string("layout");
        const unsigned int n = u(c, "node"), d = u(c, "dir"), t = u(c, "total"), e = u(c, "expect");
        if (layout == "collision")
            CHECK(CollisionAccessor::at(n, d, t) == e);
        else if (layout == "stream")
            CHECK(StreamAccessor::at(n, d, t) == e);
        else
            CHECK(BundleAccessor::at(n, d, t) == e);
        ++cases;
    }
    for (const JsonValue &c : doc.at("node_index").items())
    {
        CHECK(get_node_index(u(c, "x"), u(c, "y"), u(c, "H")) == u(c, "expect"));
        ++cases;
    }
    for (const JsonValue &c : doc.at("neighbor").items())
    {
        CHECK(get_neighbor(u(c, "node"), u(c, "dir"), u(c, "H")) == u(c, "expect"));
        ++cases;
    }
    for (const JsonValue &c : doc.at("result_index_xy").items())
    {
        CHECK(get_result_index(u(c, "x"), u(c, "y"), u(c, "H")) == u(c, "expect"));
        ++cases;
    }
    for (const JsonValue &c : doc.at("result_index_series").items())
    {
        CHECK(get_result_index(u(c, "x"), u(c, "y"), u(c, "H"), u(c, "domain_count"), u(c, "step")) == u(c, "expect"));
        ++cases;
    }
    for (const JsonValue &c : doc.at("result_index_linear").items())
    {
        CHECK(get_result_index(u(c, "node"), u(c, "H")) == u(c, "expect"));
        ++cases;
    }
    for (const JsonValue &c : doc.at("is_power_of_4").items())
    {
        CHECK(is_power_of_4(u(c, "x")) == c.at("expect").as_bool("expect"));
        ++cases;
    }
    for (const JsonValue &c : doc.at("which_power_of_4").items())
    {
        CHECK(which_power_of_4(u(c, "x")) == u(c, "expect"));
        ++cases;
    }
    for (const JsonValue &c : doc.at("which_power_of_2").items())
    {
        CHECK(which_power_of_2(u(c, "x")) == u(c, "expect"));
        ++cases;
    }
    CHECK(cases >= 70);
}
TEST_CASE("Golden: the reference's known answers (tests/golden/reference_kats.json)", golden_known_answers);

}  // namespace

int main(int argc, char **argv)
{
    if (argc > 1)
        g_settings_path = argv[1];
    for (auto &[name, fn] : registry())
    {
        g_case = name;
        try
        {
            fn();
        }
        catch (const std::exception &e)
        {
            ++g_failures;
            std::cerr << "FAILED [" << name << "] unexpected exception: " << e.what() << "\n";
        }
    }
    std::cout << registry().size() << " test cases, " << g_checks << " checks, " << g_failures << " failures\n";
    return g_failures ? 1 : 0;
}
