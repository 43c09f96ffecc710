// C ABI of the handler level (include/portlbm_c.h).
#include <cstring>
#include <memory>
#include <string>

#include "lbm/core/domain.hpp"
#include "lbm/core/domain_initialization.hpp"
#include "lbm/exceptions/exceptions.hpp"
#include "lbm/execution/hip_algorithm_handler.hpp"
#include "lbm/file_interaction/file_interaction.hpp"
#include "portlbm_c.h"

struct portlbm_handler
{
    std::unique_ptr<lbm::execution::HipAlgorithmHandler> impl;
};

namespace
{
thread_local std::string g_error;

template <class F>
int guarded(F &&f)
{
    try
    {
        f();
        return PORTLBM_OK;
    }
    catch (const lbm::exceptions::json::PropertyArgumentException &e)
    {
        g_error = e.what();
        return PORTLBM_ERR_PROPERTY_ARGUMENT;
    }
    catch (const lbm::exceptions::json::MissingPropertyException &e)
    {
        g_error = e.what();
        return PORTLBM_ERR_MISSING_PROPERTY;
    }
    catch (const lbm::exceptions::json::UnknownPropertyException &e)
    {
        g_error = e.what();
        return PORTLBM_ERR_UNKNOWN_PROPERTY;
    }
    catch (const lbm::exceptions::algorithm::WaitException &e)
    {
        g_error = e.what();
        return PORTLBM_ERR_WAIT;
    }
    catch (const lbm::exceptions::Exception &e)
    {
        g_error = e.what();
        return PORTLBM_ERR_LBM;
    }
    catch (const std::exception &e)
    {
        g_error = e.what();
        return PORTLBM_ERR_OTHER;
    }
}

lbm::core::Properties parse(const char *json_text)
{
    if (!json_text)
        throw lbm::exceptions::json::MissingPropertyException("null settings text");
    return lbm::file_interaction::json_value_to_properties(lbm::file_interaction::JsonValue::parse(json_text));
}

void need(const void *p)
{
    if (!p)
        throw lbm::exceptions::json::PropertyArgumentException("null argument");
}

void fill_extents(const lbm::core::Properties &p, const lbm::core::Domain &d, uint32_t out[9])
{
    out[0] = p.horizontal_nodes;
    out[1] = p.vertical_nodes;
    out[2] = d.horizontal_nodes;
    out[3] = d.vertical_nodes;
    out[4] = d.subdomain_horizontal_nodes;
    out[5] = d.subdomain_vertical_nodes;
    out[6] = d.subdomain_count_horizontal;
    out[7] = d.subdomain_count_vertical;
    out[8] = d.total_node_count;
}

void copy_out(const std::string &s, char *out, size_t capacity, size_t *length)
{
    if (length)
        *length = s.size();
    if (out && capacity)
    {
        const size_t n = std::min(capacity - 1, s.size());
        std::memcpy(out, s.data(), n);
        out[n] = '\0';
    }
}
}  // namespace

extern "C" {

const char *portlbm_last_error(void) { return g_error.c_str(); }

int portlbm_create_from_file(const char *settings_path, portlbm_handler **out)
{
    return guarded([&] {
        need(settings_path);
        need(out);
        auto h = std::make_unique<portlbm_handler>();
        h->impl = std::make_unique<lbm::execution::HipAlgorithmHandler>(std::string(settings_path));
        *out = h.release();
    });
}

int portlbm_create_from_json(const char *json_text, portlbm_handler **out)
{
    return guarded([&] {
        need(out);
        auto h = std::make_unique<portlbm_handler>();
        h->impl = std::make_unique<lbm::execution::HipAlgorithmHandler>(parse(json_text));
        *out = h.release();
    });
}

void portlbm_destroy(portlbm_handler *h) { delete h; }

int portlbm_initialize(portlbm_handler *h)
{
    return guarded([&] {
        need(h);
        h->impl->initialize();
    });
}

int portlbm_start(portlbm_handler *h)
{
    return guarded([&] {
        need(h);
        h->impl->start();
    });
}

int portlbm_pause(portlbm_handler *h)
{
    return guarded([&] {
        need(h);
        h->impl->pause();
    });
}

int portlbm_block_until_finished(portlbm_handler *h)
{
    return guarded([&] {
        need(h);
        h->impl->block_until_finished();
    });
}

int portlbm_is_finished(portlbm_handler *h, int *finished)
{
    return guarded([&] {
        need(h);
        need(finished);
        *finished = h->impl->is_finished() ? 1 : 0;
    });
}

int portlbm_get_field(portlbm_handler *h, int field, double *out, size_t count)
{
    return guarded([&] {
        need(h);
        need(out);
        std::vector<double> *v = nullptr;
        switch (field)
        {
        case PORTLBM_DENSITY: v = &h->impl->get_densities(); break;
        case PORTLBM_X_VELOCITY: v = &h->impl->get_x_velocities(); break;
        case PORTLBM_Y_VELOCITY: v = &h->impl->get_y_velocities(); break;
        case PORTLBM_ABS_VELOCITY: v = &h->impl->get_absolute_velocities(); break;
        default: throw lbm::exceptions::json::PropertyArgumentException("unknown field");
        }
        if (count != v->size())
            throw lbm::exceptions::json::PropertyArgumentException("count must equal the inner node count");
        std::memcpy(out, v->data(), count * sizeof(double));
    });
}

int portlbm_get_extents(portlbm_handler *h, uint32_t out[9])
{
    return guarded([&] {
        need(h);
        need(out);
        fill_extents(h->impl->get_properties(), h->impl->get_domain(), out);
    });
}

int portlbm_get_iteration(portlbm_handler *h, uint32_t *iteration)
{
    return guarded([&] {
        need(h);
        need(iteration);
        *iteration = h->impl->get_current_iteration();
    });
}

int portlbm_get_progress(portlbm_handler *h, double *progress)
{
    return guarded([&] {
        need(h);
        need(progress);
        *progress = h->impl->get_progress();
    });
}

int portlbm_get_last_frametime(portlbm_handler *h, double *ms)
{
    return guarded([&] {
        need(h);
        need(ms);
        *ms = h->impl->get_last_frametime();
    });
}

int portlbm_get_phase(portlbm_handler *h, int8_t *out, size_t count)
{
    return guarded([&] {
        need(h);
        need(out);
        const auto &phase = h->impl->get_phase();
        if (count != phase.size())
            throw lbm::exceptions::json::PropertyArgumentException("count must equal the expanded node count");
        std::memcpy(out, phase.data(), count);
    });
}

int portlbm_get_populations(portlbm_handler *h, double *out, size_t count)
{
    return guarded([&] {
        need(h);
        need(out);
        const std::vector<double> v = h->impl->get_populations();
        if (count != v.size())
            throw lbm::exceptions::json::PropertyArgumentException("count must be 9 * expanded node count");
        std::memcpy(out, v.data(), count * sizeof(double));
    });
}

int portlbm_save_snapshot(portlbm_handler *h, const char *path)
{
    return guarded([&] {
        need(h);
        need(path);
        h->impl->save_snapshot(path);
    });
}

int portlbm_load_snapshot(portlbm_handler *h, const char *path)
{
    return guarded([&] {
        need(h);
        need(path);
        h->impl->load_snapshot(path);
    });
}

int portlbm_export_csv(portlbm_handler *h, const char *path)
{
    return guarded([&] {
        need(h);
        need(path);
        h->impl->export_csv(path);
    });
}

int portlbm_debug_history(portlbm_handler *h, int field, double *out, size_t count, size_t *needed)
{
    return guarded([&] {
        need(h);
        need(needed);
        const std::vector<double> *v = nullptr;
        switch (field)
        {
        case PORTLBM_DENSITY: v = &h->impl->debug_densities(); break;
        case PORTLBM_X_VELOCITY: v = &h->impl->debug_x_velocities(); break;
        case PORTLBM_Y_VELOCITY: v = &h->impl->debug_y_velocities(); break;
        default: throw lbm::exceptions::json::PropertyArgumentException("unknown field");
        }
        *needed = v->size();
        if (out)
        {
            if (count != v->size())
                throw lbm::exceptions::json::PropertyArgumentException("count must equal the history length");
            std::memcpy(out, v->data(), count * sizeof(double));
        }
    });
}

int portlbm_validate_json(const char *json_text)
{
    return guarded([&] { parse(json_text).validate(); });
}

int portlbm_domain_extents(const char *json_text, uint32_t out[9])
{
    return guarded([&] {
        need(out);
        lbm::core::Properties p = parse(json_text);
        lbm::core::Domain d(p);
        fill_extents(p, d, out);
    });
}

int portlbm_build_phase(const char *json_text, int8_t *out, size_t count)
{
    return guarded([&] {
        need(out);
        lbm::core::Properties p = parse(json_text);
        lbm::core::Domain d(p);
        const auto phase = lbm::core::domain_initialization::build_phase(p, d);
        if (count != phase.size())
            throw lbm::exceptions::json::PropertyArgumentException("count must equal the expanded node count");
        std::memcpy(out, phase.data(), count);
    });
}

int portlbm_roundtrip_json(const char *json_text, char *out, size_t capacity, size_t *length)
{
    return guarded([&] {
        const std::string s = lbm::file_interaction::properties_to_json_value(parse(json_text)).dump(4);
        copy_out(s, out, capacity, length);
    });
}

int portlbm_properties_string(const char *json_text, char *out, size_t capacity, size_t *length)
{
    return guarded([&] { copy_out(parse(json_text).to_string(), out, capacity, length); });
}

}  // extern "C"
