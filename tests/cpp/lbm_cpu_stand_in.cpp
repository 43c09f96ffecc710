// CPU stand-in for the entry points of include/lbm_hip.h that the host
// library (libportlbm: handler, C API, benchmark driver) calls.  Test
// infrastructure only: it lets the host library run its whole handler
// lifecycle -- create, start on the worker thread, frames, pause, debug
// passes, snapshots, accessors, destroy -- under AddressSanitizer,
// UndefinedBehaviorSanitizer and ThreadSanitizer on a machine without a GPU.
//
// It keeps the GPU library's contracts (argument checks, sizes, state
// order, what is valid when) and holds every buffer in host memory.  The
// "step" is a cheap deterministic update of every population, so a restart
// from a snapshot continues bit for bit exactly when the host library hands
// the state over correctly; it is not the lattice Boltzmann update.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "lbm_hip.h"

namespace
{
thread_local std::string g_error;

int fail(int code, const std::string &msg)
{
    g_error = msg;
    return code;
}

size_t storage_bytes(int s) { return s == LBM_STORE_F64 ? 8 : s == LBM_STORE_F32 ? 4 : 2; }
}  // namespace

struct lbm_ctx
{
    lbm_desc desc;
    uint32_t N = 0, D = 0;
    bool phase_set = false, initialized = false;
    std::vector<int8_t> phase;
    std::vector<double> lat;   // 9N, the stand-in's populations
    std::vector<double> res;   // 4D: rho, jx, jy, |j|
    std::vector<double> frame; // 4D, the last frame copied by lbm_run_frames
    std::vector<double> pass;  // 9N, populations after the last split pass
    uint64_t iteration = 0, frame_iteration = ~0ull;
    double last_ms = 0.0;
};

namespace
{
void step(lbm_ctx *c, bool macros)
{
    const uint64_t t = c->iteration;
    for (size_t i = 0; i < c->lat.size(); ++i)
        c->lat[i] = c->lat[i] * 0.999 + 1e-6 * double((i * 7u + t) & 15u);
    if (macros)
        for (uint32_t r = 0; r < c->D; ++r)
        {
            const double *f = &c->lat[9ull * (r % c->N)];
            double rho = 0, jx = 0, jy = 0;
            for (int d = 0; d < 9; ++d)
            {
                rho += f[d];
                jx += f[d] * (d % 3 - 1);
                jy += f[d] * (d / 3 - 1);
            }
            c->res[r] = rho;
            c->res[c->D + r] = jx;
            c->res[2ull * c->D + r] = jy;
            c->res[3ull * c->D + r] = std::sqrt(jx * jx + jy * jy);
        }
    ++c->iteration;
}

size_t state_bytes(const lbm_ctx *c) { return 16 + (c->lat.size() + c->res.size()) * sizeof(double); }
}  // namespace

extern "C" {

const char *lbm_last_error(void) { return g_error.c_str(); }

int lbm_device_max_work_group_size(uint32_t *out)
{
    if (!out)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_device_max_work_group_size: null argument");
    *out = 1024;
    return LBM_OK;
}

int lbm_device_name(char *buf, size_t len)
{
    if (!buf || !len)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_device_name: null argument");
    const char name[] = "CPU stand-in";
    const size_t n = std::min(len - 1, sizeof name - 1);
    std::memcpy(buf, name, n);
    buf[n] = '\0';
    return LBM_OK;
}

int lbm_create(const lbm_desc *desc, lbm_ctx **out)
{
    if (!desc || !out)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_create: null argument");
    *out = nullptr;
    if (desc->horizontal_nodes < 3 || desc->vertical_nodes < 3 || desc->storage < LBM_STORE_F64 ||
        desc->storage > LBM_STORE_BF16_SHIFTED)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_create: invalid description");
    lbm_ctx *c = new lbm_ctx();
    c->desc = *desc;
    c->N = desc->expanded_horizontal_nodes * desc->expanded_vertical_nodes;
    c->D = (desc->horizontal_nodes - 2u) * (desc->vertical_nodes - 2u);
    c->phase.assign(c->N, -1);
    c->lat.assign(9ull * c->N, 0.0);
    c->res.assign(4ull * c->D, 0.0);
    std::fill(c->res.begin(), c->res.begin() + c->D, -1.0);
    c->frame.assign(4ull * c->D, 0.0);
    (void)storage_bytes(desc->storage);
    *out = c;
    return LBM_OK;
}

void lbm_destroy(lbm_ctx *c) { delete c; }

int lbm_set_phase(lbm_ctx *c, const int8_t *phase, size_t n)
{
    if (!c || !phase)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_set_phase: null argument");
    if (n != c->N)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_set_phase: node_count must equal the expanded node count");
    std::memcpy(c->phase.data(), phase, n);
    c->phase_set = true;
    c->initialized = false;
    return LBM_OK;
}

int lbm_init_populations(lbm_ctx *c)
{
    if (!c)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_init_populations: null context");
    if (!c->phase_set)
        return fail(LBM_ERR_STATE, "lbm_init_populations: lbm_set_phase must come first");
    static const double w[9] = {1.0 / 36, 1.0 / 9, 1.0 / 36, 1.0 / 9, 4.0 / 9, 1.0 / 9, 1.0 / 36, 1.0 / 9, 1.0 / 36};
    for (size_t i = 0; i < c->lat.size(); ++i)
        c->lat[i] = w[i % 9];
    c->iteration = 0;
    c->frame_iteration = ~0ull;
    c->initialized = true;
    return LBM_OK;
}

int lbm_run_frames(lbm_ctx *c, uint32_t steps, uint32_t frame_interval)
{
    if (!c)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_run_frames: null context");
    if (!c->initialized)
        return fail(LBM_ERR_STATE, "lbm_run_frames: call lbm_init_populations first");
    for (uint32_t i = 0; i < steps; ++i)
    {
        const uint64_t it = c->iteration;
        const bool frame = frame_interval ? (it % frame_interval == frame_interval - 1u) : false;
        step(c, frame || i + 1u == steps);
        if (frame)
        {
            c->frame = c->res;
            c->frame_iteration = it;
        }
    }
    c->last_ms = 1e-3 * steps;
    return LBM_OK;
}

int lbm_get_frame(lbm_ctx *c, double *rho, double *jx, double *jy, double *jabs, size_t count, uint64_t *iteration)
{
    if (!c)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_frame: null context");
    if (count != c->D)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_frame: count must equal (H-2)*(V-2)");
    if (c->frame_iteration == ~0ull)
        return fail(LBM_ERR_STATE, "lbm_get_frame: no frame has been copied (lbm_run_frames)");
    double *dst[4] = {rho, jx, jy, jabs};
    for (int k = 0; k < 4; ++k)
        if (dst[k] && count)
            std::memcpy(dst[k], c->frame.data() + size_t(k) * c->D, count * sizeof(double));
    if (iteration)
        *iteration = c->frame_iteration;
    return LBM_OK;
}

int lbm_run_split(lbm_ctx *c, uint32_t steps, lbm_pass_callback cb, void *user)
{
    if (!c)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_run_split: null context");
    if (!c->initialized)
        return fail(LBM_ERR_STATE, "lbm_run_split: call lbm_init_populations first");
    for (uint32_t i = 0; i < steps; ++i)
        for (int pass = LBM_PASS_STREAM; pass <= LBM_PASS_COLLIDE; ++pass)
        {
            if (pass == LBM_PASS_COLLIDE)
                step(c, true);
            else
                c->pass = c->lat;
            if (cb)
                cb(user, pass, c->iteration);
        }
    c->last_ms = 1e-3 * steps;
    return LBM_OK;
}

int lbm_get_pass_populations(lbm_ctx *c, double *values, size_t count)
{
    if (!c || !values)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_pass_populations: null argument");
    if (c->pass.empty())
        return fail(LBM_ERR_STATE, "lbm_get_pass_populations: no split step has run");
    if (count != 9ull * c->N)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_pass_populations: count must be 9 * expanded node count");
    std::memcpy(values, c->pass.data(), count * sizeof(double));
    return LBM_OK;
}

int lbm_last_run_ms(lbm_ctx *c, double *ms)
{
    if (!c || !ms)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_last_run_ms: null argument");
    *ms = c->last_ms;
    return LBM_OK;
}

int lbm_get_macros(lbm_ctx *c, double *rho, double *jx, double *jy, double *jabs, size_t count)
{
    if (!c)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_macros: null context");
    if (count != c->D)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_macros: count must equal (H-2)*(V-2)");
    double *dst[4] = {rho, jx, jy, jabs};
    for (int k = 0; k < 4; ++k)
        if (dst[k] && count)
            std::memcpy(dst[k], c->res.data() + size_t(k) * c->D, count * sizeof(double));
    return LBM_OK;
}

int lbm_get_populations(lbm_ctx *c, double *values, size_t count)
{
    if (!c || !values)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_populations: null argument");
    if (count != 9ull * c->N)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_populations: count must be 9 * expanded node count");
    std::memcpy(values, c->lat.data(), count * sizeof(double));
    return LBM_OK;
}

int lbm_state_size(const lbm_ctx *c, size_t *bytes)
{
    if (!c || !bytes)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_state_size: null argument");
    *bytes = state_bytes(c);
    return LBM_OK;
}

int lbm_get_state(lbm_ctx *c, void *buf, size_t bytes)
{
    if (!c || !buf)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_state: null argument");
    if (!c->initialized)
        return fail(LBM_ERR_STATE, "lbm_get_state: call lbm_init_populations first");
    if (bytes != state_bytes(c))
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_state: size must be lbm_state_size()");
    char *out = static_cast<char *>(buf);
    std::memcpy(out, &c->iteration, 8);
    std::memset(out + 8, 0, 8);
    std::memcpy(out + 16, c->lat.data(), c->lat.size() * sizeof(double));
    std::memcpy(out + 16 + c->lat.size() * sizeof(double), c->res.data(), c->res.size() * sizeof(double));
    return LBM_OK;
}

int lbm_set_state(lbm_ctx *c, const void *buf, size_t bytes)
{
    if (!c || !buf)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_set_state: null argument");
    if (!c->initialized)
        return fail(LBM_ERR_STATE, "lbm_set_state: call lbm_init_populations first");
    if (bytes != state_bytes(c))
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_set_state: size does not match this context");
    const char *in = static_cast<const char *>(buf);
    std::memcpy(&c->iteration, in, 8);
    std::memcpy(c->lat.data(), in + 16, c->lat.size() * sizeof(double));
    std::memcpy(c->res.data(), in + 16 + c->lat.size() * sizeof(double), c->res.size() * sizeof(double));
    c->frame_iteration = ~0ull;
    return LBM_OK;
}

int lbm_sync(lbm_ctx *c)
{
    if (!c)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_sync: null context");
    return LBM_OK;
}

}  // extern "C"
