// Batched domains: many independent channels of uneven size advanced by one
// persistent kernel per step (a fixed grid striding over the tiles of all
// domains), optionally with the whole run captured in one hipGraph.  Each
// domain is an unexpanded two-lattice pull in the stream layout, updated by
// the same device functions as the single-lattice kernels (pull with
// bounce-back bits, Zou-He outlet, BGK, full-segment stores).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "lbm_device.hpp"
#include "lbm_error.hpp"
#include "lbm_setup.hpp"

using namespace lbm_dev;

namespace
{

#define BATCH_TRY(expr)                                                                             \
    do                                                                                              \
    {                                                                                               \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return lbm_set_error(e_ == hipErrorOutOfMemory ? LBM_ERR_OUT_OF_MEMORY : LBM_ERR_DEVICE, \
                                 std::string(#expr) + ": " + hipGetErrorString(e_));                \
    } while (0)

constexpr uint32_t kBatchX0 = 15;  // x = 1 starts a 16-node (64 B in FP32) segment
constexpr int kTileThreads = 256;

struct BatchDom
{
    uint64_t lat;   // element offset of the domain's lattice in each arena
    uint32_t code;  // node offset in the code arena
    uint32_t res;   // offset in the result arena (4 * D doubles)
    uint32_t uy;    // offset in the outlet-history arena (2 * V doubles)
    uint32_t H, V, P, S, D;
    uint32_t tile_w;
    double omega, rho_out;
};

// Set-up geometry of one domain: an unexpanded lattice with its own pitch.
// The batched 16-bit forms store f_d - w_d (rho_ref = 1 in every column: `ones`).
inline Geom batch_geom(const BatchDom &dm, const float *ones)
{
    Geom g{dm.H, dm.H * dm.V, dm.P, kBatchX0, dm.S};
    g.ref = ones;
    g.H = dm.H;
    g.buffered = 0u;
    g.sub_h = 1u;
    g.ghost_xe = dm.H - 1u;
    return g;
}

struct BatchTile
{
    uint32_t dom;
    uint16_t x0, y0;  // first inner node of the tile
};

struct BatchArgs
{
    const BatchDom *doms;
    const BatchTile *tiles;
    uint32_t ntiles;
    const void *src;
    void *dst;
    const uint16_t *codes;
    double *res;
    double *uy;
    int parity;
    uint32_t write_macros;
    uint32_t nan_protection;
    const float *ones;  // rho_ref of the 16-bit forms: 1 in every column
};

template <int S>
__global__ __launch_bounds__(kTileThreads) void k_batch_step(BatchArgs a)
{
    using T = typename Storage<S>::T;
    using C = typename Storage<S>::C;
    const T *src = static_cast<const T *>(a.src);
    T *dst = static_cast<T *>(a.dst);
    for (uint32_t t = blockIdx.x; t < a.ntiles; t += gridDim.x)
    {
        const BatchTile tl = a.tiles[t];
        const BatchDom dm = a.doms[tl.dom];
        const uint32_t lx = threadIdx.x % dm.tile_w, ly = threadIdx.x / dm.tile_w;
        const uint32_t x = tl.x0 + lx, y = tl.y0 + ly;
        const bool inside = x + 2u <= dm.H && y + 2u <= dm.V;
        const uint32_t pn = kBatchX0 + x + y * dm.P;
        const uint32_t code = inside ? a.codes[dm.code + pn] : 0u;
        const bool fluid = code & LBM_CODE_FLUID;
        const bool segment = segment_has<T>(fluid);  // every lane takes part in the ballot
        const bool fill = (code & LBM_CODE_SOLID) && segment;
        C f[9] = {};
        if (fluid)
        {
            StepParams p{};
            p.rho = a.res + dm.res;
            p.jx = p.rho + dm.D;
            p.jy = p.rho + 2u * dm.D;
            p.jabs = p.rho + 3u * dm.D;
            p.uy_in = a.uy + dm.uy + uint32_t(a.parity) * dm.V;
            p.uy_out = a.uy + dm.uy + uint32_t(1 - a.parity) * dm.V;
            p.omega = dm.omega;
            p.rho_out = dm.rho_out;
            p.H = dm.H;
            p.write_macros = a.write_macros;
            p.nan_protection = a.nan_protection;
            pull<LBM_LAYOUT_STREAM, S>(src + dm.lat, pn, code, dm.P, dm.S, f, rt3_of<S>(a.ones, x));
            finish_node<LBM_LAYOUT_STREAM, S>(p, code, x, y, f, 1.f);
        }
        if (fluid || fill)
            store_node<LBM_LAYOUT_STREAM, S>(dst + dm.lat, pn, dm.S, f, 0u, 0.f);  // rho_ref = 1 throughout
    }
}

// Per-tile FP64 sums of |j - j_snap|^2 and |j|^2 over fluid nodes of the
// source lattice, snapshot update; one block per tile, wave64 shuffles.
template <int S>
__global__ __launch_bounds__(kTileThreads) void k_batch_residual(const BatchDom *doms, const BatchTile *tiles,
                                                                 const void *src_v, const uint16_t *codes,
                                                                 double *snap, double *partials)
{
    using T = typename Storage<S>::T;
    const T *src = static_cast<const T *>(src_v);
    const BatchTile tl = tiles[blockIdx.x];
    const BatchDom dm = doms[tl.dom];
    const uint32_t lx = threadIdx.x % dm.tile_w, ly = threadIdx.x / dm.tile_w;
    const uint32_t x = tl.x0 + lx, y = tl.y0 + ly;
    double acc0 = 0.0, acc1 = 0.0;
    if (x + 2u <= dm.H && y + 2u <= dm.V)
    {
        const uint32_t pn = kBatchX0 + x + y * dm.P;
        if (codes[dm.code + pn] & LBM_CODE_FLUID)
        {
            double jx = 0, jy = 0;
#pragma unroll
            for (int d = 0; d < 9; ++d)
            {
                const double f = Storage<S>::decode(src[dm.lat + at<LBM_LAYOUT_STREAM>(pn, uint32_t(d), dm.S)], d,
                                                    1.f);  // rho_ref = 1 in the batch
                jx += f * double(cx_of(d));
                jy += f * double(cy_of(d));
            }
            double *sp = snap + 2ull * (dm.code + pn);
            const double dx = jx - sp[0], dy = jy - sp[1];
            acc0 = dx * dx + dy * dy;
            acc1 = jx * jx + jy * jy;
            sp[0] = jx;
            sp[1] = jy;
        }
    }
    __shared__ double part[kTileThreads / 64][2];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        acc0 += __shfl_xor(acc0, off, 64);
        acc1 += __shfl_xor(acc1, off, 64);
    }
    if ((threadIdx.x & 63) == 0)
    {
        part[threadIdx.x >> 6][0] = acc0;
        part[threadIdx.x >> 6][1] = acc1;
    }
    __syncthreads();
    if (threadIdx.x < 2)
    {
        double s = 0.0;
        for (int w = 0; w < kTileThreads / 64; ++w)
            s += part[w][threadIdx.x];
        partials[2ull * blockIdx.x + threadIdx.x] = s;
    }
}

int width_class(uint32_t inner_w) { return inner_w <= 128 ? 0 : inner_w <= 256 ? 1 : inner_w <= 512 ? 2 : 3; }

// ---- scratch of lbm_batch_residuals_noalloc ---------------------------------
// A file-scope device array (each GPU holds its own copy for the life of the
// process): the per-tile partial sums (2 doubles per tile) and the per-domain
// residuals (1 double per domain).  Every call on a GPU shares that GPU's
// array; overlapping calls on one GPU race.
constexpr uint32_t kScratchTiles = 1u << 19;    // 8 MiB of partials
constexpr uint32_t kScratchDomains = 1u << 16;  // 512 KiB of residuals
constexpr size_t kScratchDoubles = 2ull * kScratchTiles + kScratchDomains;
static __device__ __align__(256) double scratch[kScratchDoubles];

// Device address of `scratch` on each GPU, looked up once per device by
// lbm_batch_create with that device current (the array's name is a host-side
// stand-in, not a device address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_error[kMaxDevices] = {};
bool g_scratch_resolved[kMaxDevices] = {};

hipError_t resolve_scratch(int dev, double **out)
{
    *out = nullptr;
    if (dev < 0 || dev >= kMaxDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_resolved[dev])
    {
        void *ptr = nullptr;
        g_scratch_error[dev] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        g_scratch_ptr[dev] = g_scratch_error[dev] == hipSuccess ? static_cast<double *>(ptr) : nullptr;
        g_scratch_resolved[dev] = true;
    }
    *out = g_scratch_ptr[dev];
    return g_scratch_error[dev];
}

// Per-domain residuals from the per-tile partials, in the order and with the
// arithmetic of lbm_batch_residuals' host loop: a domain's tiles are
// contiguous in the tile list (build_tiles), summed in list order; NaN for a
// domain whose sums are not finite.
__global__ void k_batch_residual_domains(const BatchTile *tiles, uint32_t ntiles, uint32_t ndoms,
                                         const double *partials, double *out)
{
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= ndoms)
        return;
    uint32_t lo = 0, hi = ntiles;  // first tile of domain d
    while (lo < hi)
    {
        const uint32_t mid = (lo + hi) / 2u;
        if (tiles[mid].dom < d)
            lo = mid + 1u;
        else
            hi = mid;
    }
    double num = 0.0, den = 0.0;
    for (uint32_t t = lo; t < ntiles && tiles[t].dom == d; ++t)
    {
        num += partials[2ull * t];
        den += partials[2ull * t + 1];
    }
    if (!isfinite(num) || !isfinite(den))
        out[d] = __builtin_nan("");
    else
        out[d] = den > 0.0 ? sqrt(num / den) : (num > 0.0 ? 1.0 : 0.0);
}

}  // namespace

struct lbm_batch
{
    std::vector<lbm_batch_domain> desc;
    std::vector<BatchDom> doms;
    std::vector<uint32_t> tile_first, tile_count;  // per domain, into the tile list
    uint32_t class_width[4] = {256, 256, 256, 256};
    int storage;
    int nan_protection;
    size_t elem;
    size_t arena_elems = 0, code_nodes = 0, res_doubles = 0, uy_doubles = 0, phase_nodes = 0;
    std::vector<size_t> phase_off;
    void *arena[2] = {nullptr, nullptr};
    uint16_t *codes = nullptr;
    int8_t *phase = nullptr;
    double *res = nullptr;
    double *uy = nullptr;
    BatchDom *d_doms = nullptr;
    BatchTile *d_tiles = nullptr;
    uint32_t ntiles = 0;
    uint32_t persistent_grid = 2048;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t iteration = 0;
    bool initialized = false;
    double last_ms = 0.0;
    uint64_t fluid_nodes = 0;
    hipGraphExec_t graph = nullptr;
    uint32_t graph_steps = 0;
    int graph_parity = -1;
    double *snap = nullptr;      // 2 doubles per node place (residual snapshot, allocated by lbm_batch_init)
    bool snap_valid = false;
    double *scratch = nullptr;   // device address of this GPU's `scratch` (null if the lookup failed)
    float *ones = nullptr;       // rho_ref of the 16-bit forms: max H + 64 ones
    hipError_t scratch_error = hipSuccess;
};

namespace
{
size_t store_bytes(int s) { return s == LBM_STORE_F64 ? 8 : s == LBM_STORE_F32 ? 4 : 2; }

void drop_graph(lbm_batch *b)
{
    if (b->graph)
        (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_steps = 0;
    b->graph_parity = -1;
}

// Tile list for the current class widths: per domain, rows of tiles in
// order, so that a grid striding over the list walks each domain row by row.
int build_tiles(lbm_batch *b)
{
    std::vector<BatchTile> tiles;
    for (size_t d = 0; d < b->doms.size(); ++d)
    {
        BatchDom &dm = b->doms[d];
        dm.tile_w = b->class_width[width_class(dm.H - 2)];
        const uint32_t rows = kTileThreads / dm.tile_w;
        b->tile_first[d] = uint32_t(tiles.size());
        for (uint32_t y = 1; y + 1 < dm.V; y += rows)
            for (uint32_t x = 1; x + 1 < dm.H; x += dm.tile_w)
                tiles.push_back(BatchTile{uint32_t(d), uint16_t(x), uint16_t(y)});
        b->tile_count[d] = uint32_t(tiles.size()) - b->tile_first[d];
    }
    (void)hipFree(b->d_tiles);
    b->d_tiles = nullptr;
    b->ntiles = uint32_t(tiles.size());
    BATCH_TRY(hipMalloc(&b->d_tiles, std::max<size_t>(1, tiles.size()) * sizeof(BatchTile)));
    BATCH_TRY(hipMemcpy(b->d_tiles, tiles.data(), tiles.size() * sizeof(BatchTile), hipMemcpyHostToDevice));
    BATCH_TRY(hipMemcpy(b->d_doms, b->doms.data(), b->doms.size() * sizeof(BatchDom), hipMemcpyHostToDevice));
    drop_graph(b);
    return LBM_OK;
}

template <int S>
int launch_step(lbm_batch *b, int parity, bool macros, const BatchTile *tiles, uint32_t ntiles, uint32_t grid)
{
    BatchArgs a{};
    a.ones = b->ones;
    a.doms = b->d_doms;
    a.tiles = tiles;
    a.ntiles = ntiles;
    a.src = b->arena[parity];
    a.dst = b->arena[1 - parity];
    a.codes = b->codes;
    a.res = b->res;
    a.uy = b->uy;
    a.parity = parity;
    a.write_macros = macros ? 1u : 0u;
    a.nan_protection = b->nan_protection ? 1u : 0u;
    hipLaunchKernelGGL((k_batch_step<S>), dim3(grid), dim3(kTileThreads), 0, b->stream, a);
    BATCH_TRY(hipGetLastError());
    return LBM_OK;
}

int step_dispatch(lbm_batch *b, int parity, bool macros, const BatchTile *tiles, uint32_t ntiles, uint32_t grid)
{
    switch (b->storage)
    {
    case LBM_STORE_F64: return launch_step<LBM_STORE_F64>(b, parity, macros, tiles, ntiles, grid);
    case LBM_STORE_F32: return launch_step<LBM_STORE_F32>(b, parity, macros, tiles, ntiles, grid);
    case LBM_STORE_F16_SHIFTED: return launch_step<LBM_STORE_F16_SHIFTED>(b, parity, macros, tiles, ntiles, grid);
    default: return launch_step<LBM_STORE_BF16_SHIFTED>(b, parity, macros, tiles, ntiles, grid);
    }
}

int run_persistent_steps(lbm_batch *b, uint32_t steps, uint64_t first_iteration)
{
    const uint32_t grid = std::max(1u, std::min(b->ntiles, b->persistent_grid));
    for (uint32_t i = 0; i < steps; ++i)
        if (int rc = step_dispatch(b, int((first_iteration + i) & 1u), i + 1 == steps, b->d_tiles, b->ntiles, grid))
            return rc;
    return LBM_OK;
}

// Rest populations and inlet/outlet ghost values of domain d in lattice l0.
template <int S>
int init_domain_lattice(lbm_batch *b, size_t d, typename Storage<S>::T *l0)
{
    const BatchDom &dm = b->doms[d];
    const lbm_batch_domain &ds = b->desc[d];
    const Geom g = batch_geom(dm, b->ones);
    hipLaunchKernelGGL((k_init_rest<LBM_LAYOUT_STREAM, S>), dim3((g.N + 255u) / 256u), dim3(256), 0, b->stream,
                       l0 + dm.lat, g);
    InoutValues v;
    host_equilibrium(S, ds.inlet_velocity_x, ds.inlet_velocity_y, ds.inlet_density, v.v);
    host_equilibrium(S, ds.outlet_velocity_x, ds.outlet_velocity_y, ds.outlet_density, v.v + 9);
    hipLaunchKernelGGL((k_init_inout<LBM_LAYOUT_STREAM, S>), dim3((dm.V - 2u + 255u) / 256u, 2), dim3(256), 0,
                       b->stream, l0 + dm.lat, dm.H, dm.V, g, dm.H, dm.V, 0, v);
    BATCH_TRY(hipGetLastError());
    return LBM_OK;
}

template <int S>
int init_domains(lbm_batch *b)
{
    using T = typename Storage<S>::T;
    T *l0 = static_cast<T *>(b->arena[0]);
    BATCH_TRY(hipMemsetAsync(b->arena[0], 0, b->arena_elems * sizeof(T), b->stream));
    BATCH_TRY(hipMemsetAsync(b->codes, 0, b->code_nodes * sizeof(uint16_t), b->stream));
    // results: density -1 (filled per domain below), momentum 0
    BATCH_TRY(hipMemsetAsync(b->res, 0, b->res_doubles * sizeof(double), b->stream));
    for (size_t d = 0; d < b->doms.size(); ++d)
    {
        const BatchDom &dm = b->doms[d];
        if (int rc = init_domain_lattice<S>(b, d, l0))
            return rc;
        const CodeParams cp{dm.H, dm.V, dm.H, dm.H * dm.V, dm.H, dm.V, dm.P, kBatchX0, 0, 1};
        hipLaunchKernelGGL(k_build_codes, dim3((cp.N + 255u) / 256u), dim3(256), 0, b->stream,
                           b->phase + b->phase_off[d], b->codes + dm.code, cp);
        hipLaunchKernelGGL(k_fill, dim3((dm.D + 255u) / 256u), dim3(256), 0, b->stream, b->res + dm.res,
                           size_t(dm.D), -1.0);
        BATCH_TRY(hipGetLastError());
    }
    BATCH_TRY(hipMemcpyAsync(b->arena[1], b->arena[0], b->arena_elems * sizeof(T), hipMemcpyDeviceToDevice,
                             b->stream));
    BATCH_TRY(hipMemsetAsync(b->uy, 0, b->uy_doubles * sizeof(double), b->stream));
    return LBM_OK;
}

}  // namespace

extern "C" {

void lbm_batch_destroy(lbm_batch *b)
{
    if (!b)
        return;
    if (b->stream)
        (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    (void)hipFree(b->arena[0]);
    (void)hipFree(b->arena[1]);
    (void)hipFree(b->codes);
    (void)hipFree(b->phase);
    (void)hipFree(b->res);
    (void)hipFree(b->uy);
    (void)hipFree(b->d_doms);
    (void)hipFree(b->d_tiles);
    (void)hipFree(b->snap);
    (void)hipFree(b->ones);
    if (b->ev0)
        (void)hipEventDestroy(b->ev0);
    if (b->ev1)
        (void)hipEventDestroy(b->ev1);
    if (b->stream)
        (void)hipStreamDestroy(b->stream);
    delete b;
}

int lbm_batch_create(const lbm_batch_domain *domains, uint32_t count, int32_t storage, int32_t nan_protection,
                     lbm_batch **out)
{
    if (!domains || !out || count == 0)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_create: no domains");
    if (storage < LBM_STORE_F64 || storage > LBM_STORE_BF16_SHIFTED)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_create: unknown storage type");
    *out = nullptr;
    lbm_batch *b = new lbm_batch();
    b->storage = storage;
    b->nan_protection = nan_protection;
    b->elem = store_bytes(storage);
    b->desc.assign(domains, domains + count);
    b->doms.resize(count);
    b->tile_first.resize(count);
    b->tile_count.resize(count);
    b->phase_off.resize(count);
    for (uint32_t d = 0; d < count; ++d)
    {
        const lbm_batch_domain &ds = domains[d];
        if (ds.horizontal_nodes < 4 || ds.vertical_nodes < 4 || ds.horizontal_nodes > 65535 ||
            ds.vertical_nodes > 65535 || !(ds.relaxation_time > 0.0))
        {
            delete b;
            return lbm_set_error(LBM_ERR_INVALID_ARGUMENT,
                                 "lbm_batch_create: domain " + std::to_string(d) + " has invalid extents or tau");
        }
        BatchDom &dm = b->doms[d];
        dm.H = ds.horizontal_nodes;
        dm.V = ds.vertical_nodes;
        dm.P = (kBatchX0 + dm.H + 15u) / 16u * 16u;
        dm.S = dm.P * dm.V;
        dm.D = (dm.H - 2u) * (dm.V - 2u);
        dm.lat = b->arena_elems;
        dm.code = uint32_t(b->code_nodes);
        dm.res = uint32_t(b->res_doubles);
        dm.uy = uint32_t(b->uy_doubles);
        dm.omega = host_omega(b->storage, ds.relaxation_time);
        dm.rho_out = ds.outlet_density;
        b->phase_off[d] = b->phase_nodes;
        b->arena_elems += 9ull * dm.S;
        b->code_nodes += dm.S;
        b->res_doubles += 4ull * dm.D;
        b->uy_doubles += 2ull * dm.V;
        b->phase_nodes += size_t(dm.H) * dm.V;
        if (b->code_nodes >= (1ull << 32) || b->res_doubles >= (1ull << 32) || 9ull * dm.S >= (1ull << 32))
        {
            delete b;
            return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_create: batch too large");
        }
    }
    auto cleanup = [&](int rc) {
        lbm_batch_destroy(b);
        return rc;
    };
#define CREATE_TRY(expr)                                                                                     \
    do                                                                                                       \
    {                                                                                                        \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess)                                                                                \
            return cleanup(lbm_set_error(e_ == hipErrorOutOfMemory ? LBM_ERR_OUT_OF_MEMORY : LBM_ERR_DEVICE, \
                                         std::string(#expr) + ": " + hipGetErrorString(e_)));                \
    } while (0)
    CREATE_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    CREATE_TRY(hipEventCreate(&b->ev0));
    CREATE_TRY(hipEventCreate(&b->ev1));
    CREATE_TRY(hipMalloc(&b->arena[0], b->arena_elems * b->elem));
    CREATE_TRY(hipMalloc(&b->arena[1], b->arena_elems * b->elem));
    CREATE_TRY(hipMalloc(&b->codes, b->code_nodes * sizeof(uint16_t)));
    CREATE_TRY(hipMalloc(&b->phase, b->phase_nodes));
    CREATE_TRY(hipMemset(b->phase, 0xff, b->phase_nodes));
    CREATE_TRY(hipMalloc(&b->res, b->res_doubles * sizeof(double)));
    CREATE_TRY(hipMalloc(&b->uy, b->uy_doubles * sizeof(double)));
    CREATE_TRY(hipMalloc(&b->d_doms, count * sizeof(BatchDom)));
    {
        uint32_t hmax = 0;
        for (const BatchDom &dm : b->doms)
            hmax = std::max(hmax, dm.H);
        const std::vector<float> ones(hmax + 64u, 1.f);
        CREATE_TRY(hipMalloc(&b->ones, ones.size() * sizeof(float)));
        CREATE_TRY(hipMemcpy(b->ones, ones.data(), ones.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    int dev = 0, cus = 256;
    CREATE_TRY(hipGetDevice(&dev));
    CREATE_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    b->persistent_grid = uint32_t(std::max(1, cus)) * 8u;  // 8 resident 256-item groups per CU
    b->scratch_error = resolve_scratch(dev, &b->scratch);   // reported by lbm_batch_residuals_noalloc
#undef CREATE_TRY
    if (int rc = build_tiles(b))
        return cleanup(rc);
    *out = b;
    return LBM_OK;
}

int lbm_batch_set_phase(lbm_batch *b, uint32_t domain, const int8_t *phase, size_t n)
{
    if (!b || !phase || domain >= b->doms.size())
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_set_phase: bad argument");
    if (n != size_t(b->doms[domain].H) * b->doms[domain].V)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_set_phase: node_count must be H * V");
    BATCH_TRY(hipMemcpy(b->phase + b->phase_off[domain], phase, n, hipMemcpyHostToDevice));
    b->initialized = false;
    return LBM_OK;
}

int lbm_batch_init(lbm_batch *b)
{
    if (!b)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_init: null batch");
    int rc = LBM_OK;
    switch (b->storage)
    {
    case LBM_STORE_F64: rc = init_domains<LBM_STORE_F64>(b); break;
    case LBM_STORE_F32: rc = init_domains<LBM_STORE_F32>(b); break;
    case LBM_STORE_F16_SHIFTED: rc = init_domains<LBM_STORE_F16_SHIFTED>(b); break;
    default: rc = init_domains<LBM_STORE_BF16_SHIFTED>(b); break;
    }
    if (rc)
        return rc;
    BATCH_TRY(hipStreamSynchronize(b->stream));
    // fluid node count from the codes
    std::vector<uint16_t> codes(b->code_nodes);
    BATCH_TRY(hipMemcpy(codes.data(), b->codes, codes.size() * sizeof(uint16_t), hipMemcpyDeviceToHost));
    b->fluid_nodes = uint64_t(std::count_if(codes.begin(), codes.end(), [](uint16_t c) { return c & LBM_CODE_FLUID; }));
    // the residual snapshot (2 doubles per node place) is allocated here, at
    // set-up, so that lbm_batch_residuals_noalloc never allocates
    if (!b->snap)
        BATCH_TRY(hipMalloc(&b->snap, 2ull * b->code_nodes * sizeof(double)));
    b->iteration = 0;
    b->initialized = true;
    b->snap_valid = false;
    return LBM_OK;
}

int lbm_batch_run(lbm_batch *b, uint32_t steps, int32_t mode)
{
    if (!b)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_run: null batch");
    if (!b->initialized)
        return lbm_set_error(LBM_ERR_STATE, "lbm_batch_run: call lbm_batch_init first");
    if (mode < LBM_BATCH_PER_DOMAIN || mode > LBM_BATCH_GRAPH)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_run: unknown mode");
    if (steps == 0)
        return LBM_OK;
    const int parity0 = int(b->iteration & 1u);
    if (mode == LBM_BATCH_GRAPH && (b->graph_steps != steps || b->graph_parity != parity0))
    {
        // capture the persistent launches of the whole run once per (steps, parity)
        drop_graph(b);
        hipGraph_t g = nullptr;
        BATCH_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = run_persistent_steps(b, steps, b->iteration);
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc)
            return rc;
        if (e != hipSuccess)
            return lbm_set_error(LBM_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess)
            return lbm_set_error(LBM_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        b->graph_steps = steps;
        b->graph_parity = parity0;
    }
    BATCH_TRY(hipEventRecord(b->ev0, b->stream));
    if (mode == LBM_BATCH_GRAPH)
        BATCH_TRY(hipGraphLaunch(b->graph, b->stream));
    else if (mode == LBM_BATCH_PERSISTENT)
    {
        if (int rc = run_persistent_steps(b, steps, b->iteration))
            return rc;
    }
    else
    {
        for (uint32_t i = 0; i < steps; ++i)
            for (size_t d = 0; d < b->doms.size(); ++d)
                if (b->tile_count[d])
                    if (int rc = step_dispatch(b, int((b->iteration + i) & 1u), i + 1 == steps,
                                               b->d_tiles + b->tile_first[d], b->tile_count[d], b->tile_count[d]))
                        return rc;
    }
    BATCH_TRY(hipEventRecord(b->ev1, b->stream));
    BATCH_TRY(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    BATCH_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    b->iteration += steps;
    return LBM_OK;
}

int lbm_batch_last_run_ms(lbm_batch *b, double *ms)
{
    if (!b || !ms)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_last_run_ms: null argument");
    *ms = b->last_ms;
    return LBM_OK;
}

int lbm_batch_get_macros(lbm_batch *b, uint32_t domain, double *rho, double *jx, double *jy, double *jabs,
                         size_t count)
{
    if (!b || domain >= b->doms.size())
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_get_macros: bad argument");
    const BatchDom &dm = b->doms[domain];
    if (count != dm.D)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_get_macros: count must be (H-2)*(V-2)");
    double *dst[4] = {rho, jx, jy, jabs};
    for (int k = 0; k < 4; ++k)
        if (dst[k])
            BATCH_TRY(hipMemcpyAsync(dst[k], b->res + dm.res + size_t(k) * dm.D, count * sizeof(double),
                                     hipMemcpyDeviceToHost, b->stream));
    BATCH_TRY(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

int lbm_batch_fluid_nodes(lbm_batch *b, uint64_t *count)
{
    if (!b || !count)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_fluid_nodes: null argument");
    *count = b->fluid_nodes;
    return LBM_OK;
}

int lbm_batch_set_tiles(lbm_batch *b, const uint32_t tile_width[4])
{
    if (!b || !tile_width)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_set_tiles: null argument");
    for (int c = 0; c < 4; ++c)
        if (tile_width[c] != 64 && tile_width[c] != 128 && tile_width[c] != 256)
            return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_set_tiles: widths must be 64, 128 or 256");
    std::copy(tile_width, tile_width + 4, b->class_width);
    return build_tiles(b);
}

int lbm_batch_get_tiles(lbm_batch *b, uint32_t tile_width[4])
{
    if (!b || !tile_width)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_get_tiles: null argument");
    std::copy(b->class_width, b->class_width + 4, tile_width);
    return LBM_OK;
}

int lbm_batch_autotune(lbm_batch *b, uint32_t steps_per_trial, double *best_ms_per_step)
{
    if (!b || steps_per_trial == 0)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_autotune: bad argument");
    if (!b->initialized)
        return lbm_set_error(LBM_ERR_STATE, "lbm_batch_autotune: call lbm_batch_init first");
    static const uint32_t candidates[3] = {64, 128, 256};
    bool present[4] = {false, false, false, false};
    for (const BatchDom &dm : b->doms)
        present[width_class(dm.H - 2)] = true;
    double best = 1e300;
    for (int c = 0; c < 4; ++c)
    {
        if (!present[c])
            continue;
        uint32_t best_w = b->class_width[c];
        for (uint32_t w : candidates)
        {
            uint32_t widths[4];
            std::copy(b->class_width, b->class_width + 4, widths);
            widths[c] = w;
            if (int rc = lbm_batch_set_tiles(b, widths))
                return rc;
            if (int rc = lbm_batch_run(b, 2, LBM_BATCH_PERSISTENT))  // warm-up
                return rc;
            if (int rc = lbm_batch_run(b, steps_per_trial, LBM_BATCH_PERSISTENT))
                return rc;
            const double per_step = b->last_ms / steps_per_trial;
            if (per_step < best)
            {
                best = per_step;
                best_w = w;
            }
        }
        uint32_t widths[4];
        std::copy(b->class_width, b->class_width + 4, widths);
        widths[c] = best_w;
        if (int rc = lbm_batch_set_tiles(b, widths))
            return rc;
    }
    if (best_ms_per_step)
        *best_ms_per_step = best;
    return lbm_batch_init(b);
}

int lbm_batch_residuals(lbm_batch *b, double *out, uint32_t count)
{
    if (!b || !out || count != b->doms.size())
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_residuals: count must equal the domain count");
    if (!b->initialized)
        return lbm_set_error(LBM_ERR_STATE, "lbm_batch_residuals: call lbm_batch_init first");
    if (!b->snap)
        BATCH_TRY(hipMalloc(&b->snap, 2ull * b->code_nodes * sizeof(double)));
    if (!b->snap_valid)
    {
        BATCH_TRY(hipMemsetAsync(b->snap, 0, 2ull * b->code_nodes * sizeof(double), b->stream));
        b->snap_valid = true;
    }
    double *partials = nullptr;
    BATCH_TRY(hipMalloc(&partials, 2ull * std::max(1u, b->ntiles) * sizeof(double)));
    const void *src = b->arena[b->iteration & 1u];
    switch (b->storage)
    {
    case LBM_STORE_F64:
        hipLaunchKernelGGL((k_batch_residual<LBM_STORE_F64>), dim3(b->ntiles), dim3(kTileThreads), 0, b->stream,
                           b->d_doms, b->d_tiles, src, b->codes, b->snap, partials);
        break;
    case LBM_STORE_F32:
        hipLaunchKernelGGL((k_batch_residual<LBM_STORE_F32>), dim3(b->ntiles), dim3(kTileThreads), 0, b->stream,
                           b->d_doms, b->d_tiles, src, b->codes, b->snap, partials);
        break;
    case LBM_STORE_F16_SHIFTED:
        hipLaunchKernelGGL((k_batch_residual<LBM_STORE_F16_SHIFTED>), dim3(b->ntiles), dim3(kTileThreads), 0,
                           b->stream, b->d_doms, b->d_tiles, src, b->codes, b->snap, partials);
        break;
    default:
        hipLaunchKernelGGL((k_batch_residual<LBM_STORE_BF16_SHIFTED>), dim3(b->ntiles), dim3(kTileThreads), 0,
                           b->stream, b->d_doms, b->d_tiles, src, b->codes, b->snap, partials);
        break;
    }
    hipError_t e = hipGetLastError();
    std::vector<double> host(2ull * b->ntiles);
    if (e == hipSuccess)
        e = hipMemcpyAsync(host.data(), partials, host.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(b->stream);
    (void)hipFree(partials);
    if (e != hipSuccess)
        return lbm_set_error(LBM_ERR_DEVICE, std::string("lbm_batch_residuals: ") + hipGetErrorString(e));
    for (size_t d = 0; d < b->doms.size(); ++d)
    {
        double num = 0.0, den = 0.0;
        for (uint32_t t = b->tile_first[d]; t < b->tile_first[d] + b->tile_count[d]; ++t)
        {
            num += host[2ull * t];
            den += host[2ull * t + 1];
        }
        if (!std::isfinite(num) || !std::isfinite(den))
            out[d] = std::numeric_limits<double>::quiet_NaN();  // a diverged domain never counts as converged
        else
            out[d] = den > 0.0 ? std::sqrt(num / den) : (num > 0.0 ? 1.0 : 0.0);
    }
    return LBM_OK;
}

int lbm_batch_residuals_noalloc(lbm_batch *b, double *out, uint32_t count)
{
    if (!b || !out || count != b->doms.size())
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT,
                             "lbm_batch_residuals_noalloc: count must equal the domain count");
    if (!b->initialized)
        return lbm_set_error(LBM_ERR_STATE, "lbm_batch_residuals_noalloc: call lbm_batch_init first");
    if (b->ntiles > kScratchTiles || count > kScratchDomains)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT,
                             "lbm_batch_residuals_noalloc: " + std::to_string(b->ntiles) + " tiles and " +
                                 std::to_string(count) + " domains exceed the static scratch (at most " +
                                 std::to_string(kScratchTiles) + " tiles and " + std::to_string(kScratchDomains) +
                                 " domains); use lbm_batch_residuals");
    if (!b->scratch)
        return lbm_set_error(LBM_ERR_DEVICE, std::string("lbm_batch_residuals_noalloc: hipGetSymbolAddress: ") +
                                                 hipGetErrorString(b->scratch_error));
    if (!b->snap)
        return lbm_set_error(LBM_ERR_STATE, "lbm_batch_residuals_noalloc: no residual snapshot (lbm_batch_init "
                                            "allocates it)");
    if (!b->snap_valid)
    {
        BATCH_TRY(hipMemsetAsync(b->snap, 0, 2ull * b->code_nodes * sizeof(double), b->stream));
        b->snap_valid = true;
    }
    double *partials = b->scratch;
    double *residuals = b->scratch + 2ull * kScratchTiles;
    const void *src = b->arena[b->iteration & 1u];
    switch (b->storage)
    {
    case LBM_STORE_F64:
        hipLaunchKernelGGL((k_batch_residual<LBM_STORE_F64>), dim3(b->ntiles), dim3(kTileThreads), 0, b->stream,
                           b->d_doms, b->d_tiles, src, b->codes, b->snap, partials);
        break;
    case LBM_STORE_F32:
        hipLaunchKernelGGL((k_batch_residual<LBM_STORE_F32>), dim3(b->ntiles), dim3(kTileThreads), 0, b->stream,
                           b->d_doms, b->d_tiles, src, b->codes, b->snap, partials);
        break;
    case LBM_STORE_F16_SHIFTED:
        hipLaunchKernelGGL((k_batch_residual<LBM_STORE_F16_SHIFTED>), dim3(b->ntiles), dim3(kTileThreads), 0,
                           b->stream, b->d_doms, b->d_tiles, src, b->codes, b->snap, partials);
        break;
    default:
        hipLaunchKernelGGL((k_batch_residual<LBM_STORE_BF16_SHIFTED>), dim3(b->ntiles), dim3(kTileThreads), 0,
                           b->stream, b->d_doms, b->d_tiles, src, b->codes, b->snap, partials);
        break;
    }
    BATCH_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_batch_residual_domains, dim3((count + 255u) / 256u), dim3(256), 0, b->stream, b->d_tiles,
                       b->ntiles, count, partials, residuals);
    BATCH_TRY(hipGetLastError());
    BATCH_TRY(hipMemcpyAsync(out, residuals, count * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    BATCH_TRY(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

int lbm_batch_reset_domain(lbm_batch *b, uint32_t domain)
{
    if (!b || domain >= b->doms.size())
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_reset_domain: bad domain");
    if (!b->initialized)
        return lbm_set_error(LBM_ERR_STATE, "lbm_batch_reset_domain: call lbm_batch_init first");
    const BatchDom &dm = b->doms[domain];
    void *lat = b->arena[b->iteration & 1u];
    int rc = LBM_OK;
    switch (b->storage)
    {
    case LBM_STORE_F64: rc = init_domain_lattice<LBM_STORE_F64>(b, domain, static_cast<double *>(lat)); break;
    case LBM_STORE_F32: rc = init_domain_lattice<LBM_STORE_F32>(b, domain, static_cast<float *>(lat)); break;
    case LBM_STORE_F16_SHIFTED:
        rc = init_domain_lattice<LBM_STORE_F16_SHIFTED>(b, domain, static_cast<_Float16 *>(lat));
        break;
    default: rc = init_domain_lattice<LBM_STORE_BF16_SHIFTED>(b, domain, static_cast<uint16_t *>(lat)); break;
    }
    if (rc)
        return rc;
    BATCH_TRY(hipMemsetAsync(b->uy + dm.uy, 0, 2ull * dm.V * sizeof(double), b->stream));
    BATCH_TRY(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

int lbm_batch_set_equilibrium(lbm_batch *b, uint32_t domain, const double *rho, const double *jx, const double *jy,
                              size_t count)
{
    if (!b || !rho || !jx || !jy || domain >= b->doms.size())
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_set_equilibrium: bad argument");
    if (!b->initialized)
        return lbm_set_error(LBM_ERR_STATE, "lbm_batch_set_equilibrium: call lbm_batch_init first");
    const BatchDom &dm = b->doms[domain];
    if (count != dm.D)
        return lbm_set_error(LBM_ERR_INVALID_ARGUMENT, "lbm_batch_set_equilibrium: count must be (H-2)*(V-2)");
    double *tmp = nullptr;
    BATCH_TRY(hipMalloc(&tmp, 3 * count * sizeof(double)));
    hipError_t e = hipSuccess;
    const double *src[3] = {rho, jx, jy};
    for (int k = 0; k < 3 && e == hipSuccess; ++k)
        e = hipMemcpyAsync(tmp + k * count, src[k], count * sizeof(double), hipMemcpyHostToDevice, b->stream);
    const Geom g = batch_geom(dm, b->ones);
    void *lat = b->arena[b->iteration & 1u];
    if (e == hipSuccess)
    {
        const dim3 grid((g.N + 255u) / 256u), block(256);
        switch (b->storage)
        {
        case LBM_STORE_F64:
            hipLaunchKernelGGL((k_set_equilibrium<LBM_LAYOUT_STREAM, LBM_STORE_F64>), grid, block, 0, b->stream,
                               static_cast<double *>(lat) + dm.lat, b->codes + dm.code, tmp, tmp + count,
                               tmp + 2 * count, g, dm.H, dm.H, dm.V, 0);
            break;
        case LBM_STORE_F32:
            hipLaunchKernelGGL((k_set_equilibrium<LBM_LAYOUT_STREAM, LBM_STORE_F32>), grid, block, 0, b->stream,
                               static_cast<float *>(lat) + dm.lat, b->codes + dm.code, tmp, tmp + count,
                               tmp + 2 * count, g, dm.H, dm.H, dm.V, 0);
            break;
        case LBM_STORE_F16_SHIFTED:
            hipLaunchKernelGGL((k_set_equilibrium<LBM_LAYOUT_STREAM, LBM_STORE_F16_SHIFTED>), grid, block, 0,
                               b->stream, static_cast<_Float16 *>(lat) + dm.lat, b->codes + dm.code, tmp, tmp + count,
                               tmp + 2 * count, g, dm.H, dm.H, dm.V, 0);
            break;
        default:
            hipLaunchKernelGGL((k_set_equilibrium<LBM_LAYOUT_STREAM, LBM_STORE_BF16_SHIFTED>), grid, block, 0,
                               b->stream, static_cast<uint16_t *>(lat) + dm.lat, b->codes + dm.code, tmp, tmp + count,
                               tmp + 2 * count, g, dm.H, dm.H, dm.V, 0);
            break;
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(b->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess)
        return lbm_set_error(LBM_ERR_DEVICE, std::string("lbm_batch_set_equilibrium: ") + hipGetErrorString(e));
    return LBM_OK;
}

}  // extern "C"
