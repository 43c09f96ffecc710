// MI355X (gfx950) implementation of the C ABI declared in include/lbm_hip.h.
//
// One launch per time step fuses what the reference runs as three to seven
// blocking launches (include/lbm/gpu/nptl/nptl.hpp:116-141,
// npol/npol.hpp:180-201, nsol/nsol.hpp:184-205): bounce-back is folded into
// the pull through the per-node code table, the Zou-He outlet is computed by
// the outlet nodes themselves, and macroscopic outputs are written only on
// the steps whose values the host can observe.  The one-lattice schemes keep
// the reference's buffered geometry and need one halo refresh before each
// step (two small launches).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lbm_device.hpp"
#include "lbm_error.hpp"
#include "lbm_setup.hpp"

using namespace lbm_dev;

// ---------------------------------------------------------------------------
// error handling
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

int lbm_set_error(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

static int fail(int code, const std::string &msg) { return lbm_set_error(code, msg); }

#define LBM_HIP_TRY(expr)                                                                          \
    do                                                                                             \
    {                                                                                              \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(e_ == hipErrorOutOfMemory ? LBM_ERR_OUT_OF_MEMORY : LBM_ERR_DEVICE,        \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                        \
    } while (0)

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct lbm_ctx
{
    lbm_desc desc;
    uint32_t H, V, Hx, Vx, N;
    uint32_t P, X0, Sd;       // physical placement: pitch, row offset, direction stride
    size_t lattice_elems;     // stored populations per lattice
    uint32_t sub_h, sub_v, cnt_h, cnt_v, wg;
    bool buffered;
    bool two_lattice;
    bool outlet_active;
    // npol/nsol tiles at most one wave wide run k_step_tile_sweep, which keeps
    // the buffer copies in side arrays (two of each kind, used in turn)
    bool sweep = false;
    // npol/nsol in the collision layout, tiles 8 to 64 nodes wide: k_step_tile_staged
    bool tile_staged = false;
    uint32_t side_rows[2] = {0, 0};  // element offsets from lat[0]
    uint32_t side_cols[2] = {0, 0};
    uint32_t side_parity = 0;
    bool side_valid = false;  // false: rebuild them from the lattice before the next sweep
    size_t elem;          // bytes per stored population
    void *lat[2] = {nullptr, nullptr};
    int cur = 0;          // index of the source lattice
    int8_t *phase = nullptr;
    uint16_t *code = nullptr;
    double *res = nullptr;    // 4 * D: rho, jx, jy, |j|
    double *uycol = nullptr;  // 2 * V outlet uy history (ping-pong)
    double *snap = nullptr;   // 2 * N residual snapshot (lazy)
    double *pass = nullptr;   // 9 * P * Vx FP64 pass buffer of split steps (lazy)
    double *stage = nullptr;  // population accessors' staging buffer (lazy, at most kStageDoubles)
    double *partials = nullptr;
    double *red_out = nullptr;
    uint32_t D;               // (H-2)*(V-2)
    hipStream_t stream = nullptr;
    // asynchronous frame copies (lbm_run_frames; created on first use)
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_frame = nullptr, ev_copied = nullptr;
    double *frame_host = nullptr;  // pinned, 4 * D
    bool copy_pending = false;     // a copy of the results may still be reading them
    uint64_t frame_iteration = ~0ull;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t iteration = 0;
    bool phase_set = false;
    bool initialized = false;
    bool snap_valid = false;
    double last_ms = 0.0;
    double feq_in[9], feq_out[9];
    // 16-bit forms: rho_ref per unexpanded column (Hx + 64 floats, the tail
    // 1), two tables used in turn; a step that changes the reference reads
    // with ref[ref_cur] and writes with ref[1 - ref_cur]
    float *ref[2] = {nullptr, nullptr};
    int ref_cur = 0;
    bool ref_switch = false;
    double *ref_part = nullptr;  // column sums of the update (kRefChunks * Hx * 2)
    uint32_t ref_len = 0;
};

// ---------------------------------------------------------------------------
// step kernels
// ---------------------------------------------------------------------------
constexpr int NPTL_BLOCK_X = 256;

// nptl: two lattices, 2-D grid over the expanded interior x in [1, Hx-2],
// y in [1, Vx-2] (nptl.hpp:63-67); expanded and unexpanded coordinates agree.
// A thread updates nptl_rows<S>() vertically adjacent nodes, issuing all their
// pulls before any arithmetic: with 2-byte populations one row per wave keeps
// too few bytes in flight to cover HBM latency (two rows: +5.5 % on the
// 8192 x 2048 porous channel in FP16, three: +1 % more, four: -1 %; FP32 and
// FP64 lose 2-4 % with two and keep one).
template <int S>
constexpr uint32_t nptl_rows()
{
    return sizeof(typename Storage<S>::T) == 2 ? 3u : 1u;
}

template <int L, int S>
__global__ __launch_bounds__(NPTL_BLOCK_X) void k_step_nptl(StepParams p)
{
    using T = typename Storage<S>::T;
    using C = typename Storage<S>::C;
    constexpr uint32_t R = nptl_rows<S>();
    const uint32_t x = 1u + blockIdx.x * NPTL_BLOCK_X + threadIdx.x;
    if (x > p.Hx - 2u)
        return;
    const uint32_t y0 = 1u + blockIdx.y * R;
    const Rt3 r3 = rt3_of<S>(p.ref_in, x);                  // the same column for every row
    const float orel = rt_of<S>(p.ref_out, x) - r3.own();  // stores: against the written reference
    uint32_t pn[R], code[R];
    bool fluid[R], fill[R];
    C f[R][9];
#pragma unroll
    for (uint32_t r = 0; r < R; ++r)
    {
        pn[r] = place(x, y0 + r, p.P, p.X0);
        code[r] = y0 + r <= p.Vx - 2u ? p.code[pn[r]] : 0u;
        fluid[r] = code[r] & LBM_CODE_FLUID;
        const bool segment = segment_has<T>(fluid[r]);  // every lane takes part in the ballot
        fill[r] = (code[r] & LBM_CODE_SOLID) && segment;
#pragma unroll
        for (int d = 0; d < 9; ++d)
            f[r][d] = C(0);
    }
#pragma unroll
    for (uint32_t r = 0; r < R; ++r)
        if (fluid[r])
            pull<L, S>(static_cast<const T *>(p.src), pn[r], code[r], p.P, p.S, f[r], r3);
#pragma unroll
    for (uint32_t r = 0; r < R; ++r)
    {
        if (fluid[r])
            finish_node<L, S>(p, code[r], x, y0 + r, f[r], r3.own());
        if (fluid[r] || fill[r])
            store_node<L, S>(static_cast<T *>(p.dst), pn[r], p.S, f[r], p.sr, orel);
    }
}

// nptl for the collision and bundle layouts.  There a pull instruction of a
// wave reads every 9th (3rd) element, so each touches nine (three) times the
// cache lines it uses.  Here the block first copies, with aligned 16-byte
// loads, the source ranges its row segments pull from into LDS:
//   collision: R node rows y0 .. y0+R-1 per block; range r = the nine
//              populations of row y0 - 1 + r (places x0-1 .. x0+len), R + 2
//              ranges, and row ty pulls direction d from range ty + 1 - cy.
//              Each staged row serves up to three node rows of the block, so
//              a row is fetched (R + 2) / R times instead of three times;
//   bundle:    one node row y per block; range r = array r (directions
//              3r..3r+2, cy = r - 1) of row y - cy, the only row those
//              directions are pulled from.
// Bounce-back pulls take the node's own f[8-d]: in the collision layout from
// the node's own range; in the bundle layout from range 1 when 8-d is in
// array 1, else straight from HBM (nodes next to a wall only).  The results
// go through LDS as well and leave as contiguous 16-byte stores: each node
// row's nine-population array (collision) or the row's three arrays (bundle)
// are written whole (solid nodes as zeros).  A block that holds a node that
// is neither fluid nor solid (the walls stencil's value 2 on tiny grids)
// writes node by node instead and leaves that node untouched.
// Used where it measured faster than the plain kernel on the 8192 x 2048
// porous channel (64-node segments; GFLUPS plain -> staged, range-by-range
// copy into LDS): FP64 collision 16.3 -> 21.2 (R = 8; R = 2, 4: 18.3, 20.3),
// FP32 collision 25.5 -> 45.7 (R = 4; R = 1, 2, 8: 34.0, 41.0, 34.6), FP16
// collision 43.6 -> 64.9 (R = 2; R = 1, 4, 8: 62.7, 53.3, 38.7), FP64 bundle
// 24.2 -> 28.1, FP32 bundle 42.8 -> 48.8; the flat copy below then takes FP32
// and FP16 collision to 50.2 and 67.4.  The 16-bit bundle and stream layouts
// keep the plain kernel (staging the stream layout measured 58.5 against
// 88.8 GFLUPS in FP16: the bounce-back pulls of the six directions with
// cy != 0 read rows that are not staged; staging the bundle layout 53.8, or
// 61.9 with the flat copy, against 63.7 in FP16, and 52.4 / 58.8 against
// 64.0 in BF16).  The collision layout's row segments were then narrowed to
// 32 nodes in FP64 and 16-bit storage (below).
constexpr uint32_t NPTL_STAGED_BLOCK = 64;

// Nodes per row segment of a block (a divisor of 64, so that segments start
// on 64-byte boundaries and a wave holds whole rows).  32 x R blocks measured
// faster than 64 x R in the FP64 and 16-bit collision layout on the porous
// 8192 x 2048 channel (GFLUPS): FP64 32 x 8 32.2 against 64 x 8 31.9 (32 x
// 16: 30.5); FP16 32 x 4 65.5 against 64 x 2 61.7 (32 x 2: 63.5, 32 x 8:
// 59.9); FP32 32 x 4 and 64 x 4 49.2 both (32 x 8: 47.0).
template <int L, int S>
constexpr uint32_t nptl_staged_width()
{
    if constexpr (L == LBM_LAYOUT_COLLISION && S != LBM_STORE_F32)
        return 32u;
    else
        return NPTL_STAGED_BLOCK;
}

template <int L, int S>
constexpr bool nptl_staged()
{
    if constexpr (L == LBM_LAYOUT_COLLISION)
        return true;
    else if constexpr (L == LBM_LAYOUT_BUNDLE)
        return sizeof(typename Storage<S>::T) >= 4;
    else
        return false;
}

// How the source ranges reach LDS: "flat" makes every 16-byte vector of every
// range one work item; otherwise thread i takes vector i of each range.
// Both issue all loads of a thread before its LDS writes (one wait for HBM
// per block).  Flat measured faster for the collision layout in FP32 (26.9
// -> 49.5 GFLUPS against one vector per thread and range), FP16 (59.1 ->
// 61.6) and BF16 (56.2 -> 58.2), slower for FP64 collision (31.9 -> 11.8)
// and FP64 bundle (31.8 -> 20.3).
template <int L, int S>
constexpr bool nptl_staged_flat()
{
    return L == LBM_LAYOUT_COLLISION && sizeof(typename Storage<S>::T) <= 4;
}

template <int L, int S>
constexpr uint32_t nptl_staged_rows()
{
    if constexpr (L != LBM_LAYOUT_COLLISION)
        return 1u;
    else if constexpr (S == LBM_STORE_F64)
        return 8u;
    else
        return 4u;
}

template <int L, int S>
__global__ __launch_bounds__((nptl_staged_width<L, S>() * nptl_staged_rows<L, S>())) void k_step_nptl_staged(StepParams p)
{
    using T = typename Storage<S>::T;
    using C = typename Storage<S>::C;
    using V4 = unsigned int __attribute__((ext_vector_type(4)));
    constexpr uint32_t B = nptl_staged_width<L, S>();
    constexpr uint32_t R = nptl_staged_rows<L, S>();               // node rows per block
    constexpr uint32_t NT = B * R;                                 // threads per block
    constexpr uint32_t VE = 16u / sizeof(T);                       // elements per 16-byte vector
    constexpr uint32_t K = L == LBM_LAYOUT_COLLISION ? 9u : 3u;    // staged populations per place
    constexpr uint32_t NR = L == LBM_LAYOUT_COLLISION ? R + 2u : 3u;  // staged ranges
    constexpr uint32_t ROW = (K * (B + 2u) + 2u * VE + VE - 1u) / VE * VE;
    __shared__ __align__(16) T stage[NR * ROW];
    const T *src = static_cast<const T *>(p.src);
    T *dst = static_cast<T *>(p.dst);
    const uint32_t tid = threadIdx.x, tx = tid % B, ty = tid / B;
    const uint32_t x0 = 1u + blockIdx.x * B, y0 = 1u + blockIdx.y * R;
    const uint32_t len = min(B, p.Hx - 1u - x0);
    const uint32_t rows = min(R, p.Vx - 1u - y0);                  // node rows of this block
    const uint32_t b0 = place(x0, y0, p.P, p.X0);
    // first element of range r: collision: row y0 - 1 + r; bundle (R = 1):
    // array r of row y0 + 1 - r
    auto range_start = [&](uint32_t r) -> uint32_t {
        if constexpr (L == LBM_LAYOUT_COLLISION)
            return 9u * (b0 + r * p.P - p.P - 1u);
        else
            return 3u * r * p.S + 3u * (b0 + p.P - r * p.P - 1u);
    };
    const uint32_t staged = L == LBM_LAYOUT_COLLISION ? rows + 2u : NR;
    uint32_t shift[NR];
#pragma unroll
    for (uint32_t r = 0; r < NR; ++r)
        shift[r] = range_start(r) % VE;
    if constexpr (nptl_staged_flat<L, S>())
    {
        // Every vector of every range is one work item, at most NVR per range;
        // a thread issues all of its loads before it writes any to LDS, so
        // the block waits for HBM once.
        constexpr uint32_t NVR = (K * (B + 2u) + 2u * (VE - 1u)) / VE;
        constexpr uint32_t PT = (NR * NVR + NT - 1u) / NT;
        const uint32_t nvec_max = (K * (len + 2u) + 2u * (VE - 1u)) / VE;
        V4 tmp[PT];
        uint32_t slot[PT];
#pragma unroll
        for (uint32_t j = 0; j < PT; ++j)
        {
            const uint32_t v = tid + j * NT;
            const uint32_t r = v / NVR, i = v - r * NVR;
            slot[j] = ~0u;
            if (r < staged && i < nvec_max)
            {
                const uint32_t e0 = range_start(r);
                const uint32_t ea = e0 - e0 % VE;
                if (i < (e0 - ea + K * (len + 2u) + VE - 1u) / VE)
                {
                    tmp[j] = reinterpret_cast<const V4 *>(src + ea)[i];
                    slot[j] = r * (ROW / VE) + i;
                }
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < PT; ++j)
            if (slot[j] != ~0u)
                reinterpret_cast<V4 *>(stage)[slot[j]] = tmp[j];
    }
    else
    {
        // Vector i of every range to thread i (mod NT); a thread issues the
        // loads of all its ranges before any LDS write, so the block waits
        // for HBM once (range by range it waited NR times: FP64 collision
        // 23.2 -> 31.9 GFLUPS, FP64 bundle 30.0 -> 31.8 on the porous
        // channel).
        constexpr uint32_t NVR = (K * (B + 2u) + 2u * (VE - 1u)) / VE;
        constexpr uint32_t KV = (NVR + NT - 1u) / NT;
        V4 tmp[NR][KV];
#pragma unroll
        for (uint32_t r = 0; r < NR; ++r)
        {
            const uint32_t ea = range_start(r) - shift[r];
            const uint32_t nvec = (shift[r] + K * (len + 2u) + VE - 1u) / VE;
            const V4 *g = reinterpret_cast<const V4 *>(src + ea);
#pragma unroll
            for (uint32_t k = 0; k < KV; ++k)
            {
                const uint32_t i = tid + k * NT;
                if (r < staged && i < nvec)
                    tmp[r][k] = g[i];
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < NR; ++r)
        {
            const uint32_t nvec = (shift[r] + K * (len + 2u) + VE - 1u) / VE;
            V4 *l = reinterpret_cast<V4 *>(stage + r * ROW);
#pragma unroll
            for (uint32_t k = 0; k < KV; ++k)
            {
                const uint32_t i = tid + k * NT;
                if (r < staged && i < nvec)
                    l[i] = tmp[r][k];
            }
        }
    }
    __syncthreads();
    const uint32_t y = y0 + ty;
    const uint32_t pn = b0 + ty * p.P + tx;
    const bool active = tx < len && ty < rows;
    const uint32_t code = active ? p.code[pn] : 0u;
    const bool fluid = code & LBM_CODE_FLUID;
    const bool solid = code & LBM_CODE_SOLID;
    // rho_ref of columns x - 1, x, x + 1 (16-bit forms), read and written
    const Rt3 r3 = rt3_of<S>(p.ref_in, x0 + tx);
    const float rto = rt_of<S>(p.ref_out, x0 + tx) - r3.own();  // stores: against the written reference
    C f[9] = {};
    if (fluid)
    {
#pragma unroll
        for (int d = 0; d < 9; ++d)
        {
            const bool bb = (code >> d) & 1u;
            const float rt = bb ? 0.f : r3.rel(cx_of(d));
            if constexpr (L == LBM_LAYOUT_COLLISION)
            {
                const uint32_t r = bb ? ty + 1u : uint32_t(int32_t(ty) + 1 - cy_of(d));
                const uint32_t i = bb ? tx + 1u : uint32_t(int32_t(tx) + 1 - cx_of(d));
                const int dd = bb ? 8 - d : d;
                f[d] = Storage<S>::load(stage + r * ROW + shift[r], 9u * i + uint32_t(dd), dd, rt);
            }
            else
            {
                if (!bb)
                {
                    const uint32_t r = uint32_t(d / 3);
                    const uint32_t i = uint32_t(int32_t(tx) + 1 - cx_of(d));
                    f[d] = Storage<S>::load(stage + r * ROW + shift[r], 3u * i + uint32_t(d % 3), d, rt);
                }
                else if ((8 - d) / 3 == 1)
                    f[d] = Storage<S>::load(stage + ROW + shift[1], 3u * (tx + 1u) + uint32_t((8 - d) % 3), 8 - d,
                                            rt);
                else
                    f[d] = Storage<S>::load(src, at<L>(pn, uint32_t(8 - d), p.S), 8 - d, rt);
            }
        }
        finish_node<L, S>(p, code, x0 + tx, y, f, r3.own());
    }
    // which places of each node row are fluid (a wave holds 64 / B whole rows)
    __shared__ unsigned long long fluid_row[R];
    const unsigned long long row_mask = __ballot(fluid) >> ((ty * B) % 64u);
    if (tx == 0u)
        fluid_row[ty] = B == 64u ? row_mask : row_mask & ((1ull << (B % 64u)) - 1ull);
    const bool odd = active && !fluid && !solid;
    if (__syncthreads_or(odd))
    {
        if (fluid || solid)
            store_node<L, S>(dst, pn, p.S, f, p.sr, rto);
        return;
    }
    // results through LDS (the staged ranges are free after the barrier above):
    // collision: node row ty at ty * 9B; bundle: array a at a * 3B
    T *out = stage;
    if (active)
    {
        const uint32_t h = p.sr ? mix32(p.sr ^ pn) : 0u;
        uint32_t hh = h;
#pragma unroll
        for (int d = 0; d < 9; ++d)
        {
            const uint32_t o = L == LBM_LAYOUT_COLLISION ? ty * (9u * B) + 9u * tx + uint32_t(d)
                                                         : uint32_t(d / 3) * (3u * B) + 3u * tx + uint32_t(d % 3);
            if constexpr (S == LBM_STORE_F16_SHIFTED || S == LBM_STORE_BF16_SHIFTED)
            {
                if (p.sr)
                {
                    hh = hh * 747796405u + 2891336453u;
                    Storage<S>::store_sr(out, o, f[d], d, rto, hh >> (S == LBM_STORE_F16_SHIFTED ? 19 : 16));
                    continue;
                }
            }
            Storage<S>::store(out, o, f[d], d, rto);
        }
    }
    __syncthreads();
    // contiguous output arrays: one per node row (collision), three (bundle).
    // Each starts on a 64-byte segment (b0 is a multiple of 64 places); a
    // segment that holds no fluid node is left unwritten -- solid places are
    // never read -- and every other one is written whole.
    constexpr uint32_t SE = 64u / sizeof(T);  // elements per segment
    auto segment_fluid = [&](unsigned long long m, uint32_t e) {
        const uint32_t s0 = e / SE * SE;
        const uint32_t n0 = s0 / K, n1 = min((s0 + SE - 1u) / K, B - 1u);
        const uint32_t n = n1 - n0 + 1u;
        return (m >> n0) & (n >= 64u ? ~0ull : ((1ull << n) - 1ull));
    };
    const uint32_t arrays = L == LBM_LAYOUT_COLLISION ? rows : 3u;
    const uint32_t per = K * len;                                  // elements per array
    for (uint32_t a = 0; a < arrays; ++a)
    {
        const uint32_t g0 = L == LBM_LAYOUT_COLLISION ? 9u * (b0 + a * p.P) : 3u * a * p.S + 3u * b0;
        const unsigned long long m = fluid_row[L == LBM_LAYOUT_COLLISION ? a : 0u];
        const T *o = out + a * (K * B);
        const uint32_t nvec = per / VE;
        const V4 *lo = reinterpret_cast<const V4 *>(o);
        V4 *go = reinterpret_cast<V4 *>(dst + g0);
        for (uint32_t i = tid; i < nvec; i += NT)
            if (segment_fluid(m, i * VE))
                __builtin_nontemporal_store(lo[i], go + i);
        for (uint32_t i = nvec * VE + tid; i < per; i += NT)
            if (segment_fluid(m, i))
                dst[g0 + i] = o[i];
    }
}

// lptl: two lattices, one linear range over every node of the lattice
// (lptl.hpp:63-65), work-group size from the settings.  The range runs over
// the padded placement, so the thread index is the node's physical index and
// waves stay aligned with rows; padding places carry code 0 and are skipped.
template <int L, int S>
__global__ __launch_bounds__(1024) void k_step_lptl(StepParams p)
{
    using T = typename Storage<S>::T;
    using C = typename Storage<S>::C;
    const uint32_t pn = blockIdx.x * blockDim.x + threadIdx.x;
    if (pn >= p.P * p.Vx)
        return;
    const uint32_t code = p.code[pn];
    const bool fluid = code & LBM_CODE_FLUID;
    const bool segment = segment_has<T>(fluid);  // every lane takes part in the ballot
    const bool fill = (code & LBM_CODE_SOLID) && segment;
    C f[9] = {};
    const uint32_t y = pn / p.P, x = pn - y * p.P - p.X0;  // meaningful on fluid and solid places
    if (fluid)
    {
        const Rt3 r3 = rt3_of<S>(p.ref_in, x);
        pull<L, S>(static_cast<const T *>(p.src), pn, code, p.P, p.S, f, r3);
        finish_node<L, S>(p, code, x, y, f, r3.own());
    }
    if (fluid || fill)
        store_node<L, S>(static_cast<T *>(p.dst), pn, p.S, f, p.sr, rt_of<S>(p.ref_out, x) - rt_of<S>(p.ref_in, x));
}

// npol: one lattice, one work-group per subdomain tile; every pull of the
// tile is done before the barrier and every write after it, tiles only read
// their own nodes and the buffer rows/columns around them
// (npol_kernels.hpp:89-199).
template <int L, int S>
__global__ __launch_bounds__(1024) void k_step_npol(StepParams p)
{
    using T = typename Storage<S>::T;
    using C = typename Storage<S>::C;
    T *lat = static_cast<T *>(p.dst);
    const uint32_t x = blockIdx.x * (p.sub_h + 1u) + threadIdx.x + 1u;
    const uint32_t y = blockIdx.y * (p.sub_v + 1u) + threadIdx.y + 1u;
    const uint32_t pn = place(x, y, p.P, p.X0);
    const uint32_t code = p.code[pn];
    const bool fluid = code & LBM_CODE_FLUID;
    const bool segment = segment_has<T>(fluid);  // every lane takes part in the ballot
    const bool fill = (code & LBM_CODE_SOLID) && segment;
    const uint32_t xu = x - x / (p.sub_h + 1u);
    const Rt3 r3 = rt3_of<S>(p.ref_in, xu);
    C f[9] = {};
    if (fluid)
        pull<L, S>(lat, pn, code, p.P, p.S, f, r3);
    __syncthreads();
    if (fluid)
        finish_node<L, S>(p, code, xu, y - y / (p.sub_v + 1u), f, r3.own());
    if (fluid || fill)
        store_node<L, S>(lat, pn, p.S, f, p.sr, rt_of<S>(p.ref_out, xu) - r3.own());
}

// nsol: one lattice, swap scheme (nsol_kernels.hpp:171-334).  Work-groups of
// (sub_h+2) x (sub_v+1) overlap by one buffer column.  Every node hands its
// populations 5..8 to the LDS slots of the neighbours that pull them and
// reads populations 0..3 straight from the upstream nodes in HBM.  One
// barrier separates all reads from the in-place writes (the reference uses
// two LDS slots and three barriers; four slots need one).  The LDS slots
// carry the stored bits; the node that pulls a population decodes it, with
// the rho_ref of column xu - c_x as every pull does.
template <int L, int S>
__global__ __launch_bounds__(1024) void k_step_nsol(StepParams p)
{
    using T = typename Storage<S>::T;
    using C = typename Storage<S>::C;
    extern __shared__ unsigned char lds_raw[];
    T *lds = reinterpret_cast<T *>(lds_raw);
    T *lat = static_cast<T *>(p.dst);

    const uint32_t LP = p.sub_h + 4u;
    const uint32_t plane = (p.sub_v + 2u) * LP;
    T *s5 = lds, *s6 = lds + plane, *s7 = lds + 2u * plane, *s8 = lds + 3u * plane;

    const uint32_t lx = threadIdx.x + 1u, ly = threadIdx.y;
    const uint32_t x = blockIdx.x * (p.sub_h + 1u) + threadIdx.x;
    const uint32_t y = blockIdx.y * (p.sub_v + 1u) + threadIdx.y;
    const uint32_t pn = place(x, y, p.P, p.X0);
    const uint32_t Sd = p.S, P = p.P;
    const uint32_t code = p.code[pn];
    const bool fluid = code & LBM_CODE_FLUID;
    const bool segment = segment_has<T>(fluid);  // every lane takes part in the ballot
    const bool fill = (code & LBM_CODE_SOLID) && segment;

    const uint32_t xu = x - x / (p.sub_h + 1u);
    const Rt3 r3 = rt3_of<S>(p.ref_in, xu);
    C f[9] = {};
    if (!(code & LBM_CODE_SOLID))
    {
        const T b5 = lat[at<L>(pn, 5, Sd)], b6 = lat[at<L>(pn, 6, Sd)];
        const T b7 = lat[at<L>(pn, 7, Sd)], b8 = lat[at<L>(pn, 8, Sd)];
        s5[ly * LP + lx + 1u] = b5;
        s6[(ly + 1u) * LP + lx - 1u] = b6;
        s7[(ly + 1u) * LP + lx] = b7;
        s8[(ly + 1u) * LP + lx + 1u] = b8;
        if (fluid)
        {
            f[4] = Storage<S>::load(lat, at<L>(pn, 4, Sd), 4, 0.f);
            f[3] = (code & (1u << 3)) ? Storage<S>::load(&b5, 0u, 5, 0.f)
                                      : Storage<S>::load(lat, at<L>(pn + 1u, 3, Sd), 3, r3.rel(-1));
            f[2] = (code & (1u << 2)) ? Storage<S>::load(&b6, 0u, 6, 0.f)
                                      : Storage<S>::load(lat, at<L>(pn - 1u + P, 2, Sd), 2, r3.rel(1));
            f[1] = (code & (1u << 1)) ? Storage<S>::load(&b7, 0u, 7, 0.f)
                                      : Storage<S>::load(lat, at<L>(pn + P, 1, Sd), 1, 0.f);
            f[0] = (code & (1u << 0)) ? Storage<S>::load(&b8, 0u, 8, 0.f)
                                      : Storage<S>::load(lat, at<L>(pn + 1u + P, 0, Sd), 0, r3.rel(-1));
            if (code & (1u << 5))
                f[5] = Storage<S>::load(lat, at<L>(pn, 3, Sd), 3, 0.f);
            if (code & (1u << 6))
                f[6] = Storage<S>::load(lat, at<L>(pn, 2, Sd), 2, 0.f);
            if (code & (1u << 7))
                f[7] = Storage<S>::load(lat, at<L>(pn, 1, Sd), 1, 0.f);
            if (code & (1u << 8))
                f[8] = Storage<S>::load(lat, at<L>(pn, 0, Sd), 0, 0.f);
        }
    }
    __syncthreads();
    if (fluid)
    {
        const uint32_t own = ly * LP + lx;
        if (!(code & (1u << 5)))
            f[5] = Storage<S>::load(s5, own, 5, r3.rel(1));
        if (!(code & (1u << 6)))
            f[6] = Storage<S>::load(s6, own, 6, r3.rel(-1));
        if (!(code & (1u << 7)))
            f[7] = Storage<S>::load(s7, own, 7, 0.f);
        if (!(code & (1u << 8)))
            f[8] = Storage<S>::load(s8, own, 8, r3.rel(1));
        finish_node<L, S>(p, code, xu, y - y / (p.sub_v + 1u), f, r3.own());
    }
    if (fluid || fill)
        store_node<L, S>(lat, pn, p.S, f, p.sr, rt_of<S>(p.ref_out, xu) - r3.own());
}

// npol and nsol with tiles at most one wave wide: one wave per tile, walking
// up the tile a group of rows at a time (64 / sub_h rows, one node per lane;
// twice that in FP64/FP32, two nodes per lane), with no barrier and no
// buffer-refresh launches.
//
// The buffer rows and columns of the lattice carry copies of tile-edge
// populations from the previous step (general/buffered.hpp:74-93,133-152).
// Here those copies live in compact side arrays, two of each kind used in
// turn: a step reads its neighbours' previous edge values from one and each
// tile writes its own new edge values to the other, so tiles need no order
// among themselves.  A pull through a buffer node takes what the two refresh
// passes would have put there, corners included (rows first, then columns:
// a corner carries its diagonal neighbour's population).  These steps leave
// the lattice's own buffer nodes as they are; only the split passes read
// those, and they refresh them first.
//
// Inside a tile a row group loads everything it pulls before it stores.  It
// pulls from its own rows, from rows above it (not written yet) and from the
// side arrays, except populations 6, 7, 8 of its bottom row, which come from
// the previous group's top row: that group overwrote them, so it loaded its
// own old values first and hands them down across lanes (the load that a
// hand-down replaces is issued all the same and its value dropped, so that
// the pulls of a group are one batch of loads).
//
// Side-array layouts (elements of the storage type), boundary k between
// tiles k and k + 1:
//   rows: [k < cnt_v - 1][j < 6][x < Hx], yb = (k + 1)(sub_v + 1):
//         j < 3 population j of (x, yb + 1), j >= 3 population j + 3 of (x, yb - 1)
//   cols: [k < cnt_h - 1][j < 6][y < Vx], xb = (k + 1)(sub_h + 1):
//         j < 3 population 3j + 2 of (xb - 1, y), j >= 3 population 3(j - 3) of (xb + 1, y)
constexpr uint32_t TILE_SWEEP_BLOCK = 256;  // four tiles per block
constexpr uint32_t TILE_SWEEP_MAX_WIDTH = 64;
constexpr uint32_t SWEEP_XCDS = 8;  // accelerator complex dies of an MI355X

// Does the 64-byte segment of a lane's node hold a node of `pred`?  Lanes
// lo..hi-1 hold consecutive places of one row; E elements per node are
// contiguous in an array (stream 1, bundle 3, collision 9) whose base is
// segment aligned.  Exact where the lane-group test of segment_has assumes
// rows that start on a segment.
template <class T, uint32_t E>
__device__ __forceinline__ bool row_segment_has(bool pred, uint32_t pn, int lane, int lo, int hi)
{
    constexpr uint32_t G = 64u / sizeof(T);
    const unsigned long long m = __ballot(pred);
    const uint32_t s0 = (E * pn) / G, s1 = (E * pn + E - 1u) / G;
    const uint32_t q0 = (s0 * G) / E, q1 = (s1 * G + G - 1u) / E;  // first and last node in them
    const int a = max(lo, lane - int(pn - q0));
    const int b = min(hi - 1, lane + int(q1 - pn));
    const int n = b - a + 1;
    const unsigned long long span = (n >= 64 ? ~0ull : ((1ull << n) - 1ull)) << a;
    return (m & span) != 0ull;
}

// Nodes per lane and group: FP64 and FP32 two (rows r and r + 64/sub_h), so
// that a wave has twice the loads in flight (FP64 +2 to +8 %, FP32 +2 to +6 %
// measured); 16-bit storage one (two measured -3 % to +4 %).
template <int S>
constexpr uint32_t sweep_nodes_per_lane()
{
    return sizeof(typename Storage<S>::T) >= 4 ? 2u : 1u;
}

template <int L, int S>
__global__ __launch_bounds__(TILE_SWEEP_BLOCK) void k_step_tile_sweep(StepParams p, uint32_t tiles)
{
    using T = typename Storage<S>::T;
    using C = typename Storage<S>::C;
    constexpr uint32_t E = L == LBM_LAYOUT_STREAM ? 1u : (L == LBM_LAYOUT_BUNDLE ? 3u : 9u);
    constexpr uint32_t NL = sweep_nodes_per_lane<S>();
    T *lat = static_cast<T *>(p.dst);  // the side arrays follow the lattice in its allocation
    // Blocks are dealt to the XCDs in turn.  In 16-bit storage each XCD takes
    // a contiguous run of tiles instead, so that the segments and halo rows
    // that neighbouring tiles share meet in one L2: on the porous 8192 x 2048
    // channel (nsol stream) fetch -33 %, writes -10 %, time -2 %; in FP64
    // the same mapping measured 1.4 % slower.
    uint32_t blk = blockIdx.x;
    if constexpr (sizeof(T) == 2)
    {
        const uint32_t nb = gridDim.x, q = nb / SWEEP_XCDS, rem = nb % SWEEP_XCDS, xcd = blockIdx.x % SWEEP_XCDS;
        blk = xcd * q + min(xcd, rem) + blockIdx.x / SWEEP_XCDS;  // a permutation of 0 .. nb - 1
    }
    const uint32_t tile = blk * (TILE_SWEEP_BLOCK / 64u) + threadIdx.x / 64u;
    if (tile >= tiles)
        return;  // whole waves
    const uint32_t sh = p.sub_h, sv = p.sub_v, P = p.P, Sd = p.S, Hx = p.Hx, Vx = p.Vx;
    const uint32_t R = 64u / sh;  // node rows per sub-group (one node per lane)
    const uint32_t G = NL * R;    // node rows per group
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t r = lane / sh, i = lane - r * sh;
    const uint32_t by = tile / p.cnt_h, bx = tile - by * p.cnt_h;
    const uint32_t x = bx * (sh + 1u) + 1u + i;
    const uint32_t xi = x - x / (sh + 1u);  // unexpanded column
    const uint32_t y00 = by * (sv + 1u) + 1u;
    const int lo = int(r * sh), hi = int(r * sh + sh);
    // rho_ref of unexpanded columns xi - 1, xi, xi + 1 (16-bit forms), read and written
    const Rt3 r3 = rt3_of<S>(p.ref_in, xi);
    const float rto = rt_of<S>(p.ref_out, xi) - r3.own();  // stores: against the written reference
    const bool buf_left = i == 0u && bx > 0u;                // x - 1 is a buffer column
    const bool buf_right = i == sh - 1u && bx + 1u < p.cnt_h;  // x + 1 is one
    // lanes of a group's top row that the next group's bottom-row pulls of
    // 6 (from x + 1), 7 (x) and 8 (x - 1) read
    const uint32_t top = (R - 1u) * sh;
    const int src6 = int(top + min(i + 1u, sh - 1u)), src7 = int(top + i), src8 = int(top + (i ? i - 1u : 0u));

    // this lane's node code in the sub-group starting at tile row y0 (0 outside the tile)
    auto code_at = [&](uint32_t y0) -> uint32_t {
        const uint32_t row = y0 + r;
        return (r < R && row < sv) ? uint32_t(p.code[place(x, y00 + row, P, p.X0)]) : 0u;
    };
    // All codes of the lane's nodes, loaded once into its own LDS slots (one
    // per sub-group), so that a group's pulls wait for one round trip, not two.
    extern __shared__ uint16_t sweep_codes[];
    const uint32_t groups = (sv + G - 1u) / G, slots = groups * NL;
    uint16_t *codes = sweep_codes + (threadIdx.x / 64u) * slots * 64u + lane;
    for (uint32_t s0 = 0; s0 < slots; s0 += 8u)
    {
        uint32_t v[8];
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k)
            v[k] = s0 + k < slots ? code_at((s0 + k) * R) : 0u;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k)
            if (s0 + k < slots)
                codes[(s0 + k) * 64u] = uint16_t(v[k]);
    }
    // the pulls of the sub-group starting at y0
    auto gather = [&](uint32_t y0, uint32_t code, C f[9]) {
        if (!(code & LBM_CODE_FLUID))
            return;
        const uint32_t row = y0 + r, y = y00 + row;
        const uint32_t pn = place(x, y, P, p.X0);
        const bool buf_below = row == 0u && by > 0u;             // y - 1 is a buffer row
        const bool buf_above = row == sv - 1u && by + 1u < p.cnt_v;  // y + 1 is one
#pragma unroll
        for (int d = 0; d < 9; ++d)
        {
            const int cx = cx_of(d), cy = cy_of(d);
            const bool bb = (code >> d) & 1u;
            const bool cb = (cx == 1 && buf_left) || (cx == -1 && buf_right);
            const bool rb = (cy == 1 && buf_below) || (cy == -1 && buf_above);
            uint32_t idx;
            if (bb)
                idx = at<L>(pn, uint32_t(8 - d), Sd);
            else if (rb)
            {
                const uint32_t k = cy == 1 ? by - 1u : by;
                const uint32_t xs = uint32_t(int(x) - cx - (cb ? cx : 0));
                const uint32_t j = cy == -1 ? uint32_t(d) : uint32_t(d - 3);
                idx = p.side_rows_in + (k * 6u + j) * Hx + xs;
            }
            else if (cb)
            {
                const uint32_t k = cx == 1 ? bx - 1u : bx;
                const uint32_t j = cx == 1 ? uint32_t(d - 2) / 3u : 3u + uint32_t(d) / 3u;
                idx = p.side_cols_in + (k * 6u + j) * Vx + uint32_t(int(y) - cy);
            }
            else
                idx = at<L>(uint32_t(int32_t(pn) - cy * int32_t(P) - cx), uint32_t(d), Sd);
            f[d] = Storage<S>::load(lat, idx, bb ? 8 - d : d, bb ? 0.f : r3.rel(cx));
        }
    };
    // own old populations 6, 7, 8 of the group's top row (its last
    // sub-group starts at y0), when a group follows.  Every node of the row
    // but a solid one: ghost nodes (code 0) inside a tile -- the outlet ghost
    // column when a buffer column separates it from nothing but the border
    // node -- are pulled from too.
    auto keep_top = [&](uint32_t y0, uint32_t code, C &c6, C &c7, C &c8) {
        if (r == R - 1u && y0 + R < sv && !(code & LBM_CODE_SOLID))
        {
            const uint32_t pn = place(x, y00 + y0 + r, P, p.X0);
            c6 = Storage<S>::load(lat, at<L>(pn, 6u, Sd), 6, 0.f);
            c7 = Storage<S>::load(lat, at<L>(pn, 7u, Sd), 7, 0.f);
            c8 = Storage<S>::load(lat, at<L>(pn, 8u, Sd), 8, 0.f);
        }
    };

    C c6 = C(0), c7 = C(0), c8 = C(0);  // old 6, 7, 8 of the previous group's top row
    for (uint32_t y0 = 0, g = 0; y0 < sv; y0 += G, ++g)
    {
        uint32_t code[NL];
#pragma unroll
        for (uint32_t h = 0; h < NL; ++h)
            code[h] = codes[(g * NL + h) * 64u];
        const C k6 = __shfl(c6, src6, 64), k7 = __shfl(c7, src7, 64), k8 = __shfl(c8, src8, 64);
        keep_top(y0 + (NL - 1u) * R, code[NL - 1u], c6, c7, c8);
        C f[NL][9] = {};
#pragma unroll
        for (uint32_t h = 0; h < NL; ++h)
            gather(y0 + h * R, code[h], f[h]);

        bool write[NL], fluid[NL];
#pragma unroll
        for (uint32_t h = 0; h < NL; ++h)
        {
            const uint32_t row = y0 + h * R + r, y = y00 + row;
            const uint32_t pn = place(x, y, P, p.X0);
            fluid[h] = code[h] & LBM_CODE_FLUID;
            write[h] = fluid[h] || ((code[h] & LBM_CODE_SOLID) &&
                                    row_segment_has<T, E>(fluid[h], pn, int(lane), lo, hi));
            if (fluid[h])
            {
                if (h == 0u && y0 != 0u && r == 0u)
                {
#pragma unroll
                    for (int d = 6; d < 9; ++d)
                    {
                        const int sx = int(i) - cx_of(d);
                        if (!((code[h] >> d) & 1u) && sx >= 0 && sx < int(sh))
                            f[h][d] = rebase<S>(d == 6 ? k6 : (d == 7 ? k7 : k8), d, r3.rel(cx_of(d)));
                    }
                }
                finish_node<L, S>(p, code[h], xi, y - y / (sv + 1u), f[h], r3.own());
            }
        }
        // keep_top's loads have returned before any lane overwrites what they read
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
#pragma unroll
        for (uint32_t h = 0; h < NL; ++h)
        {
            if (!write[h])
                continue;
            const uint32_t row = y0 + h * R + r, y = y00 + row;
            const uint32_t pn = place(x, y, P, p.X0);
            T t[9];
            node_bits<S>(f[h], pn, p.sr, t, rto);
#pragma unroll
            for (int d = 0; d < 9; ++d)
            {
                if constexpr (L == LBM_LAYOUT_STREAM)
                    __builtin_nontemporal_store(t[d], lat + at<L>(pn, uint32_t(d), Sd));
                else
                    lat[at<L>(pn, uint32_t(d), Sd)] = t[d];
            }
            if (fluid[h])
            {
                if (buf_right)
                    for (uint32_t j = 0; j < 3; ++j)
                        lat[p.side_cols_out + (bx * 6u + j) * Vx + y] = t[3u * j + 2u];
                if (buf_left)
                    for (uint32_t j = 0; j < 3; ++j)
                        lat[p.side_cols_out + ((bx - 1u) * 6u + 3u + j) * Vx + y] = t[3u * j];
                if (row == sv - 1u && by + 1u < p.cnt_v)
                    for (uint32_t j = 0; j < 3; ++j)
                        lat[p.side_rows_out + (by * 6u + 3u + j) * Hx + x] = t[6u + j];
                if (row == 0u && by > 0u)
                    for (uint32_t j = 0; j < 3; ++j)
                        lat[p.side_rows_out + ((by - 1u) * 6u + j) * Hx + x] = t[j];
            }
        }
    }
}

// npol and nsol, tiles 8 to 64 nodes wide: one block per tile walks up it
// CR = 256 / sub_h node rows (a chunk) at a time (npol_kernels.hpp:89-199,
// nsol_kernels.hpp:171-334 give the same fields: every pull reads the
// previous step).  The tile's rows, with the buffer (or ghost) rows and
// columns around it, pass through a ring of CR + 2 row slots in LDS, each
// lattice row loaded once with aligned 16-byte loads: a row is one range of
// 9 (sub_h + 2) elements in the collision layout, nine ranges of sub_h + 2
// (one per direction array) in the stream layout, three of 3 (sub_h + 2) in
// the bundle layout.  A chunk pulls from rows j - 1 .. j + 1 of the ring: its
// bottom neighbour row has been overwritten in HBM by the previous chunk, but
// the ring still holds its old values, and the rows above are not written
// yet.  The loads of the next chunk's CR new rows are issued before the
// current chunk's update and written to the ring slots of the rows it no
// longer needs after a barrier, so HBM latency overlaps the update.  The
// buffer places of the staged rows are then overwritten in the ring with the
// tile sweep's side arrays (what the reference's refresh would have copied
// there; through a corner, the diagonal neighbour's population), and the
// tile's edge nodes write their new edge populations to the other side
// arrays, so no refresh launches are needed.  Results leave through an LDS
// output area as 16-byte non-temporal stores; a 64-byte segment of a tile row
// with no fluid node of the tile is not written (solid places are never
// read); a chunk that holds a node that is neither fluid nor solid (the
// outlet ghost column inside the last tile of the buffered-outlet quirk)
// writes node by node and leaves that node alone.
constexpr uint32_t TILE_STAGED_THREADS = 256;
constexpr uint32_t TILE_STAGED_MIN_WIDTH = 8;
constexpr uint32_t TILE_STAGED_SIDE = 2;  // side-array values a thread loads per chunk, at most

// elements of one node in one range of a row: collision 9, stream 1, bundle 3
__host__ __device__ constexpr uint32_t tile_staged_k(int L)
{
    return L == LBM_LAYOUT_COLLISION ? 9u : (L == LBM_LAYOUT_STREAM ? 1u : 3u);
}

// 16-byte vectors a thread loads per chunk, at most (sub_h 8 .. 64, the
// first chunk's CR + 2 rows, ranges offset by up to VE - 1 elements)
template <int L, int S>
constexpr uint32_t tile_staged_vectors()
{
    constexpr uint32_t VE = 16u / sizeof(typename Storage<S>::T), NT = TILE_STAGED_THREADS;
    constexpr uint32_t K = tile_staged_k(L), NRG = 9u / K;
    uint32_t kv = 0;
    for (uint32_t sh = TILE_STAGED_MIN_WIDTH; sh <= TILE_SWEEP_MAX_WIDTH; ++sh)
    {
        const uint32_t rows = NT / sh + 2u;
        const uint32_t nvs = (VE - 1u + K * (sh + 2u) + VE - 1u) / VE;
        const uint32_t k = (rows * NRG * nvs + NT - 1u) / NT;
        kv = k > kv ? k : kv;
    }
    return kv;
}

// side-array values of one chunk, at most: six per staged row (three for
// each buffer column), 3 (sub_h + 2) for the buffer row below and above
constexpr bool tile_staged_side_fits()
{
    for (uint32_t sh = TILE_STAGED_MIN_WIDTH; sh <= TILE_SWEEP_MAX_WIDTH; ++sh)
        if (6u * (TILE_STAGED_THREADS / sh + 2u) + 6u * (sh + 2u) > TILE_STAGED_SIDE * TILE_STAGED_THREADS)
            return false;
    return true;
}
static_assert(tile_staged_side_fits(), "side values of a chunk exceed the per-thread slots");

// Elements of one range of a ring slot and of an output row (16-byte
// multiples), and the dynamic LDS bytes (ring, output rows, two sets of
// fluid masks).
__host__ __device__ inline uint32_t tile_staged_sublen(uint32_t k, uint32_t sh, uint32_t ve)
{
    return (k * (sh + 2u) + 2u * ve - 2u) / ve * ve;
}
__host__ __device__ inline uint32_t tile_staged_outsub(uint32_t k, uint32_t sh, uint32_t ve)
{
    return (k * sh + 2u * ve - 2u) / ve * ve;
}
inline size_t tile_staged_lds(int layout, uint32_t sh, size_t elem)
{
    const uint32_t ve = uint32_t(16 / elem), cr = TILE_STAGED_THREADS / sh;
    const uint32_t k = tile_staged_k(layout), nrg = 9u / k;
    return (size_t(cr + 2u) * nrg * tile_staged_sublen(k, sh, ve) + size_t(cr) * nrg * tile_staged_outsub(k, sh, ve)) *
               elem +
           2u * size_t(cr) * sizeof(unsigned long long);
}

template <int L, int S>
__global__ __launch_bounds__(TILE_STAGED_THREADS) void k_step_tile_staged(StepParams p)
{
    using T = typename Storage<S>::T;
    using C = typename Storage<S>::C;
    using V4 = unsigned int __attribute__((ext_vector_type(4)));
    constexpr uint32_t NT = TILE_STAGED_THREADS, VE = 16u / sizeof(T), KV = tile_staged_vectors<L, S>();
    constexpr uint32_t KS = TILE_STAGED_SIDE;
    constexpr uint32_t K = tile_staged_k(L), NRG = 9u / K;  // elements per node and range, ranges per row
    constexpr uint32_t SE = 64u / sizeof(T);                 // elements per 64-byte segment
    extern __shared__ __align__(16) unsigned char tile_lds[];
    const uint32_t sh = p.sub_h, sv = p.sub_v, P = p.P, Hx = p.Hx, Vx = p.Vx;
    const uint32_t CR = NT / sh, W = CR + 2u;
    const uint32_t sub = tile_staged_sublen(K, sh, VE), osub = tile_staged_outsub(K, sh, VE);
    const uint32_t rowl = NRG * sub, outl = NRG * osub;  // elements of a ring slot, of an output row
    // element distance between a row's ranges in the lattice
    const uint32_t rsl = L == LBM_LAYOUT_STREAM ? p.S : (L == LBM_LAYOUT_BUNDLE ? 3u * p.S : 0u);
    T *ring = reinterpret_cast<T *>(tile_lds);
    T *out = ring + W * rowl;
    unsigned long long *masks = reinterpret_cast<unsigned long long *>(out + CR * outl);  // [2][CR]
    T *lat = static_cast<T *>(p.dst);
    const uint32_t tid = threadIdx.x, ty = tid / sh, tx = tid - ty * sh;
    const uint32_t bx = blockIdx.x, by = blockIdx.y;
    const uint32_t x0 = bx * (sh + 1u) + 1u, y00 = by * (sv + 1u) + 1u;
    const bool buf_l = bx > 0u, buf_r = bx + 1u < p.cnt_h;  // columns x0 - 1, x0 + sh are buffers
    const bool buf_b = by > 0u, buf_t = by + 1u < p.cnt_v;  // rows y00 - 1, y00 + sv are buffers
    // window row j (0 .. sv + 1) is lattice row y00 - 1 + j; range a of it
    // starts at element a rsl + e00 + K j P, every range of every row at the
    // same offset mod VE (P is a multiple of 64 places, rsl of 16 bytes);
    // place i of a row is column x0 - 1 + i
    const uint32_t e00 = K * place(x0 - 1u, y00 - 1u, P, p.X0);
    const uint32_t shift = e00 % VE, a00 = e00 - shift;
    const uint32_t nvs = (shift + K * (sh + 2u) + VE - 1u) / VE;  // vectors per range
    const uint32_t nv = NRG * nvs;                                // per row
    const uint32_t s1 = (e00 + K) % VE;                           // offset of the tile's first node mod VE
    const uint32_t nvo = (s1 + K * sh + VE - 1u) / VE;            // output vectors per range
    const uint32_t chunks = (sv + CR - 1u) / CR;
    const uint32_t xe = x0 + tx, xu = xe - xe / (sh + 1u);
    const Rt3 r3 = rt3_of<S>(p.ref_in, xu);
    const float rto = rt_of<S>(p.ref_out, xu) - r3.own();
    // ring element of population d at place i of window row j
    auto slot = [&](uint32_t j, uint32_t i, uint32_t d) -> uint32_t {
        return (j % W) * rowl + (d / K) * sub + shift + K * i + d % K;
    };

    // row, range and vector of flat vector index w of a row batch (one range
    // per row in the collision layout: no second division)
    auto split = [&](uint32_t w, uint32_t per_row, uint32_t per_range, uint32_t &r, uint32_t &ra, uint32_t &v) {
        r = w / per_row;
        const uint32_t rem = w - r * per_row;
        if constexpr (NRG == 1)
        {
            ra = 0u;
            v = rem;
        }
        else
        {
            ra = rem / per_range;
            v = rem - ra * per_range;
        }
    };

    V4 pre[KV];
    T side[KS];
    uint32_t side_at[KS];  // ring element of each side value, ~0u: none
    // loads of window rows ja .. jb, every lane the same count (branch-free
    // addresses: lanes past the end load the first vector again), and of the
    // side values of their buffer places
    auto fetch = [&](uint32_t ja, uint32_t jb) {
        const uint32_t total = (jb - ja + 1u) * nv;
#pragma unroll
        for (uint32_t k = 0; k < KV; ++k)
        {
            const uint32_t w = tid + k * NT;
            const uint32_t wc = w < total ? w : 0u;
            uint32_t r, ra, v;
            split(wc, nv, nvs, r, ra, v);
            pre[k] = *reinterpret_cast<const V4 *>(lat + (ra * rsl + a00 + K * (ja + r) * P + v * VE));
        }
        // buffer-column places of rows that are not buffer rows: populations
        // 2, 5, 8 on the left, 0, 3, 6 on the right; buffer rows: 6, 7, 8
        // below the tile, 0, 1, 2 above it, at every place, a corner place
        // from the column beyond the buffer column
        const uint32_t ca = (ja == 0u && buf_b) ? 1u : ja, cb = (jb == sv + 1u && buf_t) ? sv : jb;
        const uint32_t nc = 6u * (cb - ca + 1u);
        const uint32_t nr = 3u * (sh + 2u);
        const uint32_t nb = (ja == 0u && buf_b) ? nr : 0u, nt = (jb == sv + 1u && buf_t) ? nr : 0u;
#pragma unroll
        for (uint32_t k = 0; k < KS; ++k)
        {
            const uint32_t w = tid + k * NT;
            side_at[k] = ~0u;
            if (w < nc)
            {
                const uint32_t j = ca + w / 6u, s = (w % 6u) / 3u, q = w % 3u;
                if (s == 0u ? buf_l : buf_r)
                {
                    const uint32_t kcol = s == 0u ? (bx - 1u) * 6u + q : bx * 6u + 3u + q;
                    side[k] = lat[p.side_cols_in + kcol * Vx + (y00 - 1u + j)];
                    side_at[k] = s == 0u ? slot(j, 0u, 3u * q + 2u) : slot(j, sh + 1u, 3u * q);
                }
            }
            else if (w < nc + nb + nt)
            {
                const uint32_t u = w - nc, s = u < nb ? 0u : 1u, v = u - s * nb;
                const uint32_t q = v / (sh + 2u), i = v - q * (sh + 2u);
                const uint32_t xs = (i == 0u && buf_l) ? x0 - 2u : ((i == sh + 1u && buf_r) ? x0 + sh + 1u : x0 - 1u + i);
                const uint32_t krow = s == 0u ? (by - 1u) * 6u + 3u + q : by * 6u + q;
                side[k] = lat[p.side_rows_in + krow * Hx + xs];
                side_at[k] = s == 0u ? slot(0u, i, 6u + q) : slot(sv + 1u, i, q);
            }
        }
    };
    auto stash = [&](uint32_t ja, uint32_t jb) {
        const uint32_t total = (jb - ja + 1u) * nv;
#pragma unroll
        for (uint32_t k = 0; k < KV; ++k)
        {
            const uint32_t w = tid + k * NT;
            if (w < total)
            {
                uint32_t r, ra, v;
                split(w, nv, nvs, r, ra, v);
                reinterpret_cast<V4 *>(ring + ((ja + r) % W) * rowl + ra * sub)[v] = pre[k];
            }
        }
    };
    auto patch = [&]() {
#pragma unroll
        for (uint32_t k = 0; k < KS; ++k)
            if (side_at[k] != ~0u)
                ring[side_at[k]] = side[k];
    };
    auto code_of = [&](uint32_t k) -> uint32_t {
        const uint32_t j = 1u + k * CR + ty;
        return (ty < CR && j <= sv) ? uint32_t(p.code[place(xe, y00 - 1u + j, P, p.X0)]) : 0u;
    };

    uint32_t code = code_of(0);
    fetch(0u, min(CR + 1u, sv + 1u));
    if (tid < 2u * CR)
        masks[tid] = 0ull;
    stash(0u, min(CR + 1u, sv + 1u));
    __syncthreads();
    patch();
    __syncthreads();
    for (uint32_t k = 0; k < chunks; ++k)
    {
        const bool more = k + 1u < chunks;
        const uint32_t na = (k + 1u) * CR + 2u, nb = min((k + 2u) * CR + 1u, sv + 1u);  // next chunk's new rows
        uint32_t code_next = 0;
        if (more)
        {
            code_next = code_of(k + 1u);
            fetch(na, nb);
        }
        unsigned long long *fluid_row = masks + (k & 1u) * CR;
        const uint32_t j = 1u + k * CR + ty;  // window row of this thread's node
        const bool active = ty < CR && j <= sv;
        const uint32_t y = y00 - 1u + j;
        const uint32_t pn = place(xe, y, P, p.X0);
        const bool fluid = code & LBM_CODE_FLUID, solid = code & LBM_CODE_SOLID;
        C f[9] = {};
        if (fluid)
        {
#pragma unroll
            for (int d = 0; d < 9; ++d)
            {
                const bool bb = (code >> d) & 1u;
                const uint32_t jr = bb ? j : uint32_t(int32_t(j) - cy_of(d));
                const uint32_t i = bb ? tx + 1u : uint32_t(int32_t(tx) + 1 - cx_of(d));
                const int dd = bb ? 8 - d : d;
                f[d] = Storage<S>::load(ring, slot(jr, i, uint32_t(dd)), dd, bb ? 0.f : r3.rel(cx_of(d)));
            }
            finish_node<L, S>(p, code, xu, y - y / (sv + 1u), f, r3.own());
            atomicOr(&fluid_row[ty], 1ull << tx);
        }
        T t[9];
        node_bits<S>(f, pn, p.sr, t, rto);
        if (fluid)
        {
            // the new edge populations the neighbouring tiles pull next step
            if (tx == sh - 1u && buf_r)
                for (uint32_t q = 0; q < 3; ++q)
                    lat[p.side_cols_out + (bx * 6u + q) * Vx + y] = t[3u * q + 2u];
            if (tx == 0u && buf_l)
                for (uint32_t q = 0; q < 3; ++q)
                    lat[p.side_cols_out + ((bx - 1u) * 6u + 3u + q) * Vx + y] = t[3u * q];
            if (j == sv && buf_t)
                for (uint32_t q = 0; q < 3; ++q)
                    lat[p.side_rows_out + (by * 6u + 3u + q) * Hx + xe] = t[6u + q];
            if (j == 1u && buf_b)
                for (uint32_t q = 0; q < 3; ++q)
                    lat[p.side_rows_out + ((by - 1u) * 6u + q) * Hx + xe] = t[q];
        }
        // every pull of the chunk is done: the ring slots of rows below j - 1
        // may take the next chunk's rows
        const bool odd = __syncthreads_or(active && !fluid && !solid);
        if (odd)
        {
            if (fluid || solid)
#pragma unroll
                for (int d = 0; d < 9; ++d)
                    lat[at<L>(pn, uint32_t(d), p.S)] = t[d];
        }
        else if (active)
        {
#pragma unroll
            for (int d = 0; d < 9; ++d)
                out[ty * outl + uint32_t(d) / K * osub + s1 + K * tx + uint32_t(d) % K] = t[d];
        }
        __syncthreads();
        if (!odd)
        {
            const uint32_t rows = min(CR, sv - k * CR);
            for (uint32_t w = tid; w < rows * NRG * nvo; w += NT)
            {
                uint32_t r, ra, v;
                split(w, NRG * nvo, nvo, r, ra, v);
                // the range's first tile element
                const uint32_t g0 = ra * rsl + e00 + K + K * (1u + k * CR + r) * P;
                const uint32_t ga = g0 - s1 + v * VE;  // 16-byte aligned
                const uint32_t seg = ga / SE * SE;
                const uint32_t lo = max(seg, g0), hi = min(seg + SE, g0 + K * sh) - 1u;
                const uint32_t n0 = (lo - g0) / K, n = (hi - g0) / K - n0 + 1u;  // n <= SE / K + 2 < 64
                if (!((fluid_row[r] >> n0) & ((1ull << n) - 1ull)))
                    continue;
                const T *o = out + r * outl + ra * osub + v * VE;
                if (v * VE >= s1 && v * VE + VE <= s1 + K * sh)
                    __builtin_nontemporal_store(*reinterpret_cast<const V4 *>(o), reinterpret_cast<V4 *>(lat + ga));
                else
                    for (uint32_t e = 0; e < VE; ++e)
                        if (v * VE + e >= s1 && v * VE + e < s1 + K * sh)
                            lat[ga + e] = o[e];
            }
        }
        if (tid < CR)
            masks[((k + 1u) & 1u) * CR + tid] = 0ull;  // the next chunk's masks (last read a chunk ago)
        // the output stores go out before the wait for the next rows
        if (more)
        {
            stash(na, nb);
            __syncthreads();
            patch();
        }
        code = code_next;
        __syncthreads();
    }
}

// The side arrays of the tile sweep for both step parities, from the
// lattice's real nodes: what k_refresh_rows / k_refresh_cols would copy into
// the buffer nodes now.
template <int L, class T>
__global__ void k_prime_side_rows(const T *lat, T *out0, T *out1, uint32_t Hx, uint32_t P, uint32_t X0, uint32_t Sd,
                                  uint32_t sub_v)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= Hx)
        return;
    const uint32_t k = blockIdx.y;
    const uint32_t yb = (k + 1u) * (sub_v + 1u);
    const uint32_t up = place(x, yb + 1u, P, X0), down = place(x, yb - 1u, P, X0);
    for (uint32_t j = 0; j < 3; ++j)
    {
        const T a = lat[at<L>(up, j, Sd)], b = lat[at<L>(down, 6u + j, Sd)];
        out0[(k * 6u + j) * Hx + x] = a;
        out1[(k * 6u + j) * Hx + x] = a;
        out0[(k * 6u + 3u + j) * Hx + x] = b;
        out1[(k * 6u + 3u + j) * Hx + x] = b;
    }
}

template <int L, class T>
__global__ void k_prime_side_cols(const T *lat, T *out0, T *out1, uint32_t Vx, uint32_t P, uint32_t X0, uint32_t Sd,
                                  uint32_t sub_h)
{
    const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= Vx)
        return;
    const uint32_t k = blockIdx.y;
    const uint32_t xb = (k + 1u) * (sub_h + 1u);
    const uint32_t left = place(xb - 1u, y, P, X0), right = place(xb + 1u, y, P, X0);
    for (uint32_t j = 0; j < 3; ++j)
    {
        const T a = lat[at<L>(left, 3u * j + 2u, Sd)], b = lat[at<L>(right, 3u * j, Sd)];
        out0[(k * 6u + j) * Vx + y] = a;
        out1[(k * 6u + j) * Vx + y] = a;
        out0[(k * 6u + 3u + j) * Vx + y] = b;
        out1[(k * 6u + 3u + j) * Vx + y] = b;
    }
}

// Buffer rows: directions 0..2 from the row above, 6..8 from the row below
// (general/buffered.hpp:74-93).  Covers the full expanded width, ghost
// columns included.
template <int L, class T>
__global__ void k_refresh_rows(T *lat, uint32_t Hx, uint32_t P, uint32_t X0, uint32_t Sd, uint32_t sub_v)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= Hx)
        return;
    const uint32_t yb = (blockIdx.y + 1u) * (sub_v + 1u);
    const uint32_t n = place(x, yb, P, X0);
    // sources and destinations never overlap: all six loads go out before
    // the first store instead of one round trip per population
    T v[6];
#pragma unroll
    for (uint32_t d = 0; d < 3; ++d)
    {
        v[d] = lat[at<L>(n + P, d, Sd)];
        v[3 + d] = lat[at<L>(n - P, 6u + d, Sd)];
    }
#pragma unroll
    for (uint32_t d = 0; d < 3; ++d)
    {
        lat[at<L>(n, d, Sd)] = v[d];
        lat[at<L>(n, 6u + d, Sd)] = v[3 + d];
    }
}

// Buffer columns: directions 2, 5, 8 from the left, 0, 3, 6 from the right
// (general/buffered.hpp:133-152).  Runs after the rows so that the corner
// nodes carry the diagonal neighbours' values.
template <int L, class T>
__global__ void k_refresh_cols(T *lat, uint32_t Vx, uint32_t P, uint32_t X0, uint32_t Sd, uint32_t sub_h)
{
    const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= Vx)
        return;
    const uint32_t xb = (blockIdx.y + 1u) * (sub_h + 1u);
    const uint32_t n = place(xb, y, P, X0);
    T v[6];  // all loads first, as in k_refresh_rows
#pragma unroll
    for (uint32_t k = 0; k < 3; ++k)
    {
        v[k] = lat[at<L>(n - 1u, 3u * k + 2u, Sd)];
        v[3 + k] = lat[at<L>(n + 1u, 3u * k, Sd)];
    }
#pragma unroll
    for (uint32_t k = 0; k < 3; ++k)
    {
        lat[at<L>(n, 3u * k + 2u, Sd)] = v[k];
        lat[at<L>(n, 3u * k, Sd)] = v[3 + k];
    }
}

// ---------------------------------------------------------------------------
// FP64 reductions (wave64 shuffles, two deterministic passes)
// ---------------------------------------------------------------------------
constexpr int RED_BLOCK = 256;
constexpr int RED_GRID = 1024;
constexpr int RED_K = 6;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

template <int K>
__device__ __forceinline__ void block_sum_store(double (&acc)[K], double *out)
{
    __shared__ double part[RED_BLOCK / 64][K];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k)
    {
        const double s = wave_sum(acc[k]);
        if (lane == 0)
            part[wave][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < K)
    {
        double s = 0.0;
        for (int w = 0; w < RED_BLOCK / 64; ++w)
            s += part[w][threadIdx.x];
        out[blockIdx.x * K + threadIdx.x] = s;
    }
}

// mode 0: sum rho, jx, jy, count, momentum-exchange force (x, y);
// mode 1: residual sums |j - j_snap|^2, |j|^2 and snapshot update.
template <int L, int S>
__global__ __launch_bounds__(RED_BLOCK) void k_reduce(const typename Storage<S>::T *lat, const uint16_t *code,
                                                      Geom g, double *partials, double *snap, int mode)
{
    double acc[RED_K] = {0, 0, 0, 0, 0, 0};
    for (uint32_t n = blockIdx.x * RED_BLOCK + threadIdx.x; n < g.N; n += RED_GRID * RED_BLOCK)
    {
        const uint32_t pn = place_logical(n, g);
        const uint32_t c = code[pn];
        if (!(c & LBM_CODE_FLUID))
            continue;
        const float rt = rt_of<S>(g.ref, unexpanded_x(n % g.Hx, g));
        double f[9];
#pragma unroll
        for (int d = 0; d < 9; ++d)
            f[d] = Storage<S>::decode(lat[at<L>(pn, uint32_t(d), g.Sd)], d, rt);
        double rho = 0, jx = 0, jy = 0;
#pragma unroll
        for (int d = 0; d < 9; ++d)
        {
            rho += f[d];
            jx += f[d] * double(cx_of(d));
            jy += f[d] * double(cy_of(d));
        }
        if (mode == 0)
        {
            acc[0] += rho;
            acc[1] += jx;
            acc[2] += jy;
            acc[3] += 1.0;
#pragma unroll
            for (int d = 0; d < 9; ++d)
            {
                // link n -> n + c_d ends in a solid iff the pull of 8-d bounces
                if (d != 4 && ((c >> (8 - d)) & 1u))
                {
                    acc[4] += 2.0 * f[d] * double(cx_of(d));
                    acc[5] += 2.0 * f[d] * double(cy_of(d));
                }
            }
        }
        else
        {
            const double dx = jx - snap[2u * n], dy = jy - snap[2u * n + 1u];
            acc[0] += dx * dx + dy * dy;
            acc[1] += jx * jx + jy * jy;
            snap[2u * n] = jx;
            snap[2u * n + 1u] = jy;
        }
    }
    block_sum_store<RED_K>(acc, partials);
}

__global__ void k_reduce_final(const double *partials, int blocks, double *out)
{
    const int k = threadIdx.x;
    if (k >= RED_K)
        return;
    double s = 0.0;
    for (int b = 0; b < blocks; ++b)
        s += partials[b * RED_K + k];
    out[k] = s;
}

// ---------------------------------------------------------------------------
// rho_ref of the 16-bit forms: the mean density of each column's fluid nodes,
// every kRefInterval steps.  Row chunks give per-(chunk, column) sums, summed
// per column in chunk order (deterministic); columns without a fluid node,
// buffer columns and the ghost columns keep their value.
// ---------------------------------------------------------------------------
constexpr uint32_t kRefInterval = 128;
constexpr uint32_t kRefChunks = 64;

template <int L, int S>
__global__ void k_column_sums(const typename Storage<S>::T *lat, const uint16_t *code, Geom g, uint32_t Vx,
                              double *part)
{
    const uint32_t xe = blockIdx.x * blockDim.x + threadIdx.x;
    if (xe >= g.Hx)
        return;
    const uint32_t rows = (Vx + kRefChunks - 1u) / kRefChunks;
    const uint32_t y0 = blockIdx.y * rows, y1 = min(Vx, y0 + rows);
    const float rt = rt_of<S>(g.ref, unexpanded_x(xe, g));
    double sum = 0.0, count = 0.0;
    for (uint32_t y = y0; y < y1; ++y)
    {
        const uint32_t pn = place(xe, y, g.P, g.X0);
        if (!(code[pn] & LBM_CODE_FLUID))
            continue;
        double rho = 0.0;
#pragma unroll
        for (int d = 0; d < 9; ++d)
            rho += Storage<S>::decode(lat[at<L>(pn, uint32_t(d), g.Sd)], d, rt);
        sum += rho;
        count += 1.0;
    }
    part[2ull * (size_t(blockIdx.y) * g.Hx + xe)] = sum;
    part[2ull * (size_t(blockIdx.y) * g.Hx + xe) + 1] = count;
}

// ref_out: a copy of the current table, then the new column means
[[maybe_unused]] static __global__ void k_column_ref(const double *part, Geom g, float *ref_out)
{
    const uint32_t xe = blockIdx.x * blockDim.x + threadIdx.x;
    if (xe >= g.Hx)
        return;
    double sum = 0.0, count = 0.0;
    for (uint32_t k = 0; k < kRefChunks; ++k)
    {
        sum += part[2ull * (size_t(k) * g.Hx + xe)];
        count += part[2ull * (size_t(k) * g.Hx + xe) + 1];
    }
    const uint32_t xu = unexpanded_x(xe, g);
    if (count > 0.0 && xu >= 1u && xu + 2u <= g.H && isfinite(sum))
        ref_out[xu] = float(sum / count);
}

__global__ void k_copy16(const uint4 *__restrict__ a, uint4 *__restrict__ b, size_t n)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        b[i] = a[i];
}

// ---------------------------------------------------------------------------
// split step (the reference's debug kernels, nptl_kernels.hpp:48-295): one
// launch per pass over every physical place, the populations between passes
// kept in an FP64 pass buffer pass[d * PV + place] (PV = P * Vx).  Each pass
// performs the fused kernel's arithmetic in the same order, so a split step
// equals a fused step bit for bit in every storage type.
// ---------------------------------------------------------------------------
// unexpanded coordinates of a physical place (buffer rows/columns mapped out)
__device__ __forceinline__ void split_xy(const StepParams &p, uint32_t pn, uint32_t buffered, uint32_t &x,
                                         uint32_t &y)
{
    y = pn / p.P;
    x = pn - y * p.P - p.X0;
    if (buffered)
    {
        x -= x / (p.sub_h + 1u);
        y -= y / (p.sub_v + 1u);
    }
}

template <int L, int S>
__global__ void k_split_stream(StepParams p, double *pass, uint32_t PV, uint32_t buffered)
{
    using T = typename Storage<S>::T;
    using C = typename Storage<S>::C;
    const uint32_t pn = blockIdx.x * blockDim.x + threadIdx.x;
    if (pn >= PV)
        return;
    const uint32_t code = p.code[pn];
    if (!(code & LBM_CODE_FLUID))
        return;
    uint32_t x, y;
    split_xy(p, pn, buffered, x, y);
    C f[9];
    pull<L, S>(static_cast<const T *>(p.src), pn, code, p.P, p.S, f, rt3_of<S>(p.ref_in, x));
#pragma unroll
    for (int d = 0; d < 9; ++d)
        pass[size_t(d) * PV + pn] = double(f[d]);  // 16-bit forms: deviations from the node's rho_ref
}

template <int S>
__global__ void k_split_outlet(StepParams p, double *pass, uint32_t PV, uint32_t buffered)
{
    using C = typename Storage<S>::C;
    const uint32_t pn = blockIdx.x * blockDim.x + threadIdx.x;
    if (pn >= PV)
        return;
    const uint32_t code = p.code[pn];
    if (!(code & LBM_CODE_FLUID) || !(code & LBM_CODE_OUTLET))
        return;
    uint32_t x, y;
    split_xy(p, pn, buffered, x, y);
    C f[9];
#pragma unroll
    for (int d = 0; d < 9; ++d)
        f[d] = C(pass[size_t(d) * PV + pn]);
    if constexpr (shifted_storage<S>())
        zou_he_outlet<C>(f, C(p.rho_out) - rt_of<S>(p.ref_in, x), C(p.uy_in[y]));
    else
        zou_he_outlet<C, Storage<S>::kPromote>(f, C(p.rho_out), C(p.uy_in[y]));
    pass[pn] = double(f[0]);
    pass[3ull * PV + pn] = double(f[3]);
    pass[6ull * PV + pn] = double(f[6]);
}

template <int S>
__global__ void k_split_moments(StepParams p, double *pass, uint32_t PV, uint32_t buffered)
{
    using C = typename Storage<S>::C;
    const uint32_t pn = blockIdx.x * blockDim.x + threadIdx.x;
    if (pn >= PV)
        return;
    const uint32_t code = p.code[pn];
    if (!(code & LBM_CODE_FLUID))
        return;
    uint32_t x, y;
    split_xy(p, pn, buffered, x, y);
    C f[9];
#pragma unroll
    for (int d = 0; d < 9; ++d)
        f[d] = C(pass[size_t(d) * PV + pn]);
    C rho, jx, jy, usq;
    moments<C, Storage<S>::kFull>(f, rho, jx, jy, usq);
    if constexpr (shifted_storage<S>())
        write_moments<C, double>(p, code, x, y, double(rt_of<S>(p.ref_in, x)) + double(rho), jx, jy, usq);
    else
        write_moments<C>(p, code, x, y, rho, jx, jy, usq);
}

template <int L, int S>
__global__ void k_split_collide(StepParams p, const double *pass, uint32_t PV, uint32_t buffered)
{
    using T = typename Storage<S>::T;
    using C = typename Storage<S>::C;
    const uint32_t pn = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t code = pn < PV ? p.code[pn] : 0u;
    const bool fluid = code & LBM_CODE_FLUID;
    const bool segment = segment_has<T>(fluid);  // every lane takes part in the ballot
    const bool fill = (code & LBM_CODE_SOLID) && segment;
    C f[9] = {};
    if (fluid)
    {
#pragma unroll
        for (int d = 0; d < 9; ++d)
            f[d] = C(pass[size_t(d) * PV + pn]);
        C rho, jx, jy, usq;
        collide<C, Storage<S>::kPromote, Storage<S>::kFull>(f, C(p.omega), rho, jx, jy, usq);
    }
    if (fluid || fill)
    {
        uint32_t x, y;
        split_xy(p, pn, buffered, x, y);
        store_node<L, S>(static_cast<T *>(p.dst), pn, p.S, f, p.sr, rt_of<S>(p.ref_out, x) - rt_of<S>(p.ref_in, x));
    }
}

// ---------------------------------------------------------------------------
// dispatch over layout x storage
// ---------------------------------------------------------------------------
inline uint32_t step_seed(uint64_t iteration)
{
    uint64_t z = iteration + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return uint32_t(z ^ (z >> 31));
}

template <int L, int S>
struct Impl
{
    using T = typename Storage<S>::T;
    using C = typename Storage<S>::C;

    static StepParams params(lbm_ctx *c, bool macros)
    {
        StepParams p{};
        p.src = c->lat[c->cur];
        p.dst = c->two_lattice ? c->lat[1 - c->cur] : c->lat[c->cur];
        p.code = c->code;
        p.rho = c->res;
        p.jx = c->res + c->D;
        p.jy = c->res + 2 * size_t(c->D);
        p.jabs = c->res + 3 * size_t(c->D);
        const int parity = int(c->iteration & 1u);
        p.uy_in = c->uycol + size_t(parity) * c->V;
        p.uy_out = c->uycol + size_t(1 - parity) * c->V;
        p.omega = host_omega(c->desc.storage, c->desc.relaxation_time);
        p.ref_in = c->ref[c->ref_cur];
        p.ref_out = c->ref_switch ? c->ref[1 - c->ref_cur] : p.ref_in;
        p.rho_out = c->desc.outlet_density;
        p.H = c->H;
        p.V = c->V;
        p.Hx = c->Hx;
        p.Vx = c->Vx;
        p.N = c->N;
        p.P = c->P;
        p.X0 = c->X0;
        p.S = c->Sd;
        p.sub_h = c->sub_h;
        p.sub_v = c->sub_v;
        p.write_macros = macros ? 1u : 0u;
        p.nan_protection = c->desc.nan_protection ? 1u : 0u;
        // per-step seed, never 0 (0 selects round-to-nearest)
        p.sr = c->desc.rounding == LBM_ROUND_STOCHASTIC ? (step_seed(c->iteration) | 1u) : 0u;
        p.cnt_h = c->cnt_h;
        p.cnt_v = c->cnt_v;
        p.side_rows_in = c->side_rows[c->side_parity];
        p.side_cols_in = c->side_cols[c->side_parity];
        p.side_rows_out = c->side_rows[c->side_parity ^ 1];
        p.side_cols_out = c->side_cols[c->side_parity ^ 1];
        return p;
    }

    // A step at an iteration that is a positive multiple of kRefInterval
    // stores with the new column means of the lattice it reads (16-bit forms).
    static int begin_reference(lbm_ctx *c)
    {
        if constexpr (shifted_storage<S>())
        {
            if (c->iteration == 0u || c->iteration % kRefInterval != 0u || c->ref_switch)
                return LBM_OK;
            float *next = c->ref[1 - c->ref_cur];
            LBM_HIP_TRY(hipMemcpyAsync(next, c->ref[c->ref_cur], c->ref_len * sizeof(float), hipMemcpyDeviceToDevice,
                                       c->stream));
            const Geom g = geom(c);
            hipLaunchKernelGGL((k_column_sums<L, S>), dim3((c->Hx + 255u) / 256u, kRefChunks), dim3(256), 0, c->stream,
                               static_cast<const T *>(c->lat[c->cur]), c->code, g, c->Vx, c->ref_part);
            hipLaunchKernelGGL(k_column_ref, dim3((c->Hx + 255u) / 256u), dim3(256), 0, c->stream, c->ref_part, g,
                               next);
            LBM_HIP_TRY(hipGetLastError());
            c->ref_switch = true;
        }
        return LBM_OK;
    }

    static void end_reference(lbm_ctx *c)
    {
        if (c->ref_switch)
        {
            c->ref_cur ^= 1;
            c->ref_switch = false;
        }
    }

    static int step(lbm_ctx *c, bool macros)
    {
        if (int rc = begin_reference(c))
            return rc;
        const StepParams p = params(c, macros);
        auto launch_staged = [&]() {
            if constexpr (nptl_staged<L, S>())
            {
                constexpr uint32_t rows = nptl_staged_rows<L, S>();
                constexpr uint32_t width = nptl_staged_width<L, S>();
                dim3 grid((c->Hx - 2u + width - 1u) / width, (c->Vx - 2u + rows - 1u) / rows);
                hipLaunchKernelGGL((k_step_nptl_staged<L, S>), grid, dim3(width * rows), 0, c->stream, p);
            }
        };
        auto launch_rows = [&]() {
            constexpr uint32_t R = nptl_rows<S>();
            dim3 grid((c->Hx - 2u + NPTL_BLOCK_X - 1u) / NPTL_BLOCK_X, (c->Vx - 2u + R - 1u) / R);
            hipLaunchKernelGGL((k_step_nptl<L, S>), grid, dim3(NPTL_BLOCK_X), 0, c->stream, p);
        };
        auto launch_sweep = [&]() {
            if (!c->side_valid)
                prime(c);
            const uint32_t tiles = c->cnt_h * c->cnt_v;
            const uint32_t per_block = TILE_SWEEP_BLOCK / 64u;
            constexpr uint32_t NL = sweep_nodes_per_lane<S>();
            const uint32_t R = 64u / c->sub_h, groups = (c->sub_v + NL * R - 1u) / (NL * R);
            const size_t lds = size_t(per_block) * groups * NL * 64u * sizeof(uint16_t);
            hipLaunchKernelGGL((k_step_tile_sweep<L, S>), dim3((tiles + per_block - 1u) / per_block),
                               dim3(TILE_SWEEP_BLOCK), lds, c->stream, p, tiles);
            c->side_parity ^= 1;
        };
        auto launch_tile_staged = [&]() {
            if (!c->side_valid)
                prime(c);
            hipLaunchKernelGGL((k_step_tile_staged<L, S>), dim3(c->cnt_h, c->cnt_v), dim3(TILE_STAGED_THREADS),
                               tile_staged_lds(L, c->sub_h, sizeof(T)), c->stream, p);
            c->side_parity ^= 1;
        };
        switch (c->desc.algorithm)
        {
        case LBM_ALG_NPTL:
        {
            if constexpr (nptl_staged<L, S>())
                launch_staged();
            else
                launch_rows();
            break;
        }
        case LBM_ALG_LPTL:
        {
            // Same two-lattice update on the same placement; where nptl's
            // kernels are faster -- the staged one (collision, FP64/FP32
            // bundle) and the two-rows-per-thread one (16-bit storage) --
            // lptl runs them too and the work-group size of the settings
            // does not apply.
            if constexpr (nptl_staged<L, S>())
            {
                launch_staged();
                break;
            }
            if constexpr (sizeof(typename Storage<S>::T) == 2)
            {
                launch_rows();
                break;
            }
            const uint32_t b = c->wg;
            const uint32_t range = c->P * c->Vx;
            hipLaunchKernelGGL((k_step_lptl<L, S>), dim3((range + b - 1u) / b), dim3(b), 0, c->stream, p);
            break;
        }
        case LBM_ALG_NPOL:
            if (c->sweep)
            {
                launch_sweep();
                break;
            }
            if (c->tile_staged)
            {
                launch_tile_staged();
                break;
            }
            refresh(c);
            hipLaunchKernelGGL((k_step_npol<L, S>), dim3(c->cnt_h, c->cnt_v), dim3(c->sub_h, c->sub_v), 0,
                               c->stream, p);
            break;
        case LBM_ALG_NSOL:
        {
            if (c->sweep)
            {
                launch_sweep();
                break;
            }
            if (c->tile_staged)
            {
                launch_tile_staged();
                break;
            }
            refresh(c);
            const size_t lds = 4u * size_t(c->sub_v + 2u) * (c->sub_h + 4u) * sizeof(C);
            hipLaunchKernelGGL((k_step_nsol<L, S>), dim3(c->cnt_h, c->cnt_v), dim3(c->sub_h + 2u, c->sub_v + 1u),
                               lds, c->stream, p);
            break;
        }
        }
        LBM_HIP_TRY(hipGetLastError());
        if (c->two_lattice)
            c->cur ^= 1;
        c->iteration++;
        end_reference(c);
        return LBM_OK;
    }

    // One pass of a split step (pass = enum lbm_pass); the collide pass ends
    // the step (lattice swap, iteration count).
    static int split(lbm_ctx *c, int pass, double *buf)
    {
        c->side_valid = false;  // the passes change the lattice behind the side arrays
        if (pass == LBM_PASS_STREAM)
            if (int rc = begin_reference(c))
                return rc;
        const StepParams p = params(c, true);
        const uint32_t PV = c->P * c->Vx;
        const dim3 grid((PV + 255u) / 256u), block(256);
        const uint32_t buffered = c->buffered ? 1u : 0u;
        switch (pass)
        {
        case LBM_PASS_STREAM:
            if (c->buffered)
                refresh(c);
            hipLaunchKernelGGL((k_split_stream<L, S>), grid, block, 0, c->stream, p, buf, PV, buffered);
            break;
        case LBM_PASS_OUTLET:
            hipLaunchKernelGGL((k_split_outlet<S>), grid, block, 0, c->stream, p, buf, PV, buffered);
            break;
        case LBM_PASS_MACROSCOPIC:
            hipLaunchKernelGGL((k_split_moments<S>), grid, block, 0, c->stream, p, buf, PV, buffered);
            break;
        default:
            hipLaunchKernelGGL((k_split_collide<L, S>), grid, block, 0, c->stream, p, buf, PV, buffered);
            LBM_HIP_TRY(hipGetLastError());
            if (c->two_lattice)
                c->cur ^= 1;
            c->iteration++;
            end_reference(c);
            return LBM_OK;
        }
        LBM_HIP_TRY(hipGetLastError());
        return LBM_OK;
    }

    static void refresh(lbm_ctx *c)
    {
        T *lat = static_cast<T *>(c->lat[c->cur]);
        if (c->cnt_v > 1)
            hipLaunchKernelGGL((k_refresh_rows<L, T>), dim3((c->Hx + 255u) / 256u, c->cnt_v - 1u), dim3(256), 0,
                               c->stream, lat, c->Hx, c->P, c->X0, c->Sd, c->sub_v);
        if (c->cnt_h > 1)
            hipLaunchKernelGGL((k_refresh_cols<L, T>), dim3((c->Vx + 255u) / 256u, c->cnt_h - 1u), dim3(256), 0,
                               c->stream, lat, c->Vx, c->P, c->X0, c->Sd, c->sub_h);
    }

    // Side arrays of the tile sweep from the current lattice (both parities).
    static void prime(lbm_ctx *c)
    {
        T *lat = static_cast<T *>(c->lat[c->cur]);
        if (c->cnt_v > 1)
            hipLaunchKernelGGL((k_prime_side_rows<L, T>), dim3((c->Hx + 255u) / 256u, c->cnt_v - 1u), dim3(256), 0,
                               c->stream, lat, lat + c->side_rows[0], lat + c->side_rows[1], c->Hx, c->P, c->X0, c->Sd,
                               c->sub_v);
        if (c->cnt_h > 1)
            hipLaunchKernelGGL((k_prime_side_cols<L, T>), dim3((c->Vx + 255u) / 256u, c->cnt_h - 1u), dim3(256), 0,
                               c->stream, lat, lat + c->side_cols[0], lat + c->side_cols[1], c->Vx, c->P, c->X0, c->Sd,
                               c->sub_h);
        c->side_valid = true;
    }

    static Geom geom(const lbm_ctx *c)
    {
        Geom g{c->Hx, c->N, c->P, c->X0, c->Sd};
        g.ref = c->ref[c->ref_cur];
        g.H = c->H;
        g.buffered = c->buffered ? 1u : 0u;
        g.sub_h = c->sub_h;
        // the outlet ghost follows the last real column (BufferedNodeAccess of H - 2, plus one)
        g.ghost_xe = c->buffered ? (c->H - 2u) + (c->H - 3u) / c->sub_h + 1u : c->H - 1u;
        return g;
    }

    static int init(lbm_ctx *c)
    {
        const uint32_t blocks = (c->N + 255u) / 256u;
        T *l0 = static_cast<T *>(c->lat[0]);
        LBM_HIP_TRY(hipMemsetAsync(l0, 0, c->lattice_elems * sizeof(T), c->stream));
        hipLaunchKernelGGL((k_init_rest<L, S>), dim3(blocks), dim3(256), 0, c->stream, l0, geom(c));
        InoutValues v;
        for (int d = 0; d < 9; ++d)
        {
            v.v[d] = c->feq_in[d];
            v.v[9 + d] = c->feq_out[d];
        }
        if (c->V > 2u)
            hipLaunchKernelGGL((k_init_inout<L, S>), dim3((c->V - 2u + 255u) / 256u, 2), dim3(256), 0, c->stream,
                               l0, c->H, c->V, geom(c), c->sub_h, c->sub_v, c->buffered ? 1 : 0, v);
        LBM_HIP_TRY(hipGetLastError());
        if (c->two_lattice)
            LBM_HIP_TRY(hipMemcpyAsync(c->lat[1], c->lat[0], c->lattice_elems * sizeof(T), hipMemcpyDeviceToDevice,
                                       c->stream));
        return LBM_OK;
    }

    // populations i0 .. i0 + count - 1 (flat layout order) to / from dev[0 .. count - 1]
    static int decode(lbm_ctx *c, double *dev_out, uint32_t i0, uint32_t count)
    {
        hipLaunchKernelGGL((k_decode<L, S>), dim3((count + 255u) / 256u), dim3(256), 0, c->stream,
                           static_cast<const T *>(c->lat[c->cur]), dev_out, geom(c), i0, count);
        LBM_HIP_TRY(hipGetLastError());
        return LBM_OK;
    }

    static int encode(lbm_ctx *c, const double *dev_in, uint32_t i0, uint32_t count)
    {
        hipLaunchKernelGGL((k_encode<L, S>), dim3((count + 255u) / 256u), dim3(256), 0, c->stream,
                           static_cast<T *>(c->lat[c->cur]), dev_in, geom(c), i0, count);
        LBM_HIP_TRY(hipGetLastError());
        return LBM_OK;
    }

    static int equilibrium(lbm_ctx *c, const double *rho, const double *jx, const double *jy)
    {
        hipLaunchKernelGGL((k_set_equilibrium<L, S>), dim3((c->N + 255u) / 256u), dim3(256), 0, c->stream,
                           static_cast<T *>(c->lat[c->cur]), c->code, rho, jx, jy, geom(c), c->H, c->sub_h,
                           c->sub_v, c->buffered ? 1 : 0);
        LBM_HIP_TRY(hipGetLastError());
        return LBM_OK;
    }

    static int reduce(lbm_ctx *c, int mode, double *dev_out6)
    {
        hipLaunchKernelGGL((k_reduce<L, S>), dim3(RED_GRID), dim3(RED_BLOCK), 0, c->stream,
                           static_cast<const T *>(c->lat[c->cur]), c->code, geom(c), c->partials, c->snap, mode);
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(64), 0, c->stream, c->partials, RED_GRID, dev_out6);
        LBM_HIP_TRY(hipGetLastError());
        return LBM_OK;
    }
};

struct VTable
{
    int (*step)(lbm_ctx *, bool);
    int (*init)(lbm_ctx *);
    int (*decode)(lbm_ctx *, double *, uint32_t, uint32_t);
    int (*encode)(lbm_ctx *, const double *, uint32_t, uint32_t);
    int (*equilibrium)(lbm_ctx *, const double *, const double *, const double *);
    int (*reduce)(lbm_ctx *, int, double *);
    int (*split)(lbm_ctx *, int, double *);
};

template <int L, int S>
static VTable make_vtable()
{
    return VTable{&Impl<L, S>::step, &Impl<L, S>::init, &Impl<L, S>::decode, &Impl<L, S>::encode,
                  &Impl<L, S>::equilibrium, &Impl<L, S>::reduce, &Impl<L, S>::split};
}

template <int L>
static VTable vtable_for_storage(int s)
{
    switch (s)
    {
    case LBM_STORE_F64: return make_vtable<L, LBM_STORE_F64>();
    case LBM_STORE_F32: return make_vtable<L, LBM_STORE_F32>();
    case LBM_STORE_F16_SHIFTED: return make_vtable<L, LBM_STORE_F16_SHIFTED>();
    default: return make_vtable<L, LBM_STORE_BF16_SHIFTED>();
    }
}

static VTable vtable(const lbm_ctx *c)
{
    switch (c->desc.layout)
    {
    case LBM_LAYOUT_COLLISION: return vtable_for_storage<LBM_LAYOUT_COLLISION>(c->desc.storage);
    case LBM_LAYOUT_STREAM: return vtable_for_storage<LBM_LAYOUT_STREAM>(c->desc.storage);
    default: return vtable_for_storage<LBM_LAYOUT_BUNDLE>(c->desc.storage);
    }
}

static size_t storage_bytes(int s)
{
    switch (s)
    {
    case LBM_STORE_F64: return 8;
    case LBM_STORE_F32: return 4;
    default: return 2;
    }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *lbm_last_error(void) { return g_last_error.c_str(); }

int lbm_device_max_work_group_size(uint32_t *out)
{
    if (!out)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_device_max_work_group_size: null output");
    int dev = 0;
    LBM_HIP_TRY(hipGetDevice(&dev));
    int v = 0;
    LBM_HIP_TRY(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxThreadsPerBlock, dev));
    *out = uint32_t(v);
    return LBM_OK;
}

int lbm_device_name(char *buf, size_t len)
{
    if (!buf || len == 0)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_device_name: empty buffer");
    int dev = 0;
    LBM_HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop{};
    LBM_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    // the marketing name can be empty (e.g. in containers); then the architecture alone
    std::string market(prop.name);
    while (!market.empty() && market.back() == ' ')
        market.pop_back();
    const std::string arch(prop.gcnArchName);
    const std::string name = market.empty() ? arch : market + " (" + arch + ")";
    const size_t n = std::min(len - 1, name.size());
    std::memcpy(buf, name.data(), n);
    buf[n] = '\0';
    return LBM_OK;
}

static int validate_desc(const lbm_desc *d)
{
    if (d->algorithm < LBM_ALG_NPTL || d->algorithm > LBM_ALG_NSOL)
        return fail(LBM_ERR_INVALID_ARGUMENT, "unknown algorithm");
    if (d->layout < LBM_LAYOUT_COLLISION || d->layout > LBM_LAYOUT_BUNDLE)
        return fail(LBM_ERR_INVALID_ARGUMENT, "unknown data layout");
    if (d->storage < LBM_STORE_F64 || d->storage > LBM_STORE_BF16_SHIFTED)
        return fail(LBM_ERR_INVALID_ARGUMENT, "unknown storage type");
    if (d->rounding != LBM_ROUND_NEAREST && d->rounding != LBM_ROUND_STOCHASTIC)
        return fail(LBM_ERR_INVALID_ARGUMENT, "unknown rounding mode");
    if (d->tile_kernel > LBM_TILE_STAGED)
        return fail(LBM_ERR_INVALID_ARGUMENT, "unknown tile kernel");
    if (d->horizontal_nodes < 4 || d->vertical_nodes < 4)
        return fail(LBM_ERR_INVALID_ARGUMENT, "horizontal_nodes and vertical_nodes must be >= 4");
    if (!(d->relaxation_time > 0.0))
        return fail(LBM_ERR_INVALID_ARGUMENT, "relaxation_time must be > 0");
    if (d->work_group_size < 1 || d->work_group_size > 1024)
        return fail(LBM_ERR_INVALID_ARGUMENT, "work_group_size must be in [1, 1024]");
    const uint64_t Hx = d->expanded_horizontal_nodes, Vx = d->expanded_vertical_nodes;
    if (d->algorithm == LBM_ALG_LPTL && (Hx != d->horizontal_nodes || Vx != d->vertical_nodes))
        return fail(LBM_ERR_INVALID_ARGUMENT, "lptl runs on the unexpanded lattice");
    if (Hx < d->horizontal_nodes || Vx < d->vertical_nodes)
        return fail(LBM_ERR_INVALID_ARGUMENT, "expanded extents smaller than the lattice");
    if (9ull * Hx * Vx >= (1ull << 32))
        return fail(LBM_ERR_INVALID_ARGUMENT, "lattice too large for 32-bit population indices");
    const uint64_t sh = d->subdomain_horizontal_nodes, sv = d->subdomain_vertical_nodes;
    const uint64_t ch = d->subdomain_count_horizontal, cv = d->subdomain_count_vertical;
    if (d->algorithm == LBM_ALG_NPOL || d->algorithm == LBM_ALG_NSOL)
    {
        if (sh < 1 || sv < 1 || ch < 1 || cv < 1)
            return fail(LBM_ERR_INVALID_ARGUMENT, "subdomain extents must be >= 1");
        if (Hx != (sh + 1) * ch + 1 || Vx != (sv + 1) * cv + 1)
            return fail(LBM_ERR_INVALID_ARGUMENT, "expanded extents do not match the buffered subdomain grid");
        if (sh * ch < d->horizontal_nodes - 2 || sv * cv < d->vertical_nodes - 2)
            return fail(LBM_ERR_INVALID_ARGUMENT, "subdomain grid does not cover the lattice");
        const uint64_t threads = d->algorithm == LBM_ALG_NPOL ? sh * sv : (sh + 2) * (sv + 1);
        if (threads > 1024)
            return fail(LBM_ERR_INVALID_ARGUMENT, "subdomain work-group exceeds 1024 work-items");
    }
    if (d->algorithm == LBM_ALG_NPTL && (Hx < 3 || Vx < 3))
        return fail(LBM_ERR_INVALID_ARGUMENT, "nptl extents too small");
    return LBM_OK;
}

void lbm_destroy(lbm_ctx *c)
{
    if (!c)
        return;
    // both streams drained before anything they may still read or write is freed
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    if (c->copy_stream)
        (void)hipStreamSynchronize(c->copy_stream);
    (void)hipFree(c->lat[0]);
    (void)hipFree(c->lat[1]);
    (void)hipFree(c->phase);
    (void)hipFree(c->code);
    (void)hipFree(c->res);
    (void)hipFree(c->uycol);
    (void)hipFree(c->snap);
    (void)hipFree(c->pass);
    (void)hipFree(c->stage);
    (void)hipFree(c->partials);
    (void)hipFree(c->red_out);
    (void)hipFree(c->ref[0]);
    (void)hipFree(c->ref[1]);
    (void)hipFree(c->ref_part);
    if (c->ev0)
        (void)hipEventDestroy(c->ev0);
    if (c->ev1)
        (void)hipEventDestroy(c->ev1);
    if (c->copy_stream)
    {
        (void)hipStreamSynchronize(c->copy_stream);
        (void)hipStreamDestroy(c->copy_stream);
    }
    if (c->ev_frame)
        (void)hipEventDestroy(c->ev_frame);
    if (c->ev_copied)
        (void)hipEventDestroy(c->ev_copied);
    if (c->frame_host)
        (void)hipHostFree(c->frame_host);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
}

int lbm_create(const lbm_desc *desc, lbm_ctx **out)
{
    if (!desc || !out)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_create: null argument");
    *out = nullptr;
    if (int rc = validate_desc(desc))
        return rc;
    lbm_ctx *c = new lbm_ctx();
    c->desc = *desc;
    c->H = desc->horizontal_nodes;
    c->V = desc->vertical_nodes;
    c->Hx = desc->expanded_horizontal_nodes;
    c->Vx = desc->expanded_vertical_nodes;
    c->N = c->Hx * c->Vx;
    c->sub_h = std::max(1u, desc->subdomain_horizontal_nodes);
    c->sub_v = std::max(1u, desc->subdomain_vertical_nodes);
    c->cnt_h = desc->subdomain_count_horizontal;
    c->cnt_v = desc->subdomain_count_vertical;
    c->wg = desc->work_group_size;
    c->buffered = desc->algorithm == LBM_ALG_NPOL || desc->algorithm == LBM_ALG_NSOL;
    c->two_lattice = !c->buffered;
    // The buffered outlet addresses the right neighbour of node H-3
    // (general/buffered.hpp:285-288); when a buffer column separates H-3 and
    // H-2 that neighbour is the buffer and the outlet never reaches the ghost
    // column, which then keeps its initial equilibrium.
    c->outlet_active = !(c->buffered && (c->H - 3u) % c->sub_h == 0u);
    c->elem = storage_bytes(desc->storage);
    c->D = (c->H - 2u) * (c->V - 2u);
    // physical placement (see Placement in lbm_device.hpp)
    c->X0 = kRowOffset;
    c->P = (c->X0 + c->Hx + 63u) / 64u * 64u;
    {
        // Direction arrays start 2 MiB (small lattices: 4 KiB) apart plus
        // 512 bytes: the nine streams a wave reads and writes then do not
        // start at the same offset modulo the HBM channel interleave
        // (stream layout, 8192 x 2048 porous: -0.5 % FP64, -1.5 % FP32,
        // -1.7 % FP16 time per step against whole 2 MiB multiples).
        const uint64_t M = uint64_t(c->P) * c->Vx;
        const uint64_t align_nodes = (M * c->elem >= (32ull << 20) ? (2ull << 20) : 4096ull) / c->elem;
        const uint64_t Sd = (M + align_nodes - 1) / align_nodes * align_nodes + 512u / c->elem;
        if (9ull * Sd >= (1ull << 32))
        {
            delete c;
            return fail(LBM_ERR_INVALID_ARGUMENT, "lattice too large for 32-bit population indices");
        }
        c->Sd = uint32_t(Sd);
        c->lattice_elems = size_t(9) * Sd;
    }
    host_equilibrium(desc->storage, desc->inlet_velocity_x, desc->inlet_velocity_y, desc->inlet_density, c->feq_in);
    host_equilibrium(desc->storage, desc->outlet_velocity_x, desc->outlet_velocity_y, desc->outlet_density,
                     c->feq_out);

    auto cleanup = [&](int rc) {
        lbm_destroy(c);
        return rc;
    };
#define CREATE_TRY(expr)                                                                                 \
    do                                                                                                   \
    {                                                                                                    \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return cleanup(fail(e_ == hipErrorOutOfMemory ? LBM_ERR_OUT_OF_MEMORY : LBM_ERR_DEVICE,       \
                                std::string(#expr) + ": " + hipGetErrorString(e_)));                      \
    } while (0)

    CREATE_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CREATE_TRY(hipEventCreate(&c->ev0));
    CREATE_TRY(hipEventCreate(&c->ev1));
    // 64 bytes of tail: the staged nptl kernel reads its source ranges in
    // aligned 16-byte vectors, up to 15 bytes past the last place of the top
    // ghost row
    const size_t lat_bytes = c->lattice_elems * c->elem;
    // npol/nsol tiles at most one wave wide: the side arrays of the tile sweep
    // and of the staged block kernel go after the lattice in the same
    // allocation (32-bit element offsets), two of each kind, each starting on
    // 256 bytes.  Where the staged kernel does not run, npol's FP64 collision
    // layout keeps the block kernel with the refresh launches: its strided
    // 72-byte nodes ran faster there than under the sweep, 15.5 against 14.6
    // GFLUPS on the porous channel (nsol's 14.8 against 14.0).
    // A lattice with fewer tiles than two per compute unit runs the block
    // kernels under LBM_TILE_AUTO: the sweep gives each tile one wave, which
    // then walks it row group by row group (npol stream FP64 per step:
    // 256 x 64, 16 tiles, 25.7 us sweep against 9.1 us block; 1024 x 256,
    // 248 tiles, 28.4 / 20.1; 2048 x 512, 992 tiles, 41.5 / 50.6; 4096 x
    // 1024, 3844 tiles, 154 / 174).
    size_t side_bytes = 0;
    int cus = 0;
    {
        int dev = 0;
        CREATE_TRY(hipGetDevice(&dev));
        CREATE_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    const bool few_tiles = uint64_t(c->cnt_h) * c->cnt_v < 2ull * uint64_t(std::max(cus, 1));
    // the staged block kernel (collision layout, FP64 and FP32) once the
    // lattice has enough tiles: each block walks its tile chunk after chunk,
    // about 16 us on any lattice, which the block kernels beat on small ones
    // (FP64 at 768 x 192, 144 tiles, 19.4 against 19.5 us; FP32 at 1024 x
    // 256, 248 tiles, 17.7 against 18.6 us).  In 16-bit storage the sweep
    // kept its place: staged 36.5 / 37.5 against 35.9 / 37.4 GFLUPS (nsol /
    // npol FP16 on the porous 8192 x 2048 channel), its side arrays costing
    // 2-byte scattered loads and stores.  So did the stream and bundle
    // layouts, whose nodes the sweep's lanes read contiguously: staged FP64
    // npol stream 14.4 against 19.0, nsol bundle 13.6 against 21.3; FP32
    // 21.2 / 26.5 against 33.1 / 33.0; FP16 26.5 / 28.8 against 48.7 / 43.5.
    const uint64_t tiles = uint64_t(c->cnt_h) * c->cnt_v;
    const bool auto_staged = c->elem >= 4 && tiles >= uint64_t(std::max(cus, 1)) / (c->elem == 8 ? 2u : 1u);
    c->tile_staged = c->buffered && c->sub_h >= TILE_STAGED_MIN_WIDTH && c->sub_h <= TILE_SWEEP_MAX_WIDTH &&
                     (desc->tile_kernel == LBM_TILE_STAGED ||
                      (desc->tile_kernel == LBM_TILE_AUTO && auto_staged && desc->layout == LBM_LAYOUT_COLLISION));
    const bool auto_sweep =
        !few_tiles && !(desc->algorithm == LBM_ALG_NPOL && desc->layout == LBM_LAYOUT_COLLISION && c->elem == 8);
    const bool want_sweep = !c->tile_staged && c->buffered && c->sub_h <= TILE_SWEEP_MAX_WIDTH &&
                            desc->tile_kernel != LBM_TILE_BLOCK && (desc->tile_kernel == LBM_TILE_SWEEP || auto_sweep);
    if (want_sweep || c->tile_staged)
    {
        const size_t align = 256 / c->elem;
        auto round_up = [&](size_t n) { return (n + align - 1) / align * align; };
        const size_t rows = round_up(size_t(c->cnt_v - 1u) * 6u * c->Hx);
        const size_t cols = round_up(size_t(c->cnt_h - 1u) * 6u * c->Vx);
        const size_t first = round_up(c->lattice_elems + 64 / c->elem);
        if (first + 2 * (rows + cols) < (1ull << 32))
        {
            c->sweep = want_sweep;
            c->side_rows[0] = uint32_t(first);
            c->side_rows[1] = uint32_t(first + rows);
            c->side_cols[0] = uint32_t(first + 2 * rows);
            c->side_cols[1] = uint32_t(first + 2 * rows + cols);
            side_bytes = (first + 2 * (rows + cols)) * c->elem - lat_bytes;
        }
        else
            c->tile_staged = false;  // the block kernels with the reference's refresh
    }
    CREATE_TRY(hipMalloc(&c->lat[0], lat_bytes + std::max<size_t>(64, side_bytes)));
    if (c->two_lattice)
        CREATE_TRY(hipMalloc(&c->lat[1], lat_bytes + 64));
    CREATE_TRY(hipMalloc(&c->phase, c->N));
    CREATE_TRY(hipMalloc(&c->code, size_t(c->Sd) * sizeof(uint16_t)));
    CREATE_TRY(hipMalloc(&c->res, 4ull * std::max(1u, c->D) * sizeof(double)));
    CREATE_TRY(hipMalloc(&c->uycol, 2ull * c->V * sizeof(double)));
    CREATE_TRY(hipMalloc(&c->partials, size_t(RED_GRID) * RED_K * sizeof(double)));
    CREATE_TRY(hipMalloc(&c->red_out, RED_K * sizeof(double)));
    if (desc->storage == LBM_STORE_F16_SHIFTED || desc->storage == LBM_STORE_BF16_SHIFTED)
    {
        // rho_ref tables (+64: the kernels read a column to either side of
        // every lane's own, padding lanes included) and the update's sums
        c->ref_len = std::max(c->Hx, c->H) + 64u;
        CREATE_TRY(hipMalloc(&c->ref[0], c->ref_len * sizeof(float)));
        CREATE_TRY(hipMalloc(&c->ref[1], c->ref_len * sizeof(float)));
        CREATE_TRY(hipMalloc(&c->ref_part, 2ull * kRefChunks * c->Hx * sizeof(double)));
    }
    // Data fills phase with -1 (simulation.cpp:424); Results fill density -1,
    // momentum 0 (simulation.cpp:407-410).
    CREATE_TRY(hipMemsetAsync(c->phase, 0xff, c->N, c->stream));
    CREATE_TRY(hipMemsetAsync(c->code, 0, size_t(c->Sd) * sizeof(uint16_t), c->stream));
    CREATE_TRY(hipMemsetAsync(c->res, 0, 4ull * std::max(1u, c->D) * sizeof(double), c->stream));
    CREATE_TRY(hipMemsetAsync(c->uycol, 0, 2ull * c->V * sizeof(double), c->stream));
    if (c->D)
    {
        hipLaunchKernelGGL(k_fill, dim3((c->D + 255u) / 256u), dim3(256), 0, c->stream, c->res, size_t(c->D), -1.0);
        CREATE_TRY(hipGetLastError());
    }
    CREATE_TRY(hipStreamSynchronize(c->stream));
#undef CREATE_TRY
    *out = c;
    return LBM_OK;
}

int lbm_set_phase(lbm_ctx *c, const int8_t *phase, size_t n)
{
    if (!c || !phase)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_set_phase: null argument");
    if (n != c->N)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_set_phase: node_count must equal the expanded node count");
    LBM_HIP_TRY(hipMemcpyAsync(c->phase, phase, n, hipMemcpyHostToDevice, c->stream));
    LBM_HIP_TRY(hipStreamSynchronize(c->stream));
    c->phase_set = true;
    c->initialized = false;
    return LBM_OK;
}

int lbm_get_phase(lbm_ctx *c, int8_t *phase, size_t n)
{
    if (!c || !phase)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_phase: null argument");
    if (n != c->N)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_phase: node_count must equal the expanded node count");
    LBM_HIP_TRY(hipMemcpyAsync(phase, c->phase, n, hipMemcpyDeviceToHost, c->stream));
    LBM_HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

// The 16-bit forms' initial rho_ref table: the rest state's 1, the ghost
// columns at rho_in and rho_out.
static int reset_reference(lbm_ctx *c)
{
    c->ref_cur = 0;
    c->ref_switch = false;
    if (!c->ref[0])
        return LBM_OK;
    std::vector<float> ref(c->ref_len, 1.f);
    host_reference(c->desc.inlet_density, c->desc.outlet_density, c->H, ref.data());
    LBM_HIP_TRY(hipMemcpyAsync(c->ref[0], ref.data(), ref.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    LBM_HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_init_populations(lbm_ctx *c)
{
    if (!c)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_init_populations: null context");
    if (!c->phase_set)
        return fail(LBM_ERR_STATE, "lbm_init_populations: lbm_set_phase must come first");
    c->cur = 0;
    c->side_valid = false;
    if (int rc = reset_reference(c))
        return rc;
    if (int rc = vtable(c).init(c))
        return rc;
    CodeParams cp{c->H,  c->V,  c->Hx, c->N, c->sub_h, c->sub_v, c->P, c->X0, c->buffered ? 1 : 0,
                  c->outlet_active ? 1 : 0};
    LBM_HIP_TRY(hipMemsetAsync(c->code, 0, size_t(c->Sd) * sizeof(uint16_t), c->stream));
    hipLaunchKernelGGL(k_build_codes, dim3((c->N + 255u) / 256u), dim3(256), 0, c->stream, c->phase, c->code, cp);
    LBM_HIP_TRY(hipGetLastError());
    LBM_HIP_TRY(hipMemsetAsync(c->uycol, 0, 2ull * c->V * sizeof(double), c->stream));
    LBM_HIP_TRY(hipStreamSynchronize(c->stream));
    c->iteration = 0;
    c->snap_valid = false;
    c->frame_iteration = ~0ull;
    c->initialized = true;
    return LBM_OK;
}

int lbm_get_node_codes(lbm_ctx *c, uint16_t *codes, size_t n)
{
    if (!c || !codes)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_node_codes: null argument");
    if (!c->initialized)
        return fail(LBM_ERR_STATE, "lbm_get_node_codes: call lbm_init_populations first");
    if (n != c->N)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_node_codes: node_count must equal the expanded node count");
    uint16_t *tmp = nullptr;
    LBM_HIP_TRY(hipMalloc(&tmp, n * sizeof(uint16_t)));
    const Geom g{c->Hx, c->N, c->P, c->X0, c->Sd};
    hipLaunchKernelGGL(k_gather_codes, dim3((c->N + 255u) / 256u), dim3(256), 0, c->stream, c->code, tmp, g);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(codes, tmp, n * sizeof(uint16_t), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess)
        return fail(LBM_ERR_DEVICE, std::string("lbm_get_node_codes: ") + hipGetErrorString(e));
    return LBM_OK;
}

int lbm_run(lbm_ctx *c, uint32_t steps, uint32_t frame_interval)
{
    if (!c)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_run: null context");
    if (!c->initialized)
        return fail(LBM_ERR_STATE, "lbm_run: call lbm_init_populations first");
    const VTable vt = vtable(c);
    LBM_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    for (uint32_t i = 0; i < steps; ++i)
    {
        const uint64_t it = c->iteration;
        const bool frame = frame_interval ? (it % frame_interval == frame_interval - 1u) : false;
        const bool macros = frame || i + 1u == steps;
        if (int rc = vt.step(c, macros))
            return rc;
    }
    LBM_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    LBM_HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    LBM_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return LBM_OK;
}

int lbm_run_frames(lbm_ctx *c, uint32_t steps, uint32_t frame_interval)
{
    if (!c)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_run_frames: null context");
    if (!c->initialized)
        return fail(LBM_ERR_STATE, "lbm_run_frames: call lbm_init_populations first");
    if (!c->copy_stream)
    {
        LBM_HIP_TRY(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
        LBM_HIP_TRY(hipEventCreateWithFlags(&c->ev_frame, hipEventDisableTiming));
        LBM_HIP_TRY(hipEventCreateWithFlags(&c->ev_copied, hipEventDisableTiming));
        LBM_HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&c->frame_host),
                                  4ull * std::max(1u, c->D) * sizeof(double), hipHostMallocDefault));
    }
    const VTable vt = vtable(c);
    uint64_t last_frame = ~0ull;
    LBM_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    for (uint32_t i = 0; i < steps; ++i)
    {
        const uint64_t it = c->iteration;
        const bool frame = frame_interval ? (it % frame_interval == frame_interval - 1u) : false;
        const bool macros = frame || i + 1u == steps;
        if (macros && c->copy_pending)
        {
            // the results must not change under the copy still reading them
            LBM_HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_copied, 0));
            c->copy_pending = false;
        }
        if (int rc = vt.step(c, macros))
            return rc;
        if (frame)
        {
            LBM_HIP_TRY(hipEventRecord(c->ev_frame, c->stream));
            LBM_HIP_TRY(hipStreamWaitEvent(c->copy_stream, c->ev_frame, 0));
            LBM_HIP_TRY(hipMemcpyAsync(c->frame_host, c->res, 4ull * c->D * sizeof(double), hipMemcpyDeviceToHost,
                                       c->copy_stream));
            LBM_HIP_TRY(hipEventRecord(c->ev_copied, c->copy_stream));
            c->copy_pending = true;
            last_frame = it;
        }
    }
    LBM_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    LBM_HIP_TRY(hipEventSynchronize(c->ev1));
    LBM_HIP_TRY(hipStreamSynchronize(c->copy_stream));
    c->copy_pending = false;
    if (last_frame != ~0ull)
        c->frame_iteration = last_frame;
    float ms = 0.f;
    LBM_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
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
            std::memcpy(dst[k], c->frame_host + size_t(k) * c->D, count * sizeof(double));
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
    if (!c->pass)
        LBM_HIP_TRY(hipMalloc(&c->pass, 9ull * c->P * c->Vx * sizeof(double)));
    const VTable vt = vtable(c);
    double total_ms = 0.0;
    for (uint32_t i = 0; i < steps; ++i)
    {
        for (int pass = LBM_PASS_STREAM; pass <= LBM_PASS_COLLIDE; ++pass)
        {
            LBM_HIP_TRY(hipEventRecord(c->ev0, c->stream));
            if (int rc = vt.split(c, pass, c->pass))
                return rc;
            LBM_HIP_TRY(hipEventRecord(c->ev1, c->stream));
            LBM_HIP_TRY(hipEventSynchronize(c->ev1));
            float ms = 0.f;
            LBM_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            total_ms += ms;
            if (cb)
                cb(user, pass, c->iteration);
        }
    }
    c->last_ms = total_ms;
    return LBM_OK;
}

int lbm_get_pass_populations(lbm_ctx *c, double *values, size_t count)
{
    if (!c || !values)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_pass_populations: null argument");
    if (!c->pass)
        return fail(LBM_ERR_STATE, "lbm_get_pass_populations: no split step has run");
    if (count != 9ull * c->N)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_pass_populations: count must be 9 * expanded node count");
    const size_t PV = size_t(c->P) * c->Vx;
    std::vector<double> phys(9 * PV);
    LBM_HIP_TRY(hipMemcpyAsync(phys.data(), c->pass, phys.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    // the 16-bit forms' passes hold deviations from each node's rho_ref: add w_d rho_ref back
    std::vector<float> ref(c->ref_len);
    if (c->ref_len)
        LBM_HIP_TRY(hipMemcpyAsync(ref.data(), c->ref[c->ref_cur], ref.size() * sizeof(float), hipMemcpyDeviceToHost,
                                   c->stream));
    LBM_HIP_TRY(hipStreamSynchronize(c->stream));
    static const float w[9] = {1.f / 36, 1.f / 9, 1.f / 36, 1.f / 9, 4.f / 9, 1.f / 9, 1.f / 36, 1.f / 9, 1.f / 36};
    for (uint32_t n = 0; n < c->N; ++n)
    {
        const uint32_t xe = n % c->Hx;
        const size_t pn = c->X0 + xe + size_t(n / c->Hx) * c->P;
        const uint32_t xu = c->buffered ? xe - xe / (c->sub_h + 1u) : xe;
        for (int d = 0; d < 9; ++d)
        {
            double v = phys[size_t(d) * PV + pn];
            if (c->ref_len && xu < c->ref_len)
                v += double(w[d] * ref[xu]);
            values[size_t(d) * c->N + n] = v;
        }
    }
    return LBM_OK;
}

int lbm_get_iteration(lbm_ctx *c, uint64_t *it)
{
    if (!c || !it)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_iteration: null argument");
    *it = c->iteration;
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
            LBM_HIP_TRY(hipMemcpyAsync(dst[k], c->res + size_t(k) * c->D, count * sizeof(double),
                                       hipMemcpyDeviceToHost, c->stream));
    LBM_HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

// iteration, lattice index, lattice count; the lattice(s); outlet history;
// results; the 16-bit forms' rho_ref table
static size_t state_bytes(const lbm_ctx *c)
{
    const size_t lattices = c->two_lattice ? 2 : 1;
    return 16 + lattices * c->lattice_elems * c->elem + 2ull * c->V * sizeof(double) + 4ull * c->D * sizeof(double) +
           size_t(c->ref_len) * sizeof(float);
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
    const uint32_t lattices = c->two_lattice ? 2u : 1u;
    const uint32_t cur = uint32_t(c->cur);
    std::memcpy(out, &c->iteration, 8);
    std::memcpy(out + 8, &cur, 4);
    std::memcpy(out + 12, &lattices, 4);
    size_t off = 16;
    const size_t lat_bytes = c->lattice_elems * c->elem;
    // the caller's buffer is pageable host memory: blocking copies, after
    // the work on the context's stream (non-blocking) has finished
    LBM_HIP_TRY(hipStreamSynchronize(c->stream));
    for (uint32_t l = 0; l < lattices; ++l, off += lat_bytes)
        LBM_HIP_TRY(hipMemcpy(out + off, c->lat[l], lat_bytes, hipMemcpyDeviceToHost));
    LBM_HIP_TRY(hipMemcpy(out + off, c->uycol, 2ull * c->V * sizeof(double), hipMemcpyDeviceToHost));
    off += 2ull * c->V * sizeof(double);
    LBM_HIP_TRY(hipMemcpy(out + off, c->res, 4ull * c->D * sizeof(double), hipMemcpyDeviceToHost));
    off += 4ull * c->D * sizeof(double);
    if (c->ref_len)
        LBM_HIP_TRY(hipMemcpy(out + off, c->ref[c->ref_cur], c->ref_len * sizeof(float), hipMemcpyDeviceToHost));
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
    uint64_t iteration = 0;
    uint32_t cur = 0, lattices = 0;
    std::memcpy(&iteration, in, 8);
    std::memcpy(&cur, in + 8, 4);
    std::memcpy(&lattices, in + 12, 4);
    if (lattices != (c->two_lattice ? 2u : 1u) || cur >= lattices)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_set_state: state of another algorithm");
    size_t off = 16;
    const size_t lat_bytes = c->lattice_elems * c->elem;
    // pageable source: blocking copies once the context's stream is idle,
    // and a device-wide wait before the next launch reads the data
    LBM_HIP_TRY(hipStreamSynchronize(c->stream));
    for (uint32_t l = 0; l < lattices; ++l, off += lat_bytes)
        LBM_HIP_TRY(hipMemcpy(c->lat[l], in + off, lat_bytes, hipMemcpyHostToDevice));
    LBM_HIP_TRY(hipMemcpy(c->uycol, in + off, 2ull * c->V * sizeof(double), hipMemcpyHostToDevice));
    off += 2ull * c->V * sizeof(double);
    LBM_HIP_TRY(hipMemcpy(c->res, in + off, 4ull * c->D * sizeof(double), hipMemcpyHostToDevice));
    off += 4ull * c->D * sizeof(double);
    c->ref_cur = 0;
    c->ref_switch = false;
    if (c->ref_len)
        LBM_HIP_TRY(hipMemcpy(c->ref[0], in + off, c->ref_len * sizeof(float), hipMemcpyHostToDevice));
    LBM_HIP_TRY(hipDeviceSynchronize());
    c->iteration = iteration;
    c->frame_iteration = ~0ull;
    c->cur = int(cur);
    c->snap_valid = false;
    c->side_valid = false;
    return LBM_OK;
}

// The population accessors convert between the storage and FP64 through a
// staging buffer of at most kStageDoubles that the context keeps after the
// first call, one chunk at a time.
constexpr uint32_t kStageDoubles = 1u << 22;  // 32 MiB

static int ensure_stage(lbm_ctx *c)
{
    if (!c->stage)
        LBM_HIP_TRY(hipMalloc(&c->stage, size_t(std::min(kStageDoubles, 9u * c->N)) * sizeof(double)));
    return LBM_OK;
}

int lbm_get_populations(lbm_ctx *c, double *values, size_t count)
{
    if (!c || !values)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_populations: null argument");
    if (count != 9ull * c->N)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_get_populations: count must be 9 * expanded node count");
    if (int rc = ensure_stage(c))
        return rc;
    const VTable vt = vtable(c);
    for (uint32_t i0 = 0; i0 < count; i0 += kStageDoubles)
    {
        const uint32_t n = uint32_t(std::min<size_t>(kStageDoubles, count - i0));
        if (int rc = vt.decode(c, c->stage, i0, n))
            return rc;
        LBM_HIP_TRY(hipMemcpyAsync(values + i0, c->stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        LBM_HIP_TRY(hipStreamSynchronize(c->stream));  // the next chunk reuses the staging buffer
    }
    return LBM_OK;
}

int lbm_set_populations(lbm_ctx *c, const double *values, size_t count)
{
    if (!c || !values)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_set_populations: null argument");
    if (!c->initialized)
        return fail(LBM_ERR_STATE, "lbm_set_populations: call lbm_init_populations first");
    if (count != 9ull * c->N)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_set_populations: count must be 9 * expanded node count");
    if (int rc = ensure_stage(c))
        return rc;
    c->side_valid = false;
    const VTable vt = vtable(c);
    for (uint32_t i0 = 0; i0 < count; i0 += kStageDoubles)
    {
        const uint32_t n = uint32_t(std::min<size_t>(kStageDoubles, count - i0));
        LBM_HIP_TRY(hipMemcpyAsync(c->stage, values + i0, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        if (int rc = vt.encode(c, c->stage, i0, n))
            return rc;
        LBM_HIP_TRY(hipStreamSynchronize(c->stream));
    }
    if (c->two_lattice)
        LBM_HIP_TRY(hipMemcpyAsync(c->lat[1 - c->cur], c->lat[c->cur], c->lattice_elems * c->elem,
                                   hipMemcpyDeviceToDevice, c->stream));
    LBM_HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_set_equilibrium_field(lbm_ctx *c, const double *rho, const double *jx, const double *jy, size_t count)
{
    if (!c || !rho || !jx || !jy)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_set_equilibrium_field: null argument");
    if (!c->initialized)
        return fail(LBM_ERR_STATE, "lbm_set_equilibrium_field: call lbm_init_populations first");
    if (count != c->D)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_set_equilibrium_field: count must equal (H-2)*(V-2)");
    double *tmp = nullptr;
    LBM_HIP_TRY(hipMalloc(&tmp, 3 * count * sizeof(double)));
    c->side_valid = false;
    hipError_t e = hipSuccess;
    const double *src[3] = {rho, jx, jy};
    for (int k = 0; k < 3 && e == hipSuccess; ++k)
        e = hipMemcpyAsync(tmp + k * count, src[k], count * sizeof(double), hipMemcpyHostToDevice, c->stream);
    int rc = e == hipSuccess ? vtable(c).equilibrium(c, tmp, tmp + count, tmp + 2 * count)
                             : fail(LBM_ERR_DEVICE, std::string("lbm_set_equilibrium_field: ") + hipGetErrorString(e));
    e = hipStreamSynchronize(c->stream);
    if (!rc && e != hipSuccess)
        rc = fail(LBM_ERR_DEVICE, std::string("lbm_set_equilibrium_field: ") + hipGetErrorString(e));
    (void)hipFree(tmp);
    c->snap_valid = false;
    return rc;
}

static int reduce_common(lbm_ctx *c, int mode, double host_out[RED_K])
{
    double *out_tmp = c->red_out;
    int rc = vtable(c).reduce(c, mode, out_tmp);
    if (!rc)
    {
        hipError_t e = hipMemcpyAsync(host_out, out_tmp, RED_K * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess)
            rc = fail(LBM_ERR_DEVICE, std::string("reduction: ") + hipGetErrorString(e));
    }
    return rc;
}

int lbm_reduce(lbm_ctx *c, double out[6])
{
    if (!c || !out)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_reduce: null argument");
    if (!c->initialized)
        return fail(LBM_ERR_STATE, "lbm_reduce: call lbm_init_populations first");
    return reduce_common(c, 0, out);
}

int lbm_residual(lbm_ctx *c, double *residual)
{
    if (!c || !residual)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_residual: null argument");
    if (!c->initialized)
        return fail(LBM_ERR_STATE, "lbm_residual: call lbm_init_populations first");
    if (!c->snap)
        LBM_HIP_TRY(hipMalloc(&c->snap, 2ull * c->N * sizeof(double)));
    if (!c->snap_valid)
    {
        LBM_HIP_TRY(hipMemsetAsync(c->snap, 0, 2ull * c->N * sizeof(double), c->stream));
        c->snap_valid = true;
    }
    double out[RED_K];
    if (int rc = reduce_common(c, 1, out))
        return rc;
    *residual = out[1] > 0.0 ? std::sqrt(out[0] / out[1]) : (out[0] > 0.0 ? 1.0 : 0.0);
    return LBM_OK;
}

int lbm_sync(lbm_ctx *c)
{
    if (!c)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_sync: null context");
    LBM_HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_measure_copy_bandwidth(size_t bytes, int repeats, double *gbps)
{
    if (!gbps || bytes < 16 || repeats < 1)
        return fail(LBM_ERR_INVALID_ARGUMENT, "lbm_measure_copy_bandwidth: bad argument");
    bytes &= ~size_t(15);
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = LBM_OK;
    auto check = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && rc == LBM_OK)
            rc = fail(LBM_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
        return e == hipSuccess;
    };
    if (check(hipMalloc(&a, bytes), "hipMalloc") && check(hipMalloc(&b, bytes), "hipMalloc") &&
        check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate") &&
        check(hipEventCreate(&e0), "hipEventCreate") && check(hipEventCreate(&e1), "hipEventCreate") &&
        check(hipMemsetAsync(a, 0, bytes, s), "hipMemsetAsync"))
    {
        const size_t n = bytes / 16;
        double best = 0.0;
        for (int r = 0; r <= repeats && rc == LBM_OK; ++r)
        {
            check(hipEventRecord(e0, s), "hipEventRecord");
            hipLaunchKernelGGL(k_copy16, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s,
                               static_cast<const uint4 *>(a), static_cast<uint4 *>(b), n);
            check(hipGetLastError(), "k_copy16");
            check(hipEventRecord(e1, s), "hipEventRecord");
            check(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0.f;
            check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            if (r > 0 && ms > 0.f)
                best = std::max(best, 2.0 * double(bytes) / (double(ms) * 1e6));
        }
        *gbps = best;
    }
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    if (s)
        (void)hipStreamDestroy(s);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

}  // extern "C"
