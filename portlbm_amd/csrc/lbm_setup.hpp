// Set-up and conversion kernels shared by the single-lattice and the batched
// contexts: initial populations, node codes, layout conversion.
#pragma once

#include "lbm_device.hpp"

namespace lbm_dev {

// ---------------------------------------------------------------------------
// set-up kernels
// ---------------------------------------------------------------------------
struct Geom
{
    uint32_t Hx, N, P, X0, Sd;
    const float *ref;   // rho_ref per unexpanded column (16-bit forms)
    uint32_t H;         // unexpanded width, ghosts included
    uint32_t buffered;  // npol / nsol: buffer columns every sub_h + 1
    uint32_t sub_h;
    uint32_t ghost_xe;  // expanded column of the outlet ghost
};

__device__ __forceinline__ uint32_t place_logical(uint32_t n, const Geom &g)
{
    return place(n % g.Hx, n / g.Hx, g.P, g.X0);
}

// Unexpanded column of expanded column xe: a buffer column maps to the real
// column on its left, the outlet ghost to H - 1 (it follows the last real
// column directly, also where that puts it on a buffer column's place).
__device__ __forceinline__ uint32_t unexpanded_x(uint32_t xe, const Geom &g)
{
    if (!g.buffered)
        return xe;
    if (xe >= g.ghost_xe)
        return g.H - 1u + (xe - g.ghost_xe);
    return xe - xe / (g.sub_h + 1u);
}

// rho_ref of the node whose population slot (xe, d) holds: a buffer column
// carries, in directions 0, 3, 6, the populations of the real node to its
// right (general/buffered.hpp:133-152), in the others those of the node to its left.
template <int S>
__device__ __forceinline__ float slot_rt(uint32_t xe, int d, const Geom &g)
{
    uint32_t xu = unexpanded_x(xe, g);
    if (g.buffered && xe != 0u && xe < g.ghost_xe && xe % (g.sub_h + 1u) == 0u && cx_of(d) == -1)
        xu += 1u;
    return rt_of<S>(g.ref, xu);
}

template <int L, int S>
__global__ void k_init_rest(typename Storage<S>::T *lat, Geom g)
{
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.N)
        return;
    const uint32_t pn = place_logical(n, g);
#pragma unroll
    for (int d = 0; d < 9; ++d)
        lat[at<L>(pn, uint32_t(d), g.Sd)] = Storage<S>::encode(weight<double>(d), d, slot_rt<S>(n % g.Hx, d, g));
}

struct InoutValues
{
    double v[18];
};

// Inlet ghost column x = 0 and outlet ghost column x = H-1, rows 1..V-2, at
// their equilibria (domain_initialization.hpp:382-398).
template <int L, int S>
__global__ void k_init_inout(typename Storage<S>::T *lat, uint32_t H, uint32_t V, Geom g, uint32_t sub_h,
                             uint32_t sub_v, int buffered, InoutValues vals)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t side = blockIdx.y;
    if (j >= V - 2u)
        return;
    const uint32_t x = 1u + side * (H - 3u);
    const uint32_t y = 1u + j;
    const uint32_t xe = buffered ? x + (x - 1u) / sub_h : x;
    const uint32_t ye = buffered ? y + (y - 1u) / sub_v : y;
    const uint32_t node = place(side ? xe + 1u : xe - 1u, ye, g.P, g.X0);
    const float rt = rt_of<S>(g.ref, side ? H - 1u : 0u);
#pragma unroll
    for (int d = 0; d < 9; ++d)
        lat[at<L>(node, uint32_t(d), g.Sd)] = Storage<S>::encode(vals.v[side * 9 + d], d, rt);
}

struct CodeParams
{
    uint32_t H, V, Hx, N, sub_h, sub_v, P, X0;
    int buffered;
    int outlet_active;
};

__device__ __forceinline__ uint32_t map_axis(uint32_t u, uint32_t last, uint32_t sub, int buffered)
{
    // unexpanded coordinate -> expanded (BufferedNodeAccess, access.hpp:292-301);
    // the high ghost sits right after the last real node
    if (!buffered)
        return u;
    if (u == 0u)
        return 0u;
    if (u == last)
        return (last - 1u) + (last - 2u) / sub + 1u;
    return u + (u - 1u) / sub;
}

// Codes of every logical node n, stored at its physical place.  Padding
// places keep 0 (not fluid, not solid).
[[maybe_unused]] static __global__ void k_build_codes(const int8_t *phase, uint16_t *code, CodeParams cp)
{
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= cp.N)
        return;
    const int8_t ph = phase[n];
    uint32_t c = 0;
    const uint32_t xe = n % cp.Hx, ye = n / cp.Hx;
    if (ph == 1)
        c = LBM_CODE_SOLID;
    else if (ph == 0)
    {
        const uint32_t x = cp.buffered ? xe - xe / (cp.sub_h + 1u) : xe;
        const uint32_t y = cp.buffered ? ye - ye / (cp.sub_v + 1u) : ye;
        c = LBM_CODE_FLUID;
        for (int d = 0; d < 9; ++d)
        {
            if (d == 4)
                continue;
            const uint32_t sx = uint32_t(int32_t(x) - cx_of(d));
            const uint32_t sy = uint32_t(int32_t(y) - cy_of(d));
            const uint32_t m = map_axis(sx, cp.H - 1u, cp.sub_h, cp.buffered) +
                               map_axis(sy, cp.V - 1u, cp.sub_v, cp.buffered) * cp.Hx;
            if (phase[m] == 1)
                c |= 1u << d;
        }
        if (cp.outlet_active && x == cp.H - 2u && y >= 2u && y + 3u <= cp.V)
            c |= LBM_CODE_OUTLET;
        if (x + 3u == cp.H && y >= 2u && y + 3u <= cp.V)
            c |= LBM_CODE_UY_SOURCE;
    }
    code[place(xe, ye, cp.P, cp.X0)] = uint16_t(c);
}

[[maybe_unused]] static __global__ void k_gather_codes(const uint16_t *code, uint16_t *out, Geom g)
{
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < g.N)
        out[n] = code[place_logical(n, g)];
}

[[maybe_unused]] static __global__ void k_fill(double *p, size_t n, double v)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// logical flat index i of a layout (reference order over N nodes) ->
// (node, direction)
template <int L>
__device__ __forceinline__ void split_flat(uint32_t i, uint32_t N, uint32_t &n, uint32_t &d)
{
    if constexpr (L == LBM_LAYOUT_COLLISION)
    {
        n = i / 9u;
        d = i % 9u;
    }
    else if constexpr (L == LBM_LAYOUT_STREAM)
    {
        d = i / N;
        n = i % N;
    }
    else
    {
        const uint32_t g = i / (3u * N), r = i % (3u * N);
        n = r / 3u;
        d = 3u * g + r % 3u;
    }
}

// Populations i0 .. i0 + count - 1 of the reference's flat layout order, to
// or from out[0 .. count - 1].
template <int L, int S>
__global__ void k_decode(const typename Storage<S>::T *lat, double *out, Geom g, uint32_t i0, uint32_t count)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count)
        return;
    uint32_t n, d;
    split_flat<L>(i0 + k, g.N, n, d);
    out[k] = Storage<S>::decode(lat[at<L>(place_logical(n, g), d, g.Sd)], int(d), slot_rt<S>(n % g.Hx, int(d), g));
}

template <int L, int S>
__global__ void k_encode(typename Storage<S>::T *lat, const double *in, Geom g, uint32_t i0, uint32_t count)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count)
        return;
    uint32_t n, d;
    split_flat<L>(i0 + k, g.N, n, d);
    lat[at<L>(place_logical(n, g), d, g.Sd)] = Storage<S>::encode(in[k], int(d), slot_rt<S>(n % g.Hx, int(d), g));
}

// Equilibrium of a given (rho, j) field on every fluid node (warm start).
template <int L, int S>
__global__ void k_set_equilibrium(typename Storage<S>::T *lat, const uint16_t *code, const double *rho,
                                  const double *jx, const double *jy, Geom g, uint32_t H, uint32_t sub_h,
                                  uint32_t sub_v, int buffered)
{
    using C = typename Storage<S>::C;
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.N)
        return;
    const uint32_t pn = place_logical(n, g);
    if (!(code[pn] & LBM_CODE_FLUID))
        return;
    const uint32_t xe = n % g.Hx, ye = n / g.Hx;
    const uint32_t x = buffered ? xe - xe / (sub_h + 1u) : xe;
    const uint32_t y = buffered ? ye - ye / (sub_v + 1u) : ye;
    const uint32_t r = (x - 1u) + (y - 1u) * (H - 2u);
    const C rh = C(rho[r]), ux = C(jx[r]), uy = C(jy[r]);
    const C usq = ux * ux + uy * uy;
    const float rt = rt_of<S>(g.ref, x);
#pragma unroll
    for (int d = 0; d < 9; ++d)
    {
        const C cu = C(cx_of(d)) * ux + C(cy_of(d)) * uy;
        const C feq = weight<C>(d) * (rh + C(3) * cu + C(4.5) * cu * cu - C(1.5) * usq);
        Storage<S>::store(lat, at<L>(pn, uint32_t(d), g.Sd), feq, d, rt);
    }
}


}  // namespace lbm_dev
