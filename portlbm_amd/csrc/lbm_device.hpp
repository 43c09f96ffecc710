// Device code shared by every D2Q9 BGK kernel of the library: population
// layouts, storage types, the pull with fused half-way bounce-back, the
// fused Zou-He pressure outlet and the BGK collision.
//
// Arithmetic follows the reference kernel text term by term so that the FP64
// path reproduces it bit for bit when no contraction is applied (the library
// is compiled with -ffp-contract=off):
//   moments + collision: include/lbm/gpu/nptl/nptl_kernels.hpp:371-437
//   outlet:              include/lbm/gpu/general/non_buffered.hpp:75-130
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <type_traits>

#include "../../include/lbm_hip.h"

namespace lbm_dev {

// ---- D2Q9 lattice (core/constants.hpp:31-39, core/access.hpp:142-152) ----
// direction d: cx = d % 3 - 1, cy = d / 3 - 1, opposite = 8 - d.
__device__ __forceinline__ constexpr int cx_of(int d) { return d % 3 - 1; }
__device__ __forceinline__ constexpr int cy_of(int d) { return d / 3 - 1; }

template <class C>
__device__ __forceinline__ C weight(int d)
{
    // 1/36, 1/9, 1/36, 1/9, 4/9, 1/9, 1/36, 1/9, 1/36 (constants.hpp:38-39)
    return (d == 4) ? C(4.0 / 9) : ((d & 1) ? C(1.0 / 9) : C(1.0 / 36));
}

// ---- population layouts (core/access.hpp:56-60, 81-85, 106-110) ---------
template <int L>
__device__ __forceinline__ uint32_t at(uint32_t node, uint32_t d, uint32_t total)
{
    if constexpr (L == LBM_LAYOUT_COLLISION)
        return 9u * node + d;
    else if constexpr (L == LBM_LAYOUT_STREAM)
        return total * d + node;
    else
        return 3u * (d / 3u) * total + (d % 3u) + 3u * node;
}

// ---- storage types ------------------------------------------------------
// T: what lives in HBM; C: the arithmetic type of the collision.
// kPromote: the arithmetic of the reference's USE_FLOAT build
// (core/constants.hpp:17-21).  Its kernels keep their double literals, so in
// the equilibrium the terms 9.0 / 2 * c.j * c.j and 3.0 / 2 * |j|^2 are
// double, the sum is rounded to float once (nptl_kernels.hpp:401-402), and
// the outlet's rebuilt populations are computed in double and rounded once
// (non_buffered.hpp:117-129).  FP32 storage follows that build bit for bit;
// the 16-bit forms have no reference counterpart and compute in float.
// kFull: the moment and equilibrium sums keep every term of the reference,
// also those whose velocity component is zero (x + 0 * f).  For finite
// populations leaving them out changes no result; once a population is
// infinite, 0 * inf is NaN in the reference, and the full sums propagate a
// blow-up exactly as it does (FP64, FP32).  The 16-bit forms leave them out.
//
// The 16-bit forms store the deviation f_d - w_d rho_ref(x) of a population
// from the rest equilibrium at a reference density of its column: x is the
// unexpanded column of the node that holds the population, and rho_ref(x)
// the mean density of the column's fluid nodes, refreshed every few hundred
// steps (a table of H floats; the ghost columns keep rho_in and rho_out).
// The stored deviation then carries the velocity and the density's departure
// from the column mean instead of the pressure drop along the channel, and
// the 16-bit grid is that much finer where the velocity signal lives.  A
// fixed reference of 1 (the shift w_d alone) leaves the density drop in the
// stored value; a fixed linear drop from rho_in to rho_out misrepresents the
// fluid still at rest while the pressure wave crosses the channel.  Measured
// in FP16 on a 1024 x 256 porous channel against FP64 (momentum L2; rounding
// to nearest / stochastic): after 500 steps w_d 5.3e-4 / 7.4e-4, linear
// 1.1e-3 / 1.4e-3, column mean 3.1e-4 / 4.2e-4; after 5000 steps w_d
// 3.1e-3 / 1.6e-3, linear 1.1e-3 / 6.3e-4, column mean 1.0e-3 / 5.1e-4.
// `rt` below is rho_ref of the node whose population is read or written;
// FP64 and FP32 ignore it.

template <int S>
struct Storage;

template <>
struct Storage<LBM_STORE_F64>
{
    using T = double;
    using C = double;
    static constexpr bool kPromote = false;
    static constexpr bool kFull = true;
    static __device__ __forceinline__ C load(const T *p, uint32_t i, int, float = 1.f) { return p[i]; }
    static __device__ __forceinline__ void store(T *p, uint32_t i, C v, int, float = 1.f) { p[i] = v; }
    static __device__ __forceinline__ void store_nt(T *p, uint32_t i, C v, int, float = 1.f)
    {
        __builtin_nontemporal_store(v, p + i);
    }
    static __device__ __forceinline__ T encode(double v, int, float = 1.f) { return v; }
    static __device__ __forceinline__ double decode(T v, int, float = 1.f) { return v; }
};

template <>
struct Storage<LBM_STORE_F32>
{
    using T = float;
    using C = float;
    static constexpr bool kPromote = true;
    static constexpr bool kFull = true;
    static __device__ __forceinline__ C load(const T *p, uint32_t i, int, float = 1.f) { return p[i]; }
    static __device__ __forceinline__ void store(T *p, uint32_t i, C v, int, float = 1.f) { p[i] = v; }
    static __device__ __forceinline__ void store_nt(T *p, uint32_t i, C v, int, float = 1.f)
    {
        __builtin_nontemporal_store(v, p + i);
    }
    static __device__ __forceinline__ T encode(double v, int, float = 1.f) { return float(v); }
    static __device__ __forceinline__ double decode(T v, int, float = 1.f) { return double(v); }
};

// 32-bit integer hash (xor-shift-multiply finaliser); random bits for
// stochastic rounding.
__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// binary32 -> binary16, stochastic: the 13 bits binary16 drops are added to
// the 13 random bits r (< 2^13) and cut, so |v| rounds up with probability
// equal to the dropped fraction.  Exact for the binary16 normal range; below
// it the conversion's own rounding applies to the remainder.  Inf/NaN pass
// through.
__device__ __forceinline__ _Float16 f32_to_f16_sr(float v, uint32_t r)
{
    uint32_t u = __float_as_uint(v);
    if ((u & 0x7f800000u) != 0x7f800000u)
        u = (u + r) & ~0x1fffu;
    return _Float16(__uint_as_float(u));
}

// The shift of direction d at a node whose rho_ref is rt (float).
__device__ __forceinline__ float shift_of(int d, float rt) { return weight<float>(d) * rt; }

// FP16 deviation storage: HBM holds f_d - w_d rho_ref(x) as IEEE binary16
// (RNE, or stochastic through store_sr).  decode gives the stored value
// exactly (binary16 and the float shift are exact in double); a kernel's
// load rounds that sum to float.
template <>
struct Storage<LBM_STORE_F16_SHIFTED>
{
    using T = _Float16;
    using C = float;
    static constexpr bool kPromote = false;
    static constexpr bool kFull = false;
    static __device__ __forceinline__ C load(const T *p, uint32_t i, int d, float rt)
    {
        return float(p[i]) + shift_of(d, rt);
    }
    static __device__ __forceinline__ void store(T *p, uint32_t i, C v, int d, float rt)
    {
        p[i] = _Float16(v - shift_of(d, rt));
    }
    static __device__ __forceinline__ void store_nt(T *p, uint32_t i, C v, int d, float rt)
    {
        __builtin_nontemporal_store(_Float16(v - shift_of(d, rt)), p + i);
    }
    static __device__ __forceinline__ void store_sr(T *p, uint32_t i, C v, int d, float rt, uint32_t r)
    {
        p[i] = f32_to_f16_sr(v - shift_of(d, rt), r);
    }
    static __device__ __forceinline__ void store_nt_sr(T *p, uint32_t i, C v, int d, float rt, uint32_t r)
    {
        __builtin_nontemporal_store(f32_to_f16_sr(v - shift_of(d, rt), r), p + i);
    }
    static __device__ __forceinline__ T encode(double v, int d, float rt)
    {
        return _Float16(float(v - double(shift_of(d, rt))));
    }
    static __device__ __forceinline__ double decode(T v, int d, float rt)
    {
        return double(float(v)) + double(shift_of(d, rt));
    }
};

// bfloat16 bit conversions with round-to-nearest-even: gfx950 converts in
// one instruction (v_cvt_pk_bf16_f32), which the compiler emits for a cast
// to __bf16.
__device__ __forceinline__ uint16_t f32_to_bf16(float f)
{
    return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// binary32 -> bfloat16, stochastic (16 dropped bits against the 16 random
// bits r < 2^16).
__device__ __forceinline__ uint16_t f32_to_bf16_sr(float f, uint32_t r)
{
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u)
        return f32_to_bf16(f);
    return uint16_t((u + r) >> 16);
}

// BF16 deviation storage: HBM holds f_d - w_d rho_ref(x) as bfloat16 (RNE,
// or stochastic through store_sr).
template <>
struct Storage<LBM_STORE_BF16_SHIFTED>
{
    using T = uint16_t;
    using C = float;
    static constexpr bool kPromote = false;
    static constexpr bool kFull = false;
    static __device__ __forceinline__ C load(const T *p, uint32_t i, int d, float rt)
    {
        return bf16_to_f32(p[i]) + shift_of(d, rt);
    }
    static __device__ __forceinline__ void store(T *p, uint32_t i, C v, int d, float rt)
    {
        p[i] = f32_to_bf16(v - shift_of(d, rt));
    }
    static __device__ __forceinline__ void store_nt(T *p, uint32_t i, C v, int d, float rt)
    {
        __builtin_nontemporal_store(f32_to_bf16(v - shift_of(d, rt)), p + i);
    }
    static __device__ __forceinline__ void store_sr(T *p, uint32_t i, C v, int d, float rt, uint32_t r)
    {
        p[i] = f32_to_bf16_sr(v - shift_of(d, rt), r);
    }
    static __device__ __forceinline__ void store_nt_sr(T *p, uint32_t i, C v, int d, float rt, uint32_t r)
    {
        __builtin_nontemporal_store(f32_to_bf16_sr(v - shift_of(d, rt), r), p + i);
    }
    static __device__ __forceinline__ T encode(double v, int d, float rt)
    {
        return f32_to_bf16(float(v - double(shift_of(d, rt))));
    }
    static __device__ __forceinline__ double decode(T v, int d, float rt)
    {
        return double(bf16_to_f32(v)) + double(shift_of(d, rt));
    }
};

// ---- physical placement of the lattice in HBM ----------------------------
// The logical node numbering is the reference's, n = x + y * Hx on the
// expanded lattice.  In HBM each row is padded to a pitch P that is a
// multiple of 64 nodes and starts X0 = 63 nodes into the row, so that the
// first interior node x = 1 of every row begins a 64-node (512 B in FP64)
// aligned segment; node (x, y) lives at p = X0 + x + y * P.  Layout formulas
// (access.hpp) are then applied to p with the per-direction stride S >= P*Vx,
// rounded up so that every direction array starts on a 2 MiB boundary.
struct Placement
{
    uint32_t P;   // row pitch (nodes)
    uint32_t X0;  // offset of x = 0 within a row (nodes)
    uint32_t S;   // stride between direction arrays (nodes)
};

constexpr uint32_t kRowOffset = 63;

__device__ __forceinline__ uint32_t place(uint32_t x, uint32_t y, uint32_t P, uint32_t X0) { return X0 + x + y * P; }

// ---- step parameters ------------------------------------------------------
struct StepParams
{
    const void *src;  // source lattice
    void *dst;        // destination lattice (== src for one-lattice schemes)
    const uint16_t *code;  // node codes, physical placement
    double *rho;      // results, (H-2)*(V-2) each
    double *jx;
    double *jy;
    double *jabs;
    const double *uy_in;  // outlet uy history, read (indexed by unexpanded y)
    double *uy_out;       // outlet uy history, written this step
    double omega;         // 1 / tau in the storage's real type (host_omega)
    double rho_out;
    uint32_t H, V;        // unexpanded extents incl. ghosts
    uint32_t Hx, Vx;      // expanded extents
    uint32_t N;           // Hx * Vx (logical node count)
    uint32_t P, X0, S;    // physical placement
    uint32_t sub_h, sub_v;
    uint32_t write_macros;
    uint32_t sr;          // stochastic-rounding seed of this step (0: round to nearest)
    uint32_t nan_protection;
    // npol/nsol tile sweep: element offsets, from the lattice base, of the
    // copies of tile-edge populations that the buffer rows/columns carry,
    // read this step and written for the next
    uint32_t side_rows_in, side_cols_in, side_rows_out, side_cols_out;
    uint32_t cnt_h, cnt_v;  // tiles per row / column
    // rho_ref per unexpanded column (16-bit forms): the source lattice was
    // stored with ref_in, this step stores with ref_out (the same table
    // except on a step that changes the reference)
    const float *ref_in;
    const float *ref_out;
};

// Is S one of the 16-bit forms (whose stored values depend on rho_ref)?
template <int S>
constexpr bool shifted_storage()
{
    return S == LBM_STORE_F16_SHIFTED || S == LBM_STORE_BF16_SHIFTED;
}

// rho_ref of unexpanded column xu for the 16-bit forms (1 otherwise, unused).
template <int S>
__device__ __forceinline__ float rt_of(const float *ref, uint32_t xu)
{
    if constexpr (shifted_storage<S>())
        return ref[xu];
    else
        return 1.f;
}

// rho_ref of columns xu - 1, xu, xu + 1 (scalars and a select, so that no
// private array reaches scratch memory).  The 16-bit forms collide in
// deviation form: a node works with g_d = f_d - w_d rho_ref(xu) of its own
// column, so a population pulled from column xu - cx is decoded against the
// difference rel(cx) = rho_ref(xu - cx) - rho_ref(xu) (exact: both are floats
// near 1), and its own slots against 0.
struct Rt3
{
    float l, c, r;
    __device__ __forceinline__ float own() const { return c; }
    __device__ __forceinline__ float src(int cx) const { return cx > 0 ? l : (cx < 0 ? r : c); }
    __device__ __forceinline__ float rel(int cx) const { return src(cx) - c; }
};

template <int S>
__device__ __forceinline__ Rt3 rt3_of(const float *ref, uint32_t xu)
{
    if constexpr (shifted_storage<S>())
        return Rt3{ref[xu ? xu - 1u : 0u], ref[xu], ref[xu + 1u]};  // xu = 0: the inlet ghost, which pulls nothing
    else
        return Rt3{1.f, 1.f, 1.f};
}

// NaN clamp of WITH_NAN_PROTECTION (nptl_kernels.hpp:413-432): NaN or values
// above FLT_MAX become 0.  Applied to outputs only.
template <class C>
__device__ __forceinline__ C clamp_output(C v, uint32_t on)
{
    if (on && (__builtin_isnan(v) || double(v) > double(__FLT_MAX__)))
        return C(0);
    return v;
}

// ---- the update of one fluid node ------------------------------------------
// Pull 9 populations.  Bit d of `code` set: the upstream node n - c_d is solid
// and half-way bounce-back supplies this node's own population 8-d.
// p is the node's physical index, P the row pitch and Sd the direction stride.
// r3: rho_ref around the node's unexpanded column xu (rt3_of): the population
// pulled for direction d is held by column xu - c_x(d) (buffer columns are
// transparent in unexpanded coordinates), or by the node itself when it
// bounces back.  The 16-bit forms return deviations from the node's own
// reference (Rt3::rel); FP64 and FP32 return the populations.
template <int L, int S>
__device__ __forceinline__ void pull(const typename Storage<S>::T *src, uint32_t p, uint32_t code,
                                     uint32_t P, uint32_t Sd, typename Storage<S>::C f[9], const Rt3 &r3)
{
#pragma unroll
    for (int d = 0; d < 9; ++d)
    {
        const bool bb = (code >> d) & 1u;
        const uint32_t m = bb ? p : uint32_t(int32_t(p) - cy_of(d) * int32_t(P) - cx_of(d));
        const int dd = bb ? 8 - d : d;
        f[d] = Storage<S>::load(src, at<L>(m, uint32_t(dd), Sd), dd, bb ? 0.f : r3.rel(cx_of(d)));
    }
}

// A population decoded against 0 by the lane that holds it, moved to a lane
// whose reference differs by rel (16-bit forms; the same operations as
// decoding it there).
template <int S>
__device__ __forceinline__ typename Storage<S>::C rebase(typename Storage<S>::C v, int d, float rel)
{
    if constexpr (shifted_storage<S>())
        return v + shift_of(d, rel);
    else
        return v;
}

// Zou-He pressure outlet at x = H-2 (non_buffered.hpp:95-129).  The three
// populations entering from the ghost column (0, 3, 6) are rebuilt from the
// six known post-stream populations, the outlet density and the previous
// step's uy of column H-3.  A: the type of the rebuild (double under kPromote).
template <class C, bool kPromote = false>
__device__ __forceinline__ void zou_he_outlet(C f[9], C rho_out, C uy_prev)
{
    using A = std::conditional_t<kPromote, double, C>;
    const C jx = -rho_out + (f[1] + f[4] + f[7] + C(2) * (f[8] + f[5] + f[2]));
    const C f8 = f[8], f5 = f[5], f2 = f[2], f1 = f[1], f7 = f[7];
    const A d17 = A(f1 - f7), ax = A(jx), auy = A(uy_prev);
    f[0] = C(A(f8) - A(0.5) * d17 - A(1.0 / 6) * ax - A(0.5) * auy);
    f[3] = C(A(f5) - A(2.0 / 3) * ax);
    f[6] = C(A(f2) + A(0.5) * d17 - A(1.0 / 6) * ax + A(0.5) * auy);
}

// Moments (rho = sum f, j = sum f c) and BGK relaxation towards the
// incompressible equilibrium feq = w (rho + 3 c.j + 4.5 (c.j)^2 - 1.5 |j|^2);
// f' = -(1/tau)(f - feq) + f.  Velocity is the momentum j (never divided by
// rho).  Term order as in nptl_kernels.hpp:382-405.  A direction and its
// opposite share 3 c.j and 4.5 (c.j)^2: c.j of the opposite is the exact
// negation of c.j for every value, infinities and NaN included, and
// (4.5 (-a)) (-a) == (4.5 a) a.  Without kFull the terms with a zero
// velocity component are left out as well (x + 0 * f == x for finite f;
// without fast-math the compiler may not drop them itself).
template <class C, bool kFull = true>
__device__ __forceinline__ void moments(const C f[9], C &rho, C &jx, C &jy, C &usq)
{
    rho = C(0);
    jx = C(0);
    jy = C(0);
#pragma unroll
    for (int d = 0; d < 9; ++d)
    {
        rho += f[d];
        if (kFull || cx_of(d) != 0)
            jx += f[d] * C(cx_of(d));
        if (kFull || cy_of(d) != 0)
            jy += f[d] * C(cy_of(d));
    }
    usq = jx * jx + jy * jy;
}

// c_d . j with the reference's term order (zero components left out unless kFull)
template <class C, bool kFull = true>
__device__ __forceinline__ C c_dot(int d, C jx, C jy)
{
    if (kFull || (cx_of(d) != 0 && cy_of(d) != 0))
        return C(cx_of(d)) * jx + C(cy_of(d)) * jy;
    else if (cx_of(d) != 0)
        return C(cx_of(d)) * jx;
    else
        return C(cy_of(d)) * jy;
}

// A: the type of the equilibrium's quadratic terms and of its sum (double
// under kPromote, rounded to C once per direction).
template <class C, bool kPromote = false, bool kFull = true>
__device__ __forceinline__ void collide(C f[9], C omega, C &rho, C &jx, C &jy, C &usq)
{
    using A = std::conditional_t<kPromote, double, C>;
    moments<C, kFull>(f, rho, jx, jy, usq);
    const A s = A(1.5) * A(usq);
    C feq[9];
#pragma unroll
    for (int d = 0; d < 4; ++d)
    {
        const C cu = c_dot<C, kFull>(d, jx, jy);
        const C t = C(3) * cu;
        const A q = A(4.5) * A(cu) * A(cu);
        feq[d] = C(A(weight<C>(d)) * (A(rho + t) + q - s));
        feq[8 - d] = C(A(weight<C>(8 - d)) * (A(rho - t) + q - s));
    }
    if constexpr (kFull)
    {
        const C cu = c_dot<C, true>(4, jx, jy);  // 0 jx + 0 jy
        const C t = C(3) * cu;
        const A q = A(4.5) * A(cu) * A(cu);
        feq[4] = C(A(weight<C>(4)) * (A(rho + t) + q - s));
    }
    else
        feq[4] = C(A(weight<C>(4)) * (A(rho) - s));
#pragma unroll
    for (int d = 0; d < 9; ++d)
        f[d] = -omega * (f[d] - feq[d]) + f[d];
}

// In the stream layout every store instruction of a wave writes whole
// contiguous segments, and the populations are written with non-temporal
// stores: nothing reads them again within the step, so they should not
// displace the pull's halo lines in L2.  In the collision and bundle layouts
// one instruction writes every 9th (3rd) element; those stores go through L2,
// where the nine (three) instructions of a node row combine into full lines.
// sr: 0 rounds to nearest; otherwise the per-step seed of stochastic rounding
// (FP16/BF16 only).  The random bits of place p: h = mix32(sr ^ p), then per
// direction one step of a 32-bit LCG (PCG's multiplier and increment), of
// which the top 13 (FP16) or 16 (BF16) bits are used.
// rt: rho_ref of the node (16-bit forms).
template <int L, int S>
__device__ __forceinline__ void store_node(typename Storage<S>::T *dst, uint32_t p, uint32_t Sd,
                                           const typename Storage<S>::C f[9], uint32_t sr = 0, float rt = 1.f)
{
    if constexpr (shifted_storage<S>())
    {
        if (sr)
        {
            constexpr int drop = S == LBM_STORE_F16_SHIFTED ? 19 : 16;
            uint32_t h = mix32(sr ^ p);
#pragma unroll
            for (int d = 0; d < 9; ++d)
            {
                h = h * 747796405u + 2891336453u;
                const uint32_t r = h >> drop;
                if constexpr (L == LBM_LAYOUT_STREAM)
                    Storage<S>::store_nt_sr(dst, at<L>(p, uint32_t(d), Sd), f[d], d, rt, r);
                else
                    Storage<S>::store_sr(dst, at<L>(p, uint32_t(d), Sd), f[d], d, rt, r);
            }
            return;
        }
    }
#pragma unroll
    for (int d = 0; d < 9; ++d)
    {
        if constexpr (L == LBM_LAYOUT_STREAM)
            Storage<S>::store_nt(dst, at<L>(p, uint32_t(d), Sd), f[d], d, rt);
        else
            Storage<S>::store(dst, at<L>(p, uint32_t(d), Sd), f[d], d, rt);
    }
}

// The bits store_node writes for a node's nine populations (same rounding,
// same random stream), for a caller that also writes them elsewhere.
template <int S>
__device__ __forceinline__ void node_bits(const typename Storage<S>::C f[9], uint32_t p, uint32_t sr,
                                          typename Storage<S>::T t[9], float rt = 1.f)
{
    if constexpr (shifted_storage<S>())
    {
        if (sr)
        {
            constexpr int drop = S == LBM_STORE_F16_SHIFTED ? 19 : 16;
            uint32_t h = mix32(sr ^ p);
#pragma unroll
            for (int d = 0; d < 9; ++d)
            {
                h = h * 747796405u + 2891336453u;
                const uint32_t r = h >> drop;
                if constexpr (S == LBM_STORE_F16_SHIFTED)
                    t[d] = f32_to_f16_sr(f[d] - shift_of(d, rt), r);
                else
                    t[d] = f32_to_bf16_sr(f[d] - shift_of(d, rt), r);
            }
            return;
        }
#pragma unroll
        for (int d = 0; d < 9; ++d)
        {
            if constexpr (S == LBM_STORE_F16_SHIFTED)
                t[d] = _Float16(f[d] - shift_of(d, rt));
            else
                t[d] = f32_to_bf16(f[d] - shift_of(d, rt));
        }
    }
    else
    {
#pragma unroll
        for (int d = 0; d < 9; ++d)
            t[d] = f[d];
    }
}

// Full-segment stores.  HBM is written in 64-byte segments; a segment that a
// wave writes only in part (a fluid/solid boundary inside it) costs a read-
// modify-write.  Solid nodes' populations are never read (half-way bounce-back
// takes the fluid node's own opposite population instead), so a solid lane
// writes zeros whenever another lane of its 64-byte segment writes, in the
// same store instructions as the fluid lanes (one store per direction after
// the branches reconverge, never one per branch).
// Ghost, buffer and padding nodes never write: they hold boundary data.
template <class T>
__device__ __forceinline__ bool segment_has(bool pred)
{
    constexpr int G = 64 / int(sizeof(T));  // lanes per 64-byte segment of one direction array
    const unsigned long long m = __ballot(pred);
    const int lane = __lane_id();
    const unsigned long long group = (G >= 64) ? ~0ull : ((1ull << G) - 1ull);
    return (m >> (lane & ~(G - 1))) & group;
}

// Finish one fluid node whose pulled populations are in f: outlet, collision,
// outlet-uy history and (on frame steps) macroscopic outputs at the inner
// index of unexpanded (x, y).  The caller stores f.
// The outlet's uy input (UY_SOURCE nodes, every step) and, on macro steps,
// the four results of inner node (x, y) (unexpanded coordinates).
// R: the density's type (the 16-bit forms pass rho_ref + sigma in double).
template <class C, class R = C>
__device__ __forceinline__ void write_moments(const StepParams &p, uint32_t code, uint32_t x, uint32_t y, R rho, C jx,
                                              C jy, C usq)
{
    if (code & LBM_CODE_UY_SOURCE)
        p.uy_out[y] = double(clamp_output(jy, p.nan_protection));
    if (p.write_macros)
    {
        const uint32_t r = (x - 1u) + (y - 1u) * (p.H - 2u);
        const C jabs = __builtin_sqrt(usq);
        p.rho[r] = double(clamp_output(rho, p.nan_protection));
        p.jx[r] = double(clamp_output(jx, p.nan_protection));
        p.jy[r] = double(clamp_output(jy, p.nan_protection));
        p.jabs[r] = double(clamp_output(jabs, p.nan_protection));
    }
}

// r0: the 16-bit forms' rho_ref of the node's column; f then holds the
// deviations g_d = f_d - w_d r0.  The equilibrium is linear in rho and
// sum_d w_d c_d = 0, so the same outlet and collision on g, with rho_out - r0
// for rho_out, give f' - w_d r0 (every operand a small deviation: FP32 then
// resolves the velocity signal instead of rounding it against |f| ~ 0.1 - 0.44),
// and the density is r0 + sum g.
template <int L, int S>
__device__ __forceinline__ void finish_node(const StepParams &p, uint32_t code, uint32_t x, uint32_t y,
                                            typename Storage<S>::C f[9], float r0 = 1.f)
{
    using C = typename Storage<S>::C;
    if constexpr (shifted_storage<S>())
    {
        if (code & LBM_CODE_OUTLET)
            zou_he_outlet<C>(f, C(p.rho_out) - r0, C(p.uy_in[y]));
        C sigma, jx, jy, usq;
        collide<C, false, false>(f, C(p.omega), sigma, jx, jy, usq);
        write_moments<C, double>(p, code, x, y, double(r0) + double(sigma), jx, jy, usq);
    }
    else
    {
        if (code & LBM_CODE_OUTLET)
            zou_he_outlet<C, Storage<S>::kPromote>(f, C(p.rho_out), C(p.uy_in[y]));
        C rho, jx, jy, usq;
        collide<C, Storage<S>::kPromote, Storage<S>::kFull>(f, C(p.omega), rho, jx, jy, usq);
        write_moments<C>(p, code, x, y, rho, jx, jy, usq);
    }
}

// ---- host side: the real-type parameters of a step -----------------------
// relaxation_time_inverse = 1 / relaxation_time in real_type
// (nptl_kernels.hpp:341); the reference reads tau from the settings as
// real_type (file_interaction.cpp:58-64), so the FP32 storage divides in float.
inline double host_omega(int storage, double tau)
{
    if (storage == LBM_STORE_F32)
        return double(1.0f / float(tau));
    return 1.0 / tau;
}

// The initial rho_ref table of the 16-bit forms (H unexpanded columns,
// ghosts included): rho_in at the inlet ghost, rho_out at the outlet ghost,
// 1 (the rest state) in between.
inline void host_reference(double rho_in, double rho_out, uint32_t H, float *out)
{
    for (uint32_t x = 0; x < H; ++x)
        out[x] = x == 0u ? float(rho_in) : (x + 1u == H ? float(rho_out) : 1.f);
}

// maxwell_boltzmann_distribution (core/simulation.hpp:333-350) in the
// storage's real type: double, or, for FP32 storage, float arguments with
// the reference's double literals and std::pow (double), rounded to float.
// The 16-bit forms take the FP64 values and round them when stored.
inline void host_equilibrium(int storage, double ux, double uy, double rho, double out[9])
{
    static const double w[9] = {1.0 / 36, 1.0 / 9, 1.0 / 36, 1.0 / 9, 4.0 / 9,
                                1.0 / 9,  1.0 / 36, 1.0 / 9, 1.0 / 36};
    for (int d = 0; d < 9; ++d)
    {
        const int cx = d % 3 - 1, cy = d / 3 - 1;
        if (storage == LBM_STORE_F32)
        {
            const float fx = float(ux), fy = float(uy), fr = float(rho), wf = float(w[d]);
            const float cu = cx * fx + cy * fy;
            out[d] = double(float(wf * (fr + 3 * cu + 9.0 / 2 * std::pow(cu, 2) - 3.0 / 2 * (fx * fx + fy * fy))));
        }
        else
            out[d] = w[d] * (rho + 3 * (cx * ux + cy * uy) + 9.0 / 2 * std::pow(cx * ux + cy * uy, 2) -
                             3.0 / 2 * (ux * ux + uy * uy));
    }
}

}  // namespace lbm_dev
