/*
 * oracle.c -- CPU restatement of SC-SGS/PortLBM's D2Q9 BGK time step.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with -ffp-contract=off so
 * that every floating-point operation is the IEEE operation the reference
 * source text writes down, in the same order.
 *
 * `real` is the reference's real_type (constants.hpp:17-21): double by
 * default (liboracle.so), float when built with -DOR_REAL=float
 * (liboracle_f32.so, the reference's USE_FLOAT build).  Expressions keep the
 * reference's literals, so C's usual arithmetic conversions promote exactly
 * what C++'s promote there: in the float build the terms 9.0 / 2 * ... and
 * 3.0 / 2 * ... of the equilibrium and the outlet's rebuilt populations are
 * evaluated in double and rounded to float on assignment, the rest in float.
 *
 * Reference files restated here (paths relative to the reference root):
 *   include/lbm/core/constants.hpp:17-21,38-39  real_type, lattice weights
 *   include/lbm/core/access.hpp:41-110          layouts (Collision/Stream/BundleAccessor::at)
 *   include/lbm/core/access.hpp:142-229         neighbour / node / result indices
 *   include/lbm/core/access.hpp:244-302         (buffered) node access
 *   include/lbm/core/simulation.hpp:333-350     host equilibrium
 *   include/lbm/core/simulation.hpp:380-444     power functions
 *   src/lbm/core/simulation.cpp:223-382         domain extents
 *   include/lbm/core/domain_initialization.hpp  scenarios and initial state
 *   include/lbm/gpu/nptl/nptl.hpp, nptl_kernels.hpp
 *   include/lbm/gpu/lptl/lptl.hpp, lptl_kernels.hpp
 *   include/lbm/gpu/npol/npol.hpp, npol_kernels.hpp
 *   include/lbm/gpu/nsol/nsol.hpp, nsol_kernels.hpp
 *   include/lbm/gpu/general/non_buffered.hpp, buffered.hpp
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef OR_REAL
#define OR_REAL double
#endif
typedef OR_REAL real;

size_t oracle_real_bytes(void) { return sizeof(real); }

/* OR_PLANT selects one deliberate mistake in the collision, for the tests
 * that show the oracle's checks catch it (never defined in a shipped build):
 *   1  relaxation factor tau instead of 1 / tau
 *   2  coefficient 3.0 instead of 9.0 / 2 on (c.j)^2
 *   3  coefficient 1.0 instead of 3.0 / 2 on |j|^2
 *   4  term 3 * c.j dropped */
#ifndef OR_PLANT
#define OR_PLANT 0
#endif

/* ---- constants.hpp:38-39 (weights are real_type) ---- */
static const real W[9] = {1.0 / 36, 1.0 / 9, 1.0 / 36, 1.0 / 9, 4.0 / 9, 1.0 / 9, 1.0 / 36, 1.0 / 9, 1.0 / 36};

/* ---- access.hpp:41-110 ---- */
static inline unsigned at(int layout, unsigned node, unsigned d, unsigned total)
{
    if (layout == OR_COLLISION)
        return 9u * node + d;                              /* access.hpp:56-60 */
    if (layout == OR_STREAM)
        return total * d + node;                           /* access.hpp:81-85 */
    return 3u * (d / 3u) * total + (d % 3u) + 3u * node;    /* access.hpp:106-110 */
}

/* ---- access.hpp:142-152 ---- */
static inline unsigned get_neighbor(unsigned node, unsigned d, unsigned h)
{
    int y_offset = (int)(d / 3) - 1;
    int x_offset = (int)(d % 3) - 1;
    return node + y_offset * h + x_offset;
}

/* ---- access.hpp:208-212 and 225-229 ---- */
static inline unsigned result_index_xy(unsigned x, unsigned y, unsigned h) { return (x - 1) + (y - 1) * (h - 2); }
static inline unsigned result_index_linear(unsigned node, unsigned h)
{
    return (((node - h) / h) * (h - 2) + (node - 1) % h);
}

/* ---- access.hpp:260-268 / 292-301 ---- */
static inline unsigned node_index(int buffered, unsigned x, unsigned y, unsigned sub_v, unsigned sub_h, unsigned hx)
{
    if (!buffered)
        return x + y * hx;
    return (x + ((x - 1) / sub_h)) + (y + ((y - 1) / sub_v)) * hx;
}

/* ---- simulation.hpp:333-350 ---- */
static void maxwell_boltzmann(real ux, real uy, real rho, real *target)
{
    int cx = 0, cy = 0;
    for (int d = 0; d < 9; ++d)
    {
        cx = (d % 3) - 1;
        cy = (d / 3) - 1;
        target[d] = W[d] * (rho + 3 * (cx * ux + cy * uy) + 9.0 / 2 * pow(cx * ux + cy * uy, 2) -
                            3.0 / 2 * (ux * ux + uy * uy));
    }
}

/* ---- simulation.hpp:401-444 ---- */
static size_t which_power_of_4(size_t x)
{
    size_t checkbit = 3, p = 0;
    if (x & 0x1)
        return 0;
    while (!(checkbit & x))
    {
        x >>= 2;
        p++;
    }
    return !(x & 2) ? p : 0;
}

static size_t which_power_of_2(size_t x)
{
    size_t checkbit = 1, p = 0;
    if (x & 0x1)
        return 0;
    while (!(checkbit & x))
    {
        x >>= 1;
        p++;
    }
    return !(x >> 1 | 0) ? p : 0;
}

/* ---- simulation.cpp:223-261 ---- */
static void power_of_two_handling(size_t wg, unsigned *sub_v, unsigned *sub_h)
{
    size_t p4 = which_power_of_4(wg), p2 = which_power_of_2(wg);
    if (p4)
    {
        size_t size = 1;
        for (size_t i = 0; i < p4; ++i)
            size *= 2;
        *sub_h = (unsigned)size;
        *sub_v = (unsigned)size;
    }
    else if (p2)
    {
        size_t ph = p2 / 2, height = 1;
        for (size_t i = 0; i < ph; ++i)
            height *= 2;
        *sub_v = (unsigned)height;
        *sub_h = (unsigned)(2 * height);
    }
    else
    {
        *sub_v = 2;
        *sub_h = (unsigned)(wg / 2);
    }
}

static unsigned count_of(unsigned inner, unsigned sub)
{
    unsigned c = (inner / sub) + 1;
    if (!(inner % sub))
        c--;
    return c;
}

/* ---- simulation.cpp:194-221, 263-382 ---- */
int oracle_domain(int alg, unsigned H, unsigned V, unsigned wg, unsigned out[7])
{
    unsigned Hx = H, Vx = V, sub_h = H, sub_v = V, cnt_h = 1, cnt_v = 1;
    if (wg < 1)
        return -1;
    if (alg == OR_NPTL || alg == OR_NPOL)
    {
        if (wg & 0x1)
        {
            sub_h = wg;
            sub_v = 1;
        }
        else
            power_of_two_handling(wg, &sub_v, &sub_h);
        cnt_v = count_of(V - 2, sub_v);
        cnt_h = count_of(H - 2, sub_h);
        if (alg == OR_NPTL)
        {
            Vx = sub_v * cnt_v + 2; /* simulation.cpp:284 */
            Hx = sub_h * cnt_h + 2; /* :291 */
        }
        else
        {
            Vx = (sub_v + 1) * cnt_v + 1; /* :317 */
            Hx = (sub_h + 1) * cnt_h + 1; /* :324 */
        }
    }
    else if (alg == OR_NSOL)
    {
        size_t w = wg;
        if (w < 6)
            return -1; /* simulation.cpp:336-353 */
        if (w & 0x1)
            w--;
        power_of_two_handling(w, &sub_v, &sub_h);
        /* which_power_of_4 reports 1 for every w = 4m with m = 1 (mod 4)
         * (e.g. 20), giving 2 x 2 tiles; without the halo columns that is a
         * tile of width 0, on which the reference divides by zero */
        if (sub_v < 2 || sub_h < 3)
            return -1;
        sub_v -= 1;
        sub_h -= 2;
        cnt_v = count_of(V - 2, sub_v);
        Vx = (sub_v + 1) * cnt_v + 1;
        cnt_h = count_of(H - 2, sub_h);
        Hx = (sub_h + 1) * cnt_h + 1;
    }
    out[0] = Hx;
    out[1] = Vx;
    out[2] = sub_h;
    out[3] = sub_v;
    out[4] = cnt_h;
    out[5] = cnt_v;
    out[6] = Hx * Vx;
    return 0;
}

/* glibc random(): TYPE_3 additive feedback generator of degree 31 and
 * separation 3, seeded by the Lehmer recurrence 16807 * r mod (2^31 - 1),
 * first 310 outputs discarded; output = r >> 1.  (Published algorithm of
 * glibc's random_r.c; std::rand() returns random().) */
void oracle_glibc_rand_seed1(int *out, size_t count)
{
    int32_t r[34];
    r[0] = 1;
    for (int i = 1; i < 31; ++i)
    {
        const int32_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
        int32_t word = 16807 * lo - 2836 * hi;
        if (word < 0)
            word += 2147483647;
        r[i] = word;
    }
    uint32_t ring[31];
    for (int i = 0; i < 31; ++i)
        ring[i] = (uint32_t)r[i];
    /* ring[k % 31] holds r[k].  r[31..33] = r[0..2] leave the ring as it is;
     * r[k] = r[k-31] + r[k-3] (mod 2^32) for k >= 34; outputs start at k = 344 */
    size_t k = 34;
    size_t produced = 0, discard = 310;
    while (produced < count)
    {
        const uint32_t v = ring[(k - 31) % 31] + ring[(k - 3) % 31];
        ring[k % 31] = v;
        ++k;
        if (discard)
        {
            --discard;
            continue;
        }
        out[produced++] = (int)(v >> 1);
    }
}

/* ---- the simulation state ---- */
struct oracle_sim
{
    int alg, layout, buffered, nan_protection;
    unsigned H, V, Hx, Vx, N, sub_h, sub_v, cnt_h, cnt_v;
    real tau, rho_in, ux_in, uy_in, rho_out, ux_out, uy_out; /* Properties, real_type */
    int8_t *phase;
    real *lat0, *lat1; /* lat0 is the source lattice (data->distribution_values_0) */
    real *rho, *jx, *jy, *jabs;
    unsigned D;
    unsigned long iteration;
};

static int g_threads = 0;
void oracle_set_threads(int t) { g_threads = t; }
int oracle_max_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

oracle_sim *oracle_create(int alg, int layout, unsigned H, unsigned V, unsigned wg, double tau, double rho_in,
                          double ux_in, double uy_in, double rho_out, double ux_out, double uy_out,
                          int nan_protection)
{
    unsigned e[7];
    if (H < 4 || V < 4 || oracle_domain(alg, H, V, wg, e) != 0)
        return NULL;
    oracle_sim *s = (oracle_sim *)calloc(1, sizeof(oracle_sim));
    s->alg = alg;
    s->layout = layout;
    s->buffered = alg == OR_NPOL || alg == OR_NSOL;
    s->nan_protection = nan_protection;
    s->H = H;
    s->V = V;
    s->Hx = e[0];
    s->Vx = e[1];
    s->sub_h = e[2];
    s->sub_v = e[3];
    s->cnt_h = e[4];
    s->cnt_v = e[5];
    s->N = e[6];
    /* json_to_properties reads these as real_type (file_interaction.cpp:58-64) */
    s->tau = (real)tau;
    s->rho_in = (real)rho_in;
    s->ux_in = (real)ux_in;
    s->uy_in = (real)uy_in;
    s->rho_out = (real)rho_out;
    s->ux_out = (real)ux_out;
    s->uy_out = (real)uy_out;
    s->phase = (int8_t *)malloc(s->N);
    memset(s->phase, 0xff, s->N); /* simulation.cpp:424 */
    s->lat0 = (real *)calloc(9ull * s->N, sizeof(real));
    if (!s->buffered)
        s->lat1 = (real *)calloc(9ull * s->N, sizeof(real));
    s->D = (H - 2) * (V - 2);
    s->rho = (real *)malloc(sizeof(real) * (s->D ? s->D : 1));
    s->jx = (real *)calloc(s->D ? s->D : 1, sizeof(real));
    s->jy = (real *)calloc(s->D ? s->D : 1, sizeof(real));
    s->jabs = (real *)calloc(s->D ? s->D : 1, sizeof(real));
    for (unsigned i = 0; i < s->D; ++i)
        s->rho[i] = -1.0; /* simulation.cpp:407 */
    return s;
}

void oracle_destroy(oracle_sim *s)
{
    if (!s)
        return;
    free(s->phase);
    free(s->lat0);
    free(s->lat1);
    free(s->rho);
    free(s->jx);
    free(s->jy);
    free(s->jabs);
    free(s);
}

size_t oracle_node_count(const oracle_sim *s) { return s->N; }
size_t oracle_result_count(const oracle_sim *s) { return s->D; }
const int8_t *oracle_phase(const oracle_sim *s) { return s->phase; }
const void *oracle_populations(const oracle_sim *s) { return s->lat0; }
void oracle_extents(const oracle_sim *s, unsigned out[7])
{
    out[0] = s->Hx;
    out[1] = s->Vx;
    out[2] = s->sub_h;
    out[3] = s->sub_v;
    out[4] = s->cnt_h;
    out[5] = s->cnt_v;
    out[6] = s->N;
}

void oracle_set_populations(oracle_sim *s, const void *v, size_t count)
{
    if (count != 9ull * s->N)
        return;
    memcpy(s->lat0, v, count * sizeof(real));
    if (s->lat1)
        memcpy(s->lat1, v, count * sizeof(real));
}

int oracle_set_phase(oracle_sim *s, const int8_t *phase, size_t n)
{
    if (n != s->N)
        return -1;
    memcpy(s->phase, phase, n);
    return 0;
}

static void widen(double *out, const real *in, unsigned n)
{
    if (out)
        for (unsigned i = 0; i < n; ++i)
            out[i] = (double)in[i];
}

void oracle_results(const oracle_sim *s, double *rho, double *jx, double *jy, double *jabs)
{
    widen(rho, s->rho, s->D);
    widen(jx, s->jx, s->D);
    widen(jy, s->jy, s->D);
    widen(jabs, s->jabs, s->D);
}

/* ---- scenarios: domain_initialization.hpp:41-192 (unsigned semantics) ---- */
static unsigned stencil(int scenario, unsigned x, unsigned y, unsigned V, unsigned H)
{
    switch (scenario)
    {
    case OR_SC_HAGEN_POISEUILLE:
        return 0; /* :43-49 */
    case OR_SC_WALLS:
    { /* :59-72 */
        unsigned cond = (x == (1 * H) / 6 || x == (3 * H) / 6 || x == (5 * H) / 6) && (y > 1) && (y < V / 2);
        cond += (x == (2 * H) / 6 || x == (4 * H) / 6) && (y < V - 2) && (y > ((V - 1) / 2));
        return cond;
    }
    case OR_SC_CIRCLE:
    { /* :80-94 ; sycl::abs of an unsigned is the value itself */
        unsigned mx = 2 * H / 10, my = V / 2, r = V / 5;
        unsigned ax = mx - x, ay = my - y;
        return (y >= my - r) && (y <= my + r) && (x >= mx - r) && (x <= mx + r) && (ax * ax + ay * ay <= r * r);
    }
    case OR_SC_SQUARE:
    { /* :102-114 */
        unsigned mx = 2 * H / 10, my = V / 2, hl = V / 10;
        return (y >= my - hl) && (y <= my + hl) && (x >= mx - hl) && (x <= mx + hl);
    }
    case OR_SC_PLATE:
    { /* :122-135 */
        unsigned mx = H / 10, my = V / 2, hh = 7 * V / 40, hw = hh / 60;
        return (y >= my - hh) && (y <= my + hh) && (x >= mx - hw) && (x <= mx + hw);
    }
    case OR_SC_SKYSCRAPER:
    { /* :144-165 */
        unsigned ha = V / 2, hwa = 1, h2 = 2 * V / 5, hw2 = V / 24, h1 = V / 4, hw1 = V / 16, mx = 3 * H / 10;
        return ((y >= 1) && (y <= h1 && x >= mx - hw1 && x <= mx + hw1)) ||
               (y <= h2 && x >= mx - hw2 && x <= mx + hw2) || (y <= ha && x >= mx - hwa && x <= mx + hwa);
    }
    case OR_SC_WING:
    { /* :173-191: int locals, unsigned x/y -> unsigned differences */
        int mx = (int)(H / 5), my = (int)(V / 2), rmax = (int)(V / 5), sy = 5;
        unsigned dx = x - (unsigned)mx, dy = y - (unsigned)my;
        unsigned sqx = dx * dx, sqy = dy * dy;
        int c1 = (0.2 * (double)sqx + (double)sqy) <= rmax;
        int c2 = (pow((double)dx, 1.4) + (double)((unsigned)(sy * sy) * sqy)) <= sy * sy * rmax;
        return c1 || c2;
    }
    default:
        return 0;
    }
}

/* splitmix64: the generator of the two extra scenarios (no reference counterpart). */
static uint64_t sm_next(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double sm_u01(uint64_t *s) { return (double)(sm_next(s) >> 11) * (1.0 / 9007199254740992.0); }

typedef struct
{
    double x, y, r;
} circle_t;

/* Seeded random sequential placement of non-overlapping discs.  Candidate:
 * r ~ U[rmin, rmax), centre x ~ U{x_lo..x_hi}, y ~ U{2..V-3}; rejected when
 * the disc leaves [x_lo, x_hi] x [3, V-4] or comes closer than one node to
 * an accepted disc ((dx^2 + dy^2) < (r1 + r2 + 1)^2).  Accepted until the
 * solid node count reaches (1 - porosity) * (H-2) * (V-4) or max_attempts candidates
 * were tried.  A node (x, y) is solid iff (x-cx)^2 + (y-cy)^2 <= r^2.
 * This is a plain O(n^2) restatement; the product uses a cell grid. */
static size_t porous_circles(unsigned H, unsigned V, const oracle_scenario_params *p, circle_t **out_c)
{
    uint64_t st = p->seed;
    const double x_lo = 1.0 + p->margin, x_hi = (double)(H - 2) - p->margin;
    const double target = (1.0 - p->porosity) * (double)(H - 2) * (double)(V - 4);
    size_t cap = 1024, n = 0;
    circle_t *c = (circle_t *)malloc(cap * sizeof(circle_t));
    double solid = 0.0;
    for (uint64_t attempt = 0; attempt < p->max_attempts && solid < target && x_hi >= x_lo; ++attempt)
    {
        const double r = p->radius_min + (p->radius_max - p->radius_min) * sm_u01(&st);
        const double cx = x_lo + floor(sm_u01(&st) * (x_hi - x_lo + 1.0));
        const double cy = 2.0 + floor(sm_u01(&st) * (double)(V - 4));
        if (cx - r < x_lo || cx + r > x_hi || cy - r < 3.0 || cy + r > (double)V - 4.0)
            continue;
        int ok = 1;
        for (size_t i = 0; i < n && ok; ++i)
        {
            const double dx = cx - c[i].x, dy = cy - c[i].y, g = r + c[i].r + 1.0;
            if (dx * dx + dy * dy < g * g)
                ok = 0;
        }
        if (!ok)
            continue;
        if (n == cap)
        {
            cap *= 2;
            c = (circle_t *)realloc(c, cap * sizeof(circle_t));
        }
        c[n].x = cx;
        c[n].y = cy;
        c[n].r = r;
        ++n;
        const long r_i = (long)ceil(r);
        for (long yy = (long)cy - r_i; yy <= (long)cy + r_i; ++yy)
            for (long xx = (long)cx - r_i; xx <= (long)cx + r_i; ++xx)
            {
                const double dx = (double)xx - cx, dy = (double)yy - cy;
                if (dx * dx + dy * dy <= r * r)
                    solid += 1.0;
            }
    }
    *out_c = c;
    return n;
}

/* domain_initialization.hpp:453-518 (phase part) */
int oracle_build_phase(oracle_sim *s, int scenario, const oracle_scenario_params *params)
{
    const unsigned H = s->H, V = s->V;
    memset(s->phase, 0xff, s->N);
#define IDX(x, y) node_index(s->buffered, (x), (y), s->sub_v, s->sub_h, s->Hx)
    if (scenario == OR_SC_POROUS)
    {
        /* cpu_phase_masks::porous_medium_mask (:265-297), std::rand() from seed 1 */
        const size_t draws = 2ull * (V > 4 ? V - 4 : 0) * (H - 2);
        int *rnd = (int *)malloc(sizeof(int) * (draws ? draws : 1));
        oracle_glibc_rand_seed1(rnd, draws);
        size_t k = 0;
        for (int y = 2; (unsigned)y < V - 2; ++y)
            for (int x = 1; (unsigned)x <= H - 2; ++x)
            {
                int one = 1 + (rnd[k++] % 10);
                int two = 1 + (rnd[k++] % 10);
                int value = (x % one == one - 1 && y % two == two - 1) ? 1 : 0;
                s->phase[IDX((unsigned)x, (unsigned)y)] = (int8_t)value;
            }
        free(rnd);
        for (int y = 1; (unsigned)y < V; y += (int)(V - 3))
            for (int x = 1; (unsigned)x <= H - 2; ++x)
            {
                /* for V = 4 the row loop reaches y = V-1; its buffered index
                 * can lie past the end of the table, which is skipped */
                const unsigned idx = IDX((unsigned)x, (unsigned)y);
                if (idx < s->N)
                    s->phase[idx] = 1;
            }
        return 0;
    }
    /* set_pipe_boundaries (:214-239): rows id0 * V + 1 - 3 * id0 = 1 and V-2 */
    for (unsigned id0 = 0; id0 < 2; ++id0)
        for (unsigned id1 = 0; id1 < H - 2; ++id1)
            s->phase[IDX(id1 + 1, id0 * V + 1 - 3 * id0)] = 1;
    if (scenario == OR_SC_POROUS_CIRCLES || scenario == OR_SC_CYLINDER)
    {
        for (unsigned y = 2; y + 2 < V; ++y)
            for (unsigned x = 1; x + 1 < H; ++x)
                s->phase[IDX(x, y)] = 0;
        if (!params)
            return -1;
        circle_t *c = NULL;
        size_t n = 0;
        if (scenario == OR_SC_CYLINDER)
        {
            c = (circle_t *)malloc(sizeof(circle_t));
            c[0].x = params->cylinder_x;
            c[0].y = (double)((V - 1) / 2);
            c[0].r = params->cylinder_radius;
            n = 1;
        }
        else
            n = porous_circles(H, V, params, &c);
        for (size_t i = 0; i < n; ++i)
        {
            /* bounding box of the disc clipped to x in [1, H-2], y in [2, V-3] */
            const double rc = ceil(c[i].r);
            const double xa = fmax(1.0, c[i].x - rc), xb = fmin((double)(H - 2), c[i].x + rc);
            const double ya = fmax(2.0, c[i].y - rc), yb = fmin((double)(V - 3), c[i].y + rc);
            for (double yy = ya; yy <= yb; yy += 1.0)
                for (double xx = xa; xx <= xb; xx += 1.0)
                {
                    const double dx = xx - c[i].x, dy = yy - c[i].y;
                    if (dx * dx + dy * dy <= c[i].r * c[i].r)
                        s->phase[IDX((unsigned)xx, (unsigned)yy)] = 1;
                }
        }
        free(c);
        return 0;
    }
    /* add_stencil (:414-440): x in [1, H-2], y in [2, V-3] */
    for (unsigned id0 = 0; id0 < V - 4; ++id0)
        for (unsigned id1 = 0; id1 < H - 2; ++id1)
        {
            unsigned x = id1 + 1, y = id0 + 2;
            s->phase[IDX(x, y)] = (int8_t)stencil(scenario, x, y, V, H);
        }
#undef IDX
    return 0;
}

/* ---- initial populations: domain_initialization.hpp:311-403, 460-467 ---- */
void oracle_init(oracle_sim *s)
{
    real rest[9], values[18];
    maxwell_boltzmann(0, 0, 1, rest);
    for (unsigned n = 0; n < s->N; ++n)
        for (int d = 0; d < 9; ++d)
            s->lat0[at(s->layout, n, (unsigned)d, s->N)] = rest[d];
    maxwell_boltzmann(s->ux_in, s->uy_in, s->rho_in, values);
    maxwell_boltzmann(s->ux_out, s->uy_out, s->rho_out, values + 9);
    const unsigned neighboring_dir[2] = {3, 5};
    for (unsigned id0 = 0; id0 < 2; ++id0)
        for (unsigned id1 = 0; id1 < s->V - 2; ++id1)
        {
            unsigned node =
                node_index(s->buffered, 1 + id0 * (s->H - 3), 1 + id1, s->sub_v, s->sub_h, s->Hx);
            node = get_neighbor(node, neighboring_dir[id0], s->Hx);
            for (int i = 0; i < 9; ++i)
                s->lat0[at(s->layout, node, (unsigned)i, s->N)] = values[i + 9 * id0];
        }
    if (s->lat1)
        memcpy(s->lat1, s->lat0, 9ull * s->N * sizeof(real));
    s->iteration = 0;
    for (unsigned i = 0; i < s->D; ++i)
    {
        s->rho[i] = -1.0;
        s->jx[i] = s->jy[i] = s->jabs[i] = 0.0;
    }
}

/* ---- the collision of one node: nptl_kernels.hpp:381-437 ---- */
static inline real clamp_out(real v, int on)
{
    if (on && (isnan(v) || v > FLT_MAX))
        return 0;
    return v;
}

static inline void collide_and_store(oracle_sim *s, real *dst, unsigned n, unsigned r, const real *f_in)
{
    real f[9];
    memcpy(f, f_in, sizeof(f));
    real density = 0, fx = 0, fy = 0, absolute = 0, result = 0, result_buf = 0;
    int cx = 0, cy = 0;
#if OR_PLANT == 1
    const real omega = s->tau;
#else
    const real omega = 1 / s->tau; /* relaxation_time_inverse, nptl_kernels.hpp:341 */
#endif
    for (int d = 0; d < 9; ++d)
    {
        density += f[d];
        cx = d % 3 - 1;
        cy = d / 3 - 1;
        fx += f[d] * cx;
        fy += f[d] * cy;
    }
    absolute = fx * fx + fy * fy;
    for (int d = 0; d < 9; ++d)
    {
        cx = (d % 3) - 1;
        cy = (d / 3) - 1;
        result_buf = cx * fx + cy * fy;
#if OR_PLANT == 2
        result = W[d] * (density + 3 * result_buf + 3.0 * result_buf * result_buf - 3.0 / 2 * absolute);
#elif OR_PLANT == 3
        result = W[d] * (density + 3 * result_buf + 9.0 / 2 * result_buf * result_buf - 1.0 * absolute);
#elif OR_PLANT == 4
        result = W[d] * (density + 9.0 / 2 * result_buf * result_buf - 3.0 / 2 * absolute);
#else
        result = W[d] * (density + 3 * result_buf + 9.0 / 2 * result_buf * result_buf - 3.0 / 2 * absolute);
#endif
        result = -omega * (f[d] - result) + f[d];
        dst[at(s->layout, n, (unsigned)d, s->N)] = result;
    }
    absolute = (real)sqrt(absolute); /* sycl::sqrt in real_type: sqrt rounded once more is exact for float */
    s->rho[r] = clamp_out(density, s->nan_protection);
    s->jx[r] = clamp_out(fx, s->nan_protection);
    s->jy[r] = clamp_out(fy, s->nan_protection);
    s->jabs[r] = clamp_out(absolute, s->nan_protection);
}

/* ---- EmplaceBounceBackKernel (nptl_kernels.hpp:471-491, lptl_kernels.hpp:451-467,
 *      npol_kernels.hpp:232-253, nsol_kernels.hpp:81-102) ---- */
static inline void emplace_one(oracle_sim *s, real *lat, unsigned n, unsigned h)
{
    if (s->phase[n] != 1)
        return;
    for (unsigned d = 0; d < 9; ++d)
    {
        if (d == 4)
            continue;
        lat[at(s->layout, n, d, s->N)] = lat[at(s->layout, get_neighbor(n, d, h), 8 - d, s->N)];
    }
}

static void bounce_back(oracle_sim *s)
{
    real *lat = s->lat0;
    if (s->alg == OR_LPTL)
    {
        for (unsigned id = 0; id < s->H * s->V; ++id)
            emplace_one(s, lat, id, s->H);
    }
    else if (s->alg == OR_NPTL)
    {
        for (unsigned gy = 0; gy < s->Vx - 2; ++gy)
            for (unsigned gx = 0; gx < s->Hx - 2; ++gx)
                emplace_one(s, lat, (gx + 1) + (gy + 1) * s->Hx, s->Hx);
    }
    else
    {
        for (unsigned gy = 0; gy < s->cnt_v * s->sub_v; ++gy)
            for (unsigned gx = 0; gx < s->cnt_h * s->sub_h; ++gx)
                emplace_one(s, lat, node_index(1, gx + 1, gy + 1, s->sub_v, s->sub_h, s->Hx), s->Hx);
    }
}

/* ---- buffered.hpp:74-93 and 133-152 ---- */
static void refresh_buffers(oracle_sim *s)
{
    real *lat = s->lat0;
    const unsigned Hx = s->Hx, N = s->N;
    if (s->cnt_v > 1)
        for (unsigned i0 = 0; i0 < s->cnt_v - 1; ++i0)
            for (unsigned x = 0; x < Hx; ++x)
            {
                const unsigned y = (i0 + 1) * (s->sub_v + 1);
                const unsigned n = x + y * Hx;
                const unsigned top = get_neighbor(n, 7, Hx), bottom = get_neighbor(n, 1, Hx);
                for (unsigned d = 0; d < 3; ++d)
                    lat[at(s->layout, n, d, N)] = lat[at(s->layout, top, d, N)];
                for (unsigned d = 6; d < 9; ++d)
                    lat[at(s->layout, n, d, N)] = lat[at(s->layout, bottom, d, N)];
            }
    if (s->cnt_h > 1)
        for (unsigned y = 0; y < s->Vx; ++y)
            for (unsigned i1 = 0; i1 < s->cnt_h - 1; ++i1)
            {
                const unsigned x = (i1 + 1) * (s->sub_h + 1);
                const unsigned n = x + y * Hx;
                const unsigned left = get_neighbor(n, 3, Hx), right = get_neighbor(n, 5, Hx);
                for (unsigned d = 2; d < 9; d += 3)
                    lat[at(s->layout, n, d, N)] = lat[at(s->layout, left, d, N)];
                for (unsigned d = 0; d < 9; d += 3)
                    lat[at(s->layout, n, d, N)] = lat[at(s->layout, right, d, N)];
            }
}

/* ---- OutletUpdateKernel: non_buffered.hpp:75-130, buffered.hpp:268-326 ---- */
static void outlet(oracle_sim *s)
{
    real *src = s->lat0;
    const unsigned Hx = s->Hx, N = s->N, H = s->H;
    const unsigned missing[3] = {0, 3, 6};
    for (unsigned id = 0; id + 4 < s->V; ++id)
    {
        const int y = (int)id + 2;
        unsigned border;
        if (s->buffered)
        {
            border = node_index(1, H - 3, (unsigned)y, s->sub_v, s->sub_h, Hx);
            border = get_neighbor(border, 5, Hx);
        }
        else
            border = (H - 2) + (unsigned)y * Hx;
        real f_inverse[3];
        const real f_1 = src[at(s->layout, get_neighbor(border, 7, Hx), 1, N)];
        const real f_4 = src[at(s->layout, border, 4, N)];
        const real f_7 = src[at(s->layout, get_neighbor(border, 1, Hx), 7, N)];
        for (int i = 0; i < 3; ++i)
            f_inverse[i] = src[at(s->layout, get_neighbor(border, missing[i], Hx), 8 - missing[i], N)];
        const real x_velocity = -s->rho_out + (f_1 + f_4 + f_7 + 2 * (f_inverse[0] + f_inverse[1] + f_inverse[2]));
        const real y_velocity = s->jy[result_index_xy(H - 3, (unsigned)y, H)];
        src[at(s->layout, get_neighbor(border, 8 - missing[0], Hx), missing[0], N)] =
            f_inverse[0] - 0.5 * (f_1 - f_7) - 1.0 / 6 * x_velocity - 0.5 * y_velocity;
        src[at(s->layout, get_neighbor(border, 8 - missing[1], Hx), missing[1], N)] =
            f_inverse[1] - (2.0 / 3) * x_velocity;
        src[at(s->layout, get_neighbor(border, 8 - missing[2], Hx), missing[2], N)] =
            f_inverse[2] + 0.5 * (f_1 - f_7) - 1.0 / 6 * x_velocity + 0.5 * y_velocity;
    }
}

/* ---- StreamCollideKernel, two-lattice pull: nptl_kernels.hpp:350-439,
 *      lptl_kernels.hpp:331-414 ---- */
static inline void pull9(const oracle_sim *s, const real *src, unsigned n, real *f)
{
    for (int d = 0; d < 9; ++d)
        f[d] = src[at(s->layout, get_neighbor(n, 8 - (unsigned)d, s->Hx), (unsigned)d, s->N)];
}

static void stream_collide_two_lattice(oracle_sim *s)
{
    const real *src = s->lat0;
    real *dst = s->lat1;
    if (s->alg == OR_NPTL)
    {
#pragma omp parallel for schedule(static) num_threads(oracle_max_threads())
        for (unsigned gy = 0; gy < s->Vx - 2; ++gy)
            for (unsigned gx = 0; gx < s->Hx - 2; ++gx)
            {
                const unsigned x = gx + 1, y = gy + 1, n = x + y * s->Hx;
                if (!s->phase[n])
                {
                    real f[9];
                    pull9(s, src, n, f);
                    collide_and_store(s, dst, n, result_index_xy(x, y, s->H), f);
                }
            }
    }
    else
    {
#pragma omp parallel for schedule(static) num_threads(oracle_max_threads())
        for (unsigned id = 0; id < s->H * s->V; ++id)
            if (!s->phase[id])
            {
                real f[9];
                pull9(s, src, id, f);
                collide_and_store(s, dst, id, result_index_linear(id, s->H), f);
            }
    }
}

/* ---- npol StreamCollideKernel (npol_kernels.hpp:89-199): per work-group,
 *      every fluid item pulls into private memory, barrier, then writes
 *      in place. ---- */
static void stream_collide_npol(oracle_sim *s)
{
    const unsigned sh = s->sub_h, sv = s->sub_v;
#pragma omp parallel for schedule(dynamic) num_threads(oracle_max_threads())
    for (unsigned wg = 0; wg < s->cnt_h * s->cnt_v; ++wg)
    {
        const unsigned bx = wg % s->cnt_h, by = wg / s->cnt_h;
        real *priv = (real *)malloc(sizeof(real) * 9 * sh * sv);
        for (unsigned ly = 0; ly < sv; ++ly)
            for (unsigned lx = 0; lx < sh; ++lx)
            {
                const unsigned gx = bx * sh + lx, gy = by * sv + ly;
                const unsigned x = gx + 1 + gx / sh, y = gy + 1 + gy / sv, n = x + y * s->Hx;
                if (!s->phase[n])
                    pull9(s, s->lat0, n, priv + 9 * (ly * sh + lx));
            }
        /* nd_item.barrier() */
        for (unsigned ly = 0; ly < sv; ++ly)
            for (unsigned lx = 0; lx < sh; ++lx)
            {
                const unsigned gx = bx * sh + lx, gy = by * sv + ly;
                const unsigned x = gx + 1 + gx / sh, y = gy + 1 + gy / sv, n = x + y * s->Hx;
                if (!s->phase[n])
                {
                    const unsigned r = result_index_xy(x - x / (sh + 1), y - y / (sv + 1), s->H);
                    collide_and_store(s, s->lat0, n, r, priv + 9 * (ly * sh + lx));
                }
            }
        free(priv);
    }
}

/* ---- nsol StreamCollideKernel (nsol_kernels.hpp:171-334), emulated work-
 *      group by work-group: the three barrier-separated phases run for all
 *      items of the group before the next phase starts.  `local` holds two
 *      values per LDS node with pitch sub_h + 4 (:161-162). ---- */
static inline real read_guarded(const oracle_sim *s, unsigned node, unsigned d)
{
    /* the reference's reads past the last node (collision layout, top-right
     * item) feed only non-fluid items, whose values are never used */
    if (node >= s->N)
        return 0.0;
    return s->lat0[at(s->layout, node, d, s->N)];
}

static void stream_collide_nsol(oracle_sim *s)
{
    const unsigned sh = s->sub_h, sv = s->sub_v;
    const unsigned lw = sh + 2, lh = sv + 1, pitch = sh + 4;
    const unsigned items = lw * lh;
#pragma omp parallel for schedule(dynamic) num_threads(oracle_max_threads())
    for (unsigned wg = 0; wg < s->cnt_h * s->cnt_v; ++wg)
    {
        const unsigned bx = wg % s->cnt_h, by = wg / s->cnt_h;
        real *local = (real *)calloc(2ull * pitch * (sv + 2), sizeof(real));
        real *priv = (real *)calloc(9ull * items, sizeof(real));
        unsigned *node = (unsigned *)malloc(sizeof(unsigned) * items);
        unsigned *own = (unsigned *)malloc(sizeof(unsigned) * items);
        for (unsigned ly = 0; ly < lh; ++ly)
            for (unsigned lx = 0; lx < lw; ++lx)
            {
                const unsigned i = ly * lw + lx;
                const unsigned g1 = bx * lw + lx;
                const unsigned x = g1 - g1 / lw;
                const unsigned y = by * lh + ly;
                node[i] = x + y * s->Hx;
                own[i] = (lx + 1) + ly * pitch;
            }
        /* phase A: directions 5 and 6 (:201-228) */
        for (unsigned i = 0; i < items; ++i)
        {
            const unsigned n = node[i];
            real *p = priv + 9 * i;
            p[4] = read_guarded(s, n, 4);
            local[2 * get_neighbor(own[i], 5, pitch)] = read_guarded(s, n, 5);
            p[3] = read_guarded(s, get_neighbor(n, 5, s->Hx), 3);
            local[2 * get_neighbor(own[i], 6, pitch) + 1] = read_guarded(s, n, 6);
            p[2] = read_guarded(s, get_neighbor(n, 6, s->Hx), 2);
        }
        /* barrier; phase B (:233-234) */
        for (unsigned i = 0; i < items; ++i)
        {
            priv[9 * i + 5] = local[2 * own[i]];
            priv[9 * i + 6] = local[2 * own[i] + 1];
        }
        /* barrier; phase C: directions 7 and 8 (:240-262) */
        for (unsigned i = 0; i < items; ++i)
        {
            const unsigned n = node[i];
            real *p = priv + 9 * i;
            local[2 * get_neighbor(own[i], 7, pitch)] = read_guarded(s, n, 7);
            p[1] = read_guarded(s, get_neighbor(n, 7, s->Hx), 1);
            local[2 * get_neighbor(own[i], 8, pitch) + 1] = read_guarded(s, n, 8);
            p[0] = read_guarded(s, get_neighbor(n, 8, s->Hx), 0);
        }
        /* barrier; phase D (:267-333) */
        for (unsigned i = 0; i < items; ++i)
        {
            priv[9 * i + 7] = local[2 * own[i]];
            priv[9 * i + 8] = local[2 * own[i] + 1];
        }
        for (unsigned i = 0; i < items; ++i)
        {
            const unsigned n = node[i];
            if (n >= s->N || s->phase[n])
                continue;
            const unsigned x = n % s->Hx, y = n / s->Hx;
            const unsigned r = result_index_xy(x - x / (sh + 1), y - y / (sv + 1), s->H);
            collide_and_store(s, s->lat0, n, r, priv + 9 * i);
        }
        free(local);
        free(priv);
        free(node);
        free(own);
    }
}

/* ---- execute() loops: nptl.hpp:121-133, lptl.hpp:117-129,
 *      npol.hpp:185-193 (emplace_bounce_back :122-159), nsol.hpp:189-197 ---- */
void oracle_run(oracle_sim *s, unsigned steps)
{
    for (unsigned t = 0; t < steps; ++t)
    {
        if (s->buffered)
        {
            refresh_buffers(s);
            bounce_back(s);
            refresh_buffers(s);
            outlet(s);
            if (s->alg == OR_NPOL)
                stream_collide_npol(s);
            else
                stream_collide_nsol(s);
        }
        else
        {
            bounce_back(s);
            outlet(s);
            stream_collide_two_lattice(s);
            real *tmp = s->lat1;
            s->lat1 = s->lat0;
            s->lat0 = tmp;
        }
        s->iteration++;
    }
}

/* ---- node codes: the per-node bounce-back and boundary facts, derived
 *      from the reference's addressing ---- */
static int is_buffer_col(const oracle_sim *s, unsigned x)
{
    return s->buffered && x != 0 && x % (s->sub_h + 1) == 0 && x + 1 < s->Hx;
}
static int is_buffer_row(const oracle_sim *s, unsigned y)
{
    return s->buffered && y != 0 && y % (s->sub_v + 1) == 0 && y + 1 < s->Vx;
}

void oracle_node_codes(const oracle_sim *s, uint16_t *codes)
{
    const unsigned H = s->H, V = s->V, Hx = s->Hx;
    int outlet_active = 1;
    if (s->buffered)
    {
        /* buffered.hpp:285-288: the outlet's border is the right neighbour of
         * node H-3; it reaches node H-2 only without a buffer in between */
        const unsigned b = get_neighbor(node_index(1, H - 3, 2, s->sub_v, s->sub_h, Hx), 5, Hx);
        outlet_active = b == node_index(1, H - 2, 2, s->sub_v, s->sub_h, Hx);
    }
    for (unsigned n = 0; n < s->N; ++n)
    {
        uint16_t c = 0;
        if (s->phase[n] == 1)
            c = 0x1000;
        else if (s->phase[n] == 0)
        {
            c = 0x0200;
            const unsigned xe = n % Hx, ye = n / Hx;
            for (unsigned d = 0; d < 9; ++d)
            {
                if (d == 4)
                    continue;
                /* the pull reads node n + c_{8-d}; a buffer there carries the
                 * value of the node one further along the same axis */
                int dx = (int)((8 - d) % 3) - 1, dy = (int)((8 - d) / 3) - 1;
                unsigned mx = xe + dx, my = ye + dy;
                if (dx && is_buffer_col(s, mx))
                    mx += dx;
                if (dy && is_buffer_row(s, my))
                    my += dy;
                if (s->phase[mx + my * Hx] == 1)
                    c |= (uint16_t)(1u << d);
            }
            const unsigned x = s->buffered ? xe - xe / (s->sub_h + 1) : xe;
            const unsigned y = s->buffered ? ye - ye / (s->sub_v + 1) : ye;
            if (outlet_active && x == H - 2 && y >= 2 && y + 3 <= V)
                c |= 0x0400;
            if (x + 3 == H && y >= 2 && y + 3 <= V)
                c |= 0x0800;
        }
        codes[n] = c;
    }
}

void oracle_reduce(const oracle_sim *s, double out[6])
{
    for (int k = 0; k < 6; ++k)
        out[k] = 0;
    uint16_t *codes = (uint16_t *)malloc(sizeof(uint16_t) * s->N);
    oracle_node_codes(s, codes);
    for (unsigned n = 0; n < s->N; ++n)
    {
        if (!(codes[n] & 0x0200))
            continue;
        double f[9], rho = 0, jx = 0, jy = 0;
        for (int d = 0; d < 9; ++d)
        {
            f[d] = (double)s->lat0[at(s->layout, n, (unsigned)d, s->N)];
            rho += f[d];
            jx += f[d] * (d % 3 - 1);
            jy += f[d] * (d / 3 - 1);
        }
        out[0] += rho;
        out[1] += jx;
        out[2] += jy;
        out[3] += 1;
        for (int d = 0; d < 9; ++d)
            if (d != 4 && ((codes[n] >> (8 - d)) & 1u))
            {
                out[4] += 2.0 * f[d] * (d % 3 - 1);
                out[5] += 2.0 * f[d] * (d / 3 - 1);
            }
    }
    free(codes);
}
