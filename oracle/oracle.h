/*
 * oracle.h -- CPU restatement of SC-SGS/PortLBM's D2Q9 BGK time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker of the GPU library:
 * only tests, the smoke check and the CPU-baseline leg of the benchmark may
 * load it.  The product (portlbm_amd) never calls it.
 *
 * Every function restates one part of the reference pass by pass, with the
 * reference's index arithmetic and term order, and names the file:line it
 * follows.  The passes are kept separate as in the reference (bounce-back
 * emplacement, buffer refreshes, outlet update, stream-collide), so the
 * fused GPU kernels are checked against the unfused algorithm.
 */
#ifndef PORTLBM_ORACLE_H
#define PORTLBM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_NPTL = 0, OR_LPTL = 1, OR_NPOL = 2, OR_NSOL = 3 };
enum { OR_COLLISION = 0, OR_STREAM = 1, OR_BUNDLE = 2 };
enum {
    OR_SC_HAGEN_POISEUILLE = 0,
    OR_SC_WALLS = 1,
    OR_SC_CIRCLE = 2,
    OR_SC_SQUARE = 3,
    OR_SC_WING = 4,
    OR_SC_SKYSCRAPER = 5,
    OR_SC_POROUS = 6,
    OR_SC_PLATE = 7,
    OR_SC_POROUS_CIRCLES = 8, /* seeded non-overlapping circles (no reference counterpart) */
    OR_SC_CYLINDER = 9        /* one disc (no reference counterpart) */
};

/* Domain extents: out = {Hx, Vx, sub_h, sub_v, cnt_h, cnt_v, N}.
 * H, V include the ghost layers.  Returns 0, or -1 for nsol with wg < 6. */
int oracle_domain(int alg, unsigned H, unsigned V, unsigned wg, unsigned out[7]);

/* glibc random() (TYPE_3 additive feedback generator) with seed 1, the
 * state std::rand() has in a fresh process.  Writes `count` outputs. */
void oracle_glibc_rand_seed1(int *out, size_t count);

/* Obstacle parameters of the two extra scenarios. */
typedef struct oracle_scenario_params {
    uint64_t seed;
    double porosity;
    double radius_min, radius_max;
    unsigned margin;
    unsigned cylinder_x;
    double cylinder_radius;
    uint64_t max_attempts;
} oracle_scenario_params;

typedef struct oracle_sim oracle_sim;

oracle_sim *oracle_create(int alg, int layout, unsigned H, unsigned V, unsigned wg, double tau, double rho_in,
                          double ux_in, double uy_in, double rho_out, double ux_out, double uy_out,
                          int nan_protection);
void oracle_destroy(oracle_sim *s);

/* Phase table of a scenario (domain_initialization.hpp:453-518).  For the
 * porous scenario the glibc sequence with seed 1 is used from its start. */
int oracle_build_phase(oracle_sim *s, int scenario, const oracle_scenario_params *params);
int oracle_set_phase(oracle_sim *s, const int8_t *phase, size_t n);
const int8_t *oracle_phase(const oracle_sim *s);
size_t oracle_node_count(const oracle_sim *s);
void oracle_extents(const oracle_sim *s, unsigned out[7]);

/* Populations at rest, inlet/outlet ghosts, copy (domain_initialization.hpp:311-403,460-467). */
void oracle_init(oracle_sim *s);

/* `steps` reference time steps (execute() loop of the algorithm). */
void oracle_run(oracle_sim *s, unsigned steps);

/* Bytes of the build's real_type: 8 (liboracle.so) or 4 (liboracle_f32.so,
 * the reference's USE_FLOAT build, constants.hpp:17-21). */
size_t oracle_real_bytes(void);

/* Current source lattice (9N values of real_type, layout order), and the 4
 * result arrays widened to double. */
const void *oracle_populations(const oracle_sim *s);
void oracle_set_populations(oracle_sim *s, const void *values, size_t count);
void oracle_results(const oracle_sim *s, double *rho, double *jx, double *jy, double *jabs);
size_t oracle_result_count(const oracle_sim *s);

/* Node-code table of the GPU library restated from the reference's own
 * neighbour arithmetic: for each fluid node the set of pulled directions
 * whose upstream node is solid (bit d), fluid/outlet/uy-source/solid flags. */
void oracle_node_codes(const oracle_sim *s, uint16_t *codes);

/* FP64 sums over fluid nodes of the source lattice: rho, jx, jy, count,
 * momentum-exchange force x, y. */
void oracle_reduce(const oracle_sim *s, double out[6]);

/* OpenMP thread count used by the step loops (0 = runtime default). */
void oracle_set_threads(int threads);
int oracle_max_threads(void);

#ifdef __cplusplus
}
#endif

#endif
