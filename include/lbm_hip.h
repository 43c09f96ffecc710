/*
 * lbm_hip.h -- C ABI of the MI355X (gfx950) D2Q9 BGK time step.
 *
 * This is the kernel-level boundary of the library: plain C types, plain
 * pointers and sizes, status codes instead of exceptions.  It replaces the
 * reference's SYCL functors and the device-memory owners they read from:
 *
 *   reference (SC-SGS/PortLBM)                           replaced by
 *   ---------------------------------------------------  ------------------------------
 *   core::Simulation / Data / Results ctors              lbm_create
 *     src/lbm/core/simulation.cpp:386-433,482-519
 *   device max_work_group_size query                     lbm_device_max_work_group_size
 *     src/lbm/core/simulation.cpp:487-499
 *   phase fills, set_pipe_boundaries, add_stencil,       lbm_set_phase
 *   porous H2D copy
 *     include/lbm/core/domain_initialization.hpp:214-239,414-518
 *   set_standstill_values, set_inout_distribution_values lbm_init_populations
 *   and the lattice 0 -> 1 copy
 *     include/lbm/core/domain_initialization.hpp:311-403,460-467
 *   NPTL/LPTL/NPOL/NSOL<A>::execute() step loop          lbm_run
 *     include/lbm/gpu/nptl/nptl.hpp:116-141, lptl/lptl.hpp:112-137,
 *     npol/npol.hpp:180-201, nsol/nsol.hpp:184-205
 *     (EmplaceBounceBackKernel, OutletUpdateKernel, StreamCollideKernel and
 *      the buffer-copy kernels of every algorithm are fused into one launch
 *      per step; the one-lattice schemes with tiles wider than 64 nodes, and
 *      npol's FP64 collision layout, add the buffer refresh: two launches)
 *   SYCLAlgorithm::copy_macroscopic_observables_to_cpu   lbm_get_macros
 *     include/lbm/execution/sycl_algorithm.hpp:90-104
 *   the same copies enqueued after each frame step       lbm_run_frames /
 *     include/lbm/gpu/nptl/nptl.hpp:71-75                  lbm_get_frame
 *   Data::distribution_values_0 (device lattice)         lbm_get_populations /
 *     include/lbm/core/simulation.hpp:240-254              lbm_set_populations
 *   Data::phase_information                              lbm_get_phase
 *   get_neighbor + EmplaceBounceBackKernel addressing    lbm_get_node_codes
 *     include/lbm/core/access.hpp:142-152,
 *     include/lbm/gpu/nptl/nptl_kernels.hpp:471-491
 *   queue->wait()                                        lbm_sync
 *   exceptions crossing the API                          int status + lbm_last_error
 *
 * The functions below that have no reference counterpart (precision storage,
 * FP64 reductions, residual, batched domains) say so in their comments.
 *
 * Threading: one lbm_ctx is used by one host thread at a time.  Every call
 * runs on the context's own HIP stream (default priority).
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define LBM_OK 0
#define LBM_ERR_INVALID_ARGUMENT 1 /* reference: exceptions::json::PropertyArgumentException */
#define LBM_ERR_DEVICE 2           /* a HIP runtime call failed */
#define LBM_ERR_OUT_OF_MEMORY 3
#define LBM_ERR_STATE 4            /* call out of order (e.g. run before init) */

/* ---- enumerations ------------------------------------------------------ */
/* core/constants.hpp:43-48 */
enum lbm_algorithm { LBM_ALG_NPTL = 0, LBM_ALG_LPTL = 1, LBM_ALG_NPOL = 2, LBM_ALG_NSOL = 3 };
/* core/constants.hpp:50, core/access.hpp:41-110 */
enum lbm_layout { LBM_LAYOUT_COLLISION = 0, LBM_LAYOUT_STREAM = 1, LBM_LAYOUT_BUNDLE = 2 };
/* Population storage type.  The reference has only the compile-time
 * real_type switch (core/constants.hpp:17-21, double or float); F32 computes
 * as its float build does.  F16/BF16 store the deviation f_i - w_i rho_ref(x)
 * from the rest state at a reference density of the population's column
 * (the mean density of the column's fluid nodes, refreshed every 128 steps;
 * rho_in and rho_out on the ghost columns; 1 for batched domains); collision
 * arithmetic then runs in FP32. */
enum lbm_storage {
    LBM_STORE_F64 = 0,
    LBM_STORE_F32 = 1,
    LBM_STORE_F16_SHIFTED = 2,
    LBM_STORE_BF16_SHIFTED = 3
};

/* Rounding of the FP16/BF16 population stores in the step kernels (FP64 and
 * FP32 ignore it).  Stochastic: rounded up with probability equal to the
 * remainder, the random bits a hash of (node place, direction, iteration) --
 * deterministic, and unbiased where round-to-nearest repeats the same error
 * every step near a steady state.  No counterpart in the reference. */
enum lbm_rounding { LBM_ROUND_NEAREST = 0, LBM_ROUND_STOCHASTIC = 1 };

/* Step kernel of npol/nsol lattices whose tiles are at most 64 nodes wide.
 * Sweep: one wave walks each tile, the tile-edge copies in side arrays (no
 * refresh launches); block: one work-group per tile with the reference's
 * buffer refresh before each step; staged (tiles at least 8 nodes wide;
 * narrower ones run the automatic choice): one work-group per tile walks it
 * a chunk of rows at a time through LDS, with the sweep's side arrays.
 * Auto: in the FP64 and FP32 collision layout the staged
 * kernel on lattices of at least half a tile (FP64) or one tile (FP32) per
 * compute unit; else the sweep when the lattice has at least two tiles per
 * compute unit and the layout is not npol's FP64 collision layout; else the
 * block kernel (a lattice of few tiles leaves most of the GPU idle under the
 * sweep: 256 x 64 with 32 x 32 tiles, 16 tiles, measured 25.7 against 9.1
 * us per step).  All give the same fields.  No counterpart in the
 * reference. */
enum lbm_tile_kernel { LBM_TILE_AUTO = 0, LBM_TILE_SWEEP = 1, LBM_TILE_BLOCK = 2, LBM_TILE_STAGED = 3 };

/* Node-code bits returned by lbm_get_node_codes (one uint16 per expanded node).
 * Bit d (0..8) set: the population pulled along direction d comes from a
 * solid node, so half-way bounce-back supplies the node's own post-collision
 * population 8-d (nptl_kernels.hpp:471-491 emplaces exactly that value). */
#define LBM_CODE_BOUNCE_MASK 0x01FFu
#define LBM_CODE_FLUID 0x0200u      /* phase == 0: the node is updated */
#define LBM_CODE_OUTLET 0x0400u     /* Zou-He pressure outlet node (x = H-2) */
#define LBM_CODE_UY_SOURCE 0x0800u  /* x = H-3: its uy feeds the outlet next step */
#define LBM_CODE_SOLID 0x1000u      /* phase == 1 */

/* ---- lattice description ----------------------------------------------- */
/* Mirrors core::Properties (core/properties.hpp:30-166) and core::Domain
 * (core/domain.hpp:29-85) after the reference's own extent arithmetic
 * (src/lbm/core/simulation.cpp:194-382).  horizontal/vertical_nodes include
 * the two ghost layers; the expanded_* / subdomain_* fields are Domain's. */
typedef struct lbm_desc {
    int32_t algorithm;      /* enum lbm_algorithm */
    int32_t layout;         /* enum lbm_layout */
    int32_t storage;        /* enum lbm_storage */
    int32_t nan_protection; /* 1: NaN or > FLT_MAX outputs are written as 0 (WITH_NAN_PROTECTION) */
    uint32_t horizontal_nodes;
    uint32_t vertical_nodes;
    uint32_t expanded_horizontal_nodes;
    uint32_t expanded_vertical_nodes;
    uint32_t subdomain_horizontal_nodes;
    uint32_t subdomain_vertical_nodes;
    uint32_t subdomain_count_horizontal;
    uint32_t subdomain_count_vertical;
    uint32_t work_group_size;
    uint32_t rounding;      /* enum lbm_rounding: stores of the FP16/BF16 forms in the step kernels */
    uint32_t tile_kernel;   /* enum lbm_tile_kernel (npol/nsol) */
    double relaxation_time;
    double inlet_density;
    double inlet_velocity_x;
    double inlet_velocity_y;
    double outlet_density;
    double outlet_velocity_x;
    double outlet_velocity_y;
} lbm_desc;

typedef struct lbm_ctx lbm_ctx;

/* Thread-local text of the last error of the calling thread ("" if none). */
const char *lbm_last_error(void);

/* Largest work-group the device accepts (1024 on gfx950).
 * Replaces sycl::info::device::max_work_group_size (simulation.cpp:443,487). */
int lbm_device_max_work_group_size(uint32_t *out);
/* Name of the current device ("<marketing name> (<gfx arch>)", or the gfx arch
 * alone when the runtime reports no marketing name), NUL-terminated,
 * truncated to len - 1 characters.  Replaces sycl::info::device::name
 * (src/lbm/benchmark/benchmark.cpp:77). */
int lbm_device_name(char *buf, size_t len);

/* Allocate the lattice(s), phase/node-code tables and result arrays.
 * Two lattices for nptl/lptl, one for npol/nsol (simulation.cpp:464-479).
 * Results start at density -1 and zero momentum (simulation.cpp:386-417). */
int lbm_create(const lbm_desc *desc, lbm_ctx **out);
void lbm_destroy(lbm_ctx *ctx);

/* Upload the node-type (phase) table: one int8 per expanded node,
 * -1 ghost/buffer/padding, 0 fluid, 1 solid (domain_initialization.hpp:453-518). */
int lbm_set_phase(lbm_ctx *ctx, const int8_t *phase, size_t node_count);
int lbm_get_phase(lbm_ctx *ctx, int8_t *phase, size_t node_count);

/* Rest equilibrium everywhere, inlet/outlet ghost columns at their
 * equilibria, copy to the second lattice, derive the node codes from the
 * phase table, zero the outlet's uy history and the step counter.
 * Requires lbm_set_phase first. */
int lbm_init_populations(lbm_ctx *ctx);

/* Node codes (LBM_CODE_* bits), one uint16 per expanded node. */
int lbm_get_node_codes(lbm_ctx *ctx, uint16_t *codes, size_t node_count);

/* Advance `steps` time steps.  Macroscopic outputs (density, momentum,
 * |momentum|) are written on the device on frame steps -- the global
 * iteration it with it % frame_interval == frame_interval - 1
 * (nptl.hpp:71-75) -- and on the last step of this call, which is when the
 * reference copies them to the host.  The outlet's uy input is kept live on
 * every step in its own column.  frame_interval 0 means "last step only". */
int lbm_run(lbm_ctx *ctx, uint32_t steps, uint32_t frame_interval);

/* Passes of a split (debug) step, the reference's debug kernels
 * (nptl.hpp:165-514, nptl_kernels.hpp:48-295): pull streaming with half-way
 * bounce-back, the Zou-He outlet, the moments (results and outlet history
 * written every step), the collision.  Between passes the populations live in
 * an FP64 pass buffer.  A split step performs the fused step's arithmetic in
 * the same order and equals it bit for bit, for every algorithm, layout and
 * storage, on every fluid node and in the results (solid slots and the
 * buffer copies of them are never read; the fused kernels zero some by wave
 * segments). */
enum lbm_pass { LBM_PASS_STREAM = 0, LBM_PASS_OUTLET = 1, LBM_PASS_MACROSCOPIC = 2, LBM_PASS_COLLIDE = 3 };
/* Called after each pass with the iteration count (after the collide pass it
 * counts the finished step); the callback may read the pass buffer
 * (lbm_get_pass_populations), the results (lbm_get_macros) and, after the
 * collide pass, the new lattice (lbm_get_populations). */
typedef void (*lbm_pass_callback)(void *user, int pass, uint64_t iteration);
int lbm_run_split(lbm_ctx *ctx, uint32_t steps, lbm_pass_callback cb, void *user);
/* The pass buffer, direction-major over logical nodes: values[d * N + n] for
 * N the expanded node count (fluid nodes only are defined). */
int lbm_get_pass_populations(lbm_ctx *ctx, double *values, size_t count);

/* Steps completed since lbm_init_populations. */
int lbm_get_iteration(lbm_ctx *ctx, uint64_t *iteration);

/* Device time of the last lbm_run call in milliseconds, measured with HIP
 * events recorded on the context's stream around its launches. */
int lbm_last_run_ms(lbm_ctx *ctx, double *ms);

/* Copy the four result arrays, each (H-2)*(V-2) values in row-major inner
 * order r = (x-1) + (y-1)*(H-2) (core/access.hpp:208-212).  Any pointer may
 * be NULL to skip that array. */
int lbm_get_macros(lbm_ctx *ctx, double *density, double *x_velocity, double *y_velocity,
                   double *absolute_velocity, size_t count);

/* lbm_run with asynchronous frame copies, the reference's non-blocking
 * queue->copy of the results after each frame step (nptl.hpp:71-75,
 * sycl_algorithm.hpp:90-104): after every frame step the four result arrays
 * are copied to pinned host memory on a second stream while the steps go on;
 * the next step that writes results waits for that copy only.  The call
 * returns when the steps and the copies are done; lbm_get_frame then hands out
 * the latest frame.  Same populations and results as lbm_run. */
int lbm_run_frames(lbm_ctx *ctx, uint32_t steps, uint32_t frame_interval);

/* The result arrays of the latest frame copied by lbm_run_frames (layout as
 * lbm_get_macros; NULL pointers skip an array) and its iteration (the frame
 * step's global iteration, counted from 0).  LBM_ERR_STATE before the first
 * frame. */
int lbm_get_frame(lbm_ctx *ctx, double *density, double *x_velocity, double *y_velocity,
                  double *absolute_velocity, size_t count, uint64_t *iteration);

/* Current source lattice: 9 * expanded node count values in the context's
 * layout order (core/access.hpp:56-60, 81-85, 106-110), widened to double;
 * shifted storage is returned as f_i (w_i times the column's reference
 * density added back).  The conversion runs in chunks through a staging
 * buffer of at most 32 MiB that the context keeps after the first call
 * (freed by lbm_destroy).  Fluid and ghost nodes
 * are current.  Solid nodes' populations are never read (half-way
 * bounce-back takes the fluid node's own), and the fused kernels leave them
 * zero or as they were.  npol/nsol with tiles at most 64 nodes wide keep
 * the buffer copies of a step outside the lattice, so the lattice's buffer
 * rows/columns hold what they held when last refreshed (a split step
 * refreshes them). */
int lbm_get_populations(lbm_ctx *ctx, double *values, size_t count);
int lbm_set_populations(lbm_ctx *ctx, const double *values, size_t count);

/* The lattice state as the device holds it, for snapshots and restarts: a
 * 16-byte head (iteration u64, source-lattice index u32, lattice count u32),
 * every lattice in its storage type and physical placement, the outlet's uy
 * history (2 V doubles), the four result arrays (4 (H-2)(V-2) doubles) and,
 * in the 16-bit forms, the column reference table (max(Hx, H) + 64 floats).
 * buf is host memory; both calls return when the copy is complete.  Loading
 * it into a context created from the same description (and node types)
 * continues the run bit for bit.  No counterpart in the reference. */
int lbm_state_size(const lbm_ctx *ctx, size_t *bytes);
int lbm_get_state(lbm_ctx *ctx, void *buf, size_t bytes);
int lbm_set_state(lbm_ctx *ctx, const void *buf, size_t bytes);

/* Overwrite every fluid node with the equilibrium of a given (density,
 * momentum) field given on the inner grid ((H-2)*(V-2) values each, same
 * order as lbm_get_macros).  Used by warm starts.  No reference counterpart. */
int lbm_set_equilibrium_field(lbm_ctx *ctx, const double *density, const double *x_velocity,
                              const double *y_velocity, size_t count);

/* FP64 reductions over fluid nodes of the current source lattice, computed
 * on the device with wave64 shuffles: out[0] = sum rho, out[1] = sum jx,
 * out[2] = sum jy, out[3] = fluid node count, out[4], out[5] = momentum-
 * exchange force (x, y) on the solid links.  No reference counterpart. */
int lbm_reduce(lbm_ctx *ctx, double out[6]);

/* Steady-state residual ||j^now - j^snap||_2 / ||j^now||_2 over fluid
 * nodes (FP64 on the device), then j^snap := j^now.  The first call after
 * init compares against zero and returns 1.  No reference counterpart. */
int lbm_residual(lbm_ctx *ctx, double *residual);

/* Block until all work on the context's stream is done (queue->wait()). */
int lbm_sync(lbm_ctx *ctx);

/* Device-to-device copy bandwidth of `bytes` bytes, best of `repeats`
 * (GB/s, read+write counted).  The measured ceiling beside the nominal
 * 8 TB/s HBM figure.  No reference counterpart. */
int lbm_measure_copy_bandwidth(size_t bytes, int repeats, double *gbps);

/* ---- batched domains (no reference counterpart) ------------------------
 * Many independent channels of different sizes and physics in one arena,
 * advanced together.  Every domain is an unexpanded lattice (the lptl
 * geometry: H x V nodes incl. ghosts) in the stream layout with two lattices;
 * its update is the same pull / bounce-back / outlet / BGK step as above.
 * Tiles of 256 work-items cover tile_width x (256 / tile_width) nodes; the
 * tile width is chosen per class of inner width (<= 128, <= 256, <= 512,
 * > 512) from {64, 128, 256}. */
typedef struct lbm_batch lbm_batch;

typedef struct lbm_batch_domain {
    uint32_t horizontal_nodes; /* H incl. ghosts */
    uint32_t vertical_nodes;   /* V incl. ghosts */
    double relaxation_time;
    double inlet_density;
    double inlet_velocity_x;
    double inlet_velocity_y;
    double outlet_density;
    double outlet_velocity_x;
    double outlet_velocity_y;
} lbm_batch_domain;

enum lbm_batch_mode {
    LBM_BATCH_PER_DOMAIN = 0, /* one launch per domain and step */
    LBM_BATCH_PERSISTENT = 1, /* one launch per step: a fixed grid strides over the tiles of all domains */
    LBM_BATCH_GRAPH = 2       /* the persistent launches of the whole run captured in one hipGraph */
};

int lbm_batch_create(const lbm_batch_domain *domains, uint32_t count, int32_t storage, int32_t nan_protection,
                     lbm_batch **out);
void lbm_batch_destroy(lbm_batch *b);
/* phase table of one domain, H * V values, row-major (same values as lbm_set_phase) */
int lbm_batch_set_phase(lbm_batch *b, uint32_t domain, const int8_t *phase, size_t node_count);
/* rest state, ghost equilibria, node codes, zero outlet history and step counter */
int lbm_batch_init(lbm_batch *b);
/* advance every domain by `steps`; macroscopic outputs are written on the last step */
int lbm_batch_run(lbm_batch *b, uint32_t steps, int32_t mode);
int lbm_batch_last_run_ms(lbm_batch *b, double *ms);
int lbm_batch_get_macros(lbm_batch *b, uint32_t domain, double *density, double *x_velocity, double *y_velocity,
                         double *absolute_velocity, size_t count);
int lbm_batch_fluid_nodes(lbm_batch *b, uint64_t *count);
int lbm_batch_set_tiles(lbm_batch *b, const uint32_t tile_width[4]);
int lbm_batch_get_tiles(lbm_batch *b, uint32_t tile_width[4]);
/* Times the persistent step for every candidate tile width of every width
 * class (coordinate descent, `steps_per_trial` steps each), keeps the fastest
 * and re-initialises the domains (trials advance the state). */
int lbm_batch_autotune(lbm_batch *b, uint32_t steps_per_trial, double *best_ms_per_step);
/* Per-domain steady-state residual ||j^now - j^snap||_2 / ||j^now||_2 over
 * fluid nodes (FP64), then j^snap := j^now; out holds one value per domain.
 * The first call after init compares against zero and returns 1; a domain
 * whose populations hold NaN or Inf yields NaN. */
int lbm_batch_residuals(lbm_batch *b, double *out, uint32_t count);
/* lbm_batch_residuals without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device or host memory in
 * their loop (serving loops, harnesses that forbid it).  The per-tile
 * partial sums and per-domain residuals live in a static device array of
 * 8.5 MiB that each GPU holds until the process ends; batches of more than
 * 524 288 tiles or 65 536 domains are rejected (LBM_ERR_INVALID_ARGUMENT).
 * Same values, bit for bit, as lbm_batch_residuals.  Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU race. */
int lbm_batch_residuals_noalloc(lbm_batch *b, double *out, uint32_t count);
/* Domain `domain` back to the at-rest start of lbm_batch_init (populations,
 * inlet/outlet ghosts, outlet history); the other domains are untouched. */
int lbm_batch_reset_domain(lbm_batch *b, uint32_t domain);
/* Equilibrium of a (density, momentum) field on every fluid node of one
 * domain ((H-2)*(V-2) values each, inner row-major) -- the warm start. */
int lbm_batch_set_equilibrium(lbm_batch *b, uint32_t domain, const double *density, const double *x_velocity,
                              const double *y_velocity, size_t count);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HIP_H */
