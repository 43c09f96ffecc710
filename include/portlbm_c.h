/*
 * portlbm_c.h -- C ABI of the handler level (libportlbm.so), for callers
 * that bind through an FFI (ctypes, cffi, JNI, cgo).
 *
 * Wraps lbm::create_handler and execution::AlgorithmHandler
 * (reference: include/lbm/portlbm_factory.hpp:59,73,
 * include/lbm/execution/algorithm_handler.hpp:40-156).  Settings are passed
 * as a settings-file path or as the JSON text of one (schema of
 * src/lbm/file_interaction/file_interaction.cpp:11-69).  Exceptions become
 * status codes; portlbm_last_error() holds the message.
 *
 * The set-up queries (portlbm_validate_json, portlbm_domain_extents,
 * portlbm_build_phase, portlbm_roundtrip_json) need no GPU.
 */
#ifndef PORTLBM_C_H
#define PORTLBM_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PORTLBM_OK 0
#define PORTLBM_ERR_PROPERTY_ARGUMENT 1 /* exceptions::json::PropertyArgumentException */
#define PORTLBM_ERR_MISSING_PROPERTY 2  /* exceptions::json::MissingPropertyException */
#define PORTLBM_ERR_UNKNOWN_PROPERTY 3  /* exceptions::json::UnknownPropertyException */
#define PORTLBM_ERR_WAIT 4              /* exceptions::algorithm::WaitException */
#define PORTLBM_ERR_LBM 5               /* any other exceptions::Exception */
#define PORTLBM_ERR_OTHER 6             /* any other std::exception */

enum portlbm_field { PORTLBM_DENSITY = 0, PORTLBM_X_VELOCITY = 1, PORTLBM_Y_VELOCITY = 2, PORTLBM_ABS_VELOCITY = 3 };

typedef struct portlbm_handler portlbm_handler;

const char *portlbm_last_error(void);

/* create_handler(settings_path) / create_handler(Properties parsed from text) */
int portlbm_create_from_file(const char *settings_path, portlbm_handler **out);
int portlbm_create_from_json(const char *json_text, portlbm_handler **out);
void portlbm_destroy(portlbm_handler *h);

int portlbm_initialize(portlbm_handler *h);
int portlbm_start(portlbm_handler *h);
int portlbm_pause(portlbm_handler *h);
int portlbm_block_until_finished(portlbm_handler *h);
int portlbm_is_finished(portlbm_handler *h, int *finished);

/* get_densities / get_x_velocities / get_y_velocities / get_absolute_velocities;
 * count must be the inner node count (H-2)*(V-2). */
int portlbm_get_field(portlbm_handler *h, int field, double *out, size_t count);

/* out = {H, V, Hx, Vx, sub_h, sub_v, cnt_h, cnt_v, N}: unexpanded extents with
 * ghosts (get_horizontal/vertical_nodes), then Domain's fields. */
int portlbm_get_extents(portlbm_handler *h, uint32_t out[9]);
int portlbm_get_iteration(portlbm_handler *h, uint32_t *iteration);
int portlbm_get_progress(portlbm_handler *h, double *progress);
int portlbm_get_last_frametime(portlbm_handler *h, double *ms);
int portlbm_get_phase(portlbm_handler *h, int8_t *out, size_t count);
int portlbm_get_populations(portlbm_handler *h, double *out, size_t count);

/* Snapshot of the handler's state to a file / restore from one (no
 * counterpart in the reference; HipAlgorithmHandler::save_snapshot /
 * load_snapshot): loading needs the same settings apart from timeSteps and
 * frameUpdateInterval and continues the run bit for bit.  Not while running. */
int portlbm_save_snapshot(portlbm_handler *h, const char *path);
int portlbm_load_snapshot(portlbm_handler *h, const char *path);
/* The four result fields as CSV (x, y, density, x_velocity, y_velocity,
 * absolute_velocity per inner node) -- the GUI's "results.csv" option
 * (include/lbm/gui/lbm_gui.hpp:106-107), which the reference leaves unwritten. */
int portlbm_export_csv(portlbm_handler *h, const char *path);

/* Debug mode (debugMode in the settings): the fields of every finished
 * iteration, (H-2)(V-2) values per step in order (the reference's
 * all_densities / all_x_velocities / all_y_velocities, nptl.hpp:170-175).
 * field: PORTLBM_DENSITY, PORTLBM_X_VELOCITY or PORTLBM_Y_VELOCITY; *needed
 * receives the value count; out may be NULL to query it. */
int portlbm_debug_history(portlbm_handler *h, int field, double *out, size_t count, size_t *needed);

/* Properties::validate() of the settings in json_text. */
int portlbm_validate_json(const char *json_text);
/* Domain extents of the settings, same order as portlbm_get_extents. */
int portlbm_domain_extents(const char *json_text, uint32_t out[9]);
/* Node-type table of the settings' scenario (count = expanded node count). */
int portlbm_build_phase(const char *json_text, int8_t *out, size_t count);
/* json -> Properties -> json; writes at most capacity bytes incl. the NUL and
 * stores the full length in *length. */
int portlbm_roundtrip_json(const char *json_text, char *out, size_t capacity, size_t *length);
/* Properties::to_string() of the settings. */
int portlbm_properties_string(const char *json_text, char *out, size_t capacity, size_t *length);

#ifdef __cplusplus
}
#endif

#endif /* PORTLBM_C_H */
