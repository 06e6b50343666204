/* zerork_cfd_plugin.h - drop-in C API of the batched reactor-chemistry path.
 *
 * Same entry points, enums, argument meaning and status codes as the
 * reference plugin header (applications/cfd_plugin/zerork_cfd_plugin.h:6-95),
 * implemented by libzerork_cfd_plugin.so of this project on MI355X (HIP).
 *
 * Entry point                               replaces (reference file:line)
 *   zerork_reactor_init                     zerork_cfd_plugin.h:35, .cpp:33-45
 *   zerork_reactor_read_options_file        .h:37, .cpp:48-56
 *   zerork_reactor_set_mechanism_files      .h:39-40, .cpp:59-75
 *   zerork_reactor_load_mechanism           .h:42, .cpp:78-83
 *   zerork_reactor_solve                    .h:48-52, .cpp:159-175
 *   zerork_reactor_set_aux_field_pointer    .h:54, .cpp:178-185
 *   zerork_reactor_set/get_*_option         .h:56-80, .cpp:86-156
 *   zerork_reactor_set_callback_fn          .h:82-84, .cpp:188-196
 *   zerork_reactor_set_reactor_ids          .h:86, .cpp:198-205
 *   zerork_reactor_free                     .h:88, .cpp:207-216
 *
 * Units and layout as the reference: T [K], P [Pa], mass fractions
 * reactor-major mf[k*num_species_stride + j] (stride option defaults to the
 * species count), integration over [0, dt] for every reactor; T and mf are
 * updated in place.  Calls on one handle are not re-entrant.
 */
#ifndef ZERORK_CFD_PLUGIN_H
#define ZERORK_CFD_PLUGIN_H

#ifdef __cplusplus
extern "C" {
#endif

#define ZERORK_CFD_PLUGIN_EXPORTS __attribute__((visibility("default")))

typedef enum _zerork_field_t {
  ZERORK_FIELD_DPDT,
  ZERORK_FIELD_COST,
  ZERORK_FIELD_GPU,
  ZERORK_FIELD_Y_SRC,
  ZERORK_FIELD_E_SRC,
  ZERORK_FIELD_IGNITION_TIME,
  ZERORK_FIELD_TEMPERATURE_DELTA
} zerork_field_t;

typedef enum _zerork_status_t {
  ZERORK_STATUS_SUCCESS = 0,
  ZERORK_STATUS_INVALID_HANDLE,
  ZERORK_STATUS_INVALID_FILENAME,
  ZERORK_STATUS_INVALID_FIELD_NAME,
  ZERORK_STATUS_INVALID_POINTER,
  ZERORK_STATUS_FAILED_MECHANISM_PARSE,
  ZERORK_STATUS_FAILED_OPTIONS_PARSE,
  ZERORK_STATUS_FAILED_SOLVE,
  ZERORK_STATUS_OPTION_NOT_DEFINED,
  ZERORK_STATUS_UNKNOWN_ERROR = 99,
} zerork_status_t;

typedef struct zerork_handle_impl* zerork_handle;

zerork_handle ZERORK_CFD_PLUGIN_EXPORTS zerork_reactor_init();

zerork_status_t ZERORK_CFD_PLUGIN_EXPORTS zerork_reactor_read_options_file(
    const char* options_filename, zerork_handle handle);

zerork_status_t ZERORK_CFD_PLUGIN_EXPORTS zerork_reactor_set_mechanism_files(
    const char* mech_file, const char* therm_file, zerork_handle handle);

zerork_status_t ZERORK_CFD_PLUGIN_EXPORTS zerork_reactor_load_mechanism(zerork_handle handle);

typedef int (*zerork_callback_fn)(int reactor_id, int steps, double time, double dt,
                                  const double* y, const double* dy, void* cb_fn_data);

zerork_status_t ZERORK_CFD_PLUGIN_EXPORTS zerork_reactor_solve(
    const int n_cycle, const double time, const double dt, const int n_reactors, double* T,
    double* P, double* mf, zerork_handle handle);

zerork_status_t ZERORK_CFD_PLUGIN_EXPORTS zerork_reactor_set_aux_field_pointer(
    zerork_field_t ft, double* field_pointer, zerork_handle handle);

zerork_status_t ZERORK_CFD_PLUGIN_EXPORTS zerork_reactor_set_int_option(
    const char* option_name_chr, int option_value, zerork_handle handle);

zerork_status_t ZERORK_CFD_PLUGIN_EXPORTS zerork_reactor_set_double_option(
    const char* option_name_chr, double option_value, zerork_handle handle);

zerork_status_t ZERORK_CFD_PLUGIN_EXPORTS zerork_reactor_set_string_option(
    const char* option_name_chr, const char* option_value, zerork_handle handle);

zerork_status_t ZERORK_CFD_PLUGIN_EXPORTS zerork_reactor_get_int_option(
    const char* option_name_chr, int* option_value, zerork_handle handle);

zerork_status_t ZERORK_CFD_PLUGIN_EXPORTS zerork_reactor_get_double_option(
    const char* option_name_chr, double* option_value, zerork_handle handle);

zerork_status_t ZERORK_CFD_PLUGIN_EXPORTS zerork_reactor_get_string_option(
    const char* option_name_chr, char** option_value, int string_length, zerork_handle handle);

/* Called for a one-reactor solve with load_balance 0, once per successful
 * CVode call (solver_cvode.cpp:247-257), after the GPU solve over step
 * records the integrator kept; a non-zero return ends the integration at
 * that call. */
zerork_status_t ZERORK_CFD_PLUGIN_EXPORTS zerork_reactor_set_callback_fn(
    zerork_callback_fn fn, void* cb_fn_data, zerork_handle handle);

zerork_status_t ZERORK_CFD_PLUGIN_EXPORTS zerork_reactor_set_reactor_ids(int* reactor_ids,
                                                                         zerork_handle handle);

zerork_status_t ZERORK_CFD_PLUGIN_EXPORTS zerork_reactor_free(zerork_handle handle);

#ifdef __cplusplus
}
#endif

#endif
