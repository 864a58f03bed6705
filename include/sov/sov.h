/* NIRNAYA stable C ABI, exported by libsov.so.
 *
 * The first 19 entry points keep the names, signatures, return conventions
 * and status codes of the reference's include/sov/sov.h:9-32 (implemented in
 * src/driver/c_api.cpp:22-177), so code bound to that header (for example the
 * ctypes wrapper python/nirnaya.py:32-62) binds to this library unchanged.
 * Continuous LPs with option method=pdhg (or gpu) are solved by the MI355X
 * PDHG engine (include/nirnaya/pdhg.h).
 *
 * Return convention: -1 on failure with sov_error_message() set, 0 on success
 * (sov_add_col / sov_add_row return the new index). sov_optimize returns 0
 * whatever the solve status; read it with sov_get_status (values of the
 * reference's Status enum, src/core/status.hpp:7-24).
 */
#ifndef SOV_H
#define SOV_H

#ifdef __cplusplus
extern "C" {
#endif

#define SOV_API __attribute__((visibility("default")))

typedef struct sov_env sov_env;
typedef struct sov_model sov_model;

/* reference sov.h:9-10 */
SOV_API sov_env* sov_env_create(void);
SOV_API void sov_env_free(sov_env* env);

/* reference sov.h:12-13 */
SOV_API sov_model* sov_model_create(sov_env* env, const char* name);
SOV_API void sov_model_free(sov_model* model);

/* reference sov.h:15-19 */
SOV_API int sov_add_col(sov_model* model, double cost, double lo, double up, int type, const char* name);
SOV_API int sov_add_row(sov_model* model, double lo, double up, int nnz, const int* idx, const double* val,
                        const char* name);
SOV_API int sov_set_sense(sov_model* model, int sense); /* 1 = Minimize, -1 = Maximize */
SOV_API int sov_set_offset(sov_model* model, double offset);
SOV_API int sov_set_quad(sov_model* model, int nnz, const int* i, const int* j, const double* v);

/* reference sov.h:21-22 */
SOV_API int sov_read_mps(sov_model* model, const char* path);
SOV_API int sov_write_mps(sov_model* model, const char* path);
/* Binary LP container (.nlpb): the model's compressed-column arrays with a
 * checksum, loaded without text parsing. No counterpart in the reference; it
 * sits beside sov_read_mps / sov_write_mps and carries the same model.
 * sov_read_lpb rejects truncated, corrupt or non-canonical files. */
SOV_API int sov_read_lpb(sov_model* model, const char* path);
SOV_API int sov_write_lpb(sov_model* model, const char* path);

/* reference sov.h:24-26 */
SOV_API int sov_set_option_string(sov_model* model, const char* name, const char* value);
SOV_API int sov_optimize(sov_model* model);
SOV_API int sov_interrupt(sov_model* model);

/* reference sov.h:28-32 */
SOV_API int sov_get_status(sov_model* model);
SOV_API double sov_get_obj(sov_model* model);
SOV_API double sov_get_bound(sov_model* model);
SOV_API int sov_get_x(sov_model* model, double* out_x);
SOV_API const char* sov_error_message(sov_model* model);

/* ---- additions (no counterpart in the reference header) ----------------- */
/* Row duals of the last solve (nrow doubles; sign convention of the
 * reference facade, include/sov/solver.hpp:270-274). */
SOV_API int sov_get_y(sov_model* model, double* out_y);
/* Iterations of the last solve (the reference's SolveResult::iterations). */
SOV_API int sov_get_iterations(sov_model* model);
/* Warm start for the PDHG engine: x (ncol) and/or y (nrow); NULL clears. */
SOV_API int sov_set_start(sov_model* model, const double* x, const double* y);
/* Number of columns / rows of the current model. */
SOV_API int sov_get_num_cols(sov_model* model);
SOV_API int sov_get_num_rows(sov_model* model);

#ifdef __cplusplus
}
#endif
#endif
