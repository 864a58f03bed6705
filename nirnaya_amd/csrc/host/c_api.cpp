// The sov_* C ABI (include/sov/sov.h). Handle ownership, error mapping and
// the NULL-handle answers follow src/driver/c_api.cpp:9-177: opaque handles
// from new/delete, a model keeps a non-owning env pointer and copies the
// env's default options at creation, exceptions become -1 with the message
// stored on the model, and option parse failures return -1 instead of
// escaping as C++ exceptions.
#include <cstring>

#include "../../../include/sov/sov.h"
#include "facade.hpp"

struct sov_env {
    nir::Options defaults;
};

struct sov_model {
    sov_env* env = nullptr;
    nir::ModelBuilder builder;
    nir::Model model;
    bool finalized = false;
    nir::Solver solver;
    nir::SolveResult result;
    std::string last_error;
    std::vector<double> start_x, start_y;

    sov_model(sov_env* e, const char* name)
        : env(e), builder(name ? name : "model"), solver(e ? e->defaults : nir::Options()) {}

    void ensure_finalized() {
        if (!finalized) {
            model = builder.finalize();
            finalized = true;
        }
    }
};

extern "C" {

SOV_API sov_env* sov_env_create(void) {
    try {
        return new sov_env();
    } catch (...) {
        return nullptr;
    }
}

SOV_API void sov_env_free(sov_env* env) { delete env; }

SOV_API sov_model* sov_model_create(sov_env* env, const char* name) {
    try {
        return new sov_model(env, name);
    } catch (...) {
        return nullptr;
    }
}

SOV_API void sov_model_free(sov_model* model) { delete model; }

SOV_API int sov_add_col(sov_model* model, double cost, double lo, double up, int type, const char* name) {
    if (!model) return -1;
    try {
        model->finalized = false;
        const nir::VarType t = type == 1 ? nir::VarType::Integer : type == 2 ? nir::VarType::Binary : nir::VarType::Continuous;
        return static_cast<int>(model->builder.add_col(cost, lo, up, t, name ? name : ""));
    } catch (const std::exception& e) {
        model->last_error = e.what();
        return -1;
    }
}

SOV_API int sov_add_row(sov_model* model, double lo, double up, int nnz, const int* idx, const double* val,
                        const char* name) {
    if (!model) return -1;
    try {
        if (nnz < 0 || (nnz > 0 && (!idx || !val))) {
            model->last_error = "sov_add_row: bad nonzero arrays";
            return -1;
        }
        model->finalized = false;
        return static_cast<int>(model->builder.add_row(lo, up, idx, val, nnz, name ? name : ""));
    } catch (const std::exception& e) {
        model->last_error = e.what();
        return -1;
    }
}

SOV_API int sov_set_sense(sov_model* model, int sense) {
    if (!model) return -1;
    model->builder.set_sense(sense < 0 ? nir::Sense::Maximize : nir::Sense::Minimize);
    return 0;
}

SOV_API int sov_set_offset(sov_model* model, double offset) {
    if (!model) return -1;
    model->builder.set_offset(offset);
    return 0;
}

SOV_API int sov_set_quad(sov_model* model, int nnz, const int* i, const int* j, const double* v) {
    if (!model) return -1;
    try {
        model->finalized = false;
        for (int k = 0; k < nnz; ++k) model->builder.add_quad(i[k], j[k], v[k]);
        return 0;
    } catch (const std::exception& e) {
        model->last_error = e.what();
        return -1;
    }
}

SOV_API int sov_read_mps(sov_model* model, const char* path) {
    if (!model || !path) return -1;
    std::string err;
    const nir::Status st = nir::read_mps(path, model->model, err);
    if (st != nir::Status::Optimal) {
        model->last_error = err;
        return -1;
    }
    model->finalized = true;
    return 0;
}

SOV_API int sov_write_mps(sov_model* model, const char* path) {
    if (!model || !path) return -1;
    try {
        model->ensure_finalized();
        return nir::write_mps(path, model->model) ? 0 : -1;
    } catch (const std::exception& e) {
        model->last_error = e.what();
        return -1;
    }
}

SOV_API int sov_read_lpb(sov_model* model, const char* path) {
    if (!model || !path) return -1;
    std::string err;
    const nir::Status st = nir::read_lpb(path, model->model, err);
    if (st != nir::Status::Optimal) {
        model->last_error = err;
        return -1;
    }
    model->finalized = true;
    return 0;
}

SOV_API int sov_write_lpb(sov_model* model, const char* path) {
    if (!model || !path) return -1;
    try {
        model->ensure_finalized();
        std::string err;
        if (!nir::write_lpb(path, model->model, err)) {
            model->last_error = err;
            return -1;
        }
        return 0;
    } catch (const std::exception& e) {
        model->last_error = e.what();
        return -1;
    }
}

SOV_API int sov_set_option_string(sov_model* model, const char* name, const char* value) {
    if (!model || !name || !value) return -1;
    return model->solver.options.set_by_name(name, value) ? 0 : -1;
}

SOV_API int sov_optimize(sov_model* model) {
    if (!model) return -1;
    try {
        model->ensure_finalized();
        model->solver.start_x = model->start_x;
        model->solver.start_y = model->start_y;
        model->result = model->solver.solve(model->model);
        model->last_error = model->result.message;
        return 0;
    } catch (const std::exception& e) {
        model->last_error = e.what();
        return -1;
    }
}

SOV_API int sov_interrupt(sov_model* model) {
    if (!model) return -1;
    model->solver.interrupt();
    return 0;
}

SOV_API int sov_get_status(sov_model* model) {
    if (!model) return static_cast<int>(nir::Status::InternalError);
    return static_cast<int>(model->result.status);
}

SOV_API double sov_get_obj(sov_model* model) { return model ? model->result.objective : 0.0; }

SOV_API double sov_get_bound(sov_model* model) { return model ? model->result.best_bound : 0.0; }

SOV_API int sov_get_x(sov_model* model, double* out_x) {
    if (!model || !out_x) return -1;
    std::copy(model->result.x.begin(), model->result.x.end(), out_x);
    return 0;
}

SOV_API const char* sov_error_message(sov_model* model) {
    if (!model) return "Null model handle";
    return model->last_error.c_str();
}

SOV_API int sov_get_y(sov_model* model, double* out_y) {
    if (!model || !out_y) return -1;
    std::copy(model->result.row_duals.begin(), model->result.row_duals.end(), out_y);
    return 0;
}

SOV_API int sov_get_iterations(sov_model* model) { return model ? model->result.iterations : -1; }

SOV_API int sov_set_start(sov_model* model, const double* x, const double* y) {
    if (!model) return -1;
    try {
        model->ensure_finalized();
        if (x) model->start_x.assign(x, x + model->model.ncol);
        else model->start_x.clear();
        if (y) model->start_y.assign(y, y + model->model.nrow);
        else model->start_y.clear();
        return 0;
    } catch (const std::exception& e) {
        model->last_error = e.what();
        return -1;
    }
}

SOV_API int sov_get_num_cols(sov_model* model) {
    if (!model) return -1;
    try {
        model->ensure_finalized();
        return model->model.ncol;
    } catch (const std::exception& e) {
        model->last_error = e.what();
        return -1;
    }
}

SOV_API int sov_get_num_rows(sov_model* model) {
    if (!model) return -1;
    try {
        model->ensure_finalized();
        return model->model.nrow;
    } catch (const std::exception& e) {
        model->last_error = e.what();
        return -1;
    }
}

}  // extern "C"
