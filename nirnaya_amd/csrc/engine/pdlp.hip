#include <stdexcept>

#include "pdlp.hpp"

namespace nir {
namespace gpu {

struct Pdlp {
    int unused = 0;
};

Pdlp* pdlp_create() { return new Pdlp; }
void PdlpDeleter::operator()(Pdlp* p) const { delete p; }

void pdlp_setup(Pdlp&, hipStream_t, const nir_lp*, const nir_params&, Csr<double>, Csr<double>, int, int) {
    throw std::runtime_error("pdlp mode is not available in this build");
}
void pdlp_run(Pdlp&, hipStream_t, const nir_params&, long long) {}
int pdlp_result(Pdlp&, hipStream_t, double*, double*, nir_info*) { return -1; }
int pdlp_time(Pdlp&, hipStream_t, long long, nir_timing*) { return -1; }

}  // namespace gpu
}  // namespace nir
