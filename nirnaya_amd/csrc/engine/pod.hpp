// Device rows of a POD basis's last decode, for engines that take warm starts
// on the device.
#pragma once

struct nir_pod;

namespace nir {
namespace gpu {
// Rows [count][len] (row-major) of the last nir_pod_decode on `device`, or
// nullptr when there are none.
const double* pod_rows(const nir_pod* p, int* count, int* len, int* device);
}  // namespace gpu
}  // namespace nir
