#pragma once
// MI355X (gfx950) execution backend: the HIP counterpart of the reference's
// make_cuda_backend() (include/vikalp/backend/CudaBackend.hpp:18-20).
// Available when VIKALP_HIP_ENABLED is defined (this library always is).
//
// All operations are FP64 and deterministic (fixed-order reductions, A' from
// a stored transpose, no atomics).  Device memory is RAII.  One backend per
// thread.  Scalar-returning ops (dot, norm_inf) synchronise; the rest are
// asynchronous on the backend's stream until synchronize()/download().

#include "vikalp/contracts/ExecutionBackend.hpp"
#include "vikalp_hip.h"

#include <memory>

#define VIKALP_HIP_ENABLED 1

namespace vikalp {

[[nodiscard]] std::unique_ptr<ExecutionBackend> make_hip_backend();

// Device-side access for the fused PDHG path: the matrix handle and the raw
// device pointer behind a HipBackend matrix/vector (nullptr when the object
// does not belong to a HipBackend).
[[nodiscard]] const vk_hip_csr *hip_matrix_handle(const BackendMatrix &matrix) noexcept;
// The same handle, writable: the fused path scales the matrix in place.
[[nodiscard]] vk_hip_csr *hip_matrix_handle(BackendMatrix &matrix) noexcept;
[[nodiscard]] double *hip_vector_data(BackendVector &vector) noexcept;

}  // namespace vikalp
