// test_abi.cpp -- the reference's test-facing launchers with their C++
// signatures (src/gpu/pdhg.cu:77-241, forward-declared by
// tests/test_gpu_pdhg.cpp:124-142 and tests/test_pdhg_kernel.cpp:11-17),
// forwarding to the C-ABI.  Errors throw std::runtime_error like CUDA_CHECK.
#include <stdexcept>
#include <string>
#include <vector>

#include "pramaan_hip.h"

namespace pramaan {
namespace gpu {
namespace {
void check(int rc) {
    if (rc != PM_OK) throw std::runtime_error(std::string("HIP error: ") + pm_last_error());
}
}  // namespace

void launch_spmv(const std::vector<int>& row_ptr, const std::vector<int>& col_idx, const std::vector<float>& values,
                 const std::vector<float>& x, std::vector<float>& y_out, int m, int n) {
    if (m == 0) {
        y_out.clear();
        return;
    }
    y_out.assign(static_cast<size_t>(m), 0.0f);
    check(pm_launch_spmv(row_ptr.data(), col_idx.data(), values.data(), static_cast<int32_t>(values.size()),
                         x.data(), y_out.data(), m, n, PM_SPMV_AUTO));
}

void launch_spmv_t(const std::vector<int>& row_ptr, const std::vector<int>& col_idx,
                   const std::vector<float>& values, const std::vector<float>& x, std::vector<float>& y_out, int m,
                   int n) {
    y_out.assign(static_cast<size_t>(n), 0.0f);
    if (n == 0) return;
    check(pm_launch_spmv_t(row_ptr.data(), col_idx.data(), values.data(), static_cast<int32_t>(values.size()),
                           x.data(), y_out.data(), m, n, PM_SPMV_AUTO));
}

void launch_pdhg_iteration(float* h_x, float* h_y, const float* h_gradient, const float* h_Ax_bar,
                           const float* h_b, const int n, const int m, const float tau, const float sigma) {
    check(pm_launch_pdhg_iteration(h_x, h_y, h_gradient, h_Ax_bar, h_b, n, m, tau, sigma));
}

}  // namespace gpu
}  // namespace pramaan
