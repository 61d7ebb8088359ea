// SPDX-License-Identifier: MIT
//
// Host-side helpers of the C ABI implementation.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/mango_amd.h"
#include "american_kernels.hpp"

namespace mango_amd {

// Final state of a set of solves copied back to the host, rows of `stride`.
struct HostState {
    int64_t stride = 0;
    std::vector<double> x, u, u_prev, c_u, c_up, dt;
    std::vector<int64_t> n_space;
};

void set_err(MangoError* err, MangoStatus code, const char* msg);
bool device_ready(MangoError* err);

}  // namespace mango_amd
