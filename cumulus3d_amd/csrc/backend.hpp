// backend.hpp — the device handle shared by the PatchStereo session (pm_session.cpp) and the
// learned initialiser (nn_net.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/c3d_pm.h"

struct c3d_pm_backend {
    int device = 0;
    std::string name;
    c3d_pm_options opts{};
    std::mutex lock;          // one PatchStereo session at a time (patchmatch.rs:96-97)
    hipStream_t stream = nullptr;
};

namespace c3d {
// Records `msg` as the calling thread's last error (c3d_pm_last_error) and returns `code`.
int set_error(int code, const std::string& msg);
}  // namespace c3d
