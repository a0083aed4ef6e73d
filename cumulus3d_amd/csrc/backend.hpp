// backend.hpp — the device handle shared by the PatchStereo session (pm_session.cpp) and the
// learned initialiser (nn_net.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/c3d_pm.h"

namespace c3d {
// Page-locked host memory (hipHostMalloc) that grows on demand and is freed with its owner.
struct PinnedBuf {
    unsigned char* p = nullptr;
    size_t n = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { release(); }
    hipError_t reserve(size_t want) {
        if (want <= n && p) return hipSuccess;
        release();
        hipError_t e = hipHostMalloc((void**)&p, want > 0 ? want : 1, hipHostMallocDefault);
        if (e == hipSuccess) n = want > 0 ? want : 1;
        else p = nullptr;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
};
}  // namespace c3d

struct c3d_pm_backend {
    int device = 0;
    std::string name;
    c3d_pm_options opts{};
    std::mutex lock;          // one PatchStereo session at a time (patchmatch.rs:96-97)
    hipStream_t stream = nullptr;
    // Lifetime: every open session and network holds the backend. c3d_pm_destroy frees it at once
    // when nothing holds it, otherwise it marks it and the last holder to close frees it.
    std::mutex life;
    int holders = 0;
    bool destroy_pending = false;
    // the two page-locked staging buffers of result downloads, used by the session that holds
    // `lock` and kept for the next one (page-locking hundreds of MB is not free)
    c3d::PinnedBuf dl_stage[2];
};

namespace c3d {
// Records `msg` as the calling thread's last error (c3d_pm_last_error) and returns `code`.
int set_error(int code, const std::string& msg);
// Takes a hold on `be` for a new session or network; fails when c3d_pm_destroy was already called.
int backend_hold(c3d_pm_backend* be);
// Drops a hold; frees the backend when it was the last one and c3d_pm_destroy was called.
void backend_release(c3d_pm_backend* be);
}  // namespace c3d
