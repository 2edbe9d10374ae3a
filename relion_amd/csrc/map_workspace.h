// Work buffers and 3-D FFT plans of the map-level stages (reference preparation and reconstruction), kept per
// device between calls so that a call does not allocate gigabytes and re-plan its transforms every time: buffers
// only grow, plans are cached by (size, type). One mutex per device serialises the calls that share them. The
// workspaces live until the process ends (they are never freed: the HIP runtime may be gone by the time static
// destructors would run).
#pragma once

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <mutex>
#include <vector>

#include "host_common.h"

namespace rln {

struct MapWorkspace {
    std::mutex mtx;
    DevBuf buf[12];
    struct Plan {
        int n;
        hipfftType type;
        hipfftHandle h;
    };
    std::vector<Plan> plans;

    // A cached n^3 plan of the given type, bound to stream st.
    int plan3d(int n, hipfftType type, hipStream_t st, hipfftHandle *out) {
        hipfftHandle h = 0;
        for (const Plan &p : plans)
            if (p.n == n && p.type == type) h = p.h;
        if (!h) {
            if (hipfftPlan3d(&h, n, n, n, type) != HIPFFT_SUCCESS)
                RLN_FAIL(RLN_ERR_HIP, "hipfftPlan3d of a %d^3 box (type %d) failed", n, (int)type);
            plans.push_back(Plan{n, type, h});
        }
        if (hipfftSetStream(h, st) != HIPFFT_SUCCESS) RLN_FAIL(RLN_ERR_HIP, "hipfftSetStream failed");
        *out = h;
        return RLN_OK;
    }
};

inline MapWorkspace &map_workspace(int device) {
    static std::mutex guard;
    static std::vector<MapWorkspace *> ws;
    std::lock_guard<std::mutex> lk(guard);
    if ((int)ws.size() <= device) ws.resize(device + 1, nullptr);
    if (!ws[device]) ws[device] = new MapWorkspace();
    return *ws[device];
}

}  // namespace rln
