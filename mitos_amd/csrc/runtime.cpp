#include "../../include/mitos_hip.h"
#include "runtime.h"

#include <map>
#include <mutex>

namespace mitos {

static std::mutex g_mu;
static std::map<int, hipStream_t> g_streams;
static std::map<std::pair<int, int>, std::pair<void*, size_t>> g_ws;

hipStream_t api_stream() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_streams.find(dev);
    if (it != g_streams.end()) return it->second;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
    g_streams[dev] = s;
    return s;
}

int api_sync() {
    MITOS_HIP_CHECK(hipStreamSynchronize(api_stream()));
    MITOS_HIP_CHECK(hipGetLastError());
    return MITOS_OK;
}

void* workspace(int slot, size_t bytes) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    auto key = std::make_pair(dev, slot);
    auto it = g_ws.find(key);
    if (it != g_ws.end() && it->second.second >= bytes) return it->second.first;
    if (it != g_ws.end()) {
        hipDeviceSynchronize();
        hipFree(it->second.first);
        g_ws.erase(it);
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    g_ws[key] = std::make_pair(p, bytes);
    return p;
}

void release_workspaces() {
    std::lock_guard<std::mutex> lock(g_mu);
    hipDeviceSynchronize();
    for (auto& kv : g_ws) hipFree(kv.second.first);
    g_ws.clear();
}

}  // namespace mitos

extern "C" int mitos_release_workspace(void) {
    mitos::release_workspaces();
    return MITOS_OK;
}
