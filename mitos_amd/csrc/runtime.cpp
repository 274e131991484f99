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

// pooled blocks: live ones (pointer -> device, size) and cached free ones
static std::map<void*, std::pair<int, size_t>> g_live;
static std::multimap<std::pair<int, size_t>, void*> g_free;

static void release_free_blocks_locked() {
    for (auto& kv : g_free) hipFree(kv.second);
    g_free.clear();
}

void* pool_alloc(size_t bytes) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    const auto key = std::make_pair(dev, bytes);
    auto it = g_free.find(key);
    if (it != g_free.end()) {
        void* p = it->second;
        g_free.erase(it);
        g_live[p] = key;
        return p;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        hipDeviceSynchronize();
        release_free_blocks_locked();
        p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    g_live[p] = key;
    return p;
}

void pool_free(void* p) {
    if (!p) return;
    hipDeviceSynchronize();
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_live.find(p);
    if (it == g_live.end()) {
        hipFree(p);  // not from the pool
        return;
    }
    g_free.emplace(it->second, p);
    g_live.erase(it);
}

void pool_trim() {
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_free.empty()) return;
    hipDeviceSynchronize();
    release_free_blocks_locked();
}

void release_workspaces() {
    std::lock_guard<std::mutex> lock(g_mu);
    hipDeviceSynchronize();
    for (auto& kv : g_ws) hipFree(kv.second.first);
    g_ws.clear();
    release_free_blocks_locked();
}

}  // namespace mitos

extern "C" int mitos_release_workspace(void) {
    mitos::release_workspaces();
    return MITOS_OK;
}
