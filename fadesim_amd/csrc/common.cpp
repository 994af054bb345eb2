// Error reporting and workspace management for the fadesim HIP library.
#include "common.h"

#include <mutex>
#include <vector>

#include "fadesim.h"

namespace fadesim {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

namespace {
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_ws_mutex;
std::vector<DevBuf> g_ws;     // indexed by device ordinal
std::vector<DevBuf> g_ws_aux; // indexed by device ordinal
std::vector<DevBuf> g_ws_warm; // indexed by device ordinal

void *grow(std::vector<DevBuf> &pool, size_t bytes);
} // namespace

void *workspace(size_t bytes) { return grow(g_ws, bytes); }
void *workspace_aux(size_t bytes) { return grow(g_ws_aux, bytes); }
void *workspace_warm(size_t bytes) { return grow(g_ws_warm, bytes); }

namespace {
void *grow(std::vector<DevBuf> &pool, size_t bytes)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error(FADESIM_E_HIP, "hipGetDevice failed");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_ws_mutex);
    if ((int)pool.size() <= dev)
        pool.resize(dev + 1);
    DevBuf &b = pool[dev];
    if (b.bytes >= bytes && b.ptr)
        return b.ptr;
    if (b.ptr) {
        // the previous buffer may still be in use by queued work
        if (hipDeviceSynchronize() != hipSuccess || hipFree(b.ptr) != hipSuccess) {
            set_error(FADESIM_E_HIP, "workspace release failed");
            return nullptr;
        }
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes < (size_t)(1 << 20) ? (size_t)(1 << 20) : bytes;
    if (hipMalloc(&b.ptr, want) != hipSuccess) {
        b.ptr = nullptr;
        set_error(FADESIM_E_NOMEM, "workspace hipMalloc of " + std::to_string(want) + " bytes failed");
        return nullptr;
    }
    b.bytes = want;
    return b.ptr;
}
} // namespace

} // namespace fadesim

extern "C" const char *fadesim_last_error(void) { return fadesim::g_last_error.c_str(); }

extern "C" int fadesim_version(void) { return 100; }
