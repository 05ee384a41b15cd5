// Device handle, streams, events, staging copies and error reporting.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "nlhip_internal.h"

namespace nlhip {

namespace {
thread_local std::string g_last_error;
std::atomic<int> g_injected_failures{0};
}  // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }

void log_error(const std::string& msg) {
    set_last_error(msg);
    std::fprintf(stderr, "[NL] %s\n", msg.c_str());
    std::fflush(stderr);
}

[[noreturn]] void throw_hip(hipError_t e, const char* expr, const char* file, int line) {
    (void)hipGetLastError();
    const int code = (e == hipErrorOutOfMemory) ? NLHIP_ERR_OOM : NLHIP_ERR_HIP;
    throw Error(code, std::string(hipGetErrorString(e)) + " in " + expr + " (" + file + ":" +
                          std::to_string(line) + ")");
}

bool take_injected_failure() {
    int v = g_injected_failures.load();
    while (v > 0) {
        if (g_injected_failures.compare_exchange_weak(v, v - 1)) return true;
    }
    return false;
}

static Device init_device() {
    Device d;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        (void)hipGetLastError();
        throw Error(NLHIP_ERR_NO_DEVICE, "no HIP device available");
    }
    int id = 0;
    NLHIP_CHECK(hipGetDevice(&id));
    hipDeviceProp_t prop;
    NLHIP_CHECK(hipGetDeviceProperties(&prop, id));
    d.id = id;
    d.cu_count = prop.multiProcessorCount;
    d.hbm_bytes = prop.totalGlobalMem;
    d.name = prop.name;
    d.arch = prop.gcnArchName;
    if (d.arch.rfind("gfx950", 0) != 0) {
        throw Error(NLHIP_ERR_NO_DEVICE,
                    "device " + d.arch + " is not gfx950; this library is built for gfx950 only");
    }
    return d;
}

const Device& device() {
    static std::mutex mu;
    static bool done = false;
    static Device dev;
    static int err_code = 0;
    static std::string err_msg;
    std::lock_guard<std::mutex> lock(mu);
    if (!done) {
        try {
            dev = init_device();
        } catch (const Error& e) {
            err_code = e.code;
            err_msg = e.what();
        }
        done = true;
    }
    if (err_code != 0) throw Error(err_code, err_msg);
    NLHIP_CHECK(hipSetDevice(dev.id));
    return dev;
}

bool device_usable() {
    try {
        device();
        return true;
    } catch (...) {
        return false;
    }
}

bool is_device_pointer(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

void copy_in(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return;
    NLHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s));
}

void copy_out(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return;
    NLHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s));
}

bool direct_launches() {
    static const bool on = [] {
        const char* v = std::getenv("NLHIP_DIRECT_LAUNCH");
        return v != nullptr && v[0] != '\0' && std::strcmp(v, "0") != 0;
    }();
    return on;
}

std::shared_mutex& capture_mutex() {
    static std::shared_mutex* m = new std::shared_mutex();  // process lifetime
    return *m;
}

Stream::Stream() {
    device();
    auto g = no_capture_guard();
    NLHIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
}

Stream::~Stream() {
    if (s_) {
        auto g = no_capture_guard();
        (void)hipStreamDestroy(s_);
    }
}

Event::Event() {
    auto g = no_capture_guard();
    NLHIP_CHECK(hipEventCreate(&e_));
}
Event::~Event() {
    if (e_) {
        auto g = no_capture_guard();
        (void)hipEventDestroy(e_);
    }
}

float elapsed_ms(const Event& a, const Event& b) {
    float ms = 0.f;
    NLHIP_CHECK(hipEventElapsedTime(&ms, a.get(), b.get()));
    return ms;
}

}  // namespace nlhip

using namespace nlhip;

extern "C" {

NLAPI const char* NLAPIENTRY nlhipLastError(void) { return g_last_error.c_str(); }

NLAPI int NLAPIENTRY nlhipDeviceInfo(char* name, size_t name_len, int* cu_count,
                                     uint64_t* hbm_bytes) {
    return guarded("nlhipDeviceInfo", [&] {
        const Device& d = device();
        if (name && name_len) {
            std::snprintf(name, name_len, "%s (%s)", d.name.c_str(), d.arch.c_str());
        }
        if (cu_count) *cu_count = d.cu_count;
        if (hbm_bytes) *hbm_bytes = d.hbm_bytes;
    });
}

// Test hook: the next `count` solves fail with NLHIP_ERR_INJECTED.
NLAPI void NLAPIENTRY nlhipInjectFailures(int count) { g_injected_failures.store(count); }

}  // extern "C"
