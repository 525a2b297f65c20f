// Host-side infrastructure for the TurboNeRF MI355X C++ layer: error
// handling, RAII device/pinned buffers and the single device stream.
// Errors surface as std::runtime_error, as the reference's CUDA_CHECK_THROW
// and trainer exceptions do (reference src/common.h:28-36).
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_version.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "turbonerf/tn_c_api.h"

namespace turbo {

inline const char* tn_status_name(int rc) {
    switch (rc) {
        case TN_OK: return "TN_OK";
        case TN_ERR_INVALID_ARGUMENT: return "invalid argument";
        case TN_ERR_LAUNCH: return "kernel launch failed";
        case TN_ERR_UNSUPPORTED: return "unsupported configuration";
        case TN_ERR_OUT_OF_MEMORY: return "workspace too small";
        case TN_ERR_NO_RAYS: return "No rays were generated for this training batch!";
        case TN_ERR_NO_SAMPLES: return "No samples were generated for this training batch!";
        default: return "unknown error";
    }
}

#define TN_HIP(expr)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            throw std::runtime_error(std::string(#expr) + " failed: " + hipGetErrorString(e_));            \
    } while (0)

#define TN_TRY(expr)                                                                                       \
    do {                                                                                                   \
        int rc_ = (expr);                                                                                  \
        if (rc_ != TN_OK) throw std::runtime_error(std::string(#expr) + ": " + ::turbo::tn_status_name(rc_)); \
    } while (0)

inline hipStream_t as_hip(tn_stream_t s) { return reinterpret_cast<hipStream_t>(s); }
inline tn_stream_t as_tn(hipStream_t s) { return reinterpret_cast<tn_stream_t>(s); }

// Owning device allocation of n elements of T (hipMalloc / hipFree).
template <typename T>
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) { resize(n); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), n_(o.n_) { o.ptr_ = nullptr; o.n_ = 0; }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            release();
            ptr_ = o.ptr_; n_ = o.n_;
            o.ptr_ = nullptr; o.n_ = 0;
        }
        return *this;
    }

    // (re)allocate to exactly n elements; contents are not preserved
    void resize(size_t n) {
        if (n == n_) return;
        release();
        if (n > 0) {
            void* p = nullptr;
            hipError_t e = hipMalloc(&p, n * sizeof(T));
            if (e != hipSuccess)
                throw std::runtime_error("hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed: " +
                                         hipGetErrorString(e));
            ptr_ = static_cast<T*>(p);
            n_ = n;
        }
    }
    // grow only
    void reserve(size_t n) { if (n > n_) resize(n); }
    void release() {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        n_ = 0;
    }
    void zero(hipStream_t s) { if (n_) TN_HIP(hipMemsetAsync(ptr_, 0, n_ * sizeof(T), s)); }
    void fill_bytes(int v, hipStream_t s) { if (n_) TN_HIP(hipMemsetAsync(ptr_, v, n_ * sizeof(T), s)); }

    T* data() const { return ptr_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }

private:
    T* ptr_ = nullptr;
    size_t n_ = 0;
};

// Pinned host allocation (hipHostMalloc) for asynchronous device-to-host reads.
template <typename T>
class PinnedBuffer {
public:
    explicit PinnedBuffer(size_t n = 1) {
        void* p = nullptr;
        TN_HIP(hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault));
        ptr_ = static_cast<T*>(p);
        n_ = n;
    }
    ~PinnedBuffer() { if (ptr_) (void)hipHostFree(ptr_); }
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    T* data() const { return ptr_; }
    T& operator[](size_t i) { return ptr_[i]; }

private:
    T* ptr_ = nullptr;
    size_t n_ = 0;
};

// HIP runtime version as major.minor.patch (HIP_VERSION's encoding,
// major * 10^7 + minor * 10^5 + patch).
struct RuntimeVersion {
    int major = 0, minor = 0, patch = 0;
    static RuntimeVersion from_hip(int v) { return {v / 10000000, (v / 100000) % 100, v % 100000}; }
    static RuntimeVersion compiled() { return from_hip(HIP_VERSION); }
    bool operator>=(const RuntimeVersion& o) const {
        if (major != o.major) return major > o.major;
        if (minor != o.minor) return minor > o.minor;
        return patch >= o.patch;
    }
    std::string to_string() const {
        return std::to_string(major) + "." + std::to_string(minor) + "." + std::to_string(patch);
    }
};

// Runtime and device checks (reference RuntimeManager,
// src/services/runtime-manager.cuh:94-202, which checks the CUDA driver
// version and each device's compute capability against the compiled
// architectures). The kernels here are built for one architecture, gfx950
// (MI355X). The HIP runtime keeps its interface within a major version (the
// library's soname), and a process may load another minor release than the
// one the build compiled against (PyTorch ships its own), so the runtime must
// be of the build's major version.
class RuntimeManager {
public:
    const std::vector<std::string> gpu_archs{"gfx950"};

    static RuntimeVersion compiled_runtime_version() { return RuntimeVersion::compiled(); }
    static RuntimeVersion required_runtime_version() { return {RuntimeVersion::compiled().major, 0, 0}; }
    static RuntimeVersion current_runtime_version() {
        int v = 0;
        TN_HIP(hipRuntimeGetVersion(&v));
        return RuntimeVersion::from_hip(v);
    }
    static bool is_driver_version_sufficient() {
        const RuntimeVersion c = current_runtime_version(), r = required_runtime_version();
        return c.major == r.major && c >= r;
    }
    // the device's architecture name without target features ("gfx950" of "gfx950:sramecc+:xnack-")
    static std::string get_device_architecture(int device_id) {
        hipDeviceProp_t p;
        TN_HIP(hipGetDeviceProperties(&p, device_id));
        std::string a(p.gcnArchName);
        return a.substr(0, a.find(':'));
    }
    static bool is_architecture_supported(const std::string& arch) { return arch == "gfx950"; }

    // Prints what it finds, as the reference does; false when the runtime is
    // older than the build's or a device is not gfx950.
    bool check_runtime() const {
        std::printf("Checking HIP runtime and device capabilities...\n");
        if (!is_driver_version_sufficient()) {
            std::printf("HIP runtime version is %s which is not sufficient.  This binary was compiled for HIP "
                        "runtime version %s.\n", current_runtime_version().to_string().c_str(),
                        compiled_runtime_version().to_string().c_str());
            return false;
        }
        int n = 0;
        TN_HIP(hipGetDeviceCount(&n));
        for (int d = 0; d < n; ++d) {
            const std::string a = get_device_architecture(d);
            if (!is_architecture_supported(a)) {
                std::printf("Device %d (%s) is not supported.\nSupported architecture(s) are: \ngfx950 (MI355X)\n",
                            d, a.c_str());
                return false;
            }
            std::printf("Device %d (%s) is supported!\n", d, a.c_str());
        }
        std::printf("HIP runtime is %s which is sufficient to run this binary which was built for %s. We're good "
                    "to go!\n", current_runtime_version().to_string().c_str(),
                    compiled_runtime_version().to_string().c_str());
        return true;
    }

    // Throws unless the current device is gfx950: the kernel library holds
    // no code for any other architecture, so a launch there would fail.
    static void require_current_device() {
        int d = 0;
        TN_HIP(hipGetDevice(&d));
        const std::string a = get_device_architecture(d);
        if (!is_architecture_supported(a))
            throw std::runtime_error("device " + std::to_string(d) + " is " + a +
                                     "; this build's kernels are compiled for gfx950 (MI355X) only");
    }
};

// One device and one stream (reference DeviceManager,
// src/services/device-manager.cuh:9-113, which also fixes the device count to 1).
// Creating the stream checks that the device is gfx950.
class DeviceManager {
public:
    static int get_device_count() { return 1; }
    static hipStream_t stream() {
        static hipStream_t s = [] {
            RuntimeManager::require_current_device();
            hipStream_t st;
            TN_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            return st;
        }();
        return s;
    }
    static void synchronize() { TN_HIP(hipStreamSynchronize(stream())); }
    static void teardown() { TN_HIP(hipDeviceSynchronize()); }
};

}  // namespace turbo
