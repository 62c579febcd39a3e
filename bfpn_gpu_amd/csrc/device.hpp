// device.hpp — device buffers and the split operand sets of the MFMA kernels (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "tables.hpp"

namespace bfpn {

class DeviceBuffer {
public:
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void release() {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        bytes_ = 0;
    }
    void ensure(size_t bytes) {
        if (bytes <= bytes_) return;
        release();
        if (bytes == 0) return;
        hipError_t e = hipMalloc(&ptr_, bytes);
        if (e != hipSuccess) {
            ptr_ = nullptr;
            (void)hipGetLastError();
            throw Error(BFPN_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                          hipGetErrorString(e));
        }
        bytes_ = bytes;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(ptr_);
    }
    size_t bytes() const { return bytes_; }
    void swap(DeviceBuffer& o) {
        std::swap(ptr_, o.ptr_);
        std::swap(bytes_, o.bytes_);
    }

private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

// Host -> device copy of a (usually temporary) host vector; complete on return.
template <typename T>
void upload(DeviceBuffer& b, const std::vector<T>& host) {
    b.ensure(host.size() * sizeof(T));
    if (!host.empty())
        BFPN_HIP(hipMemcpy(b.as<void>(), host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
}

template <typename T>
std::vector<T> cast_vector(const std::vector<double>& v) {
    return std::vector<T>(v.begin(), v.end());
}

std::vector<double> yz_weights(int Ny, int Nz, bool trapezoidal);

// Device storage of a SourcePlan (FP32 source on the matrix cores).
struct SourcePlanBuffers {
    DeviceBuffer kth, ktl, e1h, e1l, e2h, e2l, band, kt_count, kt_list, glist, gcount, amask,
        steps, scratch;
    SourcePlan plan;
    // kv_f32: device FP32 Kv [Ng][nang][nang]; e1/e2/beg/end: host tables.
    void build(const float* kv_f32, const std::vector<double>& e1, const std::vector<double>& e2,
               const std::vector<int>& beg, const std::vector<int>& end, int Ng, int nang,
               double dg, hipStream_t s);
    size_t bytes() const {
        return kth.bytes() + ktl.bytes() + e1h.bytes() + e1l.bytes() + e2h.bytes() + e2l.bytes() +
               band.bytes() + kt_count.bytes() + kt_list.bytes() + glist.bytes() +
               gcount.bytes() + amask.bytes() + steps.bytes();
    }
};

// Split-FP16 sine bases for the MFMA DST (FP32 state, Nmu, Nom <= 32).
struct AnglePlanBuffers {
    DeviceBuffer h, l;
    AnglePlan plan;
    // The MFMA DST holds planes of at most 32 x 32 directions and addresses a
    // launch's planes with 32-bit offsets (launch_angle_mfma); other grids
    // take the vector path.
    static bool usable(int Nmu, int Nom, long long planes) {
        return Nmu <= 32 && Nom <= 32 && planes < (1LL << 30) &&
               planes * Nmu * Nom < (1LL << 31);
    }
    void build(int Nmu, int Nom, hipStream_t s);
};

}  // namespace bfpn
