// Shared definitions of the gfx950 NonbondedForce/PME library.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "status.h"

#define OMM_HIP(call)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            throw ::ommhip::Error(OMM_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace ommhip {

// 1/(4 pi eps0) in kJ/mol nm / e^2 (nonbonded.rs:16).
constexpr double ONE_4PI_EPS0 = 138.93545764438198;
constexpr double PI = 3.14159265358979323846;

// Fixed-point scales of the int64 accumulators.  Forces and energies carry
// 32 fractional bits (resolution 2.3e-10 kJ/mol[/nm], range +-2.1e9); the
// charge grid carries 40 (resolution 9e-13 e, range +-8.4e6 e).
constexpr double FORCE_SCALE = 4294967296.0;
constexpr double ENERGY_SCALE = 4294967296.0;
constexpr double GRID_SCALE = 1099511627776.0;

// Status bits raised on the device and read back by the host.
enum : unsigned {
    ST_NONFINITE = 1u,
    ST_OVERFLOW = 2u,
    ST_LIST_OVERFLOW = 4u,
};

// Energy slots of the device energy accumulator (fixed point).
enum : int {
    E_DIRECT = 0,
    E_EXCEPT = 1,
    E_RECIP = 2,
    E_BOND = 3,
    E_ANGLE = 4,
    E_SLOTS = 8,
};

// Periodic box in the reference's layout (rows a, b, c) plus its reciprocal
// rows and volume (pme.rs:176-199).
struct Box {
    double v[9];
    double recip[9];
    double volume;
    int periodic;
    int triclinic;
};

Box make_box(const double* box9);

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        if (count) OMM_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
        n = count;
    }
    // Host-to-device copy ordered on stream s; returns once the copy is done,
    // so the host buffer may be released immediately.
    void upload(const T* host, size_t count, hipStream_t s) {
        alloc(count);
        if (count) {
            OMM_HIP(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
            OMM_HIP(hipStreamSynchronize(s));
        }
    }
    void upload(const std::vector<T>& host, hipStream_t s) { upload(host.data(), host.size(), s); }
    void download(T* host, size_t count, hipStream_t s) const {
        if (count) OMM_HIP(hipMemcpyAsync(host, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

// Device code helpers ----------------------------------------------------------

__device__ __forceinline__ long long to_fixed(double v, double scale, unsigned* status) {
    double s = v * scale;
    if (!(fabs(s) < 9.0e18)) {
        if (status) atomicOr(status, (s != s) ? ST_NONFINITE : ST_OVERFLOW);
        return 0;
    }
    return (long long)s;
}

__device__ __forceinline__ void add_fixed(long long* dst, double v, double scale, unsigned* status) {
    long long q = to_fixed(v, scale, status);
    if (q) atomicAdd(reinterpret_cast<unsigned long long*>(dst), (unsigned long long)q);
}

// Accumulation into the mixed form's int64 fixed-point buffers or the plain
// form's FP64 buffers.
__device__ __forceinline__ void acc_add(long long* p, double v, double scale, unsigned* st) {
    add_fixed(p, v, scale, st);
}
__device__ __forceinline__ void acc_add(double* p, double v, double, unsigned*) {
    if (v != 0.0) unsafeAtomicAdd(p, v);
}

// Fill kernels (graph-safe replacements for memset).
template <class T>
__global__ void k_fill(T* p, size_t n, T v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

template <class T>
inline void fill_async(T* p, size_t n, T v, hipStream_t s) {
    if (!n) return;
    int blocks = (int)std::min<size_t>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_fill<T>, dim3(blocks), dim3(256), 0, s, p, n, v);
}

// Nearest image in the reduced box (c, then b, then a) with round-to-nearest
// shifts; used where any consistent image will do (binning, block bounds).
__device__ __forceinline__ void wrap_nearest(double d[3], const Box& b) {
    for (int k = 2; k >= 0; --k) {
        double s = rint(d[k] / b.v[k * 4]);
        for (int a = 0; a <= k; ++a) d[a] -= s * b.v[k * 3 + a];
    }
}

// Order-independent hash of a pair set: the splitmix64 finaliser of
// (min << 32 | max), summed with wrap-around over the set.
__host__ __device__ inline uint64_t pair_key(int32_t i, int32_t j) {
    uint64_t z = ((uint64_t)(uint32_t)i << 32) | (uint64_t)(uint32_t)j;
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Reduced-triclinic minimum image in the reference's operation order
// (harmonic_bond.rs:152-159): c, then b, then a; no fused multiply-add.
__device__ __forceinline__ void min_image_exact(double d[3], const double* bv) {
#pragma clang fp contract(off)
    for (int dd = 2; dd >= 0; --dd) {
        double shift = floor(d[dd] / bv[dd * 3 + dd] + 0.5);
        for (int axis = 0; axis <= dd; ++axis) {
            double t = shift * bv[dd * 3 + axis];
            d[axis] = d[axis] - t;
        }
    }
}

}  // namespace ommhip
