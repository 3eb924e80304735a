// Shared definitions of the gfx950 NonbondedForce/PME library.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "status.h"

#include <rocprofiler-sdk-roctx/roctx.h>

#define OMM_HIP(call)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            throw ::ommhip::Error(OMM_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace ommhip {

// A roctx range over a host-side scope (rocprofv3 --marker-trace shows them
// beside the kernels): step chunks, graph capture and replay, overflow
// replays, preparation, autotuning, evaluations and queries.  Without a
// profiler attached a range costs two calls that return at once.
struct TraceRange {
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

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
    ST_BIN_OVERFLOW = 8u,  // a PME spread tile received more atoms than its bin holds
    ST_CUTOFF_BOX = 16u,   // a box edge fell below twice the cutoff
    ST_CONSTRAINTS = 32u,  // a constraint projection was singular or did not converge
    ST_VBOX = 64u,         // an atom left the region its non-periodic system's virtual box serves
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
// Checks box vectors the way the reference does (system.rs:16-39): finite, a
// along x, b in the x-y plane, reduced form.
void validate_box(const double* v);

// Recomputes the reciprocal rows, volume and triclinic flag of a box from its
// vectors (pme.rs:176-199); used where a barostat rescales a box on the device.
__host__ __device__ inline void box_refresh(Box& b) {
    const double* a = b.v;
    const double* bb = b.v + 3;
    const double* c = b.v + 6;
    const double bc[3] = {bb[1] * c[2] - bb[2] * c[1], bb[2] * c[0] - bb[0] * c[2], bb[0] * c[1] - bb[1] * c[0]};
    const double ca[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
    const double ab[3] = {a[1] * bb[2] - a[2] * bb[1], a[2] * bb[0] - a[0] * bb[2], a[0] * bb[1] - a[1] * bb[0]};
    const double vol = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
    for (int i = 0; i < 3; ++i) {
        b.recip[i] = bc[i] / vol;
        b.recip[3 + i] = ca[i] / vol;
        b.recip[6 + i] = ab[i] / vol;
    }
    b.volume = vol;
    b.periodic = 1;
    b.triclinic = (b.v[3] != 0.0 || b.v[6] != 0.0 || b.v[7] != 0.0) ? 1 : 0;
}

constexpr int MAX_STAGES = 12;  // FFT stages per axis (radices, 0-terminated)

// One system of a batch as the kernels see it.  A batch concatenates the
// particle arrays of its systems; every index the kernels exchange (atoms,
// sorted slots, list blocks, cells, grid points) is global, and this record
// gives the ranges that belong to system s:
//   atoms   [atom_off, atom_off + n)
//   slots   [slot_off, slot_off + 32 nblocks)   (slots past n hold -1)
//   blocks  [block_off, block_off + nblocks)
//   cells   [cell_off, cell_off + ncell[0] ncell[1] ncell[2])
//   chunks  [chunk_off, chunk_off + ceil(32 nblocks / 64))  (PME spread/gather)
//   grid    qgrid_off (real points), cgrid_off (complex half-spectrum points)
//   tiles   [tile_off, tile_off + tiles[0] tiles[1] tiles[2])  (PME spread)
// The box is device state: a barostat rescales it in place.
struct SysInfo {
    Box box;
    int atom_off, n;
    int slot_off, block_off, nblocks;
    int cell_off, ncell[3];
    int chunk_off;
    int grid[3], nzh;
    long long qgrid_off, cgrid_off;
    int tiles[3], tile_off;  // PME spread tiles per axis, first tile
    int ptile;               // PME spread tile edge (6 or 8 grid points)
    int dispersion;          // this table serves the LJPME dispersion mesh (its influence function)
    int roots_off[3], moduli_off[3];
    int fft_lines[2][3];  // lines per FFT workgroup along each axis, [0] FP32 / [1] FP64 transforms
    // Plane transforms (one workgroup per x plane): z lines and y lines per
    // fft_lds call, [0] FP32 / [1] FP64.
    int plane_lines[2][2];
    int radix[3][MAX_STAGES];
    double alpha;
    // Volume-dependent host terms evaluated on the device:
    // E_self = self_const + neutral_coeff / V, E_dispersion = disp_coeff / V.
    double self_const, neutral_coeff, disp_coeff;
    unsigned long long seed;  // integrator noise stream of this system
    // Non-periodic methods on a virtual box (vcheck = 1): the integrators
    // raise ST_VBOX when an atom leaves [vlo, vhi], the region within which
    // every displacement the list and kernels see is the true one.
    int vcheck;
    double vlo[3], vhi[3];
};

// Largest s in [0, count) with off[s] <= x, for x in [off[0], off[count]).
__device__ __forceinline__ int find_sys(const int* __restrict__ off, int count, int x) {
    int lo = 0, hi = count;
    while (hi - lo > 1) {
        const int m = (lo + hi) >> 1;
        if (off[m] <= x) lo = m;
        else hi = m;
    }
    return lo;
}

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
// recip[4k] is 1 / v[4k] for a reduced (lower-triangular) box.
__device__ __forceinline__ void wrap_nearest(double d[3], const Box& b) {
    for (int k = 2; k >= 0; --k) {
        double s = rint(d[k] * b.recip[k * 4]);
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

// Raises the rebuild flags of a neighbour list for the lanes with moved set:
// flag[0] (some system) once per wave and flag[1 + si] per system, with one
// atomic when every moved lane of the wave belongs to the same system.
// Called by all 64 lanes.
__device__ __forceinline__ void wave_raise_flags(int* flag, bool moved, int si) {
    const unsigned long long m = __ballot(moved);
    if (!m) return;
    const int first = __ffsll((long long)m) - 1;
    const int lead = __shfl(si, first, 64);
    const bool mixed = __ballot(moved && si != lead) != 0ull;
    const int lane = threadIdx.x & 63;
    // The flags only go from 0 to 1 here: a plain read skips the atomic when
    // another wave has already raised it (contended device-scope atomics are
    // slow, and a whole system's atoms may cross the skin in the same step).
    if (lane == first) {
        if (!__atomic_load_n(flag, __ATOMIC_RELAXED)) atomicOr(flag, 1);
        if (!__atomic_load_n(flag + 1 + lead, __ATOMIC_RELAXED)) atomicOr(flag + 1 + lead, 1);
    }
    if (mixed && moved && si != lead && !__atomic_load_n(flag + 1 + si, __ATOMIC_RELAXED)) atomicOr(flag + 1 + si, 1);
}

// Adds each lane's energy v to slot `slot` of its system si (si < 0: lane
// idle) in the per-system energy array: one atomic per wave when the active
// lanes share a system, one per lane otherwise.  Called by all 64 lanes; a
// null energy array (force-only evaluations) adds nothing.
template <class A>
__device__ __forceinline__ void wave_energy_add(A* energy, int si, int slot, double v, unsigned* status) {
    if (!energy) return;  // force-only evaluation
    const int lead = __shfl(si, 0, 64);  // lane 0 holds the lowest index, so it is active if any lane is
    const bool uniform = __ballot(si >= 0 && si != lead) == 0ull;
    if (uniform) {
        const double t = wave_sum(v);
        if ((threadIdx.x & 63) == 0 && lead >= 0 && t != 0.0) acc_add(energy + lead * E_SLOTS + slot, t, ENERGY_SCALE, status);
    } else if (si >= 0 && v != 0.0) {
        acc_add(energy + si * E_SLOTS + slot, v, ENERGY_SCALE, status);
    }
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
