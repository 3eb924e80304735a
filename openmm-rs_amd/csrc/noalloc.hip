// The charge-mesh derivatives of PmeCache::charge_derivatives (pme.rs:88-138)
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops): inputs and outputs are the caller's device
// buffers, every launch goes to the caller's stream, and the only other memory
// is the file-scope `scratch` array below, one FP64 complex grid of at most
// NM_MAX_POINTS points per device, held until the process ends.
//
// The mesh is the reference's algorithm on the full complex grid: spread of
// the charges with order-5 B-splines, forward transform along z, y and x (a
// direct DFT per line, roots from sincospi), the influence function with the
// B-spline moduli (pme.rs:239-259 and 332-372, k = 0 cleared), the
// unnormalised inverse transform, and the potential at each listed particle's
// stencil added to its derivative.  All FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/ommhip.h"
#include "spline.h"

namespace ommhip {
namespace {

constexpr int NM_MAX_EDGE = 64;
constexpr long long NM_MAX_POINTS = (long long)NM_MAX_EDGE * NM_MAX_EDGE * NM_MAX_EDGE;
constexpr int NM_MAX_DEVICES = 64;

}  // namespace
}  // namespace ommhip

// One grid per device: 64^3 FP64 complex values, 4 MiB.
static __device__ __align__(256) double2 scratch[ommhip::NM_MAX_POINTS];

namespace ommhip {
namespace {

std::mutex g_scratch_mutex;
double2* g_scratch[NM_MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array at its own address).
double2* scratch_ptr() {
    int dev = 0;
    OMM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NM_MAX_DEVICES) throw Error(OMM_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(OMM_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<double2*>(p);
    }
    return g_scratch[dev];
}

__global__ void k_nm_clear(double2* g, long long npts) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < npts;
         i += (long long)gridDim.x * blockDim.x)
        g[i] = make_double2(0.0, 0.0);
}

// Spread (pme.rs:315-330): five threads per particle, one per x plane of its
// stencil; uncharged particles do not spread.
__global__ void k_nm_spread(int n, const double* __restrict__ pos, const double* __restrict__ charge, Box b,
                            int nx, int ny, int nz, double2* g) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= 5LL * n) return;
    const int a = (int)(t / 5), ix = (int)(t - 5LL * a);
    const double q = charge[a];
    if (q == 0.0) return;
    const int dims[3] = {nx, ny, nz};
    int base[3];
    double w[3][5], dw[3][5];
    stencil<double>(pos + 3 * (size_t)a, b, dims, base, w, dw);
    const int gx = (base[0] + ix) % nx;
    for (int iy = 0; iy < 5; ++iy) {
        const int gy = (base[1] + iy) % ny;
        double2* row = g + ((size_t)gx * ny + gy) * nz;
        for (int iz = 0; iz < 5; ++iz)
            atomicAdd(&row[(base[2] + iz) % nz].x, q * w[0][ix] * w[1][iy] * w[2][iz]);
    }
}

// One DFT pass along `axis`, in place, one workgroup per line of n <= 64
// points: out[j] = sum_k in[k] e^{-+2 pi i j k / n} (unnormalised inverse).
__global__ __launch_bounds__(NM_MAX_EDGE) void k_nm_dft(double2* g, int nx, int ny, int nz, int axis, int inverse) {
    __shared__ double2 line[NM_MAX_EDGE], root[NM_MAX_EDGE];
    const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const long long L = blockIdx.x;
    long long base, stride;
    if (axis == 2) {
        base = L * nz;
        stride = 1;
    } else if (axis == 1) {
        base = (L / nz) * (long long)ny * nz + L % nz;
        stride = nz;
    } else {
        base = L;
        stride = (long long)ny * nz;
    }
    const int t = threadIdx.x;
    if (t < n) {
        line[t] = g[base + t * stride];
        double s, c;
        sincospi(2.0 * t / n, &s, &c);
        root[t] = make_double2(c, inverse ? s : -s);
    }
    __syncthreads();
    if (t < n) {
        double ax = 0.0, ay = 0.0;
        int idx = 0;
        for (int k = 0; k < n; ++k) {
            const double2 w = root[idx], v = line[k];
            ax += v.x * w.x - v.y * w.y;
            ay += v.x * w.y + v.y * w.x;
            idx += t;
            if (idx >= n) idx -= n;
        }
        g[base + t * stride] = make_double2(ax, ay);
    }
}

// |sum_i M5(i+1) e^{2 pi i k i / n}|^2 (pme.rs:239-259) before the
// replacement of near-zero values.
__device__ double nm_modulus_raw(int k, int n) {
    const double m5[4] = {1.0 / 24.0, 11.0 / 24.0, 11.0 / 24.0, 1.0 / 24.0};
    double re = 0.0, im = 0.0;
    for (int i = 0; i < 4; ++i) {
        double s, c;
        sincospi(2.0 * (double)((long long)k * i % n) / n, &s, &c);
        re += m5[i] * c;
        im += m5[i] * s;
    }
    return re * re + im * im;
}

__device__ double nm_modulus(int k, int n) {
    const double v = nm_modulus_raw(k, n);
    if (v >= 1e-7) return v;
    return 0.5 * (nm_modulus_raw((k + n - 1) % n, n) + nm_modulus_raw((k + 1) % n, n));
}

// Reciprocal convolution (pme.rs:332-372) over the full grid, k = 0 cleared.
__global__ void k_nm_convolve(double2* g, int nx, int ny, int nz, Box b, double alpha) {
    const long long npts = (long long)nx * ny * nz;
    const double pi2_a2 = PI * PI / (alpha * alpha);
    const double pref = ONE_4PI_EPS0 / (PI * b.volume);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < npts;
         i += (long long)gridDim.x * blockDim.x) {
        if (i == 0) {
            g[0] = make_double2(0.0, 0.0);
            continue;
        }
        const int z = (int)(i % nz), y = (int)((i / nz) % ny), x = (int)(i / ((long long)ny * nz));
        const double m[3] = {(double)(x < (nx + 1) / 2 ? x : x - nx), (double)(y < (ny + 1) / 2 ? y : y - ny),
                             (double)(z < (nz + 1) / 2 ? z : z - nz)};
        double k2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            const double kv = m[0] * b.recip[a] + m[1] * b.recip[3 + a] + m[2] * b.recip[6 + a];
            k2 += kv * kv;
        }
        const double mod = nm_modulus(x, nx) * nm_modulus(y, ny) * nm_modulus(z, nz);
        const double infl = pref * exp(-pi2_a2 * k2) / (k2 * mod);
        g[i] = make_double2(g[i].x * infl, g[i].y * infl);
    }
}

// derivatives[k] += the potential at the stencil of particle atoms[k]
// (pme.rs:117-137); an index outside [0, n) gives NaN.
__global__ void k_nm_potential(int count, const int* __restrict__ atoms, int n, const double* __restrict__ pos, Box b,
                               int nx, int ny, int nz, const double2* __restrict__ g, double* out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int a = atoms[k];
    if (a < 0 || a >= n) {
        out[k] = __longlong_as_double(0x7ff8000000000000ll);
        return;
    }
    const int dims[3] = {nx, ny, nz};
    int base[3];
    double w[3][5], dw[3][5];
    stencil<double>(pos + 3 * (size_t)a, b, dims, base, w, dw);
    double sum = 0.0;
    for (int ix = 0; ix < 5; ++ix) {
        const int gx = (base[0] + ix) % nx;
        for (int iy = 0; iy < 5; ++iy) {
            const int gy = (base[1] + iy) % ny;
            const double2* row = g + ((size_t)gx * ny + gy) * nz;
            for (int iz = 0; iz < 5; ++iz) sum += row[(base[2] + iz) % nz].x * w[0][ix] * w[1][iy] * w[2][iz];
        }
    }
    out[k] += sum;
}

}  // namespace
}  // namespace ommhip

using ommhip::Error;

extern "C" omm_status omm_pme_charge_derivatives_noalloc(int32_t num_particles, const double* pos,
                                                         const double* charges, const double* box, double alpha,
                                                         const int32_t grid[3], const int32_t* indices, int32_t count,
                                                         double* derivatives, void* stream) {
    return ommhip::guarded([&]() -> omm_status {
        using namespace ommhip;
        if (num_particles < 0 || count < 0) throw Error(OMM_ERR_INVALID, "negative particle or index count");
        if (!box || !grid) throw Error(OMM_ERR_INVALID, "null box or grid");
        if (num_particles > 0 && (!pos || !charges)) throw Error(OMM_ERR_INVALID, "null positions or charges");
        if (count > 0 && (!indices || !derivatives)) throw Error(OMM_ERR_INVALID, "null indices or derivatives");
        if (!(std::isfinite(alpha) && alpha > 0.0)) throw Error(OMM_ERR_INVALID, "alpha must be positive and finite");
        for (int a = 0; a < 3; ++a)
            if (grid[a] < 5 || grid[a] > NM_MAX_EDGE)
                throw Error(OMM_ERR_INVALID, "grid dimensions must lie in [5, 64] for the allocation-free mesh");
        validate_box(box);
        if (count == 0) return OMM_OK;
        const Box b = make_box(box);
        const int nx = grid[0], ny = grid[1], nz = grid[2];
        const long long npts = (long long)nx * ny * nz;
        hipStream_t s = static_cast<hipStream_t>(stream);
        double2* g = scratch_ptr();
        const int fill_blocks = (int)std::min<long long>((npts + 255) / 256, 1024);
        hipLaunchKernelGGL(k_nm_clear, dim3(fill_blocks), dim3(256), 0, s, g, npts);
        if (num_particles > 0)
            hipLaunchKernelGGL(k_nm_spread, dim3((unsigned)((5LL * num_particles + 255) / 256)), dim3(256), 0, s,
                               num_particles, pos, charges, b, nx, ny, nz, g);
        hipLaunchKernelGGL(k_nm_dft, dim3((unsigned)(nx * ny)), dim3(NM_MAX_EDGE), 0, s, g, nx, ny, nz, 2, 0);
        hipLaunchKernelGGL(k_nm_dft, dim3((unsigned)(nx * nz)), dim3(NM_MAX_EDGE), 0, s, g, nx, ny, nz, 1, 0);
        hipLaunchKernelGGL(k_nm_dft, dim3((unsigned)(ny * nz)), dim3(NM_MAX_EDGE), 0, s, g, nx, ny, nz, 0, 0);
        hipLaunchKernelGGL(k_nm_convolve, dim3(fill_blocks), dim3(256), 0, s, g, nx, ny, nz, b, alpha);
        hipLaunchKernelGGL(k_nm_dft, dim3((unsigned)(ny * nz)), dim3(NM_MAX_EDGE), 0, s, g, nx, ny, nz, 0, 1);
        hipLaunchKernelGGL(k_nm_dft, dim3((unsigned)(nx * nz)), dim3(NM_MAX_EDGE), 0, s, g, nx, ny, nz, 1, 1);
        hipLaunchKernelGGL(k_nm_dft, dim3((unsigned)(nx * ny)), dim3(NM_MAX_EDGE), 0, s, g, nx, ny, nz, 2, 1);
        hipLaunchKernelGGL(k_nm_potential, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, count, indices,
                           num_particles, pos, b, nx, ny, nz, g, derivatives);
        OMM_HIP(hipGetLastError());
        return OMM_OK;
    });
}
