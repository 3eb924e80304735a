// Plain Ewald reciprocal space on the device (ewald.rs:22-122).
#include "ewald.h"

#include <cmath>

namespace ommhip {

namespace {

constexpr int EW_THREADS = 256;
constexpr int ZCHUNK = 16;  // wave vectors along z per structure-factor workgroup

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// e^{2 pi i m x_a / L_a} for m in [0, kmax), every atom and axis.
__global__ void k_ewald_phases(int n, int kmax, const double* __restrict__ pos, const SysInfo* __restrict__ sys,
                               double2* __restrict__ phase) {
    const size_t total = (size_t)3 * kmax * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % n);
        const int m = (int)((e / n) % kmax);
        const int a = (int)(e / ((size_t)n * kmax));
        const double f = pos[3 * i + a] / sys[0].box.v[a * 4];
        double sv, cv;
        sincospi(2.0 * m * f, &sv, &cv);
        phase[e] = double2{cv, sv};
    }
}

__device__ __forceinline__ double2 phase_of(const double2* __restrict__ phase, int n, int kmax, int axis, int m,
                                            int i) {
    double2 v = phase[((size_t)axis * kmax + (m < 0 ? -m : m)) * n + i];
    if (m < 0) v.y = -v.y;
    return v;
}

// One workgroup per (x, y, ZCHUNK wave vectors along z): S(k) over all atoms
// (each thread a strided subset, then a fixed-order tree), stored as I(k) S(k);
// the energy I(k) |S(k)|^2 goes to slot E_RECIP.
template <class A>
__global__ __launch_bounds__(EW_THREADS) void k_ewald_structure(int n, int kmax, int kz,
                                                                const int4* __restrict__ chunks,
                                                                const double2* __restrict__ phase,
                                                                const double* __restrict__ charge,
                                                                const SysInfo* __restrict__ sys, double alpha,
                                                                double2* __restrict__ sfac, A* energy,
                                                                unsigned* status) {
    const int4 c = chunks[blockIdx.x];  // (x, y, first z, index of the first wave vector)
    const int count = min(ZCHUNK, kz - c.z);
    double re[ZCHUNK], im[ZCHUNK];
#pragma unroll
    for (int k = 0; k < ZCHUNK; ++k) re[k] = im[k] = 0.0;
    for (int i = threadIdx.x; i < n; i += EW_THREADS) {
        const double q = charge[i];
        const double2 xy = cmul(phase_of(phase, n, kmax, 0, c.x, i), phase_of(phase, n, kmax, 1, c.y, i));
        const double2 step = phase_of(phase, n, kmax, 2, 1, i);
        double2 v = cmul(xy, phase_of(phase, n, kmax, 2, c.z, i));
        v.x *= q;
        v.y *= q;
#pragma unroll
        for (int k = 0; k < ZCHUNK; ++k) {
            if (k < count) {
                re[k] += v.x;
                im[k] += v.y;
                v = cmul(v, step);
            }
        }
    }
    __shared__ double s_re[EW_THREADS / 64][ZCHUNK], s_im[EW_THREADS / 64][ZCHUNK];
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < ZCHUNK; ++k) {
        const double r = wave_sum(re[k]), m = wave_sum(im[k]);
        if ((threadIdx.x & 63) == 0) {
            s_re[wave][k] = r;
            s_im[wave][k] = m;
        }
    }
    __syncthreads();
    if (threadIdx.x < count) {
        const int k = threadIdx.x;
        const double sr = (s_re[0][k] + s_re[1][k]) + (s_re[2][k] + s_re[3][k]);
        const double si = (s_im[0][k] + s_im[1][k]) + (s_im[2][k] + s_im[3][k]);
        const Box& b = sys[0].box;
        const double wx = 2.0 * PI * c.x / b.v[0], wy = 2.0 * PI * c.y / b.v[4], wz = 2.0 * PI * (c.z + k) / b.v[8];
        const double k2 = wx * wx + wy * wy + wz * wz;
        const double coef = ONE_4PI_EPS0 * 4.0 * PI / (b.v[0] * b.v[4] * b.v[8]);
        const double infl = coef * exp(-k2 / (4.0 * alpha * alpha)) / k2;
        sfac[c.w + k] = double2{infl * sr, infl * si};
        if (energy) acc_add(energy + E_RECIP, infl * (sr * sr + si * si), ENERGY_SCALE, status);
    }
}

// One thread per atom: F_i = sum_k 2 Im(conj(I S(k)) q_i e^{i k.r_i}) k.
template <class A>
__global__ __launch_bounds__(EW_THREADS) void k_ewald_forces(int n, int kmax, int kz, int nrows,
                                                             const int4* __restrict__ rows,
                                                             const double2* __restrict__ phase,
                                                             const double* __restrict__ charge,
                                                             const SysInfo* __restrict__ sys,
                                                             const double2* __restrict__ sfac, A* force,
                                                             unsigned* status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double q = charge[i];
    if (q == 0.0) return;
    const double2 step = phase_of(phase, n, kmax, 2, 1, i);
    double gx = 0.0, gy = 0.0, gz = 0.0;  // sum of scale * (x, y, z)
    for (int r = 0; r < nrows; ++r) {
        const int4 row = rows[r];  // (x, y, first z, index of the first wave vector)
        const double2 xy = cmul(phase_of(phase, n, kmax, 0, row.x, i), phase_of(phase, n, kmax, 1, row.y, i));
        double2 v = cmul(xy, phase_of(phase, n, kmax, 2, row.z, i));
        v.x *= q;
        v.y *= q;
        double sx = 0.0, sz = 0.0;
        for (int z = row.z; z < kz; ++z) {
            const double2 S = sfac[row.w + (z - row.z)];
            const double scale = 2.0 * (S.x * v.y - S.y * v.x);
            sx += scale;
            sz += scale * z;
            v = cmul(v, step);
        }
        gx += sx * row.x;
        gy += sx * row.y;
        gz += sz;
    }
    const Box& b = sys[0].box;
    acc_add(force + 3 * i, gx * (2.0 * PI / b.v[0]), FORCE_SCALE, status);
    acc_add(force + 3 * i + 1, gy * (2.0 * PI / b.v[4]), FORCE_SCALE, status);
    acc_add(force + 3 * i + 2, gz * (2.0 * PI / b.v[8]), FORCE_SCALE, status);
}

}  // namespace

void ewald_parameters(const double widths[3], double cutoff, double tolerance, double* alpha, int limits[3]) {
    const double a = std::sqrt(-std::log(2.0 * tolerance)) / cutoff;
    *alpha = a;
    for (int axis = 0; axis < 3; ++axis) {
        const double wa = widths[axis] * a;
        // Estimated error of the truncated sum with limit k.
        auto excess = [&](int k) { return tolerance - 0.05 * std::sqrt(wa) * k * std::exp(-std::pow(k * PI / wa, 2)); };
        int k = 10;
        if (excess(k) > 0.0) {
            while (k > 0 && excess(k) > 0.0) --k;
            ++k;
        } else {
            while (excess(k) < 0.0) ++k;
        }
        limits[axis] = k | 1;
    }
}

namespace {
struct EwaldLayout {
    std::vector<int4> rows, chunks;
    int nk = 0;
};

EwaldLayout ewald_layout(const int lim[3]) {
    EwaldLayout L;
    for (int x = 0; x < lim[0]; ++x)
        for (int y = x == 0 ? 0 : 1 - lim[1]; y < lim[1]; ++y) {
            const int zlo = (x == 0 && y == 0) ? 1 : 1 - lim[2];
            if (zlo >= lim[2]) continue;
            L.rows.push_back(int4{x, y, zlo, L.nk});
            for (int z = zlo; z < lim[2]; z += ZCHUNK) L.chunks.push_back(int4{x, y, z, L.nk + (z - zlo)});
            L.nk += lim[2] - zlo;
        }
    return L;
}
}  // namespace

void ewald_setup(EwaldDev& ew, int n, const int limits[3], hipStream_t s) {
    for (int a = 0; a < 3; ++a)
        if (limits[a] < 1 || limits[a] > 4096) throw Error(OMM_ERR_INVALID, "Ewald wave-vector limits out of range");
    ew.n = n;
    for (int a = 0; a < 3; ++a) ew.limits[a] = limits[a];
    ew.kmax = std::max(std::max(limits[0], limits[1]), std::max(limits[2], 2));
    EwaldLayout L = ewald_layout(limits);
    ew.nk = L.nk;
    std::vector<int4> both(L.rows);
    both.insert(both.end(), L.chunks.begin(), L.chunks.end());
    ew.kvec.upload(both, s);  // rows, then structure-factor chunks
    ew.phase.alloc((size_t)3 * ew.kmax * n);
    ew.sfac.alloc(std::max(1, ew.nk));
}

void ewald_run(EwaldDev& ew, const double* pos, const double* charge, const SysInfo* sys_dev, double alpha,
               long long* force_fx, double* force_d, long long* energy_fx, double* energy_d, unsigned* status,
               bool mixed, hipStream_t s) {
    if (ew.n == 0 || ew.nk == 0) return;
    EwaldLayout L = ewald_layout(ew.limits);
    const int nrows = (int)L.rows.size(), nchunks = (int)L.chunks.size();
    const size_t total = (size_t)3 * ew.kmax * ew.n;
    const int pb = (int)std::min<size_t>((total + 255) / 256, 65535);
    hipLaunchKernelGGL(k_ewald_phases, dim3(pb), dim3(256), 0, s, ew.n, ew.kmax, pos, sys_dev, ew.phase.p);
    const int4* rows = ew.kvec.p;
    const int4* chunks = ew.kvec.p + nrows;
    const int fb = (ew.n + EW_THREADS - 1) / EW_THREADS;
    if (mixed) {
        hipLaunchKernelGGL(k_ewald_structure<long long>, dim3(nchunks), dim3(EW_THREADS), 0, s, ew.n, ew.kmax,
                           ew.limits[2], chunks, ew.phase.p, charge, sys_dev, alpha, ew.sfac.p, energy_fx, status);
        hipLaunchKernelGGL(k_ewald_forces<long long>, dim3(fb), dim3(EW_THREADS), 0, s, ew.n, ew.kmax, ew.limits[2],
                           nrows, rows, ew.phase.p, charge, sys_dev, ew.sfac.p, force_fx, status);
    } else {
        hipLaunchKernelGGL(k_ewald_structure<double>, dim3(nchunks), dim3(EW_THREADS), 0, s, ew.n, ew.kmax,
                           ew.limits[2], chunks, ew.phase.p, charge, sys_dev, alpha, ew.sfac.p, energy_d, status);
        hipLaunchKernelGGL(k_ewald_forces<double>, dim3(fb), dim3(EW_THREADS), 0, s, ew.n, ew.kmax, ew.limits[2],
                           nrows, rows, ew.phase.p, charge, sys_dev, ew.sfac.p, force_d, status);
    }
    OMM_HIP(hipGetLastError());
}

}  // namespace ommhip
