// Smooth PME on the device: B-spline spread into a fixed-point (mixed) or
// FP64 (plain) grid, a real-to-complex 3D FFT as three passes of batched 1D
// mixed-radix Stockham transforms held in LDS, the reciprocal convolution and
// its energy fused into the x pass (forward x, multiply, inverse x in one
// kernel), the inverse passes and the force gather.
//
// The reference transforms the full complex grid (pme.rs:271-303) and sums the
// energy over all of it (pme.rs:339-372); for a real charge grid S(-k) =
// conj(S(k)) and the influence function is even, so the half spectrum
// kz <= nz/2 with weights 1 (kz = 0, nz/2) and 2 (otherwise) carries the same
// sum.  Any grid size is supported: radices are the prime factors of n (8 and
// 4 preferred), each stage a direct DFT of its radix.
#include <cmath>
#include <complex>

#include "pme.h"

namespace ommhip {

namespace {

constexpr int FFT_THREADS = 256;
constexpr int MAX_STAGES = 12;

template <class R>
struct Cx {
    R x, y;
};

// Order-5 cardinal B-spline weights and derivatives (pme.rs:151-174).
template <class R>
__device__ __forceinline__ void spline5(R f, R w[5], R dw[5]) {
    w[0] = (R)1 - f;
    w[1] = f;
    w[2] = f * w[1] / (R)2;
    w[1] = ((f + (R)1) * w[0] + ((R)2 - f) * w[1]) / (R)2;
    w[0] *= ((R)1 - f) / (R)2;
    w[3] = f * w[2] / (R)3;
    w[2] = ((f + (R)1) * w[1] + ((R)3 - f) * w[2]) / (R)3;
    w[1] = ((f + (R)2) * w[0] + ((R)2 - f) * w[1]) / (R)3;
    w[0] *= ((R)1 - f) / (R)3;
    dw[0] = -w[0];
    dw[1] = w[0] - w[1];
    dw[2] = w[1] - w[2];
    dw[3] = w[2] - w[3];
    dw[4] = w[3];
    w[4] = f * w[3] / (R)4;
    w[3] = ((f + (R)1) * w[2] + ((R)4 - f) * w[3]) / (R)4;
    w[2] = ((f + (R)2) * w[1] + ((R)3 - f) * w[2]) / (R)4;
    w[1] = ((f + (R)3) * w[0] + ((R)2 - f) * w[1]) / (R)4;
    w[0] *= ((R)1 - f) / (R)4;
}

// pme.rs:201-218: base index and spline of the fractional grid coordinate.
template <class R>
__device__ __forceinline__ void stencil(const double* p, const Box& b, const int dims[3], int idx[3], R w[3][5],
                                        R dw[3][5]) {
    for (int a = 0; a < 3; ++a) {
        double f = p[0] * b.recip[a * 3] + p[1] * b.recip[a * 3 + 1] + p[2] * b.recip[a * 3 + 2];
        double c = (f - floor(f)) * (double)dims[a];
        double fl = floor(c);
        idx[a] = ((int)fl) % dims[a];
        spline5<R>((R)(c - fl), w[a], dw[a]);
    }
}

template <class R, class G>
__global__ void k_spread(int n, const double* __restrict__ pos, const double* __restrict__ charge,
                         const Box* __restrict__ box, int nx, int ny, int nz, G* grid, unsigned* status) {
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    int atom = gid / 5, ix = gid - atom * 5;
    if (atom >= n) return;
    double q = charge[atom];
    if (q == 0.0) return;
    const int dims[3] = {nx, ny, nz};
    int idx[3];
    R w[3][5], dw[3][5];
    stencil<R>(pos + 3 * atom, *box, dims, idx, w, dw);
    const R qw = (R)q * w[0][ix];
    const int gx = (idx[0] + ix) % nx;
    for (int iy = 0; iy < 5; ++iy) {
        const int gy = (idx[1] + iy) % ny;
        const R qwy = qw * w[1][iy];
        G* row = grid + ((size_t)gx * ny + gy) * nz;
        for (int iz = 0; iz < 5; ++iz) {
            const int gz = (idx[2] + iz) % nz;
            acc_add(row + gz, (double)(qwy * w[2][iz]), GRID_SCALE, status);
        }
    }
}

template <class R, class A>
__global__ void k_gather(int n, const double* __restrict__ pos, const double* __restrict__ charge,
                         const Box* __restrict__ box, int nx, int ny, int nz, const R* __restrict__ phi, A* force,
                         unsigned* status) {
    int atom = blockIdx.x * blockDim.x + threadIdx.x;
    if (atom >= n) return;
    double q = charge[atom];
    if (q == 0.0) return;
    const Box& b = *box;
    const int dims[3] = {nx, ny, nz};
    int idx[3];
    R w[3][5], dw[3][5];
    stencil<R>(pos + 3 * atom, b, dims, idx, w, dw);
    R g0 = 0, g1 = 0, g2 = 0;
    for (int ix = 0; ix < 5; ++ix) {
        const int gx = (idx[0] + ix) % nx;
        for (int iy = 0; iy < 5; ++iy) {
            const int gy = (idx[1] + iy) % ny;
            const R wxy = w[0][ix] * w[1][iy];
            const R d0 = dw[0][ix] * w[1][iy];
            const R d1 = w[0][ix] * dw[1][iy];
            const R* row = phi + ((size_t)gx * ny + gy) * nz;
            for (int iz = 0; iz < 5; ++iz) {
                const R p = row[(idx[2] + iz) % nz];
                g0 += p * d0 * w[2][iz];
                g1 += p * d1 * w[2][iz];
                g2 += p * wxy * dw[2][iz];
            }
        }
    }
    // pme.rs:417-422
    for (int a = 0; a < 3; ++a) {
        double s = (double)g0 * nx * b.recip[0 * 3 + a] + (double)g1 * ny * b.recip[1 * 3 + a] +
                   (double)g2 * nz * b.recip[2 * 3 + a];
        acc_add(force + 3 * atom + a, -q * s, FORCE_SCALE, status);
    }
}

enum : int { MODE_Z_R2C = 0, MODE_Y_FWD = 1, MODE_X_CONV = 2, MODE_Y_INV = 3, MODE_Z_C2R = 4 };

struct FftConst {
    int nx, ny, nz, nzh;
    int lines_per_block;  // power of two dividing FFT_THREADS
    double alpha;
};

__device__ __forceinline__ double grid_real(long long v) { return (double)v * (1.0 / GRID_SCALE); }
__device__ __forceinline__ double grid_real(double v) { return v; }

// Batched Stockham FFT of `lines` lines of length n held line-major in a/b.
// Each stage of radix R is evaluated as a direct DFT per output:
// out[(j/Ns)*Ns*R + k + q*Ns] = sum_r in[j + r*m] W_n^{r (k n/(Ns R) + q m)}.
template <class R>
__device__ Cx<R>* fft_lds(Cx<R>* a, Cx<R>* b, int n, const int* radices, const Cx<R>* roots, bool inverse, int T) {
    const int li = threadIdx.x / T, t = threadIdx.x % T;
    int Ns = 1;
    for (int s = 0; s < MAX_STAGES && radices[s]; ++s) {
        const int Rr = radices[s];
        const int m = n / Rr;
        const int step = n / (Ns * Rr);
        const Cx<R>* in = a + li * n;
        Cx<R>* out = b + li * n;
        for (int o = t; o < n; o += T) {
            const int q = o / m, j = o - q * m;
            const int k = j % Ns;
            const int base = (k * step + q * m) % n;
            R ax = 0, ay = 0;
            int idx = 0;
            for (int r = 0; r < Rr; ++r) {
                Cx<R> w = roots[idx];
                if (inverse) w.y = -w.y;
                const Cx<R> v = in[j + r * m];
                ax += v.x * w.x - v.y * w.y;
                ay += v.x * w.y + v.y * w.x;
                idx += base;
                if (idx >= n) idx -= n;
            }
            out[(j / Ns) * Ns * Rr + k + q * Ns] = Cx<R>{ax, ay};
        }
        __syncthreads();
        Cx<R>* tmp = a;
        a = b;
        b = tmp;
        Ns *= Rr;
    }
    return a;
}

template <class R, class G, class A, int MODE>
__global__ __launch_bounds__(FFT_THREADS) void k_fft(FftConst C, const G* __restrict__ qgrid, Cx<R>* cgrid, R* phi,
                                                     const Cx<R>* __restrict__ roots_all,
                                                     const R* __restrict__ moduli_all, const int* __restrict__ radix_all,
                                                     const Box* __restrict__ box, A* energy, unsigned* status) {
    extern __shared__ unsigned char smem_raw[];
    const bool zmode = MODE == MODE_Z_R2C || MODE == MODE_Z_C2R;
    const bool ymode = MODE == MODE_Y_FWD || MODE == MODE_Y_INV;
    const int axis = zmode ? 2 : (ymode ? 1 : 0);
    const int n = zmode ? C.nz : (ymode ? C.ny : C.nx);
    const int LINES = C.lines_per_block;
    const int T = FFT_THREADS / LINES;
    const int nlines = zmode ? C.nx * C.ny : (ymode ? C.nx * C.nzh : C.ny * C.nzh);
    const int line0 = blockIdx.x * LINES;
    const int root_off = axis == 0 ? 0 : (axis == 1 ? C.nx : C.nx + C.ny);

    Cx<R>* bufA = reinterpret_cast<Cx<R>*>(smem_raw);
    Cx<R>* bufB = bufA + LINES * n;
    Cx<R>* roots = bufB + LINES * n;
    __shared__ int s_radix[MAX_STAGES];
    if (threadIdx.x < MAX_STAGES) s_radix[threadIdx.x] = radix_all[axis * MAX_STAGES + threadIdx.x];
    for (int i = threadIdx.x; i < n; i += FFT_THREADS) roots[i] = roots_all[root_off + i];

    // Load.
    for (int e = threadIdx.x; e < LINES * n; e += FFT_THREADS) {
        int li, el;
        if (zmode) {
            li = e / n;
            el = e - li * n;
        } else {
            li = e % LINES;
            el = e / LINES;
        }
        const int L = line0 + li;
        Cx<R> v = {0, 0};
        if (L < nlines) {
            if (MODE == MODE_Z_R2C) {
                v.x = (R)grid_real(qgrid[(size_t)L * C.nz + el]);
            } else if (MODE == MODE_Z_C2R) {
                if (el < C.nzh) {
                    v = cgrid[(size_t)L * C.nzh + el];
                } else {
                    v = cgrid[(size_t)L * C.nzh + (C.nz - el)];
                    v.y = -v.y;
                }
            } else if (ymode) {
                const int x = L / C.nzh, kz = L - x * C.nzh;
                v = cgrid[((size_t)x * C.ny + el) * C.nzh + kz];
            } else {
                const int y = L / C.nzh, kz = L - y * C.nzh;
                v = cgrid[((size_t)el * C.ny + y) * C.nzh + kz];
            }
        }
        bufA[li * n + el] = v;
    }
    __syncthreads();
    const int* rad = s_radix;
    const bool inverse_first = MODE == MODE_Y_INV || MODE == MODE_Z_C2R;
    Cx<R>* res = fft_lds<R>(bufA, bufB, n, rad, roots, inverse_first, T);

    if (MODE == MODE_X_CONV) {
        // Reciprocal convolution and energy (pme.rs:332-372) on the half spectrum.
        const Box& b = *box;
        const R* mod_x = moduli_all;
        const R* mod_y = moduli_all + C.nx;
        const R* mod_z = moduli_all + C.nx + C.ny;
        const double pi2_a2 = PI * PI / (C.alpha * C.alpha);
        const double pref = ONE_4PI_EPS0 / (PI * b.volume);
        double eacc = 0.0;
        for (int e = threadIdx.x; e < LINES * n; e += FFT_THREADS) {
            const int li = e % LINES, kx = e / LINES;
            const int L = line0 + li;
            if (L >= nlines) continue;
            const int y = L / C.nzh, kz = L - y * C.nzh;
            Cx<R>& S = res[li * n + kx];
            if (kx == 0 && y == 0 && kz == 0) {
                S = Cx<R>{0, 0};
                continue;
            }
            const double mx = kx < (C.nx + 1) / 2 ? kx : kx - C.nx;
            const double my = y < (C.ny + 1) / 2 ? y : y - C.ny;
            const double mz = kz < (C.nz + 1) / 2 ? kz : kz - C.nz;
            double k2 = 0.0;
            for (int a = 0; a < 3; ++a) {
                double kv = mx * b.recip[0 * 3 + a] + my * b.recip[1 * 3 + a] + mz * b.recip[2 * 3 + a];
                k2 += kv * kv;
            }
            const double mod = (double)mod_x[kx] * (double)mod_y[y] * (double)mod_z[kz];
            const double infl = pref * exp(-pi2_a2 * k2) / (k2 * mod);
            const double wz = (kz == 0 || 2 * kz == C.nz) ? 1.0 : 2.0;
            const double s2 = (double)S.x * S.x + (double)S.y * S.y;
            eacc += wz * 0.5 * infl * s2;
            S.x = (R)(S.x * infl);
            S.y = (R)(S.y * infl);
        }
        double et = wave_sum(eacc);
        if ((threadIdx.x & 63) == 0) acc_add(energy + E_RECIP, et, ENERGY_SCALE, status);
        __syncthreads();
        Cx<R>* other = res == bufA ? bufB : bufA;
        res = fft_lds<R>(res, other, n, rad, roots, true, T);
    }

    // Store.
    for (int e = threadIdx.x; e < LINES * n; e += FFT_THREADS) {
        int li, el;
        if (zmode) {
            li = e / n;
            el = e - li * n;
        } else {
            li = e % LINES;
            el = e / LINES;
        }
        const int L = line0 + li;
        if (L >= nlines) continue;
        const Cx<R> v = res[li * n + el];
        if (MODE == MODE_Z_R2C) {
            if (el < C.nzh) cgrid[(size_t)L * C.nzh + el] = v;
        } else if (MODE == MODE_Z_C2R) {
            phi[(size_t)L * C.nz + el] = v.x;
        } else if (ymode) {
            const int x = L / C.nzh, kz = L - x * C.nzh;
            cgrid[((size_t)x * C.ny + el) * C.nzh + kz] = v;
        } else {
            const int y = L / C.nzh, kz = L - y * C.nzh;
            cgrid[((size_t)el * C.ny + y) * C.nzh + kz] = v;
        }
    }
}

std::vector<int> factorize(int n) {
    std::vector<int> r;
    while (n % 8 == 0) {
        r.push_back(8);
        n /= 8;
    }
    while (n % 4 == 0) {
        r.push_back(4);
        n /= 4;
    }
    while (n % 2 == 0) {
        r.push_back(2);
        n /= 2;
    }
    for (int p = 3; p * p <= n; p += 2)
        while (n % p == 0) {
            r.push_back(p);
            n /= p;
        }
    if (n > 1) r.push_back(n);
    return r;
}

int lines_for(int n, size_t csize) {
    int lines = 16;
    while (lines > 1 && (2 * (size_t)lines * n + n) * csize > 60000) lines >>= 1;
    return lines;
}

template <class R, class G, class A, int MODE>
void launch_fft(const PmeDev& pm, FftConst C, const G* qgrid, Cx<R>* cgrid, R* phi, const R* roots, const R* moduli,
                const Box* box, A* energy, unsigned* status, hipStream_t s) {
    const bool zmode = MODE == MODE_Z_R2C || MODE == MODE_Z_C2R;
    const bool ymode = MODE == MODE_Y_FWD || MODE == MODE_Y_INV;
    const int n = zmode ? C.nz : (ymode ? C.ny : C.nx);
    const int nlines = zmode ? C.nx * C.ny : (ymode ? C.nx * C.nzh : C.ny * C.nzh);
    C.lines_per_block = lines_for(n, sizeof(Cx<R>));
    const int blocks = (nlines + C.lines_per_block - 1) / C.lines_per_block;
    const size_t shmem = (2 * (size_t)C.lines_per_block * n + n) * sizeof(Cx<R>);
    hipLaunchKernelGGL((k_fft<R, G, A, MODE>), dim3(blocks), dim3(FFT_THREADS), shmem, s, C, qgrid, cgrid, phi,
                       reinterpret_cast<const Cx<R>*>(roots), moduli, pm.radix_dev.p, box, energy, status);
}

template <class R, class G, class A>
void run_t(PmeDev& pm, int n, const double* pos, const double* charge, const Box* box, G* qgrid, R* cgrid, R* phi,
           const R* roots, const R* moduli, A* force, A* energy, unsigned* status, int phases, hipStream_t s) {
    const size_t G3 = (size_t)pm.nx * pm.ny * pm.nz;
    if (phases & PME_SPREAD) {
        fill_async<G>(qgrid, G3, (G)0, s);
        hipLaunchKernelGGL((k_spread<R, G>), dim3((5 * n + 255) / 256), dim3(256), 0, s, n, pos, charge, box, pm.nx,
                           pm.ny, pm.nz, qgrid, status);
    }
    if (phases & PME_FFT) {
        FftConst C{pm.nx, pm.ny, pm.nz, pm.nzh, 16, pm.alpha};
        Cx<R>* cg = reinterpret_cast<Cx<R>*>(cgrid);
        launch_fft<R, G, A, MODE_Z_R2C>(pm, C, qgrid, cg, phi, roots, moduli, box, energy, status, s);
        launch_fft<R, G, A, MODE_Y_FWD>(pm, C, qgrid, cg, phi, roots, moduli, box, energy, status, s);
        launch_fft<R, G, A, MODE_X_CONV>(pm, C, qgrid, cg, phi, roots, moduli, box, energy, status, s);
        launch_fft<R, G, A, MODE_Y_INV>(pm, C, qgrid, cg, phi, roots, moduli, box, energy, status, s);
        launch_fft<R, G, A, MODE_Z_C2R>(pm, C, qgrid, cg, phi, roots, moduli, box, energy, status, s);
    }
    if (phases & PME_GATHER)
        hipLaunchKernelGGL((k_gather<R, A>), dim3((n + 127) / 128), dim3(128), 0, s, n, pos, charge, box, pm.nx,
                           pm.ny, pm.nz, phi, force, status);
    OMM_HIP(hipGetLastError());
}

}  // namespace

std::vector<double> bspline_moduli(int n) {
    // Order-5 spline at integer points: M5(1..4) = 1/24, 11/24, 11/24, 1/24
    // (the weights of a zero fractional coordinate).
    const double w[5] = {1.0 / 24.0, 11.0 / 24.0, 11.0 / 24.0, 1.0 / 24.0, 0.0};
    std::vector<double> values(n);
    for (int k = 0; k < n; ++k) {
        double re = 0.0, im = 0.0;
        for (int i = 0; i < 5; ++i) {
            double t = 2.0 * PI * (double)k * (double)i / (double)n;
            re += w[i] * std::cos(t);
            im += w[i] * std::sin(t);
        }
        values[k] = re * re + im * im;
    }
    for (int k = 0; k < n; ++k)
        if (values[k] < 1e-7) values[k] = 0.5 * (values[(k + n - 1) % n] + values[(k + 1) % n]);
    return values;
}

void pme_setup(PmeDev& pm, double alpha, const int dims[3], bool mixed, hipStream_t s) {
    for (int a = 0; a < 3; ++a)
        if (dims[a] < 5 || dims[a] > 1024) throw Error(OMM_ERR_INVALID, "PME grid dimensions must be in [5, 1024]");
    pm.nx = dims[0];
    pm.ny = dims[1];
    pm.nz = dims[2];
    pm.nzh = pm.nz / 2 + 1;
    pm.alpha = alpha;
    std::vector<int> rad(3 * MAX_STAGES, 0);
    for (int a = 0; a < 3; ++a) {
        pm.radix[a] = factorize(dims[a]);
        if ((int)pm.radix[a].size() >= MAX_STAGES) throw Error(OMM_ERR_INVALID, "PME grid dimension has too many factors");
        for (size_t k = 0; k < pm.radix[a].size(); ++k) rad[a * MAX_STAGES + k] = pm.radix[a][k];
    }
    pm.radix_dev.upload(rad, s);
    std::vector<double> roots, moduli;
    for (int a = 0; a < 3; ++a) {
        int n = dims[a];
        for (int m = 0; m < n; ++m) {
            double t = -2.0 * PI * (double)m / (double)n;
            roots.push_back(std::cos(t));
            roots.push_back(std::sin(t));
        }
        auto mod = bspline_moduli(n);
        moduli.insert(moduli.end(), mod.begin(), mod.end());
    }
    const size_t G3 = (size_t)pm.nx * pm.ny * pm.nz;
    const size_t GH = (size_t)pm.nx * pm.ny * pm.nzh;
    pm.roots_d.upload(roots, s);
    pm.moduli_d.upload(moduli, s);
    pm.cgrid_d.alloc(2 * GH);
    pm.phi_d.alloc(G3);
    if (mixed) {
        std::vector<float> rf(roots.begin(), roots.end()), mf(moduli.begin(), moduli.end());
        pm.roots_f.upload(rf, s);
        pm.moduli_f.upload(mf, s);
        pm.qgrid_fx.alloc(G3);
        pm.cgrid_f.alloc(2 * GH);
        pm.phi_f.alloc(G3);
    } else {
        pm.qgrid_d.alloc(G3);
    }
}

void pme_run(PmeDev& pm, int mode, int n, const double* pos, const double* charge, const Box* box_dev,
             long long* force_fx, double* force_d, long long* energy_fx, double* energy_d, unsigned* status,
             hipStream_t s, int phases) {
    if (mode == MESH_F32)
        run_t<float, long long, long long>(pm, n, pos, charge, box_dev, pm.qgrid_fx.p, pm.cgrid_f.p, pm.phi_f.p,
                                           pm.roots_f.p, pm.moduli_f.p, force_fx, energy_fx, status, phases, s);
    else if (mode == MESH_F64_FIXED)
        run_t<double, long long, long long>(pm, n, pos, charge, box_dev, pm.qgrid_fx.p, pm.cgrid_d.p, pm.phi_d.p,
                                            pm.roots_d.p, pm.moduli_d.p, force_fx, energy_fx, status, phases, s);
    else
        run_t<double, double, double>(pm, n, pos, charge, box_dev, pm.qgrid_d.p, pm.cgrid_d.p, pm.phi_d.p,
                                      pm.roots_d.p, pm.moduli_d.p, force_d, energy_d, status, phases, s);
}

}  // namespace ommhip
