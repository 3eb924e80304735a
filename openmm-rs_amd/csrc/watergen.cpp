// Seeded synthetic flexible 3-site water boxes (host only).
//
// The workloads are water boxes built the same way at every size: oxygens on a
// jittered cubic lattice (a seeded random subset of the m^3 sites when M < m^3,
// so vacancies are spread through the box rather than left in one slab) and a
// uniformly random orientation per molecule.  Molecules are kept whole, since
// exceptions are evaluated without periodic wrapping by default
// (nonbonded.rs:146, 1006-1010).
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/ommhip.h"
#include "status.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// splitmix64 (Steele, Lea, Flood 2014) to expand the seed.
uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// xoshiro256** (Blackman, Vigna 2018).
struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        for (auto& v : s) v = splitmix64(seed);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t result = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    // Uniform in [0, 1) with 53 random bits.
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    uint64_t below(uint64_t bound) { return (uint64_t)(uniform() * (double)bound); }
};

}  // namespace

extern "C" omm_status omm_water_box(int32_t num_molecules, double box_edge, uint64_t seed, double* positions) {
    return ommhip::guarded([&]() -> omm_status {
        if (num_molecules <= 0 || !(box_edge > 0.0) || !std::isfinite(box_edge) || !positions)
            return ommhip::fail(OMM_ERR_INVALID, "omm_water_box: need num_molecules > 0, finite box_edge > 0 and an output buffer");
        Xoshiro rng(seed);
        int64_t m = (int64_t)std::ceil(std::cbrt((double)num_molecules));
        while (m * m * m < num_molecules) ++m;
        while (m > 1 && (m - 1) * (m - 1) * (m - 1) >= num_molecules) --m;
        const int64_t sites = m * m * m;
        // Seeded choice of num_molecules distinct sites (partial Fisher-Yates),
        // then kept in lattice order.
        std::vector<int64_t> order((size_t)sites);
        for (int64_t i = 0; i < sites; ++i) order[(size_t)i] = i;
        for (int64_t i = 0; i < num_molecules; ++i) {
            int64_t j = i + (int64_t)rng.below((uint64_t)(sites - i));
            std::swap(order[(size_t)i], order[(size_t)j]);
        }
        std::vector<char> used((size_t)sites, 0);
        for (int64_t i = 0; i < num_molecules; ++i) used[(size_t)order[(size_t)i]] = 1;
        const double spacing = box_edge / (double)m;
        const double r_oh = 0.09572;
        const double half_angle = 0.5 * 104.52 * kPi / 180.0;
        const double hx = r_oh * std::sin(half_angle), hy = r_oh * std::cos(half_angle);
        int64_t mol = 0;
        for (int64_t site = 0; site < sites && mol < num_molecules; ++site) {
            if (!used[(size_t)site]) continue;
            int64_t ix = site / (m * m), iy = (site / m) % m, iz = site % m;
            double o[3] = {((double)ix + 0.5) * spacing, ((double)iy + 0.5) * spacing, ((double)iz + 0.5) * spacing};
            for (double& v : o) v += 0.04 * (rng.uniform() - 0.5);
            // Shoemake uniform random rotation quaternion.
            double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
            double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
            double qx = a * std::sin(2.0 * kPi * u2), qy = a * std::cos(2.0 * kPi * u2);
            double qz = b * std::sin(2.0 * kPi * u3), qw = b * std::cos(2.0 * kPi * u3);
            double R[3][3] = {
                {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
                {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
                {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
            double local[2][3] = {{hx, hy, 0.0}, {-hx, hy, 0.0}};
            double* out = positions + 9 * mol;
            for (int d = 0; d < 3; ++d) out[d] = o[d];
            for (int h = 0; h < 2; ++h)
                for (int d = 0; d < 3; ++d)
                    out[3 * (h + 1) + d] = o[d] + R[d][0] * local[h][0] + R[d][1] * local[h][1] + R[d][2] * local[h][2];
            ++mol;
        }
        return OMM_OK;
    });
}
