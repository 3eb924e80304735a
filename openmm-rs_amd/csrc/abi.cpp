// extern "C" entry points of include/ommhip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <memory>
#include <mutex>

#include "../../include/ommhip.h"
#include "md.h"

namespace ommhip {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& message) { g_last_error = message; }
const char* last_error_cstr() { return g_last_error.c_str(); }

}  // namespace ommhip

using ommhip::Error;
using ommhip::guarded;

struct omm_nb {
    ommhip::NbOperator op;
    double parts[5] = {0, 0, 0, 0, 0};
    omm_nb(const omm_nb_desc* d, int precision) : op(&d, 1, precision) {}
};

struct omm_ctx {
    ommhip::MdContext ctx;
    omm_ctx(const omm_md_desc* d, int precision) : ctx(d, 1, precision) {}
};

// The charge mesh of PmeCache (pme.rs:8-138): plain-FP64 PME operators keyed
// by particle count and grid, most recently used first.  Two are kept: the
// constant-potential solver alternates between its system and a single unit
// charge (its preconditioner, constant_potential/solver.rs:219-226).
struct omm_pme {
    struct Mesh {
        int n;
        std::array<int, 3> grid;
        std::unique_ptr<ommhip::NbOperator> op;
    };
    std::vector<Mesh> meshes;
    std::mutex mutex;

    ommhip::NbOperator& mesh(int n, const int32_t* grid, double alpha, const double* box) {
        if (n <= 0) throw Error(OMM_ERR_INVALID, "the charge mesh needs particles");
        if (!grid || !box) throw Error(OMM_ERR_INVALID, "null grid or box");
        const std::array<int, 3> g = {grid[0], grid[1], grid[2]};
        for (size_t k = 0; k < meshes.size(); ++k)
            if (meshes[k].n == n && meshes[k].grid == g) {
                std::rotate(meshes.begin(), meshes.begin() + k, meshes.begin() + k + 1);
                return *meshes[0].op;
            }
        // A reciprocal-only PME operator with the explicit (alpha, grid); its
        // particles are placeholders, every call brings the charges.  The
        // cutoff only has to fit the box.
        std::vector<double> zero(n, 0.0), one(n, 1.0);
        omm_nb_desc d;
        omm_nb_desc_defaults(&d);
        d.num_particles = n;
        d.charge = zero.data();
        d.sigma = one.data();
        d.epsilon = zero.data();
        d.method = OMM_NB_PME;
        d.cutoff = std::min(0.3, 0.5 * std::min(box[0], std::min(box[4], box[8])));
        d.pme_alpha = alpha;
        for (int a = 0; a < 3; ++a) d.pme_grid[a] = grid[a];
        d.use_dispersion_correction = 0;
        d.include_direct_space = 0;
        const omm_nb_desc* dp = &d;
        std::unique_ptr<ommhip::NbOperator> op(new ommhip::NbOperator(&dp, 1, OMM_PRECISION_DOUBLE));
        op->initialize(n, box);
        meshes.insert(meshes.begin(), Mesh{n, g, std::move(op)});
        if (meshes.size() > 2) meshes.pop_back();
        return *meshes[0].op;
    }
};

struct omm_batch {
    ommhip::MdContext ctx;
    omm_batch(const omm_md_desc* d, int count, int precision) : ctx(d, count, precision) {}
};

extern "C" {

const char* omm_last_error(void) { return ommhip::last_error_cstr(); }

int32_t omm_abi_version(void) { return OMMHIP_ABI_VERSION; }

int32_t omm_device_count(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) return 0;
    int usable = 0;
    for (int d = 0; d < count; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) == hipSuccess && std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
            ++usable;
    }
    return usable;
}

uint64_t omm_pair_key(int32_t i, int32_t j) { return ommhip::pair_key(i, j); }

void omm_nb_desc_defaults(omm_nb_desc* d) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->method = OMM_NB_NO_CUTOFF;
    d->cutoff = 1.0;
    d->reaction_field_dielectric = 78.3;
    d->ewald_error_tolerance = 5e-4;
    d->pme_alpha = 0.0;
    d->use_dispersion_correction = 1;
    d->exceptions_use_pbc = 0;
    d->use_switching_function = 0;
    d->switching_distance = -1.0;
    d->force_group = 0;
    d->reciprocal_force_group = -1;
    d->include_direct_space = 1;
}

omm_status omm_nb_create(const omm_nb_desc* desc, int32_t precision, omm_nb** out) {
    return guarded([&]() -> omm_status {
        if (!out) throw Error(OMM_ERR_INVALID, "null output handle");
        *out = nullptr;
        *out = new omm_nb(desc, precision);
        return OMM_OK;
    });
}

omm_status omm_nb_initialize(omm_nb* nb, int32_t num_particles, const double* box) {
    return guarded([&]() -> omm_status {
        if (!nb) throw Error(OMM_ERR_INVALID, "null handle");
        std::lock_guard<std::mutex> lock(nb->op.mutex);
        nb->op.initialize(num_particles, box);
        return OMM_OK;
    });
}

omm_status omm_nb_get_pme_parameters(const omm_nb* nb, double* alpha, int32_t grid[3]) {
    return guarded([&]() -> omm_status {
        if (!nb) throw Error(OMM_ERR_INVALID, "null handle");
        if (!nb->op.initialized()) throw Error(OMM_ERR_STATE, "NonbondedForce not initialized");
        if (alpha) *alpha = nb->op.alpha();
        if (grid)
            for (int a = 0; a < 3; ++a) grid[a] = nb->op.grid()[a];
        return OMM_OK;
    });
}

omm_status omm_nb_get_ljpme_parameters(const omm_nb* nb, double* alpha, int32_t grid[3]) {
    return guarded([&]() -> omm_status {
        if (!nb) throw Error(OMM_ERR_INVALID, "null handle");
        if (!nb->op.initialized()) throw Error(OMM_ERR_STATE, "NonbondedForce not initialized");
        if (alpha) *alpha = nb->op.lj_alpha();
        if (grid)
            for (int a = 0; a < 3; ++a) grid[a] = nb->op.lj_grid()[a];
        return OMM_OK;
    });
}

omm_status omm_nb_calc_forces_and_energy(omm_nb* nb, const double* pos, const double* box, double* forces_out,
                                         double* energy_out) {
    return guarded([&]() -> omm_status {
        if (!nb) throw Error(OMM_ERR_INVALID, "null handle");
        std::lock_guard<std::mutex> lock(nb->op.mutex);
        nb->op.compute(pos, box, nb->op.settings().include_direct, true, forces_out, energy_out, nb->parts, nullptr,
                       nullptr);
        return OMM_OK;
    });
}

omm_status omm_nb_calc_forces_and_energy_groups(omm_nb* nb, const double* pos, const double* box, uint32_t groups,
                                                double* forces_out, double* energy_out, int32_t* evaluated) {
    return guarded([&]() -> omm_status {
        if (!nb) throw Error(OMM_ERR_INVALID, "null handle");
        std::lock_guard<std::mutex> lock(nb->op.mutex);
        const ommhip::NbSettings& c = nb->op.settings();
        // nonbonded.rs:1139-1149
        bool direct = c.include_direct && (groups & (1u << c.force_group)) != 0;
        uint32_t rg = c.reciprocal_group < 0 ? (uint32_t)c.force_group : (uint32_t)c.reciprocal_group;
        bool recip = (c.method == OMM_NB_EWALD || c.method == OMM_NB_PME || c.method == OMM_NB_LJPME) &&
                     (groups & (1u << rg)) != 0;
        if (evaluated) *evaluated = (direct || recip) ? 1 : 0;
        if (!(direct || recip)) return OMM_OK;
        nb->op.compute(pos, box, direct, recip, forces_out, energy_out, nb->parts, nullptr, nullptr);
        return OMM_OK;
    });
}

omm_status omm_nb_calc_forces_and_energy_device(omm_nb* nb, const double* d_pos, const double* box, uint32_t groups,
                                                double* d_forces, double* energy_out, int32_t* evaluated,
                                                void* stream) {
    return guarded([&]() -> omm_status {
        if (!nb) throw Error(OMM_ERR_INVALID, "null handle");
        if (!d_pos) throw Error(OMM_ERR_INVALID, "positions must not be null");
        std::lock_guard<std::mutex> lock(nb->op.mutex);
        const ommhip::NbSettings& c = nb->op.settings();
        // nonbonded.rs:1139-1149, as omm_nb_calc_forces_and_energy_groups.
        bool direct = c.include_direct && (groups & (1u << c.force_group)) != 0;
        uint32_t rg = c.reciprocal_group < 0 ? (uint32_t)c.force_group : (uint32_t)c.reciprocal_group;
        bool recip = (c.method == OMM_NB_EWALD || c.method == OMM_NB_PME || c.method == OMM_NB_LJPME) &&
                     (groups & (1u << rg)) != 0;
        if (evaluated) *evaluated = (direct || recip) ? 1 : 0;
        if (!(direct || recip)) return OMM_OK;
        nb->op.compute(nullptr, box, direct, recip, nullptr, energy_out, nb->parts, nullptr, nullptr, d_pos, d_forces,
                       static_cast<hipStream_t>(stream));
        return OMM_OK;
    });
}

omm_status omm_nb_energy_parts(omm_nb* nb, const double* pos, const double* box, double parts[5]) {
    return guarded([&]() -> omm_status {
        if (!nb || !parts) throw Error(OMM_ERR_INVALID, "null argument");
        std::lock_guard<std::mutex> lock(nb->op.mutex);
        nb->op.compute(pos, box, true, true, nullptr, nullptr, parts, nullptr, nullptr);
        return OMM_OK;
    });
}

omm_status omm_nb_pair_stats(omm_nb* nb, const double* pos, const double* box, int64_t* num_pairs,
                             uint64_t* pair_hash) {
    return guarded([&]() -> omm_status {
        if (!nb) throw Error(OMM_ERR_INVALID, "null handle");
        std::lock_guard<std::mutex> lock(nb->op.mutex);
        int64_t np = 0;
        uint64_t h = 0;
        nb->op.compute(pos, box, true, false, nullptr, nullptr, nullptr, &np, &h);
        if (num_pairs) *num_pairs = np;
        if (pair_hash) *pair_hash = h;
        return OMM_OK;
    });
}

omm_status omm_nb_set_global_parameter(omm_nb* nb, const char* name, double value, int32_t* recognised) {
    return guarded([&]() -> omm_status {
        if (!nb || !name) throw Error(OMM_ERR_INVALID, "null argument");
        std::lock_guard<std::mutex> lock(nb->op.mutex);
        const bool found = nb->op.set_global_parameter(name, value);
        if (recognised) *recognised = found ? 1 : 0;
        return OMM_OK;
    });
}

omm_status omm_nb_get_global_parameter(omm_nb* nb, const char* name, double* value, int32_t* recognised) {
    return guarded([&]() -> omm_status {
        if (!nb || !name) throw Error(OMM_ERR_INVALID, "null argument");
        std::lock_guard<std::mutex> lock(nb->op.mutex);
        const bool found = nb->op.get_global_parameter(name, value);
        if (recognised) *recognised = found ? 1 : 0;
        return OMM_OK;
    });
}

omm_status omm_nb_update_parameters(omm_nb* nb, const omm_nb_desc* desc) {
    return guarded([&]() -> omm_status {
        if (!nb) throw Error(OMM_ERR_INVALID, "null handle");
        std::lock_guard<std::mutex> lock(nb->op.mutex);
        nb->op.update_parameters(desc);
        return OMM_OK;
    });
}

omm_status omm_nb_destroy(omm_nb* nb) {
    return guarded([&]() -> omm_status {
        delete nb;
        return OMM_OK;
    });
}

omm_status omm_ctx_create(const omm_md_desc* desc, int32_t precision, omm_ctx** out) {
    return guarded([&]() -> omm_status {
        if (!out) throw Error(OMM_ERR_INVALID, "null output handle");
        *out = nullptr;
        *out = new omm_ctx(desc, precision);
        return OMM_OK;
    });
}

#define CTX_CALL(body)                                                      \
    return guarded([&]() -> omm_status {                                    \
        if (!ctx) throw Error(OMM_ERR_INVALID, "null context");             \
        std::lock_guard<std::mutex> lock(ctx->ctx.mutex);                   \
        body;                                                               \
        return OMM_OK;                                                      \
    })

omm_status omm_ctx_set_positions(omm_ctx* ctx, const double* pos) {
    CTX_CALL(if (!pos) throw Error(OMM_ERR_INVALID, "null positions"); ctx->ctx.set_positions(0, pos));
}

omm_status omm_ctx_set_velocities(omm_ctx* ctx, const double* vel) {
    CTX_CALL(if (!vel) throw Error(OMM_ERR_INVALID, "null velocities"); ctx->ctx.set_velocities(0, vel));
}

omm_status omm_ctx_apply_constraints(omm_ctx* ctx, double tolerance) {
    CTX_CALL(ctx->ctx.project_constraints(tolerance, false));
}

omm_status omm_ctx_apply_velocity_constraints(omm_ctx* ctx, double tolerance) {
    CTX_CALL(ctx->ctx.project_constraints(tolerance, true));
}

omm_status omm_ctx_set_velocities_to_temperature(omm_ctx* ctx, double temperature, uint64_t seed) {
    CTX_CALL(ctx->ctx.set_velocities_to_temperature(temperature, seed));
}

omm_status omm_ctx_step(omm_ctx* ctx, int32_t nsteps) { CTX_CALL(ctx->ctx.step(nsteps)); }

omm_status omm_ctx_set_integrator(omm_ctx* ctx, double temperature, double friction, double step_size) {
    CTX_CALL(ctx->ctx.set_integrator(temperature, friction, step_size));
}

omm_status omm_ctx_set_graph_steps(omm_ctx* ctx, int32_t steps) { CTX_CALL(ctx->ctx.set_graph_steps(steps)); }

omm_status omm_ctx_set_parameter(omm_ctx* ctx, const char* name, double value, int32_t* recognised) {
    CTX_CALL(if (!name) throw Error(OMM_ERR_INVALID, "null parameter name");
             const bool found = ctx->ctx.set_parameter(name, value); if (recognised) *recognised = found ? 1 : 0);
}

static void fill_info(const ommhip::BaroState& st, omm_barostat_info* info) {
    info->volume_scale = st.volume_scale;
    info->attempted = st.attempted;
    info->accepted = st.accepted;
    info->total_attempted = st.total_attempted;
    info->total_accepted = st.total_accepted;
    info->volume_old = st.v_old;
    info->volume_new = st.v_new;
    info->energy_old = st.e_old;
    info->energy_new = st.e_new;
    info->work = st.work;
    info->accepted_last = st.accepted_last;
    info->skipped = st.skipped;
    info->axis = st.axis;
    for (int a = 0; a < 3; ++a) info->scale[a] = st.scale[a];
    info->area_change = st.area_change;
    info->log_shape = st.log_shape;
    for (int k = 0; k < 9; ++k) info->box_new[k] = st.box_new[k];
}

omm_status omm_ctx_set_barostat(omm_ctx* ctx, double pressure, double temperature, int32_t frequency,
                                uint64_t seed) {
    CTX_CALL(ctx->ctx.set_barostat(pressure, temperature, frequency, seed));
}

omm_status omm_ctx_set_anisotropic_barostat(omm_ctx* ctx, const double pressure[3], double temperature,
                                            int32_t scale_axes, int32_t frequency, uint64_t seed) {
    CTX_CALL(if (!pressure) throw Error(OMM_ERR_INVALID, "null pressure");
             ctx->ctx.set_anisotropic_barostat(pressure, temperature, scale_axes, frequency, seed));
}

omm_status omm_ctx_set_flexible_barostat(omm_ctx* ctx, double pressure, double temperature, int32_t frequency,
                                         uint64_t seed) {
    CTX_CALL(ctx->ctx.set_flexible_barostat(pressure, temperature, frequency, seed));
}

omm_status omm_ctx_set_membrane_barostat(omm_ctx* ctx, double pressure, double surface_tension, double temperature,
                                         int32_t xy_mode, int32_t z_mode, int32_t frequency, uint64_t seed) {
    CTX_CALL(ctx->ctx.set_membrane_barostat(pressure, surface_tension, temperature, xy_mode, z_mode, frequency, seed));
}

omm_status omm_ctx_barostat_info(omm_ctx* ctx, omm_barostat_info* info) {
    CTX_CALL(if (!info) throw Error(OMM_ERR_INVALID, "null output"); fill_info(ctx->ctx.barostat_state(0), info));
}

omm_status omm_ctx_barostat_attempt(omm_ctx* ctx) { CTX_CALL(ctx->ctx.barostat_attempt()); }

#define NULL_OUT(p) \
    if (!(p)) throw Error(OMM_ERR_INVALID, "null output")

omm_status omm_ctx_compute_current_pressure(omm_ctx* ctx, double* pressure) {
    CTX_CALL(NULL_OUT(pressure); ctx->ctx.current_pressure(pressure));
}

omm_status omm_ctx_compute_current_anisotropic_pressure(omm_ctx* ctx, double pressure[3]) {
    CTX_CALL(NULL_OUT(pressure); ctx->ctx.axis_pressures(1, pressure));
}

omm_status omm_ctx_compute_current_membrane_pressure(omm_ctx* ctx, double pressure[3]) {
    CTX_CALL(NULL_OUT(pressure); ctx->ctx.axis_pressures(2, pressure));
}

omm_status omm_ctx_compute_current_flexible_pressure(omm_ctx* ctx, double pressure[6]) {
    CTX_CALL(NULL_OUT(pressure); ctx->ctx.flexible_pressure(pressure));
}

omm_status omm_ctx_compute_stress_tensor(omm_ctx* ctx, int32_t include_kinetic, double stress[6]) {
    CTX_CALL(NULL_OUT(stress); ctx->ctx.stress_tensor(include_kinetic != 0, stress));
}

omm_status omm_ctx_get_box(omm_ctx* ctx, double box[9]) {
    CTX_CALL(if (!box) throw Error(OMM_ERR_INVALID, "null output"); ctx->ctx.get_box(0, box));
}

omm_status omm_ctx_list_stats(omm_ctx* ctx, int64_t* tiles, int64_t* entries) {
    CTX_CALL(ctx->ctx.list_stats(tiles, entries));
}

namespace {
void fill_tuning(const ommhip::NbOperator::Tuning& t, omm_launch_tuning* out) {
    if (!out) throw ommhip::Error(OMM_ERR_INVALID, "null output");
    out->spread_tile_edge = t.spread_edge;
    out->fft_lines_per_workgroup = t.fft_cap;
    out->pair_grid = t.pair_grid;
    out->spread_ms = t.spread_ms;
    out->fft_ms = t.fft_ms;
    out->pair_ms = t.pair_ms;
    out->fft_plane = t.fft_plane;
    out->pair_scale = t.pair_scale;
    out->from_cache = t.from_cache;
}

void read_tuning(const omm_launch_tuning* in, ommhip::NbOperator::Tuning* t) {
    if (!in) throw ommhip::Error(OMM_ERR_INVALID, "null tuning");
    t->spread_edge = in->spread_tile_edge;
    t->fft_cap = in->fft_lines_per_workgroup;
    t->fft_plane = in->fft_plane;
    t->pair_scale = in->pair_scale;
}
}  // namespace

omm_status omm_ctx_set_launch_tuning(omm_ctx* ctx, const omm_launch_tuning* tuning) {
    CTX_CALL(ommhip::NbOperator::Tuning t; read_tuning(tuning, &t); ctx->ctx.set_tuning(t));
}

omm_status omm_set_tuning_cache(const char* path) {
    ommhip::set_tuning_cache_path(path);
    return OMM_OK;
}

omm_status omm_ctx_launch_tuning(omm_ctx* ctx, omm_launch_tuning* out) {
    CTX_CALL(fill_tuning(ctx->ctx.tuning(), out));
}

omm_status omm_ctx_get_state(omm_ctx* ctx, double* pos, double* vel, double* forces, double* potential_energy,
                             int64_t* step_count) {
    CTX_CALL(ctx->ctx.get_state(0, pos, vel, forces, potential_energy, step_count));
}

omm_status omm_ctx_set_state(omm_ctx* ctx, int64_t step_count, const double* box, const double* pos,
                             const double* vel, const double* list_positions) {
    CTX_CALL(ctx->ctx.set_state(0, step_count, box, pos, vel, list_positions));
}

omm_status omm_ctx_set_capacity(omm_ctx* ctx, int32_t list_cap, int32_t bin_cap) {
    CTX_CALL(ctx->ctx.set_capacity(list_cap, bin_cap));
}

omm_status omm_ctx_capacity(omm_ctx* ctx, int32_t* list_cap, int32_t* bin_cap, int64_t* overflow_replays) {
    CTX_CALL(if (!list_cap || !bin_cap || !overflow_replays) throw Error(OMM_ERR_INVALID, "null output");
             ctx->ctx.capacity(list_cap, bin_cap, overflow_replays));
}

omm_status omm_ctx_get_list_positions(omm_ctx* ctx, double* list_positions) {
    CTX_CALL(if (!list_positions) throw Error(OMM_ERR_INVALID, "null output");
             ctx->ctx.list_positions(0, list_positions));
}

omm_status omm_ctx_kinetic_energy(omm_ctx* ctx, double* kinetic_energy) {
    CTX_CALL(if (!kinetic_energy) throw Error(OMM_ERR_INVALID, "null output"); *kinetic_energy = ctx->ctx.kinetic_energy(0));
}

omm_status omm_ctx_profile(omm_ctx* ctx, int32_t reps, double phase_ms[8], int64_t* num_pairs) {
    CTX_CALL(if (!phase_ms) throw Error(OMM_ERR_INVALID, "null output"); ctx->ctx.profile(reps, phase_ms, num_pairs));
}

omm_status omm_ctx_destroy(omm_ctx* ctx) {
    return guarded([&]() -> omm_status {
        delete ctx;
        return OMM_OK;
    });
}

omm_status omm_pme_create(omm_pme** out) {
    return guarded([&]() -> omm_status {
        if (!out) throw Error(OMM_ERR_INVALID, "null output handle");
        *out = new omm_pme();
        return OMM_OK;
    });
}

omm_status omm_pme_evaluate_charges(omm_pme* pme, int32_t num_particles, const double* pos, const double* charges,
                                    const double* box, double alpha, const int32_t grid[3], double* forces,
                                    double* energy) {
    return guarded([&]() -> omm_status {
        if (!pme) throw Error(OMM_ERR_INVALID, "null handle");
        if (!energy) throw Error(OMM_ERR_INVALID, "null energy output");
        std::lock_guard<std::mutex> lock(pme->mutex);
        ommhip::NbOperator& op = pme->mesh(num_particles, grid, alpha, box);
        *energy = op.mesh_charges(pos, box, alpha, charges, forces, nullptr, 0, nullptr);
        return OMM_OK;
    });
}

omm_status omm_pme_charge_derivatives(omm_pme* pme, int32_t num_particles, const double* pos, const double* charges,
                                      const double* box, double alpha, const int32_t grid[3],
                                      const int32_t* indices, int32_t count, double* derivatives) {
    return guarded([&]() -> omm_status {
        if (!pme) throw Error(OMM_ERR_INVALID, "null handle");
        if (count < 0) throw Error(OMM_ERR_INVALID, "negative index count");
        if (count == 0) return OMM_OK;
        std::lock_guard<std::mutex> lock(pme->mutex);
        ommhip::NbOperator& op = pme->mesh(num_particles, grid, alpha, box);
        op.mesh_charges(pos, box, alpha, charges, nullptr, indices, count, derivatives);
        return OMM_OK;
    });
}

omm_status omm_pme_destroy(omm_pme* pme) {
    return guarded([&]() -> omm_status {
        delete pme;
        return OMM_OK;
    });
}

omm_status omm_batch_create(const omm_md_desc* replicas, int32_t count, int32_t precision, omm_batch** out) {
    return guarded([&]() -> omm_status {
        if (!out) throw Error(OMM_ERR_INVALID, "null output handle");
        *out = nullptr;
        if (!replicas || count <= 0) throw Error(OMM_ERR_INVALID, "a batch needs at least one replica");
        *out = new omm_batch(replicas, count, precision);
        return OMM_OK;
    });
}

#define BATCH_CALL(body)                                                    \
    return guarded([&]() -> omm_status {                                    \
        if (!b) throw Error(OMM_ERR_INVALID, "null batch");                 \
        std::lock_guard<std::mutex> lock(b->ctx.mutex);                     \
        body;                                                               \
        return OMM_OK;                                                      \
    })

omm_status omm_batch_size(omm_batch* b, int32_t* count) {
    BATCH_CALL(if (!count) throw Error(OMM_ERR_INVALID, "null output"); *count = b->ctx.nsys());
}

omm_status omm_batch_set_positions(omm_batch* b, int32_t replica, const double* pos) {
    BATCH_CALL(if (!pos) throw Error(OMM_ERR_INVALID, "null positions"); b->ctx.set_positions(replica, pos));
}

omm_status omm_batch_set_velocities(omm_batch* b, int32_t replica, const double* vel) {
    BATCH_CALL(if (!vel) throw Error(OMM_ERR_INVALID, "null velocities"); b->ctx.set_velocities(replica, vel));
}

omm_status omm_batch_set_velocities_to_temperature(omm_batch* b, double temperature, uint64_t seed) {
    BATCH_CALL(b->ctx.set_velocities_to_temperature(temperature, seed));
}

omm_status omm_batch_set_integrator(omm_batch* b, double temperature, double friction, double step_size) {
    BATCH_CALL(b->ctx.set_integrator(temperature, friction, step_size));
}

omm_status omm_batch_step(omm_batch* b, int32_t nsteps) { BATCH_CALL(b->ctx.step(nsteps)); }

omm_status omm_batch_set_graph_steps(omm_batch* b, int32_t steps) { BATCH_CALL(b->ctx.set_graph_steps(steps)); }

omm_status omm_batch_set_barostat(omm_batch* b, double pressure, double temperature, int32_t frequency,
                                  uint64_t seed) {
    BATCH_CALL(b->ctx.set_barostat(pressure, temperature, frequency, seed));
}

omm_status omm_batch_set_anisotropic_barostat(omm_batch* b, const double pressure[3], double temperature,
                                              int32_t scale_axes, int32_t frequency, uint64_t seed) {
    BATCH_CALL(if (!pressure) throw Error(OMM_ERR_INVALID, "null pressure");
               b->ctx.set_anisotropic_barostat(pressure, temperature, scale_axes, frequency, seed));
}

omm_status omm_batch_set_flexible_barostat(omm_batch* b, double pressure, double temperature, int32_t frequency,
                                           uint64_t seed) {
    BATCH_CALL(b->ctx.set_flexible_barostat(pressure, temperature, frequency, seed));
}

omm_status omm_batch_set_membrane_barostat(omm_batch* b, double pressure, double surface_tension,
                                           double temperature, int32_t xy_mode, int32_t z_mode, int32_t frequency,
                                           uint64_t seed) {
    BATCH_CALL(b->ctx.set_membrane_barostat(pressure, surface_tension, temperature, xy_mode, z_mode, frequency, seed));
}

omm_status omm_batch_barostat_info(omm_batch* b, int32_t replica, omm_barostat_info* info) {
    BATCH_CALL(if (!info) throw Error(OMM_ERR_INVALID, "null output");
               fill_info(b->ctx.barostat_state(replica), info));
}

omm_status omm_batch_barostat_attempt(omm_batch* b) { BATCH_CALL(b->ctx.barostat_attempt()); }

omm_status omm_batch_compute_current_pressure(omm_batch* b, double* pressure) {
    BATCH_CALL(NULL_OUT(pressure); b->ctx.current_pressure(pressure));
}

omm_status omm_batch_compute_current_anisotropic_pressure(omm_batch* b, double* pressure) {
    BATCH_CALL(NULL_OUT(pressure); b->ctx.axis_pressures(1, pressure));
}

omm_status omm_batch_compute_current_membrane_pressure(omm_batch* b, double* pressure) {
    BATCH_CALL(NULL_OUT(pressure); b->ctx.axis_pressures(2, pressure));
}

omm_status omm_batch_compute_current_flexible_pressure(omm_batch* b, double* pressure) {
    BATCH_CALL(NULL_OUT(pressure); b->ctx.flexible_pressure(pressure));
}

omm_status omm_batch_compute_stress_tensor(omm_batch* b, int32_t include_kinetic, double* stress) {
    BATCH_CALL(NULL_OUT(stress); b->ctx.stress_tensor(include_kinetic != 0, stress));
}

omm_status omm_batch_get_state(omm_batch* b, int32_t replica, double* pos, double* vel, double* forces,
                               double* potential_energy, int64_t* step_count) {
    BATCH_CALL(b->ctx.get_state(replica, pos, vel, forces, potential_energy, step_count));
}

omm_status omm_batch_set_state(omm_batch* b, int32_t replica, int64_t step_count, const double* box,
                               const double* pos, const double* vel, const double* list_positions) {
    BATCH_CALL(b->ctx.set_state(replica, step_count, box, pos, vel, list_positions));
}

omm_status omm_batch_get_list_positions(omm_batch* b, int32_t replica, double* list_positions) {
    BATCH_CALL(if (!list_positions) throw Error(OMM_ERR_INVALID, "null output");
               b->ctx.list_positions(replica, list_positions));
}

omm_status omm_batch_energies(omm_batch* b, double* potential_energies) {
    BATCH_CALL(if (!potential_energies) throw Error(OMM_ERR_INVALID, "null output");
               b->ctx.energies(potential_energies));
}

omm_status omm_batch_kinetic_energy(omm_batch* b, int32_t replica, double* kinetic_energy) {
    BATCH_CALL(if (!kinetic_energy) throw Error(OMM_ERR_INVALID, "null output");
               *kinetic_energy = b->ctx.kinetic_energy(replica));
}

omm_status omm_batch_get_box(omm_batch* b, int32_t replica, double box[9]) {
    BATCH_CALL(if (!box) throw Error(OMM_ERR_INVALID, "null output"); b->ctx.get_box(replica, box));
}

omm_status omm_batch_list_stats(omm_batch* b, int64_t* tiles, int64_t* entries) {
    BATCH_CALL(b->ctx.list_stats(tiles, entries));
}

omm_status omm_batch_launch_tuning(omm_batch* b, omm_launch_tuning* out) {
    BATCH_CALL(fill_tuning(b->ctx.tuning(), out));
}

omm_status omm_batch_set_launch_tuning(omm_batch* b, const omm_launch_tuning* tuning) {
    BATCH_CALL(ommhip::NbOperator::Tuning t; read_tuning(tuning, &t); b->ctx.set_tuning(t));
}

omm_status omm_batch_profile(omm_batch* b, int32_t reps, double phase_ms[8], int64_t* num_pairs) {
    BATCH_CALL(if (!phase_ms) throw Error(OMM_ERR_INVALID, "null output"); b->ctx.profile(reps, phase_ms, num_pairs));
}

omm_status omm_batch_destroy(omm_batch* b) {
    return guarded([&]() -> omm_status {
        delete b;
        return OMM_OK;
    });
}

}  // extern "C"
