// NonbondedForce operator: validation, parameter freezing and the evaluation
// sequence (nonbonded.rs:771-1155) on one device.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>

#include "nbforce.h"

namespace ommhip {

// Neighbour-list skin (nm): the list holds the pairs within cutoff + skin and
// is rebuilt once an atom has moved skin/2.  At 300 K water rebuilds every
// ~8 steps at 0.10 nm; 0.13 nm measured 1.3% faster steps at 23,550 atoms and
// 0.8% at 300,000 atoms and for the 64-replica batch (0.08 and 0.16 nm: no
// gain).  Results do not depend on it beyond FP32 partial-sum order.
constexpr double LIST_SKIN = 0.13;

namespace {

__global__ void k_fixed_to_double(const long long* __restrict__ fx, double* __restrict__ out, size_t n3) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n3) out[i] = (double)fx[i] * (1.0 / FORCE_SCALE);
}

void validate_particle(double q, double s, double e) {
    // nonbonded.rs:1252-1261
    if (!(std::isfinite(q) && std::isfinite(s) && std::isfinite(e) && s >= 0.0 && e >= 0.0))
        throw Error(OMM_ERR_INVALID, "nonbonded parameters must be finite with nonnegative sigma and epsilon");
}


// nonbonded.rs:1290-1336 — adaptive Simpson for the switched tail integral.
template <class F>
double simpson(const F& f, double a, double b, double fa, double fm, double fb, double est, double tol, int depth) {
    double m = 0.5 * (a + b);
    double fl = f(0.5 * (a + m)), fr = f(0.5 * (m + b));
    double left = (m - a) / 6.0 * (fa + 4.0 * fl + fm);
    double right = (b - m) / 6.0 * (fm + 4.0 * fr + fb);
    double diff = left + right - est;
    if (depth == 0 || std::fabs(diff) <= 15.0 * tol) return left + right + diff / 15.0;
    return simpson(f, a, m, fa, fl, fm, left, 0.5 * tol, depth - 1) +
           simpson(f, m, b, fm, fr, fb, right, 0.5 * tol, depth - 1);
}

}  // namespace

// system.rs:16-39
void validate_box(const double* v) {
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(v[i])) throw Error(OMM_ERR_NONFINITE, "periodic box vectors must be finite");
    if (!(v[1] == 0.0 && v[2] == 0.0)) throw Error(OMM_ERR_INVALID, "first periodic box vector must be parallel to x");
    if (!(v[5] == 0.0)) throw Error(OMM_ERR_INVALID, "second periodic box vector must be in the x-y plane");
    if (!(v[0] > 0.0 && v[4] > 0.0 && v[8] > 0.0 && v[0] >= 2.0 * std::fabs(v[3]) && v[0] >= 2.0 * std::fabs(v[6]) &&
          v[4] >= 2.0 * std::fabs(v[7])))
        throw Error(OMM_ERR_INVALID, "periodic box vectors must be in reduced form");
}

Box make_box(const double* v) {
    Box b{};
    if (!v) {
        b.periodic = 0;
        return b;
    }
    for (int i = 0; i < 9; ++i) b.v[i] = v[i];
    box_refresh(b);
    if (!(std::isfinite(b.volume) && b.volume > 0.0))
        throw Error(OMM_ERR_INVALID, "periodic box volume must be positive and finite");
    return b;
}

void launch_fixed_to_double(const long long* fx, double* out, size_t n3, hipStream_t s) {
    if (!n3) return;
    hipLaunchKernelGGL(k_fixed_to_double, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, fx, out, n3);
    OMM_HIP(hipGetLastError());
}

NbSettings settings_from_desc(const omm_nb_desc* d) {
    if (!d) throw Error(OMM_ERR_INVALID, "null NonbondedForce descriptor");
    NbSettings c;
    if (d->num_particles < 0 || d->num_exceptions < 0) throw Error(OMM_ERR_INVALID, "negative particle/exception count");
    if (d->num_particles > 0 && (!d->charge || !d->sigma || !d->epsilon))
        throw Error(OMM_ERR_INVALID, "particle arrays must not be null");
    if (d->num_exceptions > 0 && (!d->exception_particle1 || !d->exception_particle2 || !d->exception_charge_prod ||
                                  !d->exception_sigma || !d->exception_epsilon))
        throw Error(OMM_ERR_INVALID, "exception arrays must not be null");
    for (int i = 0; i < d->num_particles; ++i) {
        validate_particle(d->charge[i], d->sigma[i], d->epsilon[i]);
        c.charge.push_back(d->charge[i]);
        c.sigma.push_back(d->sigma[i]);
        c.epsilon.push_back(d->epsilon[i]);
    }
    for (int i = 0; i < d->num_exceptions; ++i) {
        validate_particle(d->exception_charge_prod[i], d->exception_sigma[i], d->exception_epsilon[i]);
        c.exc1.push_back(d->exception_particle1[i]);
        c.exc2.push_back(d->exception_particle2[i]);
        c.exc_qq.push_back(d->exception_charge_prod[i]);
        c.exc_sig.push_back(d->exception_sigma[i]);
        c.exc_eps.push_back(d->exception_epsilon[i]);
    }
    if (d->method < OMM_NB_NO_CUTOFF || d->method > OMM_NB_LJPME) throw Error(OMM_ERR_INVALID, "unknown nonbonded method");
    c.method = d->method;
    if (!(std::isfinite(d->cutoff) && d->cutoff > 0.0)) throw Error(OMM_ERR_INVALID, "cutoff must be positive and finite");
    c.cutoff = d->cutoff;
    if (std::isnan(d->reaction_field_dielectric) || d->reaction_field_dielectric < 1.0)
        throw Error(OMM_ERR_INVALID, "reaction field dielectric must be at least one");
    c.rf_dielectric = d->reaction_field_dielectric;
    if (!(std::isfinite(d->ewald_error_tolerance) && d->ewald_error_tolerance > 0.0 && d->ewald_error_tolerance < 0.5))
        throw Error(OMM_ERR_INVALID, "Ewald tolerance must be between zero and one half");
    c.tolerance = d->ewald_error_tolerance;
    if (!(std::isfinite(d->pme_alpha) && d->pme_alpha >= 0.0))
        throw Error(OMM_ERR_INVALID, "PME alpha must be nonnegative and finite");
    if (d->pme_alpha > 0.0)
        for (int a = 0; a < 3; ++a)
            if (d->pme_grid[a] < 5) throw Error(OMM_ERR_INVALID, "PME grid dimensions must be at least five");
    c.pme_alpha = d->pme_alpha;
    for (int a = 0; a < 3; ++a) c.pme_grid[a] = d->pme_grid[a];
    // nonbonded.rs:615-630
    if (!(std::isfinite(d->ljpme_alpha) && d->ljpme_alpha >= 0.0))
        throw Error(OMM_ERR_INVALID, "LJPME alpha must be nonnegative and finite");
    if (d->ljpme_alpha > 0.0)
        for (int a = 0; a < 3; ++a)
            if (d->ljpme_grid[a] < 5) throw Error(OMM_ERR_INVALID, "LJPME grid dimensions must be at least five");
    c.ljpme_alpha = d->ljpme_alpha;
    for (int a = 0; a < 3; ++a) c.ljpme_grid[a] = d->ljpme_alpha > 0.0 ? d->ljpme_grid[a] : 0;
    c.dispersion_correction = d->use_dispersion_correction != 0;
    c.exceptions_pbc = d->exceptions_use_pbc != 0;
    c.switching = d->use_switching_function != 0;
    c.switching_distance = d->switching_distance;
    if (c.switching && !(std::isfinite(c.switching_distance) && c.switching_distance >= 0.0))
        throw Error(OMM_ERR_INVALID, "switching distance must be nonnegative and finite");
    if (d->force_group < 0 || d->force_group > 31) throw Error(OMM_ERR_INVALID, "force group must be between zero and 31");
    if (d->reciprocal_force_group < -1 || d->reciprocal_force_group > 31)
        throw Error(OMM_ERR_INVALID, "reciprocal force group must be between -1 and 31");
    c.force_group = d->force_group;
    c.reciprocal_group = d->reciprocal_force_group;
    c.include_direct = d->include_direct_space != 0;
    // Global parameters and offsets (nonbonded.rs:247-303).
    if (d->num_global_parameters < 0 || d->num_particle_offsets < 0 || d->num_exception_offsets < 0)
        throw Error(OMM_ERR_INVALID, "negative parameter or offset count");
    if (d->num_global_parameters > 0 && (!d->global_parameter_names || !d->global_parameter_defaults))
        throw Error(OMM_ERR_INVALID, "global parameter arrays must not be null");
    for (int k = 0; k < d->num_global_parameters; ++k) {
        const char* name = d->global_parameter_names[k];
        const double v = d->global_parameter_defaults[k];
        if (!name || !*name || !std::isfinite(v))
            throw Error(OMM_ERR_INVALID, "global parameters require a name and a finite default value");
        for (const std::string& other : c.gp_names)
            if (other == name) throw Error(OMM_ERR_INVALID, "duplicate global parameter name");
        c.gp_names.emplace_back(name);
        c.gp_defaults.push_back(v);
        c.gp_values.push_back(v);
    }
    auto offsets = [&](int count, const int32_t* param, const int32_t* index, const double* scales, int targets,
                       std::vector<ParamOffset>& out) {
        if (count > 0 && (!param || !index || !scales)) throw Error(OMM_ERR_INVALID, "offset arrays must not be null");
        for (int k = 0; k < count; ++k) {
            ParamOffset o;
            o.parameter = param[k];
            o.index = index[k];
            if (o.parameter < 0 || o.parameter >= (int)c.gp_names.size())
                throw Error(OMM_ERR_INVALID, "unknown offset parameter");
            if (o.index < 0 || o.index >= targets) throw Error(OMM_ERR_INVALID, "offset index out of range");
            for (int a = 0; a < 3; ++a) {
                o.scale[a] = scales[3 * k + a];
                if (!std::isfinite(o.scale[a])) throw Error(OMM_ERR_INVALID, "offset scales must be finite");
            }
            out.push_back(o);
        }
    };
    offsets(d->num_particle_offsets, d->particle_offset_parameter, d->particle_offset_index,
            d->particle_offset_scales, d->num_particles, c.particle_offsets);
    offsets(d->num_exception_offsets, d->exception_offset_parameter, d->exception_offset_index,
            d->exception_offset_scales, d->num_exceptions, c.exception_offsets);
    c.base_charge = c.charge;
    c.base_sigma = c.sigma;
    c.base_epsilon = c.epsilon;
    c.base_qq = c.exc_qq;
    c.base_sig = c.exc_sig;
    c.base_eps = c.exc_eps;
    c.apply_offsets();
    return c;
}

void NbSettings::apply_offsets() {
    // nonbonded.rs:395-430: offsets applied last-first, the first seen of each
    // (parameter, target) wins, i.e. a later offset replaces an earlier one.
    charge = base_charge;
    sigma = base_sigma;
    epsilon = base_epsilon;
    exc_qq = base_qq;
    exc_sig = base_sig;
    exc_eps = base_eps;
    std::set<std::pair<int, int>> seen;
    for (size_t k = particle_offsets.size(); k-- > 0;) {
        const ParamOffset& o = particle_offsets[k];
        if (!seen.insert({o.parameter, o.index}).second) continue;
        const double v = gp_values[o.parameter];
        charge[o.index] += v * o.scale[0];
        sigma[o.index] += v * o.scale[1];
        epsilon[o.index] += v * o.scale[2];
    }
    seen.clear();
    for (size_t k = exception_offsets.size(); k-- > 0;) {
        const ParamOffset& o = exception_offsets[k];
        if (!seen.insert({o.parameter, o.index}).second) continue;
        const double v = gp_values[o.parameter];
        exc_qq[o.index] += v * o.scale[0];
        exc_sig[o.index] += v * o.scale[1];
        exc_eps[o.index] += v * o.scale[2];
    }
}

NbOperator::NbOperator(const omm_nb_desc* const* descs, int nsys, int precision) {
    if (!descs || nsys <= 0) throw Error(OMM_ERR_INVALID, "a NonbondedForce batch needs at least one descriptor");
    if (precision != OMM_PRECISION_MIXED && precision != OMM_PRECISION_DOUBLE)
        throw Error(OMM_ERR_INVALID, "precision must be OMM_PRECISION_MIXED or OMM_PRECISION_DOUBLE");
    mixed_ = precision == OMM_PRECISION_MIXED;
    for (int s = 0; s < nsys; ++s) {
        cfg_.push_back(settings_from_desc(descs[s]));
        const NbSettings& a = cfg_[0];
        const NbSettings& b = cfg_.back();
        if (a.method != b.method || a.cutoff != b.cutoff || a.rf_dielectric != b.rf_dielectric ||
            a.tolerance != b.tolerance || a.pme_alpha != b.pme_alpha || a.ljpme_alpha != b.ljpme_alpha ||
            a.dispersion_correction != b.dispersion_correction || a.exceptions_pbc != b.exceptions_pbc ||
            a.switching != b.switching || a.switching_distance != b.switching_distance ||
            a.force_group != b.force_group || a.reciprocal_group != b.reciprocal_group ||
            a.include_direct != b.include_direct)
            throw Error(OMM_ERR_INVALID,
                        "the systems of a batch must share the nonbonded method, cutoff, tolerance and options");
        n_.push_back((int)b.charge.size());
    }
    alpha_.assign(nsys, 0.0);
    grid_.assign(nsys, {0, 0, 0});
    lj_alpha_.assign(nsys, 0.0);
    lj_grid_.assign(nsys, {0, 0, 0});
    sum_q2_.assign(nsys, 0.0);
    sum_q_.assign(nsys, 0.0);
    disp_coeff_.assign(nsys, 0.0);
    disp_cached_.assign(nsys, 0);
}

NbOperator::~NbOperator() {
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    if (ev_join_) (void)hipEventDestroy(ev_join_);
    if (ev_list_) (void)hipEventDestroy(ev_list_);
    if (ev_caller_) (void)hipEventDestroy(ev_caller_);
    if (side_) (void)hipStreamDestroy(side_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void NbOperator::check_box(const double* box) const {
    if (!cfg_[0].periodic()) return;  // box vectors are ignored (nonbonded.rs:802-806)
    if (!box) throw Error(OMM_ERR_INVALID, "periodic nonbonded forces require box vectors");
    validate_box(box);
    // nonbonded.rs:806-812
    for (int a = 0; a < 3; ++a)
        if (!(2.0 * cfg_[0].cutoff <= box[a * 4] + 1e-12))
            throw Error(OMM_ERR_CUTOFF_BOX, "cutoff must not exceed half the periodic box size");
}

void NbOperator::initialize(const int* nlist, const double* const* boxes) {
    const int nsys = (int)cfg_.size();
    ntotal_ = 0;
    for (int si = 0; si < nsys; ++si) {
        NbSettings& c = cfg_[si];
        const int n = nlist[si];
        const double* box = boxes[si];
        // nonbonded.rs:1055-1107
        if (n != (int)c.charge.size())
            throw Error(OMM_ERR_INVALID, "NonbondedForce must have one parameter record per particle");
        std::set<std::pair<int, int>> pairs;
        for (size_t e = 0; e < c.exc1.size(); ++e) {
            int a = c.exc1[e], b = c.exc2[e];
            if (a < 0 || b < 0 || a >= n || b >= n)
                throw Error(OMM_ERR_INVALID, "exception particle index out of range");
            if (!pairs.insert({std::min(a, b), std::max(a, b)}).second)
                throw Error(OMM_ERR_INVALID, "duplicate exception pair");
        }
        if (c.use_cutoff() && c.switching && !(c.switching_distance >= 0.0 && c.switching_distance < c.cutoff))
            throw Error(OMM_ERR_INVALID, "switching distance must be between zero and cutoff");
        if (!c.periodic()) {
            // Non-periodic: exceptions never use the box (nonbonded.rs:1006), and
            // the virtual box is chosen from the first positions.
            if (nsys != 1)
                throw Error(OMM_ERR_INVALID, "non-periodic methods run on a one-system operator");
            c.exceptions_pbc = false;
            n_[si] = n;
            ntotal_ += n;
            continue;
        }
        if (!box) throw Error(OMM_ERR_INVALID, "periodic nonbonded forces require box vectors");
        validate_box(box);
        for (int a = 0; a < 3; ++a)
            if (!(c.cutoff <= 0.5 * box[a * 4]))
                throw Error(OMM_ERR_CUTOFF_BOX, "cutoff must not exceed half the periodic box size");
        if (c.method == OMM_NB_EWALD) {
            // nonbonded.rs:1087-1097: rectangular boxes; explicit PME
            // parameters do not apply.
            if (!(box[3] == 0.0 && box[6] == 0.0 && box[7] == 0.0))
                throw Error(OMM_ERR_INVALID, "Ewald requires a rectangular periodic box; use PME for triclinic boxes");
            if (nsys != 1) throw Error(OMM_ERR_INVALID, "Ewald runs on a one-system operator");
            const double widths[3] = {box[0], box[4], box[8]};
            ewald_parameters(widths, c.cutoff, c.tolerance, &alpha_[si], grid_[si].data());
        }
        if (c.pme_mesh()) {
            if (c.pme_alpha > 0.0) {
                alpha_[si] = c.pme_alpha;
                for (int a = 0; a < 3; ++a) grid_[si][a] = c.pme_grid[a];
            } else {
                // nonbonded.rs:647-658
                alpha_[si] = std::sqrt(-std::log(2.0 * c.tolerance)) / c.cutoff;
                for (int a = 0; a < 3; ++a) {
                    double v = std::ceil(2.0 * alpha_[si] * box[a * 4] / (3.0 * std::pow(c.tolerance, 0.2)));
                    grid_[si][a] = (int)std::max(v, 6.0);
                }
            }
        }
        if (c.method == OMM_NB_LJPME) {
            if (c.ljpme_alpha > 0.0) {
                lj_alpha_[si] = c.ljpme_alpha;
                for (int a = 0; a < 3; ++a) lj_grid_[si][a] = c.ljpme_grid[a];
            } else {
                // nonbonded.rs:635-645: the Coulomb rule with half the grid.
                lj_alpha_[si] = std::sqrt(-std::log(2.0 * c.tolerance)) / c.cutoff;
                for (int a = 0; a < 3; ++a) {
                    double v = std::ceil(lj_alpha_[si] * box[a * 4] / (3.0 * std::pow(c.tolerance, 0.2)));
                    lj_grid_[si][a] = (int)std::max(v, 6.0);
                }
            }
        }
        n_[si] = n;
        ntotal_ += n;
    }
    device_ready_ = false;
    virtual_ = !cfg_[0].periodic();
    if (nsys == 1 && n_[0] == 0) {
        initialized_ = true;
        return;
    }
    for (int si = 0; si < nsys; ++si)
        if (alpha_[si] != alpha_[0] || lj_alpha_[si] != lj_alpha_[0])
            throw Error(OMM_ERR_INVALID, "the systems of a batch must share the Ewald alpha");
    if (virtual_) {
        initialized_ = true;  // set up at the first evaluation, when positions are known
        return;
    }
    std::vector<Box> bx(nsys);
    for (int si = 0; si < nsys; ++si) bx[si] = make_box(boxes[si]);
    setup_device(bx);
}

void NbOperator::setup_device(const std::vector<Box>& boxes) {
    const int nsys = (int)cfg_.size();
    int count = 0;
    OMM_HIP(hipGetDeviceCount(&count));
    if (count <= 0) throw Error(OMM_ERR_DEVICE, "no HIP device available");
    if (!stream_) OMM_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    if (!side_) OMM_HIP(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    if (!ev_fork_) OMM_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
    if (!ev_join_) OMM_HIP(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming));
    if (!ev_list_) OMM_HIP(hipEventCreateWithFlags(&ev_list_, hipEventDisableTiming));
    if (!ev_caller_) OMM_HIP(hipEventCreateWithFlags(&ev_caller_, hipEventDisableTiming));
    hipStream_t s = stream_;
    const int N = ntotal_;
    std::vector<int> atom_sys(N);
    std::vector<int> p1, p2;
    std::vector<unsigned long long> seeds(nsys, 0ull);
    for (int si = 0; si < nsys && si < (int)sys_h_.size(); ++si) seeds[si] = sys_h_[si].seed;
    sys_h_.assign(nsys, SysInfo{});
    for (int si = 0; si < nsys; ++si) sys_h_[si].seed = seeds[si];
    std::vector<NlistSystem> nls(nsys);
    int off = 0;
    for (int si = 0; si < nsys; ++si) {
        const NbSettings& c = cfg_[si];
        for (int i = 0; i < n_[si]; ++i) atom_sys[off + i] = si;
        for (size_t e = 0; e < c.exc1.size(); ++e) {
            p1.push_back(off + c.exc1[e]);
            p2.push_back(off + c.exc2[e]);
        }
        SysInfo& S = sys_h_[si];
        S.box = boxes[si];
        S.alpha = alpha_[si];
        for (int a = 0; a < 3; ++a) S.grid[a] = grid_[si][a];
        nls[si].n = n_[si];
        nls[si].exc1 = &c.exc1;
        nls[si].exc2 = &c.exc2;
        nls[si].box = S.box;
        off += n_[si];
    }
    part_.atom_sys.upload(atom_sys, s);
    exc_.count = (int)p1.size();
    exc_.p1.upload(p1, s);
    exc_.p2.upload(p2, s);
    upload_parameters(s);

    const NbSettings& c0 = cfg_[0];
    // NoCutoff lists hold every later atom of the system.
    nlist_setup(nl_, nls, sys_h_, c0.use_cutoff() ? c0.cutoff : INFINITY, LIST_SKIN, s);
    if (c0.pme_mesh()) pme_setup(pme_, sys_h_, N, part_.atom_sys.p, mixed_, s);
    if (c0.method == OMM_NB_EWALD) ewald_setup(ewald_, N, grid_[0].data(), s);
    if (c0.method == OMM_NB_LJPME) {
        // The dispersion mesh: the same atoms and sort order, its own grid.
        sys_lj_h_ = sys_h_;
        for (int si = 0; si < nsys; ++si) {
            SysInfo& L = sys_lj_h_[si];
            L.alpha = lj_alpha_[si];
            for (int a = 0; a < 3; ++a) L.grid[a] = lj_grid_[si][a];
            L.dispersion = 1;
        }
        pme_setup(pme_lj_, sys_lj_h_, N, part_.atom_sys.p, mixed_, s);
        sys_lj_dev_.upload(sys_lj_h_, s);
    }
    sys_dev_.upload(sys_h_, s);

    pp_ = PairParams{};
    pp_.nblocks = nl_.nblocks;
    pp_.nslots = nl_.nslots;
    pp_.cap = nl_.cap;
    pp_.periodic = c0.periodic() ? 1 : 0;
    pp_.use_cutoff = c0.use_cutoff() ? 1 : 0;
    pp_.cutoff = c0.cutoff;
    pp_.alpha = alpha_[0];
    // nonbonded.rs:887-890: switching applies with a cutoff only.
    const bool switching = c0.switching && c0.use_cutoff() && c0.method != OMM_NB_LJPME;
    pp_.switching = switching ? 1 : 0;
    pp_.switch_distance = c0.switching_distance;
    pp_.inv_switch_width = switching ? 1.0 / (c0.cutoff - c0.switching_distance) : 0.0;
    pp_.ljpme = c0.method == OMM_NB_LJPME ? 1 : 0;
    pp_.lj_alpha = lj_alpha_[0];
    if (c0.screened()) {
        pp_.elec = ELEC_EWALD;
    } else if (c0.method == OMM_NB_NO_CUTOFF) {
        // Bare Coulomb (nonbonded.rs:1219-1228) is the reaction-field form
        // with k_rf = c_rf = 0.
        pp_.elec = ELEC_REACTION_FIELD;
        pp_.krf = 0.0;
        pp_.crf = 0.0;
    } else {
        pp_.elec = ELEC_REACTION_FIELD;
        double rc = c0.cutoff, eps = c0.rf_dielectric;
        if (std::isinf(eps)) {
            pp_.krf = 0.5 / (rc * rc * rc);
            pp_.crf = 1.5 / rc;
        } else {
            pp_.krf = (eps - 1.0) / (2.0 * eps + 1.0) / (rc * rc * rc);
            pp_.crf = (3.0 * eps) / (2.0 * eps + 1.0) / rc;
        }
    }
    pair_grid_ = std::max(64, std::min(8192, nl_.nblocks * 4));

    if (nsys == 1) {
        pos_dev_.alloc(3 * (size_t)N);
        if (mixed_) {
            force_fx_.alloc(3 * (size_t)N);
            energy_fx_.alloc(E_SLOTS);
        } else {
            force_d_.alloc(3 * (size_t)N);
            energy_d_.alloc(E_SLOTS);
        }
        status_.alloc(1);
        stats_.alloc(2);
        out_forces_.alloc(3 * (size_t)N);
    }
    OMM_HIP(hipStreamSynchronize(s));
    have_box_ = false;
    initialized_ = true;
    device_ready_ = true;
}

Box NbOperator::virtual_box_for(const double* pos, int n, bool* changed) {
    const NbSettings& c = cfg_[0];
    double lo[3] = {pos[0], pos[1], pos[2]}, hi[3] = {pos[0], pos[1], pos[2]};
    for (int i = 1; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], pos[3 * i + a]);
            hi[a] = std::max(hi[a], pos[3 * i + a]);
        }
    // Every true displacement d has |d_a| <= extent_a.  With a cutoff the box
    // needs L_a > 2 reach (pairs in reach keep their displacement) and
    // L_a >= extent_a + reach (an imaged far pair stays out of reach); with
    // no cutoff every pair is in the list, so L_a > 2 extent_a.  The new box
    // leaves room for the atoms to spread before it must change.
    const double reach = c.cutoff + LIST_SKIN;
    double need[3], want[3];
    for (int a = 0; a < 3; ++a) {
        const double e = hi[a] - lo[a];
        need[a] = c.use_cutoff() ? std::max(2.0 * reach, e + reach) * (1.0 + 1e-9) + 1e-9 : 2.0 * e * (1.0 + 1e-9) + 1e-9;
        want[a] = c.use_cutoff() ? 1.5 * e + 2.0 * reach + 1.0 : 3.0 * e + 1.0;
    }
    bool fits = device_ready_;
    for (int a = 0; a < 3 && fits; ++a) fits = vbox_.v[a * 4] > need[a];
    if (fits) {
        *changed = false;
        return vbox_;
    }
    double v[9] = {want[0], 0, 0, 0, want[1], 0, 0, 0, want[2]};
    vbox_ = make_box(v);
    *changed = true;
    return vbox_;
}

bool NbOperator::fit_virtual_box(const double* pos, int count) {
    if (!virtual_) return false;
    for (int i = 0; i < 3 * count; ++i)
        if (!std::isfinite(pos[i])) throw Error(OMM_ERR_NONFINITE, "particle positions must be finite");
    const NbSettings& c = cfg_[0];
    const double reach = c.cutoff + LIST_SKIN;
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = hi[a] = pos[a];
        for (int i = 1; i < count; ++i) {
            lo[a] = std::min(lo[a], pos[3 * i + a]);
            hi[a] = std::max(hi[a], pos[3 * i + a]);
        }
    }
    // A box that still fits but leaves the atoms less than 0.2 nm to spread
    // in some direction is replaced now rather than a few steps later.
    if (device_ready_)
        for (int a = 0; a < 3; ++a) {
            const double L = vbox_.v[a * 4];
            if ((c.use_cutoff() ? L - reach : 0.5 * L) - (hi[a] - lo[a]) < 0.4) vbox_.v[a * 4] = 0.0;
        }
    bool changed = false;
    const Box b = virtual_box_for(pos, count, &changed);
    const bool setup = changed || !device_ready_;
    if (setup) setup_device(std::vector<Box>{b});
    // The integrators' region: an interval of width L - reach (cutoff) or
    // L / 2 (no cutoff) per axis, centred on these positions, in which any
    // two atoms keep the true displacement and a far pair stays out of reach.
    SysInfo& S = sys_h_[0];
    for (int a = 0; a < 3; ++a) {
        const double L = S.box.v[a * 4];
        const double w = (c.use_cutoff() ? L - reach : 0.5 * L) * (1.0 - 1e-9);
        const double mid = 0.5 * (lo[a] + hi[a]);
        S.vlo[a] = mid - 0.5 * w;
        S.vhi[a] = mid + 0.5 * w;
    }
    S.vcheck = 1;
    sys_dev_.upload(sys_h_, stream_);
    return setup;
}

void NbOperator::upload_systems(hipStream_t s) { sys_dev_.upload(sys_h_, s); }

// Effective per-particle and exception parameters of every system in the
// kernels' form, the charge sums and the energy constants of sys_h_.
void NbOperator::upload_parameters(hipStream_t s) {
    const int nsys = (int)cfg_.size();
    const int N = ntotal_;
    std::vector<float> qk(N), sh(N), e2(N);
    std::vector<double> charge(N), qkd(N), shd(N), e2d(N), amp;
    std::vector<double> eqq, esig, eeps, ec6;
    const bool ljpme = cfg_[0].method == OMM_NB_LJPME;
    if (ljpme) amp.resize(N);
    const double sk = std::sqrt(ONE_4PI_EPS0);
    int off = 0;
    for (int si = 0; si < nsys; ++si) {
        const NbSettings& c = cfg_[si];
        sum_q2_[si] = 0.0;
        sum_q_[si] = 0.0;
        for (int i = 0; i < n_[si]; ++i) {
            if (!(std::isfinite(c.charge[i]) && c.sigma[i] >= 0.0 && c.epsilon[i] >= 0.0 && std::isfinite(c.sigma[i]) &&
                  std::isfinite(c.epsilon[i])))
                throw Error(OMM_ERR_INVALID, "effective particle parameters must be finite with nonnegative sigma "
                                             "and epsilon");
            const int g = off + i;
            charge[g] = c.charge[i];
            qkd[g] = c.charge[i] * sk;
            shd[g] = 0.5 * c.sigma[i];
            e2d[g] = 2.0 * std::sqrt(c.epsilon[i]);
            qk[g] = (float)qkd[g];
            sh[g] = (float)shd[g];
            e2[g] = (float)e2d[g];
            sum_q2_[si] += c.charge[i] * c.charge[i];
            sum_q_[si] += c.charge[i];
            if (ljpme) amp[g] = 2.0 * std::sqrt(c.epsilon[i]) * std::pow(c.sigma[i], 3);  // pme.rs:50-51
        }
        if (ljpme) {
            const std::vector<double> v = exception_c6(c);
            ec6.insert(ec6.end(), v.begin(), v.end());
        }
        for (size_t e = 0; e < c.exc1.size(); ++e) {
            eqq.push_back(c.exc_qq[e]);
            esig.push_back(c.exc_sig[e]);
            eeps.push_back(c.exc_eps[e]);
        }
        disp_cached_[si] = 0;
        SysInfo& S = sys_h_[si];
        S.self_const = S.neutral_coeff = S.disp_coeff = 0.0;
        if (c.screened()) {
            S.self_const = -ONE_4PI_EPS0 * alpha_[si] / std::sqrt(PI) * sum_q2_[si];
            S.neutral_coeff = -ONE_4PI_EPS0 * PI * sum_q_[si] * sum_q_[si] / (2.0 * alpha_[si] * alpha_[si]);
        }
        if (c.method == OMM_NB_LJPME) S.self_const += lj_self_energy(si);
        if (c.periodic() && c.dispersion_correction && c.method != OMM_NB_LJPME)
            S.disp_coeff = dispersion_coefficient(si);
        off += n_[si];
    }
    part_.charge.upload(charge, s);
    pme_.max_charge = 0.0;
    for (double q : charge) pme_.max_charge = std::max(pme_.max_charge, std::fabs(q));
    if (ljpme) {
        part_.c6amp.upload(amp, s);
        pme_lj_.max_charge = 0.0;
        for (double b : amp) pme_lj_.max_charge = std::max(pme_lj_.max_charge, std::fabs(b));
        if (ec6.empty()) ec6.push_back(0.0);
        exc_.c6.upload(ec6, s);
    }
    if (mixed_) {
        part_.qk.upload(qk, s);
        part_.sh.upload(sh, s);
        part_.e2.upload(e2, s);
    }
    part_.qk_d.upload(qkd, s);
    part_.sh_d.upload(shd, s);
    part_.e2_d.upload(e2d, s);
    exc_.qq.upload(eqq, s);
    exc_.sig.upload(esig, s);
    exc_.eps.upload(eeps, s);
}

void NbOperator::refresh_parameters() {
    hipStream_t s = stream_;
    upload_parameters(s);
    // Patch the energy constants into the device table, whose boxes may have
    // moved on the device (barostat).
    std::vector<SysInfo> dev(sys_h_.size());
    OMM_HIP(hipMemcpyAsync(dev.data(), sys_dev_.p, dev.size() * sizeof(SysInfo), hipMemcpyDeviceToHost, s));
    OMM_HIP(hipStreamSynchronize(s));
    for (size_t k = 0; k < dev.size(); ++k) {
        dev[k].self_const = sys_h_[k].self_const;
        dev[k].neutral_coeff = sys_h_[k].neutral_coeff;
        dev[k].disp_coeff = sys_h_[k].disp_coeff;
    }
    sys_dev_.upload(dev, s);
}

bool NbOperator::set_global_parameter(const std::string& name, double value, bool* changed) {
    if (!std::isfinite(value)) throw Error(OMM_ERR_INVALID, "global parameter values must be finite");
    bool found = false, any = false;
    for (NbSettings& c : cfg_)
        for (size_t k = 0; k < c.gp_names.size(); ++k)
            if (c.gp_names[k] == name) {
                found = true;
                if (c.gp_values[k] != value) {
                    c.gp_values[k] = value;
                    c.apply_offsets();
                    any = true;
                }
            }
    if (any && device_ready_ && ntotal_ > 0) refresh_parameters();
    if (changed) *changed = any;
    return found;
}

bool NbOperator::get_global_parameter(const std::string& name, double* value) const {
    for (const NbSettings& c : cfg_)
        for (size_t k = 0; k < c.gp_names.size(); ++k)
            if (c.gp_names[k] == name) {
                if (value) *value = c.gp_values[k];
                return true;
            }
    return false;
}

void NbOperator::update_parameters(const omm_nb_desc* desc) {
    // nonbonded.rs:721-767
    if (cfg_.size() != 1) throw Error(OMM_ERR_STATE, "parameter updates need a one-system operator");
    NbSettings n = settings_from_desc(desc);
    NbSettings& c = cfg_[0];
    if (n.base_charge.size() != c.base_charge.size()) throw Error(OMM_ERR_INVALID, "particle count must match");
    if (n.exc1.size() != c.exc1.size()) throw Error(OMM_ERR_INVALID, "exception count must match");
    if (n.gp_names.size() != c.gp_names.size())
        throw Error(OMM_ERR_INVALID, "global parameter names must not change");
    std::vector<int> remap(n.gp_names.size());
    for (size_t k = 0; k < n.gp_names.size(); ++k) {
        auto it = std::find(c.gp_names.begin(), c.gp_names.end(), n.gp_names[k]);
        if (it == c.gp_names.end()) throw Error(OMM_ERR_INVALID, "global parameter names must not change");
        remap[k] = (int)(it - c.gp_names.begin());
    }
    auto same_targets = [&](std::vector<ParamOffset>& fresh, const std::vector<ParamOffset>& old) {
        if (fresh.size() != old.size()) throw Error(OMM_ERR_INVALID, "parameter offset count must match");
        for (ParamOffset& o : fresh) {
            o.parameter = remap[o.parameter];
            bool hit = false;
            for (const ParamOffset& p : old) hit = hit || (p.parameter == o.parameter && p.index == o.index);
            if (!hit) throw Error(OMM_ERR_INVALID, "parameter offset names and target indices must not change");
        }
    };
    same_targets(n.particle_offsets, c.particle_offsets);
    same_targets(n.exception_offsets, c.exception_offsets);
    for (size_t e = 0; e < c.exc1.size(); ++e) {
        if (n.exc1[e] != c.exc1[e] || n.exc2[e] != c.exc2[e])
            throw Error(OMM_ERR_INVALID, "exception particle indices must not change");
        bool has_offset = false;
        for (const ParamOffset& o : c.exception_offsets) has_offset = has_offset || o.index == (int)e;
        if (!(c.base_qq[e] != 0.0 || c.base_eps[e] != 0.0 || has_offset || (n.base_qq[e] == 0.0 && n.base_eps[e] == 0.0)))
            throw Error(OMM_ERR_INVALID, "excluded exceptions cannot become non-excluded without reinitialization");
    }
    c.base_charge = n.base_charge;
    c.base_sigma = n.base_sigma;
    c.base_epsilon = n.base_epsilon;
    c.base_qq = n.base_qq;
    c.base_sig = n.base_sig;
    c.base_eps = n.base_eps;
    c.particle_offsets = n.particle_offsets;
    c.exception_offsets = n.exception_offsets;
    for (size_t k = 0; k < n.gp_names.size(); ++k) c.gp_defaults[remap[k]] = n.gp_defaults[k];
    c.apply_offsets();
    if (device_ready_ && ntotal_ > 0) refresh_parameters();
}

double NbOperator::dispersion_coefficient(int si) const {
    // nonbonded.rs:660-719
    if (disp_cached_[si]) return disp_coeff_[si];
    const NbSettings& c = cfg_[si];
    double moments[13] = {0}, diagonal[13] = {0};
    for (size_t i = 0; i < c.sigma.size(); ++i) {
        double power = 1.0;
        for (int k = 0; k <= 12; ++k) {
            moments[k] += std::sqrt(c.epsilon[i]) * power;
            diagonal[k] += c.epsilon[i] * power;
            power *= c.sigma[i];
        }
    }
    auto pair_moment = [&](int p) {
        double coef = 1.0, sum = 0.0;
        for (int k = 0; k <= p; ++k) {
            sum += coef * moments[k] * moments[p - k];
            coef *= (double)(p - k) / (double)(k + 1);
        }
        return sum / std::pow(2.0, p) + diagonal[p];
    };
    auto integral = [&](int p) {
        double tail = std::pow(c.cutoff, 3 - p) / (double)(p - 3);
        if (!c.switching) return tail;
        double sw = c.switching_distance, rc = c.cutoff;
        auto f = [&](double r) {
            double t = (r - sw) / (rc - sw);
            return std::pow(t, 3) * (10.0 + t * (-15.0 + 6.0 * t)) * std::pow(r, 2 - p);
        };
        double v0 = f(sw), v1 = f(0.5 * (sw + rc)), v2 = f(rc);
        double est = (rc - sw) / 6.0 * (v0 + 4.0 * v1 + v2);
        return tail + simpson(f, sw, rc, v0, v1, v2, est, 1e-11 * std::max(std::fabs(est), 1e-12), 20);
    };
    double n = (double)c.sigma.size();
    disp_coeff_[si] = 8.0 * PI * n / (n + 1.0) * (pair_moment(12) * integral(12) - pair_moment(6) * integral(6));
    disp_cached_[si] = 1;
    return disp_coeff_[si];
}

double NbOperator::lj_self_energy(int si) const {
    const NbSettings& c = cfg_[si];
    double sum = 0.0;
    for (size_t i = 0; i < c.sigma.size(); ++i) sum += c.epsilon[i] * std::pow(c.sigma[i], 6);
    return sum * std::pow(lj_alpha_[si], 6) / 3.0;
}

std::vector<double> exception_c6(const NbSettings& c) {
    std::vector<double> out(c.exc1.size(), 0.0);
    if (c.method != OMM_NB_LJPME) return out;
    for (size_t e = 0; e < out.size(); ++e) {
        const int i = c.exc1[e], j = c.exc2[e];
        out[e] = 4.0 * std::sqrt(c.epsilon[i] * c.epsilon[j]) * std::pow(c.sigma[i] * c.sigma[j], 3);
    }
    return out;
}

HostTerms NbOperator::host_terms(int si, double volume, bool direct, bool recip) const {
    HostTerms t;
    const NbSettings& c = cfg_[si];
    const double a = alpha_[si];
    if (recip && c.method == OMM_NB_LJPME) t.self += lj_self_energy(si);
    if (recip && c.screened())
        t.self += -ONE_4PI_EPS0 * (a / std::sqrt(PI) * sum_q2_[si] + PI * sum_q_[si] * sum_q_[si] / (2.0 * a * a * volume));
    if (direct && c.periodic() && c.dispersion_correction && c.method != OMM_NB_LJPME)
        t.dispersion = dispersion_coefficient(si) / volume;
    return t;
}

void NbOperator::grow(unsigned status) {
    if (status & ST_LIST_OVERFLOW) {
        const int limit = std::min(255 * JTILE, (nl_.nslots + 2 * JTILE) / JTILE * JTILE);
        if (nl_.cap >= limit) throw Error(OMM_ERR_OVERFLOW, "neighbour list needs more than 255 tiles per cluster");
        nlist_resize(nl_, std::min(limit, 2 * nl_.cap));
        pp_.cap = nl_.cap;
    }
    if ((status & ST_BIN_OVERFLOW) && cfg_[0].pme_mesh()) pme_grow_bins(pme_);
    if ((status & ST_BIN_OVERFLOW) && ljpme()) pme_grow_bins(pme_lj_);
}

void NbOperator::fit_bins(const double* pos_dev, unsigned* status, hipStream_t s) {
    if (!cfg_[0].pme_mesh()) return;
    for (int mesh = 0; mesh < (ljpme() ? 2 : 1); ++mesh) {
        for (int attempt = 0; attempt < 8; ++attempt) {
            fill_async<unsigned>(status, 1, 0u, s);
            if (mesh == 0)
                pme_run(pme_, mesh_mode(false), nl_, pos_dev, part_.charge.p, sys_dev_.p, nullptr, nullptr, nullptr,
                        nullptr, status, s, PME_SPREAD);
            else
                lj_mesh(mesh_mode(false), pos_dev, nullptr, nullptr, nullptr, nullptr, status, s, PME_SPREAD);
            unsigned st = 0;
            OMM_HIP(hipMemcpyAsync(&st, status, sizeof(unsigned), hipMemcpyDeviceToHost, s));
            OMM_HIP(hipStreamSynchronize(s));
            if (!(st & ST_BIN_OVERFLOW)) break;
            pme_grow_bins(mesh == 0 ? pme_ : pme_lj_);
        }
    }
    fill_async<unsigned>(status, 1, 0u, s);
}

namespace {
__global__ void k_copy_boxes(int nsys, const SysInfo* __restrict__ src, SysInfo* dst) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsys) dst[s].box = src[s].box;
}
}  // namespace

void NbOperator::lj_mesh(int mode, const double* pos_dev, long long* force_fx, double* force_d, long long* energy_fx,
                         double* energy_d, unsigned* status, hipStream_t s, int phases, bool atom_order) {
    // The boxes live in the Coulomb table (a barostat rescales them there).
    const int nsys = (int)sys_lj_h_.size();
    if (phases & PME_SPREAD) {
        hipLaunchKernelGGL(k_copy_boxes, dim3((nsys + 63) / 64), dim3(64), 0, s, nsys, sys_dev_.p, sys_lj_dev_.p);
        OMM_HIP(hipGetLastError());
    }
    pme_run(pme_lj_, mode, nl_, pos_dev, part_.c6amp.p, sys_lj_dev_.p, force_fx, force_d, energy_fx, energy_d, status,
            s, phases, atom_order);
}

void NbOperator::patch_layout(hipStream_t s) {
    std::vector<SysInfo> dev(sys_h_.size());
    OMM_HIP(hipMemcpyAsync(dev.data(), sys_dev_.p, dev.size() * sizeof(SysInfo), hipMemcpyDeviceToHost, s));
    OMM_HIP(hipStreamSynchronize(s));
    for (size_t k = 0; k < dev.size(); ++k) {
        const SysInfo& h = sys_h_[k];
        SysInfo& d = dev[k];
        d.ptile = h.ptile;
        d.tile_off = h.tile_off;
        for (int a = 0; a < 3; ++a) {
            d.tiles[a] = h.tiles[a];
            for (int p = 0; p < 2; ++p) d.fft_lines[p][a] = h.fft_lines[p][a];
        }
        for (int p = 0; p < 2; ++p)
            for (int k = 0; k < 2; ++k) d.plane_lines[p][k] = h.plane_lines[p][k];
    }
    sys_dev_.upload(dev, s);
}

int NbOperator::pair_grid_for(long long tasks, int scale) {
    const long long scaled = scale > 1 ? tasks * scale : (scale < 0 ? tasks / 2 : tasks);
    return (int)std::max<long long>(64, std::min<long long>(65535, scaled));
}

namespace {

// The tuning cache: a text file of "key<TAB>edge cap plane scale spread_ms
// fft_ms pair_ms" lines, appended by every timed autotune; the last line of
// a key wins.  Its path is set by omm_set_tuning_cache, or taken from the
// environment variable OMMHIP_TUNING_CACHE; empty: no cache.
std::mutex g_cache_mutex;
std::string g_cache_path;
bool g_cache_path_set = false;

std::string cache_path_locked() {
    if (!g_cache_path_set) {
        const char* env = std::getenv("OMMHIP_TUNING_CACHE");
        g_cache_path = env ? env : "";
        g_cache_path_set = true;
    }
    return g_cache_path;
}

bool cache_lookup(const std::string& key, NbOperator::Tuning& t) {
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    const std::string path = cache_path_locked();
    if (path.empty()) return false;
    std::FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    bool found = false;
    char line[1024];
    while (std::fgets(line, sizeof line, f)) {
        char* tab = std::strchr(line, '\t');
        if (!tab || std::string(line, tab) != key) continue;
        NbOperator::Tuning v;
        if (std::sscanf(tab + 1, "%d %d %d %d %lf %lf %lf", &v.spread_edge, &v.fft_cap, &v.fft_plane, &v.pair_scale,
                        &v.spread_ms, &v.fft_ms, &v.pair_ms) == 7) {
            t = v;
            found = true;
        }
    }
    std::fclose(f);
    return found;
}

void cache_store(const std::string& key, const NbOperator::Tuning& t) {
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    const std::string path = cache_path_locked();
    if (path.empty()) return;
    std::FILE* f = std::fopen(path.c_str(), "a");
    if (!f) return;  // an unwritable cache only costs the next process a tuning run
    std::fprintf(f, "%s\t%d %d %d %d %.6f %.6f %.6f\n", key.c_str(), t.spread_edge, t.fft_cap, t.fft_plane,
                 t.pair_scale, t.spread_ms, t.fft_ms, t.pair_ms);
    std::fclose(f);
}

}  // namespace

void set_tuning_cache_path(const char* path) {
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    g_cache_path = path ? path : "";
    g_cache_path_set = true;
}

std::string NbOperator::tuning_key() const {
    int dev = 0;
    OMM_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    OMM_HIP(hipGetDeviceProperties(&prop, dev));
    // FNV-1a over the per-system sizes and grids.
    unsigned long long h = 1469598103934665603ull;
    auto mix = [&](long long v) {
        for (int b = 0; b < 8; ++b) {
            h ^= (unsigned long long)((v >> (8 * b)) & 0xff);
            h *= 1099511628211ull;
        }
    };
    for (size_t k = 0; k < sys_h_.size(); ++k) {
        mix(sys_h_[k].n);
        for (int a = 0; a < 3; ++a) mix(sys_h_[k].grid[a]);
    }
    char buf[512];
    std::snprintf(buf, sizeof buf, "%s/%d mixed=%d method=%d cutoff=%.6g systems=%zu atoms=%d layout=%016llx",
                  prop.gcnArchName, prop.multiProcessorCount, mixed_ ? 1 : 0, (int)cfg_[0].method, cfg_[0].cutoff,
                  sys_h_.size(), ntotal_, h);
    return buf;
}

void NbOperator::apply_tuning(const Tuning& t, const double* pos_dev, unsigned* status, hipStream_t s) {
    const NbSettings& c = cfg_[0];
    // Shapes of phases this operator does not run are ignored.
    if (c.pme_mesh()) {
        if (t.spread_edge != 0 && t.spread_edge != 6 && t.spread_edge != 8)
            throw Error(OMM_ERR_INVALID, "spread tile edge must be 0 (atom-owned spread), 6 or 8");
        if (t.spread_edge == 0 && !mixed_)
            throw Error(OMM_ERR_INVALID, "the atom-owned spread serves the mixed-precision mesh only");
        if (t.fft_cap != 16 && t.fft_cap != 8 && t.fft_cap != 4 && t.fft_cap != 2)
            throw Error(OMM_ERR_INVALID, "FFT lines per workgroup must be 16, 8, 4 or 2");
        if (t.fft_plane != 0 && t.fft_plane != 1) throw Error(OMM_ERR_INVALID, "FFT decomposition must be 0 or 1");
    }
    if (c.include_direct && t.pair_scale != 1 && t.pair_scale != 2 && t.pair_scale != -2)
        throw Error(OMM_ERR_INVALID, "pair scale must be 1, 2 or -2");
    if (c.pme_mesh()) {
        const int prec = mesh_mode(false) == MESH_F32 ? 0 : 1;
        if (!pme_set_fft_plane(pme_, sys_h_, prec, t.fft_plane) && t.fft_plane) {
            pme_set_fft_plane(pme_, sys_h_, prec, tuning_.fft_plane);
            throw Error(OMM_ERR_INVALID, "plane FFT passes do not fit this workload's grids in LDS");
        }
        pme_set_fft_lines(pme_, sys_h_, prec, t.fft_cap, s);
        pme_.spread_atoms = t.spread_edge == 0;
        if (t.spread_edge != 0) pme_set_spread_edge(pme_, sys_h_, t.spread_edge, s);
        patch_layout(s);
        fit_bins(pos_dev, status, s);
    }
    Tuning v = t;
    if (!c.include_direct) v.pair_scale = 1;
    v.pair_grid = pair_grid_ = pair_grid_for(calibrated_tasks_, v.pair_scale);
    tuning_ = v;
    tuned_ = true;
    fill_async<unsigned>(status, 1, 0u, s);
    OMM_HIP(hipStreamSynchronize(s));
}

void NbOperator::autotune(const double* pos_dev, unsigned* status, hipStream_t s) {
    TraceRange range("omm.autotune");
    tuned_ = true;
    {
        const std::string key = tuning_key();
        Tuning cached;
        if (cache_lookup(key, cached)) {
            cached.from_cache = 1;
            bool ok = true;
            try {
                apply_tuning(cached, pos_dev, status, s);
            } catch (const Error&) {
                ok = false;  // an entry this build cannot apply: tune afresh
            }
            if (ok) {
                tuning_.from_cache = 1;
                return;
            }
        }
    }
    hipEvent_t t0, t1;
    OMM_HIP(hipEventCreate(&t0));
    OMM_HIP(hipEventCreate(&t1));
    auto timed = [&](auto&& body) {
        body();  // warm-up
        OMM_HIP(hipEventRecord(t0, s));
        for (int r = 0; r < 5; ++r) body();
        OMM_HIP(hipEventRecord(t1, s));
        OMM_HIP(hipEventSynchronize(t1));
        float v = 0.f;
        OMM_HIP(hipEventElapsedTime(&v, t0, t1));
        return (double)v / 5;
    };
    const NbSettings& c = cfg_[0];
    // Scratch accumulators: the FP64 mesh's convolution always sums the
    // energy, and the pair timings add forces.
    const size_t eslots = (size_t)E_SLOTS * sys_h_.size();
    DevBuf<long long> f_fx, e_fx;
    DevBuf<double> f_d, e_d;
    if (mixed_) {
        f_fx.alloc(3 * (size_t)ntotal_);
        e_fx.alloc(eslots);
    } else {
        f_d.alloc(3 * (size_t)ntotal_);
        e_d.alloc(eslots);
    }
    if (c.pme_mesh()) {
        const int mode = mesh_mode(false);
        const int prec = mode == MESH_F32 ? 0 : 1;
        auto phase = [&](int ph) {
            return [&, ph] {
                pme_run(pme_, mode, nl_, pos_dev, part_.charge.p, sys_dev_.p, f_fx.p, f_d.p, e_fx.p, e_d.p, status,
                        s, ph);
            };
        };
        // Spread: tiles of edge 8 or 6, or atom-owned chunks (edge 0).
        double best = 1e30;
        for (int e : {8, 6, 0}) {
            if (e == 0 && !mixed_) continue;  // fixed-point grids only
            pme_.spread_atoms = e == 0;
            if (e != 0) {
                pme_set_spread_edge(pme_, sys_h_, e, s);
                patch_layout(s);
                fit_bins(pos_dev, status, s);
            }
            const double t = timed(phase(PME_SPREAD));
            if (t < best) {
                best = t;
                tuning_.spread_edge = e;
            }
        }
        tuning_.spread_ms = best;
        pme_.spread_atoms = tuning_.spread_edge == 0;
        if (tuning_.spread_edge != 0) pme_set_spread_edge(pme_, sys_h_, tuning_.spread_edge, s);
        patch_layout(s);
        fit_bins(pos_dev, status, s);
        // FFT decomposition (five line passes, or plane passes around the x
        // pass) and lines per workgroup of the line passes: every candidate
        // computes each line with the same arithmetic, so the results are
        // the same bits.
        best = 1e30;
        for (int plane : {0, 1}) {
            if (!pme_set_fft_plane(pme_, sys_h_, prec, plane) && plane) continue;
            for (int cap : {16, 8, 4, 2}) {
                pme_set_fft_lines(pme_, sys_h_, prec, cap, s);
                patch_layout(s);
                const double t = timed(phase(PME_FFT));
                if (t < best) {
                    best = t;
                    tuning_.fft_cap = cap;
                    tuning_.fft_plane = plane;
                }
            }
        }
        tuning_.fft_ms = best;
        pme_set_fft_plane(pme_, sys_h_, prec, tuning_.fft_plane);
        pme_set_fft_lines(pme_, sys_h_, prec, tuning_.fft_cap, s);
        patch_layout(s);
    }
    if (c.include_direct) {
        // Pair kernel grid: waves per task of the calibrated list, halved or doubled.
        const int base = pair_grid_;
        double best = 1e30;
        for (int scale : {1, 2, -2}) {
            const int g = scale == 1 ? base : (scale == 2 ? std::min(65535, 2 * base) : std::max(64, base / 2));
            const double t = timed([&] {
                launch_pairs(mixed_, false, nl_, pp_, part_, pos_dev, sys_dev_.p, f_fx.p, f_d.p, e_fx.p, e_d.p, status,
                             nullptr, g, s);
            });
            if (t < best) {
                best = t;
                tuning_.pair_grid = g;
                tuning_.pair_scale = scale;
            }
        }
        tuning_.pair_ms = best;
        pair_grid_ = tuning_.pair_grid;
    }
    tuning_.from_cache = 0;
    cache_store(tuning_key(), tuning_);
    fill_async<unsigned>(status, 1, 0u, s);
    OMM_HIP(hipStreamSynchronize(s));
    OMM_HIP(hipEventDestroy(t0));
    OMM_HIP(hipEventDestroy(t1));
}

bool NbOperator::calibrate(const double* pos_dev, unsigned* status, hipStream_t s) {
    TraceRange range("omm.calibrate");
    if (ntotal_ == 0) return false;
    fit_bins(pos_dev, status, s);
    if (!cfg_[0].include_direct) return false;
    int max_npad = 0;
    for (const SysInfo& S : sys_h_) max_npad = std::max(max_npad, S.nblocks * IBLOCK);
    bool changed = false;
    std::string history;
    for (int attempt = 0; attempt < 8; ++attempt) {
        fill_async<unsigned>(status, 1, 0u, s);
        nlist_build(nl_, pos_dev, sys_dev_.p, status, s, true);
        std::vector<int> ccounts(nl_.nclusters);
        unsigned st = 0;
        nl_.ccount.download(ccounts.data(), ccounts.size(), s);
        OMM_HIP(hipMemcpyAsync(&st, status, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        OMM_HIP(hipStreamSynchronize(s));
        int mx = 0;
        long long tasks = 0;
        for (int c : ccounts) {
            mx = std::max(mx, c);
            tasks += ((c + JTILE - 1) / JTILE + TPT - 1) / TPT;
        }
        // One wave per task of the calibrated list (grid-stride beyond it):
        // enough resident waves to hide the latency of the pair loop.
        calibrated_tasks_ = tasks;
        const int grid = pair_grid_for(tasks, tuning_.pair_scale);
        if (grid != pair_grid_) {
            pair_grid_ = grid;
            changed = true;  // captured graphs hold the old launch shape
        }
        long long want = std::min<long long>((long long)max_npad + JTILE, 2LL * mx + 2 * JTILE);
        want = (want + JTILE - 1) / JTILE * JTILE;
        if (!(st & ST_LIST_OVERFLOW) && want <= nl_.cap) break;
        history += " [longest list " + std::to_string(mx) + ", capacity " + std::to_string(nl_.cap) + ", status " +
                   std::to_string(st) + "]";
        const long long limit = std::min<long long>(255 * JTILE, (max_npad + 2 * JTILE) / JTILE * JTILE);
        if (want > limit)
            throw Error(OMM_ERR_OVERFLOW, "neighbour list needs more than 255 tiles per cluster:" + history);
        nlist_resize(nl_, (int)std::min<long long>(limit, std::max<long long>(want, 2LL * nl_.cap)));
        pp_.cap = nl_.cap;
        changed = true;
    }
    take_mesh_order(s);
    if (!tuned_) {
        autotune(pos_dev, status, s);
        changed = true;
    }
    fill_async<unsigned>(status, 1, 0u, s);
    OMM_HIP(hipStreamSynchronize(s));
    return changed;
}

void NbOperator::take_mesh_order(hipStream_t s) {
    if (!cfg_[0].pme_mesh()) return;
    pme_take_order(pme_, nl_, s);
    if (ljpme()) pme_take_order(pme_lj_, nl_, s);
}

void NbOperator::list_stats(int64_t* tiles, int64_t* entries) {
    if (ntotal_ == 0) {
        if (tiles) *tiles = 0;
        if (entries) *entries = 0;
        return;
    }
    std::vector<int> counts(nl_.nclusters);
    nl_.ccount.download(counts.data(), counts.size(), stream_);
    OMM_HIP(hipStreamSynchronize(stream_));
    int64_t e = 0, total = 0;
    for (int c : counts) {
        e += c;
        total += (c + JTILE - 1) / JTILE;
    }
    if (tiles) *tiles = total;
    if (entries) *entries = e;
}

void NbOperator::force_rebuild(hipStream_t s) { fill_async<int>(nl_.flag.p, (size_t)nl_.nsys + 1, 1, s); }

void NbOperator::recheck_list(const double* pos_dev, hipStream_t s) {
    if (ntotal_ == 0 || !cfg_[0].include_direct) return;
    fill_async<int>(nl_.flag.p, (size_t)nl_.nsys + 1, 0, s);
    nlist_check(nl_, pos_dev, 0, s);
}

void NbOperator::set_capacity(int list_cap, int bin_cap, const double* src, unsigned* status, hipStream_t s) {
    if (list_cap < 0 || bin_cap < 0) throw Error(OMM_ERR_INVALID, "capacities must be nonnegative");
    OMM_HIP(hipStreamSynchronize(s));
    if (list_cap > 0 && cfg_[0].include_direct) {
        const int cap = (list_cap + JTILE - 1) / JTILE * JTILE;
        nlist_resize(nl_, cap);
        pp_.cap = nl_.cap;
        nlist_build(nl_, src, sys_dev_.p, status, s, true);
        take_mesh_order(s);
    }
    if (bin_cap > 0 && cfg_[0].pme_mesh()) {
        pm_set_bin_cap(pme_, bin_cap);
        if (ljpme()) pm_set_bin_cap(pme_lj_, bin_cap);
    }
    OMM_HIP(hipStreamSynchronize(s));
}

double NbOperator::mesh_charges(const double* pos, const double* box, double alpha, const double* charges,
                                double* forces, const int* atoms, int count, double* derivatives) {
    TraceRange range(count > 0 ? "omm.pme_charge_derivatives" : "omm.pme_evaluate_charges");
    if (!initialized_ || !device_ready_) throw Error(OMM_ERR_STATE, "PME operator not initialized");
    if (cfg_.size() != 1 || cfg_[0].method != OMM_NB_PME || mixed_)
        throw Error(OMM_ERR_INVALID, "the charge mesh runs on a one-system plain-FP64 PME operator");
    if (!pos || !box || !charges) throw Error(OMM_ERR_INVALID, "null positions, box or charges");
    if (!(std::isfinite(alpha) && alpha > 0.0)) throw Error(OMM_ERR_INVALID, "alpha must be positive and finite");
    if (count < 0 || (count > 0 && (!atoms || !derivatives)))
        throw Error(OMM_ERR_INVALID, "charge derivatives need particle indices and an output array");
    validate_box(box);
    const int n = n_[0];
    if (n <= 0) throw Error(OMM_ERR_INVALID, "the charge mesh needs particles");
    double qmax = 0.0;
    for (int i = 0; i < n; ++i) {
        if (!std::isfinite(charges[i])) throw Error(OMM_ERR_NONFINITE, "charges must be finite");
        qmax = std::max(qmax, std::fabs(charges[i]));
    }
    for (int i = 0; i < 3 * n; ++i)
        if (!std::isfinite(pos[i])) throw Error(OMM_ERR_NONFINITE, "particle positions must be finite");
    for (int k = 0; k < count; ++k)
        if (atoms[k] < 0 || atoms[k] >= n) throw Error(OMM_ERR_INVALID, "particle index out of range");
    hipStream_t s = stream_;
    const Box b = make_box(box);
    if (!have_box_ || std::memcmp(b.v, last_box_.v, sizeof(b.v)) != 0 || sys_h_[0].alpha != alpha) {
        sys_h_[0].box = b;
        sys_h_[0].alpha = alpha;
        upload_systems(s);
        last_box_ = b;
        have_box_ = true;
    }
    pos_dev_.upload(pos, 3 * (size_t)n, s);
    mesh_q_.upload(charges, (size_t)n, s);
    const double saved_max = pme_.max_charge;
    pme_.max_charge = qmax;  // selects the spread's staging; the FP64 grid is exact either way
    const int phases = count > 0 ? (PME_SPREAD | PME_FFT) : PME_ALL;
    unsigned st = 0;
    for (int attempt = 0; attempt < 8; ++attempt) {
        fill_async<unsigned>(status_.p, 1, 0u, s);
        fill_async<double>(force_d_.p, 3 * (size_t)n, 0.0, s);
        fill_async<double>(energy_d_.p, E_SLOTS, 0.0, s);
        pme_run(pme_, MESH_F64, nl_, pos_dev_.p, mesh_q_.p, sys_dev_.p, nullptr, force_d_.p, nullptr, energy_d_.p,
                status_.p, s, phases);
        status_.download(&st, 1, s);
        OMM_HIP(hipStreamSynchronize(s));
        if (!(st & ST_BIN_OVERFLOW)) break;
        grow(st);
    }
    pme_.max_charge = saved_max;
    if (st & ST_BIN_OVERFLOW) throw Error(OMM_ERR_OVERFLOW, "PME spread bins kept overflowing");
    if (st & ST_NONFINITE) throw Error(OMM_ERR_NONFINITE, "non-finite mesh result");
    if (count > 0) {
        mesh_atoms_.upload(atoms, (size_t)count, s);
        mesh_out_.alloc((size_t)count);
        pme_potential_at(pme_, sys_dev_.p, pos_dev_.p, mesh_atoms_.p, count, mesh_out_.p, s);
        std::vector<double> pot(count);
        mesh_out_.download(pot.data(), (size_t)count, s);
        OMM_HIP(hipStreamSynchronize(s));
        for (int k = 0; k < count; ++k) derivatives[k] += pot[k];
        return 0.0;
    }
    double e[E_SLOTS];
    energy_d_.download(e, E_SLOTS, s);
    std::vector<double> f(forces ? 3 * (size_t)n : 0);
    if (forces) force_d_.download(f.data(), 3 * (size_t)n, s);
    OMM_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < f.size(); ++i) forces[i] += f[i];
    return e[E_RECIP];
}

void NbOperator::set_box(int s, const double* box9, hipStream_t st) {
    if (virtual_) throw Error(OMM_ERR_INVALID, "a non-periodic system takes no box vectors");
    if (s < 0 || s >= (int)sys_h_.size()) throw Error(OMM_ERR_INVALID, "system index out of range");
    if (!cfg_[0].periodic()) throw Error(OMM_ERR_INVALID, "non-periodic methods have no box vectors");
    if (!box9) throw Error(OMM_ERR_INVALID, "null box vectors");
    validate_box(box9);
    for (int a = 0; a < 3; ++a)
        if (!(cfg_[s].cutoff <= 0.5 * box9[a * 4]))
            throw Error(OMM_ERR_CUTOFF_BOX, "cutoff must not exceed half the periodic box size");
    if (cfg_[0].method == OMM_NB_EWALD && !(box9[3] == 0.0 && box9[6] == 0.0 && box9[7] == 0.0))
        throw Error(OMM_ERR_INVALID, "Ewald requires a rectangular periodic box; use PME for triclinic boxes");
    const Box b = make_box(box9);
    sys_h_[s].box = b;
    OMM_HIP(hipMemcpyAsync(&sys_dev_.p[s].box, &b, sizeof(Box), hipMemcpyHostToDevice, st));
    OMM_HIP(hipStreamSynchronize(st));
}

int NbOperator::mesh_mode(bool energy) const {
    if (!mixed_) return MESH_F64;
    return energy ? MESH_F64_FIXED : MESH_F32;
}

void NbOperator::enqueue(const double* pos_dev, bool direct, bool recip, bool energy, int rebuild, bool exceptions,
                         long long* force_fx, double* force_d, long long* energy_fx, double* energy_d,
                         unsigned* status, hipStream_t s, bool join_now) {
    const NbSettings& c = cfg_[0];
    // The mesh runs on the side stream, concurrent with the list build and
    // the direct-space pairs; join() makes s wait for it.  It is forked
    // before the list build: the spread and the gather visit the atoms in the
    // meshes' own copy of an earlier build's order (any order of the same
    // slot layout gives the same sums), and the pair forces' fold copies this
    // step's order into the meshes' other half for the next step.  Only
    // without such a copy (a new layout) does the gather wait for the build
    // and read its order (join() then takes the copy).
    const bool mesh = recip && c.screened();
    const bool own_order = mesh_order_ready();
    if (mesh) {
        OMM_HIP(hipEventRecord(ev_fork_, s));
        OMM_HIP(hipStreamWaitEvent(side_, ev_fork_, 0));
        if (c.pme_mesh()) {
            pme_run(pme_, mesh_mode(energy), nl_, pos_dev, part_.charge.p, sys_dev_.p, force_fx, force_d, energy_fx,
                    energy_d, status, side_, PME_SPREAD | PME_FFT, true);
            if (ljpme())
                lj_mesh(mesh_mode(energy), pos_dev, force_fx, force_d, energy_fx, energy_d, status, side_,
                        PME_SPREAD | PME_FFT, true);
        } else
            ewald_run(ewald_, pos_dev, part_.charge.p, sys_dev_.p, alpha_[0], force_fx, force_d,
                      energy ? energy_fx : nullptr, energy ? energy_d : nullptr, status, mixed_, side_);
    }
    if (direct) {
        if (rebuild == REBUILD_CHECK) {
            nlist_check(nl_, pos_dev, 0, s);
            nlist_build(nl_, pos_dev, sys_dev_.p, status, s, false);
        } else if (rebuild == REBUILD_FLAG) {
            nlist_build(nl_, pos_dev, sys_dev_.p, status, s, false);
        } else if (rebuild == REBUILD_FORCED) {
            nlist_build(nl_, pos_dev, sys_dev_.p, status, s, true);
        }
    }
    if (mesh) {
        if (c.pme_mesh()) {
            if (!own_order) {
                OMM_HIP(hipEventRecord(ev_list_, s));
                OMM_HIP(hipStreamWaitEvent(side_, ev_list_, 0));
            }
            pme_run(pme_, mesh_mode(energy), nl_, pos_dev, part_.charge.p, sys_dev_.p, force_fx, force_d, energy_fx,
                    energy_d, status, side_, PME_GATHER, own_order);
            if (ljpme())
                lj_mesh(mesh_mode(energy), pos_dev, force_fx, force_d, energy_fx, energy_d, status, side_, PME_GATHER,
                        own_order);
            if (!own_order) follow_order_ = true;
        }
        OMM_HIP(hipEventRecord(ev_join_, side_));
        joined_ = false;
    }
    if (direct) {
        int* next[2] = {nullptr, nullptr};
        if (mesh && own_order && c.pme_mesh()) {
            next[0] = pme_next_order(pme_, nl_);
            if (ljpme()) next[1] = pme_next_order(pme_lj_, nl_);
        }
        launch_pairs(mixed_, energy, nl_, pp_, part_, pos_dev, sys_dev_.p, force_fx, force_d, energy_fx, energy_d,
                     status, nullptr, pair_grid_, s, true, next[0], next[1]);
        if (exceptions)
            launch_exceptions(mixed_, exc_, part_, pos_dev, sys_dev_.p, c.exceptions_pbc ? 1 : 0,
                              c.screened() ? 1 : 0, alpha_[0], ljpme() ? lj_alpha_[0] : 0.0, force_fx, force_d,
                              energy_fx, energy_d, status, s);
    }
    if (join_now) join(s);
}

void NbOperator::join(hipStream_t s) {
    if (joined_) return;
    OMM_HIP(hipStreamWaitEvent(s, ev_join_, 0));
    joined_ = true;
    if (follow_order_) {
        pme_follow_order(pme_, ljpme() ? &pme_lj_ : nullptr, nl_, s);
        follow_order_ = false;
    }
}

bool NbOperator::mesh_order_ready() const {
    return pme_order_ready(pme_, nl_) && (!ljpme() || pme_order_ready(pme_lj_, nl_));
}

void NbOperator::profile(const double* pos_dev, int reps, long long* force_fx, double* force_d, long long* energy_fx,
                         double* energy_d, unsigned* status, double ms[6], int64_t* num_pairs, hipStream_t s) {
    hipEvent_t t0, t1;
    OMM_HIP(hipEventCreate(&t0));
    OMM_HIP(hipEventCreate(&t1));
    auto timed = [&](auto&& body) {
        body();  // warm-up
        OMM_HIP(hipEventRecord(t0, s));
        for (int r = 0; r < reps; ++r) body();
        OMM_HIP(hipEventRecord(t1, s));
        OMM_HIP(hipEventSynchronize(t1));
        float v = 0.f;
        OMM_HIP(hipEventElapsedTime(&v, t0, t1));
        return (double)v / reps;
    };
    const NbSettings& c = cfg_[0];
    const bool pme = c.pme_mesh();
    const bool screened = c.screened();
    const double lj_a = ljpme() ? lj_alpha_[0] : 0.0;
    ms[0] = timed([&] { nlist_build(nl_, pos_dev, sys_dev_.p, status, s, true); });
    // The pair kernel alone; its slot-ordered sums are folded in once after.
    ms[1] = timed([&] {
        launch_pairs(mixed_, false, nl_, pp_, part_, pos_dev, sys_dev_.p, force_fx, force_d, energy_fx, energy_d,
                     status, nullptr, pair_grid_, s, false);
    });
    launch_slot_forces(mixed_, nl_, force_fx, force_d, s);
    ms[2] = timed([&] {
        launch_exceptions(mixed_, exc_, part_, pos_dev, sys_dev_.p, c.exceptions_pbc ? 1 : 0, screened ? 1 : 0,
                          alpha_[0], lj_a, force_fx, force_d, energy_fx, energy_d, status, s);
    });
    const int phases[3] = {PME_SPREAD, PME_FFT, PME_GATHER};
    for (int k = 0; k < 3; ++k)
        ms[3 + k] = pme ? timed([&] {
            pme_run(pme_, mesh_mode(false), nl_, pos_dev, part_.charge.p, sys_dev_.p, force_fx, force_d, energy_fx,
                    energy_d, status, s, phases[k]);
            if (ljpme()) lj_mesh(mesh_mode(false), pos_dev, force_fx, force_d, energy_fx, energy_d, status, s, phases[k]);
        })
                        : 0.0;
    if (c.method == OMM_NB_EWALD)
        ms[4] = timed([&] {
            ewald_run(ewald_, pos_dev, part_.charge.p, sys_dev_.p, alpha_[0], force_fx, force_d, nullptr, nullptr,
                      status, mixed_, s);
        });
    if (num_pairs) {
        DevBuf<unsigned long long> st_dev;
        st_dev.alloc(2 * sys_h_.size());
        fill_async<unsigned long long>(st_dev.p, 2 * sys_h_.size(), 0ull, s);
        launch_pairs(mixed_, false, nl_, pp_, part_, pos_dev, sys_dev_.p, force_fx, force_d, energy_fx, energy_d,
                     status, st_dev.p, pair_grid_, s);
        std::vector<unsigned long long> st(2 * sys_h_.size());
        st_dev.download(st.data(), st.size(), s);
        OMM_HIP(hipStreamSynchronize(s));
        int64_t total = 0;
        for (size_t k = 0; k < sys_h_.size(); ++k) total += (int64_t)st[2 * k];
        *num_pairs = total;
    }
    OMM_HIP(hipEventDestroy(t0));
    OMM_HIP(hipEventDestroy(t1));
}

void NbOperator::compute(const double* pos, const double* box, bool direct, bool recip, double* forces,
                         double* energy, double parts[5], int64_t* num_pairs, uint64_t* pair_hash,
                         const double* pos_dev_in, double* forces_dev_out, hipStream_t caller) {
    TraceRange range(pos_dev_in ? "omm.nb_evaluate_device" : "omm.nb_evaluate");
    if (!initialized_) throw Error(OMM_ERR_STATE, "NonbondedForce used before initialize");
    if (cfg_.size() != 1) throw Error(OMM_ERR_STATE, "host-pointer evaluation needs a one-system operator");
    if (!pos && !pos_dev_in) throw Error(OMM_ERR_INVALID, "positions must not be null");
    check_box(box);
    const int n = n_[0];
    // Device positions: a non-periodic method chooses its virtual box from a
    // host copy; otherwise non-finite positions surface as non-finite forces.
    std::vector<double> host_pos;
    if (pos_dev_in && virtual_ && n > 0) {
        host_pos.resize(3 * (size_t)n);
        OMM_HIP(hipMemcpyAsync(host_pos.data(), pos_dev_in, host_pos.size() * sizeof(double), hipMemcpyDeviceToHost,
                               caller));
        OMM_HIP(hipStreamSynchronize(caller));
        pos = host_pos.data();
    }
    if (pos)
        for (int i = 0; i < 3 * n; ++i)
            if (!std::isfinite(pos[i])) throw Error(OMM_ERR_NONFINITE, "particle positions must be finite");
    Box b{};
    bool vchanged = false;
    if (n > 0 && virtual_) {
        b = virtual_box_for(pos, n, &vchanged);
        if (!device_ready_) setup_device(std::vector<Box>{b});
    } else {
        b = make_box(box);
    }
    hipStream_t s = stream_;
    if (n == 0) {
        if (energy) *energy = 0.0;
        if (parts)
            for (int k = 0; k < 5; ++k) parts[k] = 0.0;
        if (num_pairs) *num_pairs = 0;
        if (pair_hash) *pair_hash = 0;
        return;
    }
    bool box_changed = vchanged || !have_box_ || std::memcmp(b.v, last_box_.v, sizeof(b.v)) != 0;
    if (box_changed) {
        sys_h_[0].box = b;
        upload_systems(s);
        last_box_ = b;
        have_box_ = true;
    }
    if (pos_dev_in) {
        // Ordered after the caller's work on its stream; the call returns
        // once the results are complete, as the host form does.
        OMM_HIP(hipEventRecord(ev_caller_, caller));
        OMM_HIP(hipStreamWaitEvent(s, ev_caller_, 0));
        OMM_HIP(hipMemcpyAsync(pos_dev_.p, pos_dev_in, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
        pos_dev_.upload(pos, 3 * (size_t)n, s);
    }
    const bool want_stats = num_pairs || pair_hash;
    // Pair statistics are gathered by a force-only pass; energies by the
    // energy pass (FP64 pair energies and FP64 mesh in the mixed form).
    const bool want_energy = !want_stats && (energy || parts);
    const NbSettings& c = cfg_[0];
    for (int attempt = 0; attempt < 6; ++attempt) {
        fill_async<unsigned>(status_.p, 1, 0u, s);
        if (want_stats) fill_async<unsigned long long>(stats_.p, 2, 0ull, s);
        if (mixed_) {
            fill_async<long long>(force_fx_.p, 3 * (size_t)n, 0ll, s);
            fill_async<long long>(energy_fx_.p, E_SLOTS, 0ll, s);
        } else {
            fill_async<double>(force_d_.p, 3 * (size_t)n, 0.0, s);
            fill_async<double>(energy_d_.p, E_SLOTS, 0.0, s);
        }
        if (direct) {
            nlist_check(nl_, pos_dev_.p, (box_changed || attempt > 0) ? 1 : 0, s);
            nlist_build(nl_, pos_dev_.p, sys_dev_.p, status_.p, s);
            launch_pairs(mixed_, want_energy, nl_, pp_, part_, pos_dev_.p, sys_dev_.p, force_fx_.p, force_d_.p,
                         energy_fx_.p, energy_d_.p, status_.p, want_stats ? stats_.p : nullptr, pair_grid_, s);
            launch_exceptions(mixed_, exc_, part_, pos_dev_.p, sys_dev_.p, c.exceptions_pbc ? 1 : 0,
                              c.screened() ? 1 : 0, alpha_[0], ljpme() ? lj_alpha_[0] : 0.0, force_fx_.p,
                              force_d_.p, energy_fx_.p, energy_d_.p, status_.p, s);
        }
        if (recip && c.pme_mesh())
            pme_run(pme_, mesh_mode(want_energy), nl_, pos_dev_.p, part_.charge.p, sys_dev_.p, force_fx_.p,
                    force_d_.p, energy_fx_.p, energy_d_.p, status_.p, s);
        if (recip && ljpme())
            lj_mesh(mesh_mode(want_energy), pos_dev_.p, force_fx_.p, force_d_.p, energy_fx_.p, energy_d_.p, status_.p, s);
        if (recip && c.method == OMM_NB_EWALD)
            ewald_run(ewald_, pos_dev_.p, part_.charge.p, sys_dev_.p, alpha_[0], force_fx_.p, force_d_.p,
                      energy_fx_.p, energy_d_.p, status_.p, mixed_, s);
        unsigned st = 0;
        status_.download(&st, 1, s);
        OMM_HIP(hipStreamSynchronize(s));
        if (st & (ST_LIST_OVERFLOW | ST_BIN_OVERFLOW)) {
            grow(st);
            continue;
        }
        if (st & ST_NONFINITE) throw Error(OMM_ERR_NONFINITE, "non-finite force or energy on the device");
        if (st & ST_OVERFLOW) throw Error(OMM_ERR_OVERFLOW, "fixed-point force/energy accumulator overflow");
        double e[E_SLOTS];
        if (mixed_) {
            long long efx[E_SLOTS];
            energy_fx_.download(efx, E_SLOTS, s);
            launch_fixed_to_double(force_fx_.p, forces_dev_out ? forces_dev_out : out_forces_.p, 3 * (size_t)n, s);
            if (forces) out_forces_.download(forces, 3 * (size_t)n, s);
            OMM_HIP(hipStreamSynchronize(s));
            for (int k = 0; k < E_SLOTS; ++k) e[k] = (double)efx[k] * (1.0 / ENERGY_SCALE);
        } else {
            energy_d_.download(e, E_SLOTS, s);
            if (forces) force_d_.download(forces, 3 * (size_t)n, s);
            if (forces_dev_out)
                OMM_HIP(hipMemcpyAsync(forces_dev_out, force_d_.p, 3 * (size_t)n * sizeof(double),
                                       hipMemcpyDeviceToDevice, s));
            OMM_HIP(hipStreamSynchronize(s));
        }
        if (want_stats) {
            unsigned long long st2[2];
            stats_.download(st2, 2, s);
            OMM_HIP(hipStreamSynchronize(s));
            if (num_pairs) *num_pairs = (int64_t)st2[0];
            if (pair_hash) *pair_hash = st2[1];
        }
        HostTerms ht = host_terms(0, b.volume, direct, recip);
        double p[5] = {e[E_DIRECT], e[E_EXCEPT], e[E_RECIP], ht.self, ht.dispersion};
        if (parts)
            for (int k = 0; k < 5; ++k) parts[k] = p[k];
        // Order of nonbonded.rs:942-1036: direct, reciprocal, self, dispersion, exceptions.
        if (energy) *energy = p[0] + p[2] + p[3] + p[4] + p[1];
        return;
    }
    throw Error(OMM_ERR_OVERFLOW, "neighbour list capacity could not be grown enough");
}

}  // namespace ommhip
