// settings.cpp -- the reference's parameter files as engine requests.
//
// Reads a SimulationManagerSettings parameter file (Settings/SimulationManagerSettings.h:34-186)
// and the per-particle files its Particle_Provider lists (Provider/ParticleProvider.h:230-295)
// and produces one sp_request plus per-particle property arrays, i.e. what the
// reference's SimulationManager builds one Problem object per particle from
// (SimulationManager.h:373-478).  Key names follow the serialize() functions of
// Settings/*.h and Properties/*.h; a few older keys of Examples/*.ini are accepted too.
//
// The particle files become a table of particle types; particle i of the ensemble is
// provide_particle(i) of sp_provider.h, run here on the host when the arrays are asked
// for (sp_settings_request) or by the provider kernel on the device
// (sp_settings_prepare, sp_settings_submit).  MyCEL's own parameterisation of the
// lognormal distribution is not available to compare with.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../sp_engine.h"
#include "../sp_host.h"
#include "../sp_provider.h"
#include "../sp_rng.h"
#include "ini.h"

namespace sp::frontend {

struct Dist {
    std::string type = "delta";
    double width = 0.0;
    bool relative = true;
};

struct ParticleFile {
    double T = 295.0, radius = 0, Ms = 0, alpha = 0, gamma = 0, K = 0, r_h = 0, visc = 0;
    double K2 = 0;                           // uniaxialcubic: K_cubic (K is K_uniaxial)
    double urot[3] = {0, 0, 0};              // uniaxialcubic: Uniaxial_Rotation (Euler angles)
    bool rand_urot = false;                  // uniaxialcubic: UseRandomCubicOrientation
    int aniso = SP_ANISO_UNIAXIAL;
    bool rand_orient = false, rand_mag = false;
    double euler[3] = {0, 0, 0}, mag[3] = {0, 0, 1};
    Dist rad, shell, aniso_dist, aniso_dist2;
    uint64_t count = 0;
};

struct Settings {
    sp_request rq{};
    bool discrete = false;
    std::vector<ParticleFile> particles;
    std::string savefile, single_savefile, single_prefix = "Single";
    std::shared_ptr<const ProvSpec> spec;   // the particle types as the provider draws them
    // host arrays, materialised on first use (under mu: the C API hands out const handles, which callers
    // may share between threads)
    std::mutex mu;
    bool materialised = false;
    std::vector<double> radius, Ms, alpha, gamma, T, K, r_h, visc, axes, y0, K2, urot;
};

static const std::string* opt(const ini::File& f, const std::string& sec, std::initializer_list<const char*> keys) {
    for (const char* k : keys)
        if (const std::string* v = f.find(sec, k)) return v;
    return nullptr;
}
static double num(const ini::File& f, const std::string& sec, std::initializer_list<const char*> keys, double dflt, bool required = false) {
    if (const std::string* v = opt(f, sec, keys)) return ini::to_double(*v);
    if (required) throw std::runtime_error(std::string("missing key ") + *keys.begin() + " in [" + sec + "]");
    return dflt;
}
static bool flag(const ini::File& f, const std::string& sec, std::initializer_list<const char*> keys, bool dflt) {
    if (const std::string* v = opt(f, sec, keys)) return ini::to_bool(*v);
    return dflt;
}
static std::vector<double> vec(const ini::File& f, const std::string& sec, std::initializer_list<const char*> keys) {
    if (const std::string* v = opt(f, sec, keys)) return ini::to_vector(*v);
    return {};
}
static int parse_enum_or_throw(const char* kind, const std::string& name) {
    int32_t v = 0;
    if (sp_parse_enum(kind, name.c_str(), &v) != SP_OK) throw std::runtime_error(std::string("unknown ") + kind + " '" + name + "'");
    return v;
}

// One particle type's parameters (Parameters::ParticleSimulationParameters) from the sections under `root` of f
// ("" for a particle file of its own).
static ParticleFile load_particle_from(const ini::File& f, const std::string& root) {
    ParticleFile p;
    const std::string R = root.empty() ? std::string() : root + ".";
    const std::string PP = R + "Particle_Properties";
    p.T = num(f, PP, {"Temperature"}, 295.0, true);
    const std::string MP = PP + ".Magnetic_Properties";
    p.radius = num(f, MP, {"Magnetic_radius"}, 0, true);
    p.Ms = num(f, MP, {"Saturation_magnetisation"}, 0, true);
    p.alpha = num(f, MP, {"Damping_constant"}, 0, true);
    p.gamma = num(f, MP, {"Gyromagnetic_ratio"}, 0, true);
    const std::string* at = opt(f, MP, {"Type_of_anisotropy"});
    p.aniso = at ? parse_enum_or_throw("anisotropy", *at) : SP_ANISO_UNIAXIAL;
    if (p.aniso == SP_ANISO_UNIAXIALCUBIC) {
        // Anisotropy_Properties.{Uniaxial_Part, Cubic_Part}, Uniaxial_Rotation (UniaxialCubic.hpp:45-50)
        p.K = num(f, PP, {"K_uniaxial"}, 0, true);
        p.K2 = num(f, PP, {"K_cubic"}, 0, true);
        const auto ur = vec(f, PP, {"Uniaxial_Rotation"});
        for (int i = 0; i < 3 && i < (int)ur.size(); ++i) p.urot[i] = ur[i];
    } else {
        p.K = num(f, PP, {p.aniso == SP_ANISO_CUBIC ? "K_cubic" : "K_uniaxial", "Anisotropy_1"}, 0, true);
    }
    const std::string HP = PP + ".Hydrodynamic_Properties";
    p.r_h = num(f, HP, {"Hydrodynamic_Radius"}, 0.0);
    p.visc = num(f, HP, {"Viscosity"}, 0.0);
    const std::string IN = R + "Particle_Simulation_Initialization";   // ParticleSimulationInitSettings.h:70-91
    p.rand_orient = flag(f, IN, {"Use_random_initial_particle_orientation"}, false);
    p.rand_mag = flag(f, IN, {"Use_random_initial_particle_magnetisation"}, false);
    auto e = vec(f, IN, {"Initial_particle_orientation"});
    auto m = vec(f, IN, {"Initial_magnetisation_direction"});
    for (int i = 0; i < 3; ++i) { if (i < (int)e.size()) p.euler[i] = e[i]; if (i < (int)m.size()) p.mag[i] = m[i]; }
    const std::string SS = R + "Particle_Simulation_Settings";         // ParticleSimulationSettings.h:156-200
    if (const std::string* v = opt(f, SS, {"Magnetic_radius_distribution_type"})) p.rad.type = *v;
    p.rad.width = num(f, SS, {"Magnetic_radius_distribution_width"}, 0.0);
    p.rad.relative = flag(f, SS, {"Use_relative_distribution_width_for_magnetic_radius"}, true);
    if (const std::string* v = opt(f, SS, {"Hydrodynamic_shell_distribution_type"})) p.shell.type = *v;
    p.shell.width = num(f, SS, {"Hydrodynamic_shell_distribution_width"}, 0.0);
    p.shell.relative = flag(f, SS, {"Use_relative_distribution_width_for_hydrodynamic_shell"}, true);
    if (p.aniso == SP_ANISO_UNIAXIALCUBIC) {
        // Anisotropy_Distribution.{Uniaxial_Distribution, Cubic_Distribution}, UseRandomCubicOrientation
        // (UniaxialCubic.hpp:81-86, Uniaxial.hpp:86-101, Cubic.hpp:79-94)
        const std::string AD = SS + ".Anisotropy_Distribution";
        auto read = [&](const std::string& sec, const char* sigma, Dist& d) {
            if (const std::string* v = f.find(sec, "Distribution_type")) d.type = *v;
            if (const std::string* v = f.find(sec, sigma)) d.width = ini::to_double(*v);
            if (const std::string* v = f.find(sec, "Use_relative_distribution_width")) d.relative = ini::to_bool(*v);
        };
        read(AD + ".Uniaxial_Distribution", "Sigma_K_uniaxial", p.aniso_dist);
        read(AD + ".Cubic_Distribution", "Sigma_K_cubic", p.aniso_dist2);
        p.rand_urot = flag(f, AD, {"UseRandomCubicOrientation"}, false);
        return p;
    }
    if (const std::string* v = opt(f, SS, {"Distribution_type", "Anisotropy_constants_distribution_type"})) p.aniso_dist.type = *v;
    const auto w = vec(f, SS, {"Sigma_K_uniaxial", "Sigma_K_cubic", "Anisotropy_constants_distribution_width"});
    p.aniso_dist.width = w.empty() ? 0.0 : w[0];
    p.aniso_dist.relative = flag(f, SS, {"Use_relative_distribution_width", "Use_relative_distribution_width_for_anisotropy_constants"}, true);
    return p;
}

static bool file_exists(const std::string& path) {
    std::ifstream in(path);
    return in.good();
}

// The particle type `name` of a parameter file (ParticleProvider.h:267-282): from the particle file its
// Particle_List entry names -- a relative path is taken relative to the parameter file's directory when the file
// is there, else relative to the working directory, as the reference opens it -- or, for an empty entry, from the
// sections under [name] of the parameter file itself (or under [Particle_Provider.name], where the reference's
// writer puts the particles it saves in the same file, ParticleProvider.h:126-138).
static ParticleFile load_particle(const ini::File& f, const std::string& name, const std::string& file) {
    if (file.empty()) {
        for (const std::string root : {name, std::string("Particle_Provider.") + name})
            if (f.has_section(root + ".Particle_Properties")) return load_particle_from(f, root);
        throw std::runtime_error("particle type '" + name + "' has no file and no [" + name + ".Particle_Properties]");
    }
    std::string path = file;
    if (file[0] != '/') {
        const std::string beside = f.path + "/" + file;
        path = file_exists(beside) || !file_exists(file) ? beside : file;
    }
    return load_particle_from(ini::parse_file(path), "");
}

// ---- particle types for the provider
static ProvDist prov_dist(const Dist& d, double mean) {
    ProvDist p{};
    p.mean = mean;
    if (d.type == "delta" || d.width == 0.0) { p.type = PD_DELTA; return p; }
    const double sd = d.relative ? d.width * mean : d.width;
    if (d.type == "normal") { p.type = PD_NORMAL; p.sd = sd; return p; }
    if (d.type == "lognormal") {
        const double s2 = std::log1p((sd * sd) / (mean * mean));
        p.type = PD_LOGNORMAL;
        p.sigma = std::sqrt(s2);
        p.half_s2 = 0.5 * s2;
        return p;
    }
    throw std::runtime_error("unknown distribution type '" + d.type + "'");
}

static std::shared_ptr<const ProvSpec> build_spec(const Settings& s) {
    auto spec = std::make_shared<ProvSpec>();
    ProvDesc& d = spec->base;
    d = ProvDesc{};
    d.seed = s.rq.seed;
    d.N = s.rq.n_particles;
    d.problem = s.rq.problem;
    d.aniso = s.particles[0].aniso;
    d.n_types = (uint32_t)s.particles.size();
    d.discrete = s.discrete ? 1u : 0u;
    double acc = 0.0;
    uint64_t acc_n = 0;
    for (const ParticleFile& pf : s.particles) {
        if (pf.aniso != d.aniso) throw std::runtime_error("all particle types of one run must share the anisotropy type");
        ProvType t{};
        t.rad = prov_dist(pf.rad, pf.radius);
        t.shell = prov_dist(pf.shell, pf.r_h - pf.radius);
        t.K = prov_dist(pf.aniso_dist, pf.K);
        t.K2 = prov_dist(pf.aniso_dist2, pf.K2);
        t.Ms = pf.Ms; t.alpha = pf.alpha; t.gamma = pf.gamma; t.T = pf.T; t.visc = pf.visc;
        for (int k = 0; k < 3; ++k) { t.euler[k] = pf.euler[k]; t.mag[k] = pf.mag[k]; t.urot[k] = pf.urot[k]; }
        t.rand_orient = pf.rand_orient; t.rand_mag = pf.rand_mag; t.rand_urot = pf.rand_urot;
        acc += (double)pf.count;
        acc_n += pf.count;
        t.cum = acc;
        t.cum_count = acc_n;
        spec->types.push_back(t);
    }
    d.total_count = acc;
    return spec;
}

// the N particles as host arrays (Provider::getProvidedObject, ParticleSimulationSettings::applySettingsToParticleProperties)
static void materialise(Settings& s) {
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.materialised) return;
    const uint64_t N = s.rq.n_particles;
    uint32_t dy = 0, dw = 0, dout = 0;
    request_dims(s.rq.problem, &dy, &dw, &dout);
    s.radius.resize(N); s.Ms.resize(N); s.alpha.resize(N); s.gamma.resize(N); s.T.resize(N); s.K.resize(N);
    s.r_h.resize(N); s.visc.resize(N); s.axes.resize(N * 9); s.y0.resize(N * dy);
    const bool uc = s.spec->base.aniso == SP_ANISO_UNIAXIALCUBIC;
    if (uc) { s.K2.resize(N); s.urot.resize(N * 3); }
    ProvDesc d = s.spec->base;
    ParticleProps q;
    for (uint64_t i = 0; i < N; ++i) {
        provide_particle(d, s.spec->types.data(), i, q);
        s.radius[i] = q.radius; s.Ms[i] = q.Ms; s.alpha[i] = q.alpha; s.gamma[i] = q.gamma; s.T[i] = q.T;
        s.K[i] = q.K; s.r_h[i] = q.r_h; s.visc[i] = q.visc;
        if (uc) { s.K2[i] = q.K2; for (int k = 0; k < 3; ++k) s.urot[i * 3 + k] = q.urot[k]; }
        for (int k = 0; k < 9; ++k) s.axes[i * 9 + k] = q.axes[k];
        for (uint32_t k = 0; k < dy; ++k) s.y0[i * dy + k] = q.y0[k];
    }
    s.materialised = true;
}

// particle p's properties: from the host arrays when they exist, else drawn on the host for this particle
static ParticleProps particle_at(Settings& s, uint64_t p) {
    ParticleProps q{};
    bool have;
    {
        std::lock_guard<std::mutex> lk(s.mu);
        have = s.materialised;
    }
    if (have) {
        q.radius = s.radius[p]; q.Ms = s.Ms[p]; q.alpha = s.alpha[p]; q.gamma = s.gamma[p]; q.T = s.T[p];
        q.K = s.K[p]; q.r_h = s.r_h[p]; q.visc = s.visc[p];
    } else {
        provide_particle(s.spec->base, s.spec->types.data(), p, q);
    }
    return q;
}

Settings* load(const std::string& path) {
    const ini::File f = ini::parse_file(path);
    std::unique_ptr<Settings> s(new Settings());
    sp_request& r = s->rq;
    std::memset(&r, 0, sizeof(r));
    // [Solver_Settings] (SolverSettings.h:106-142)
    const std::string* sv = opt(f, "Solver_Settings", {"Solver"});
    if (!sv) throw std::runtime_error("missing [Solver_Settings] Solver");
    r.solver = parse_enum_or_throw("solver", *sv);
    r.newton_max_iter = (uint32_t)num(f, "Solver_Settings", {"Max_Iterations"}, 100);
    r.newton_tol = num(f, "Solver_Settings", {"Accuracy_Goal"}, 1e-10);
    // [Simulation_Settings] (SimulationSettings.h:67-79)
    const std::string SIM = "Simulation_Settings";
    r.dt = num(f, SIM, {"Timestep"}, 0, true);
    r.n_steps = (uint64_t)num(f, SIM, {"Number_of_steps"}, 0, true);
    r.oversampling = (uint64_t)num(f, SIM, {"Oversampling"}, 1);
    r.n_particles = (uint64_t)num(f, SIM, {"Number_of_simulations"}, 0, true);
    r.seed = (uint64_t)num(f, SIM, {"Seed"}, 1);
    r.precision = SP_PREC_FP64;
    if (const std::string* pv = opt(f, SIM, {"Precision"})) r.precision = (*pv == "mixed") ? SP_PREC_MIXED : SP_PREC_FP64;
    r.check_tol = 1e-3;
    // [Problem_Settings] (ProblemSettings.h:55-237)
    const std::string PS = "Problem_Settings";
    const std::string* pt = opt(f, PS, {"Type_of_problem", "Problem_type"});
    if (!pt) throw std::runtime_error("missing [Problem_Settings] Type_of_problem");
    r.problem = parse_enum_or_throw("problem", *pt);
    if (flag(f, PS, {"Use_simple_model"}, false)) r.flags |= SP_FLAG_BN_SIMPLE_MODEL;
    if (flag(f, PS, {"Use_coordinate_transformation", "Use_spherical_coordinate_transformation"}, false))
        r.flags |= SP_FLAG_COORD_TRANSFORM;
    r.min_angle = num(f, PS, {"Min_angle_transformation", "Neel_Min_angle_before_transformation"}, 0.0);
    // BrownAndNeelEulerSpherical (ProblemSettings.h:146-159): the Brown part's Euler chart
    if (flag(f, PS, {"Use_euler_coordinate_transformation"}, false)) r.flags |= SP_FLAG_EULER_TRANSFORM;
    r.brown_min_angle = num(f, PS, {"Brown_Min_angle_before_transformation"}, 0.0);
    // [Field_Parameters] (FieldProperties.h:61, Properties/Fields/*.hpp)
    std::string FS = f.has_section("Field_Parameters") ? "Field_Parameters" : "Field_Properties";
    const std::string* ft = opt(f, FS, {"Type_of_field"});
    r.field = ft ? parse_enum_or_throw("field", *ft) : SP_FIELD_ZERO;
    auto fill3 = [](double* dst, const std::vector<double>& v, bool broadcast) {
        for (int i = 0; i < 3; ++i) dst[i] = i < (int)v.size() ? v[i] : (broadcast && v.size() == 1 ? v[0] : 0.0);
    };
    if (r.field == SP_FIELD_SINUSOIDAL || r.field == SP_FIELD_LISSAJOUS) {
        fill3(r.field_offset, vec(f, FS, {"Offset_Field"}), false);
        fill3(r.field_amplitude, vec(f, FS, {"Amplitude"}), false);
        if (r.field == SP_FIELD_SINUSOIDAL) {
            r.field_freq[0] = num(f, FS, {"Frequency"}, 0.0);
            r.field_phase[0] = num(f, FS, {"Phase"}, 0.0);
        } else {
            const auto fr = vec(f, FS, {"Frequencies"});
            if (fr.size() != 3) throw std::runtime_error("LissajousField: Frequencies needs 3 values");   // LissajousField.h:54-57
            fill3(r.field_freq, fr, false);
            const auto ph = vec(f, FS, {"Phases"});
            if (!(ph.size() == 1 || ph.size() == 3)) throw std::runtime_error("LissajousField: Phases needs 1 or 3 values");
            fill3(r.field_phase, ph, true);
        }
    } else if (r.field == SP_FIELD_CONSTANT) {                                   // Properties/Fields/Constant.hpp:29-32
        fill3(r.field_offset, vec(f, FS, {"Offset_Field"}), false);
    } else if (r.field == SP_FIELD_TRIANGULAR || r.field == SP_FIELD_RECTANGULAR || r.field == SP_FIELD_SINC) {
        // Triangular.hpp:30-36, Rectangular.hpp:32-40, Sinc.hpp:30-36
        fill3(r.field_offset, vec(f, FS, {"Offset_Field"}), false);
        fill3(r.field_amplitude, vec(f, FS, {"Amplitude"}), false);
        r.field_period = num(f, FS, {"Periode"}, 0.0);
        r.field_time_offset = num(f, FS, {"TimeOffset"}, 0.0);
        if (r.field == SP_FIELD_RECTANGULAR) {
            r.field_tau = num(f, FS, {"Tau"}, 0.0);
            r.field_alternating = flag(f, FS, {"Alternating"}, false) ? 1u : 0u;
        }
    }
    // [Result_Settings] (ResultSettings.h:90-118)
    const std::string RS = "Result_Settings";
    if (const std::string* sf = opt(f, RS, {"Savefile"})) s->savefile = *sf;
    if (flag(f, RS, {"Save_single_simulations"}, false)) {
        r.save_interval = (uint32_t)std::max(1.0, num(f, RS, {"Save_intervall"}, 1.0));
        if (const std::string* v = opt(f, RS, {"Single_simulations_savefile"})) s->single_savefile = *v;
        if (const std::string* v = opt(f, RS, {"Save_single_name_prefix"})) s->single_prefix = *v;
    }
    // [Particle_Provider] (ParticleProvider.h:230-295)
    const std::string PP = "Particle_Provider";
    s->discrete = flag(f, PP, {"Use_discrete_distribution_to_select_particle"}, false);
    auto list = f.sections.find(PP + ".Particle_List");
    auto nums = f.sections.find(PP + ".Particle_Numbers");
    if (list == f.sections.end()) throw std::runtime_error("missing [Particle_Provider.Particle_List]");
    for (auto& [name, file] : list->second) {
        ParticleFile pf = load_particle(f, name, file);
        if (nums != f.sections.end() && nums->second.count(name)) pf.count = (uint64_t)ini::to_double(nums->second.at(name));
        s->particles.push_back(pf);
    }
    if (s->particles.empty()) throw std::runtime_error("empty particle list");
    uint32_t dy = 0, dw = 0, dout = 0;
    if (request_dims(r.problem, &dy, &dw, &dout)) throw std::runtime_error("problem type not available on the GPU backend");
    s->spec = build_spec(*s);
    r.anisotropy = s->spec->base.aniso;
    return s.release();
}

static void write_array(std::ostream& o, const double* v, size_t n) {
    o << '[';
    char buf[40];
    for (size_t i = 0; i < n; ++i) { std::snprintf(buf, sizeof buf, "%.17g", v[i]); o << (i ? "," : "") << buf; }
    o << ']';
}

// Mean_Results {Time, Field, Results, Weight, Mean_Properties} (MeanSimulationResult.h:131-141), as JSON
int write_result(const Settings& s, const sp_result& res, const std::string& path) {
    std::ofstream o(path);
    if (!o) return set_error(SP_ERR_INVALID_ARG, "cannot write " + path);
    o << "{\"Mean_Results\":{\"Time\":";
    write_array(o, res.time, res.n_out);
    o << ",\"Field\":[";
    for (uint64_t k = 0; k < res.n_out; ++k) { if (k) o << ','; write_array(o, res.field + 3 * k, 3); }
    o << "],\"Results\":[";
    for (uint64_t k = 0; k < res.n_out; ++k) { if (k) o << ','; write_array(o, res.mean + res.dim_out * k, res.dim_out); }
    o << "],\"Weight\":";
    write_array(o, res.weight, res.dim_out);
    char buf[512];
    std::snprintf(buf, sizeof buf,
                  ",\"Mean_Properties\":{\"Temperature\":%.17g,\"Magnetic_radius\":%.17g,\"Saturation_magnetisation\":%.17g,"
                  "\"Damping_constant\":%.17g,\"Gyromagnetic_ratio\":%.17g,\"Anisotropy_constant\":%.17g,"
                  "\"Hydrodynamic_Radius\":%.17g,\"Viscosity\":%.17g}}",
                  res.particle_means[0], res.particle_means[1], res.particle_means[2], res.particle_means[3],
                  res.particle_means[4], res.particle_means[5], res.particle_means[6], res.particle_means[7]);
    o << buf;
    std::snprintf(buf, sizeof buf,
                  ",\"Engine\":{\"device_ms\":%.6f,\"particle_steps\":%.17g,\"precision_used\":%u,\"check_error\":%.6g,"
                  "\"check_noise\":%.6g,\"check_bias\":%.6g,\"check_bias_se\":%.6g,\"used_fallback\":%u}}",
                  res.device_ms, res.particle_steps, res.precision_used, res.check_error, res.check_noise, res.check_bias,
                  res.check_bias_se, res.used_fallback);
    o << buf << '\n';
    return o ? SP_OK : set_error(SP_ERR_INVALID_ARG, "write failed: " + path);
}

// Single-simulation results (SimulationResultManager.h:115-149, SingleSimulationResult::save :96-104):
// one named entry per kept particle with its properties, Time, Weight, Field and its own (unweighted)
// window means.  The entry of particle p is named <Save_single_name_prefix>_<p + 1>: the reference numbers
// its entries by completion order, which has no fixed meaning; here the number is the particle's.
int write_single_results(Settings& s, const sp_result& res, const std::string& path) {
    if (!res.trajectories || res.n_saved == 0) return set_error(SP_ERR_INVALID_ARG, "the result holds no single-simulation outputs");
    std::ofstream o(path);
    if (!o) return set_error(SP_ERR_INVALID_ARG, "cannot write " + path);
    const uint64_t k = s.rq.save_interval, S = res.n_saved, D = res.dim_out;
    char buf[512];
    o << '{';
    std::vector<double> row(D);
    for (uint64_t j = 0; j < S; ++j) {
        const uint64_t p = (j + 1) * k - 1;
        const ParticleProps q = particle_at(s, p);
        const double V = 4.0 / 3.0 * 3.14159265358979323846 * q.radius * q.radius * q.radius;
        const double mu = q.Ms * V;
        std::snprintf(buf, sizeof buf,
                      "%s\"%s_%llu\":{\"UsedProperties\":{\"Temperature\":%.17g,\"Magnetic_radius\":%.17g,"
                      "\"Saturation_magnetisation\":%.17g,\"Damping_constant\":%.17g,\"Gyromagnetic_ratio\":%.17g,"
                      "\"Anisotropy_constant\":%.17g,\"Hydrodynamic_Radius\":%.17g,\"Viscosity\":%.17g},\"Time\":",
                      j ? "," : "", s.single_prefix.c_str(), (unsigned long long)(p + 1), q.T, q.radius, q.Ms,
                      q.alpha, q.gamma, q.K, q.r_h, q.visc);
        o << buf;
        write_array(o, res.time, res.n_out);
        for (uint64_t c = 0; c < D; ++c) row[c] = (D == 6 && c < 3) ? 1.0 : mu;   // getWeighting of the problem
        o << ",\"Weight\":";
        write_array(o, row.data(), D);
        o << ",\"Field\":[";
        for (uint64_t w = 0; w < res.n_out; ++w) { if (w) o << ','; write_array(o, res.field + 3 * w, 3); }
        o << "],\"Results\":[";
        for (uint64_t w = 0; w < res.n_out; ++w) {
            for (uint64_t c = 0; c < D; ++c) row[c] = res.trajectories[(w * D + c) * S + j];
            if (w) o << ',';
            write_array(o, row.data(), D);
        }
        o << "]}";
    }
    o << "}\n";
    return o ? SP_OK : set_error(SP_ERR_INVALID_ARG, "write failed: " + path);
}

}  // namespace sp::frontend

struct sp_settings {
    sp::frontend::Settings* s = nullptr;
};

extern "C" {

int sp_settings_load(const char* ini_path, sp_settings** out) {
    if (!ini_path || !out) return sp::set_error(SP_ERR_INVALID_ARG, "NULL argument");
    try {
        sp_settings* h = new sp_settings();
        h->s = sp::frontend::load(ini_path);
        *out = h;
        return SP_OK;
    } catch (const std::exception& e) {
        return sp::set_error(SP_ERR_INVALID_ARG, e.what());
    }
}

int sp_settings_request(const sp_settings* h, sp_request* req, sp_particles* parts) {
    if (!h || !h->s) return sp::set_error(SP_ERR_INVALID_ARG, "NULL settings");
    auto& s = *h->s;
    if (req) *req = s.rq;
    if (parts) {
        sp::frontend::materialise(s);
        parts->magnetic_radius = s.radius.data();
        parts->saturation_magnetisation = s.Ms.data();
        parts->damping = s.alpha.data();
        parts->gyromagnetic_ratio = s.gamma.data();
        parts->temperature = s.T.data();
        parts->anisotropy_constant = s.K.data();
        parts->hydrodynamic_radius = s.r_h.data();
        parts->viscosity = s.visc.data();
        parts->axes = s.axes.data();
        parts->initial_state = s.y0.data();
        parts->anisotropy_constant_cubic = s.K2.empty() ? nullptr : s.K2.data();
        parts->uniaxial_rotation = s.urot.empty() ? nullptr : s.urot.data();
    }
    return SP_OK;
}

int sp_settings_prepare(sp_ctx* ctx, const sp_settings* h, sp_prepared** out) {
    if (!ctx || !h || !h->s || !out) return sp::set_error(SP_ERR_INVALID_ARG, "NULL argument");
    const auto& s = *h->s;
    std::unique_ptr<sp_prepared> p(new sp_prepared());
    p->ctx = ctx;
    if (int e = sp::prepare_provided(&s.rq, s.spec, p->pr.hr)) return e;
    p->pr.keep_final = (s.rq.flags & SP_FLAG_KEEP_FINAL_STATE) != 0;
    {
        std::lock_guard<std::mutex> lk(ctx->run_mu);
        if (int e = sp::prepare_plain(ctx, p->pr)) return e;
    }
    *out = p.release();
    return SP_OK;
}

int sp_settings_submit(sp_ctx* ctx, const sp_settings* h, uint64_t* id) {
    if (!ctx || !h || !h->s || !id) return sp::set_error(SP_ERR_INVALID_ARG, "NULL argument");
    const auto& s = *h->s;
    sp::HostRequest hr;
    if (int e = sp::prepare_provided(&s.rq, s.spec, hr)) return e;
    return sp::submit_request(ctx, std::move(hr), (s.rq.flags & SP_FLAG_KEEP_FINAL_STATE) != 0, id);
}

const char* sp_settings_single_savefile(const sp_settings* h) { return (h && h->s) ? h->s->single_savefile.c_str() : nullptr; }

int sp_settings_write_single_results(const sp_settings* h, const sp_result* res, const char* path) {
    if (!h || !h->s || !res || !path) return sp::set_error(SP_ERR_INVALID_ARG, "NULL argument");
    return sp::frontend::write_single_results(*h->s, *res, path);
}

const char* sp_settings_savefile(const sp_settings* h) { return (h && h->s) ? h->s->savefile.c_str() : nullptr; }

int sp_settings_write_result(const sp_settings* h, const sp_result* res, const char* path) {
    if (!h || !h->s || !res || !path) return sp::set_error(SP_ERR_INVALID_ARG, "NULL argument");
    return sp::frontend::write_result(*h->s, *res, path);
}

void sp_settings_free(sp_settings* h) {
    if (!h) return;
    delete h->s;
    delete h;
}

}  // extern "C"
