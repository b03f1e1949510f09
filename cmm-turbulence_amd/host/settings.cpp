// Settings, COMMAND=VALUE parsing and the parameter file of a CMM run.
// Keys, presets and formats follow Cuda-CMM/src/ui/settings.cu and Cuda-CMM/src/ui/cmm-param.cu.
#include "settings.hpp"

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

namespace cmmhost {

void fail(const std::string &msg, int code) {
    std::cout << msg << std::endl;
    std::exit(code);
}

bool parse_bool(const std::string &v) { return v == "true" || v == "True" || v == "1"; }

std::string fmt16(double v) {
    std::ostringstream os;
    os.precision(16);
    os << v;
    return os.str();
}

std::string fmt3(double v) {
    std::ostringstream os;
    os.precision(3);
    os << v;
    return os.str();
}

namespace {

template <class T>
std::string s16(const T &v) {
    std::ostringstream os;
    os.precision(16);
    os << v;
    return os.str();
}

// "{a=1,b=2}" or "\"a=1,b=2\"" -> calls f(key, value) for each comma-separated key=value
template <class F>
void for_each_pair(std::string s, F f) {
    if (s.size() < 2) return;
    bool braces = s.front() == '{' && s.back() == '}';
    bool quotes = s.front() == '"' && s.back() == '"';
    if (!braces && !quotes) return;
    s = s.substr(1, s.size() - 2);
    size_t pos = 0;
    while (true) {
        size_t next = s.find(',', pos);
        std::string item = s.substr(pos, next == std::string::npos ? std::string::npos : next - pos);
        size_t eq = item.find('=');
        if (eq != std::string::npos) f(item.substr(0, eq), item.substr(eq + 1));
        if (next == std::string::npos) break;
        pos = next + 1;
    }
}

// "{v1,v2,...}" -> up to n doubles (Cuda-CMM/src/ui/settings.h:29-52)
void parse_array(std::string s, double *out, int n) {
    if (s.empty() || (s.front() != '{' && s.front() != '"') || (s.back() != '}' && s.back() != '"')) return;
    s = s.substr(1, s.size() - 2);
    std::stringstream ss(s);
    double v;
    int i = 0;
    while (i < n && ss >> v) {
        out[i++] = v;
        if (ss.peek() == ',') ss.ignore();
    }
}

std::string array_string(const double *a, int n) {
    std::stringstream ss;
    ss << "{";
    for (int i = 0; i < n; ++i) ss << a[i] << (i < n - 1 ? "," : "");
    ss << "}";
    return ss.str();
}

int particle_scheme(const std::string &v) {
    // particle names, Cuda-CMM/src/ui/settings.cu:733-744 (RK2 and Heun are 20 for particles)
    if (v == "EulerExp" || v == "RK1") return 10;
    if (v == "Heun" || v == "RK2") return 20;
    if (v == "RK3") return 30;
    if (v == "RK4") return 40;
    if (v == "RK3Mod") return 31;
    if (v == "RK4Mod") return 41;
    return -1;
}

int particle_init(const std::string &v) {
    if (v == "uniform") return 0;
    if (v == "normal" || v == "gaussian") return 1;
    if (v == "circular_ring") return 2;
    if (v == "uniform_grid") return 3;
    if (v == "sine_sheets") return 4;
    return -1;
}

}  // namespace

void SaveComputational::set_all(const std::string &b) {
    for_each_pair(b, [&](const std::string &k, const std::string &v) {
        if (k == "is_instant") is_instant = parse_bool(v);
        else if (k == "time_start") time_start = std::stod(v);
        else if (k == "time_end") time_end = std::stod(v);
        else if (k == "time_step") time_step = std::stod(v);
        else if (k == "var") var = v;
        else if (k == "conv") conv = parse_bool(v);
    });
}
std::string SaveComputational::str() const {
    std::ostringstream os;
    os << "{is_instant=" << s16(is_instant) << ",time_start=" << s16(time_start);
    if (!is_instant) os << ",time_end=" << s16(time_end) << ",time_step=" << s16(time_step);
    os << ",var=" << var << ",conv=" << s16(conv) << "}";
    return os.str();
}

void SaveSample::set_all(const std::string &b) {
    for_each_pair(b, [&](const std::string &k, const std::string &v) {
        if (k == "is_instant") is_instant = parse_bool(v);
        else if (k == "time_start") time_start = std::stod(v);
        else if (k == "time_end") time_end = std::stod(v);
        else if (k == "time_step") time_step = std::stod(v);
        else if (k == "var") var = v;
        else if (k == "grid") grid = std::stoi(v);
    });
}
std::string SaveSample::str() const {
    std::ostringstream os;
    os << "{is_instant=" << s16(is_instant) << ",time_start=" << s16(time_start);
    if (!is_instant) os << ",time_end=" << s16(time_end) << ",time_step=" << s16(time_step);
    os << ",var=" << var << ",grid=" << s16(grid) << "}";
    return os.str();
}

void SaveZoom::set_all(const std::string &b) {
    for_each_pair(b, [&](const std::string &k, const std::string &v) {
        if (k == "is_instant") is_instant = parse_bool(v);
        else if (k == "time_start") time_start = std::stod(v);
        else if (k == "time_end") time_end = std::stod(v);
        else if (k == "time_step") time_step = std::stod(v);
        else if (k == "var") var = v;
        else if (k == "grid") grid = std::stoi(v);
        else if (k == "pos_x") pos_x = std::stod(v);
        else if (k == "pos_y") pos_y = std::stod(v);
        else if (k == "width_x") width_x = std::stod(v);
        else if (k == "width_y") width_y = std::stod(v);
        else if (k == "rep") rep = std::stoi(v);
        else if (k == "rep_fac") rep_fac = std::stod(v);
    });
}
std::string SaveZoom::str() const {
    std::ostringstream os;
    os << "{is_instant=" << s16(is_instant) << ",time_start=" << s16(time_start);
    if (!is_instant) os << ",time_end=" << s16(time_end) << ",time_step=" << s16(time_step);
    os << ",var=" << var << ",grid=" << s16(grid);
    // the reference prints pos_y with the stream's default precision
    os << ",pos_x=" << s16(pos_x) << ",pos_y=" << pos_y << ",width_x=" << s16(width_x) << ",width_y=" << s16(width_y);
    os << ",rep=" << s16(rep) << ",rep_fac=" << s16(rep_fac) << "}";
    return os.str();
}

void ParticlesAdvected::set_all(const std::string &b) {
    for_each_pair(b, [&](const std::string &k, const std::string &v) {
        if (k == "num") num = std::stoi(v);
        else if (k == "tau") tau = std::stod(v);
        else if (k == "seed") seed = std::stoull(v);
        else if (k == "time_integration") { time_integration = v; time_integration_num = particle_scheme(v); }
        else if (k == "init_name") { init_name = v; init_num = particle_init(v); }
        else if (k == "init_time") init_time = std::stod(v);
        else if (k == "init_vel") init_vel = parse_bool(v);
        else if (k == "init_param_1") p1 = std::stod(v);
        else if (k == "init_param_2") p2 = std::stod(v);
        else if (k == "init_param_3") p3 = std::stod(v);
        else if (k == "init_param_4") p4 = std::stod(v);
    });
}
std::string ParticlesAdvected::str() const {
    std::ostringstream os;
    os << "{num=" << s16(num) << ",tau=" << s16(tau) << ",seed=" << s16(seed);
    os << ",time_integration=" << time_integration << ",init_name=" << init_name << ",init_time=" << s16(init_time)
       << ",init_vel=" << s16(init_vel);
    os << ",init_param_1=" << s16(p1) << ",init_param_2=" << s16(p2) << ",init_param_3=" << s16(p3)
       << ",init_param_4=" << s16(p4) << "}";
    return os.str();
}

void ParticlesForwarded::set_all(const std::string &b) {
    for_each_pair(b, [&](const std::string &k, const std::string &v) {
        if (k == "num") num = std::stoi(v);
        else if (k == "seed") seed = std::stoull(v);
        else if (k == "init_name") { init_name = v; init_num = particle_init(v); }
        else if (k == "init_time") init_time = std::stod(v);
        else if (k == "init_param_1") p1 = std::stod(v);
        else if (k == "init_param_2") p2 = std::stod(v);
        else if (k == "init_param_3") p3 = std::stod(v);
        else if (k == "init_param_4") p4 = std::stod(v);
    });
}
std::string ParticlesForwarded::str() const {
    std::ostringstream os;
    os << "{num=" << s16(num) << ",seed=" << s16(seed) << ",init_name=" << init_name << ",init_time=" << s16(init_time);
    os << ",init_param_1=" << s16(p1) << ",init_param_2=" << s16(p2) << ",init_param_3=" << s16(p3)
       << ",init_param_4=" << s16(p4) << "}";
    return os.str();
}

void Settings::set_time_integration(const std::string &v) {
    // Cuda-CMM/src/ui/settings.h:386-398: the Lagrange order is raised to the scheme's need
    time_integration = v;
    int need = 0;
    if (v == "EulerExp" || v == "RK1") { time_integration_num = 10; need = 1; }
    else if (v == "Heun" || v == "RK2") { time_integration_num = 21; need = 2; }
    else if (v == "AB2") { time_integration_num = 20; need = 2; }
    else if (v == "RK3") { time_integration_num = 30; need = 3; }
    else if (v == "RK4") { time_integration_num = 40; need = 4; }
    else if (v == "RK3Mod") { time_integration_num = 31; need = 3; }
    else if (v == "RK4Mod") { time_integration_num = 41; need = 4; }
    else time_integration_num = -1;
    if (lagrange_order < need) lagrange_order = need;
}

void Settings::set_initial_condition(const std::string &v) {
    // Cuda-CMM/src/ui/settings.h:297-332, plus the conditions of this engine (100, 101)
    initial_condition = v;
    if (simulation_type == "cmm_vlasov_poisson_1d") {
        if (v == "landau_damping") initial_condition_num = 0;
        else if (v == "two_stream") initial_condition_num = 1;
        else if (v == "nonlin_landau_damping") initial_condition_num = 2;
        else {
            initial_condition = "zero";
            initial_condition_num = -1;
            fail("Ok, try it again please with an existing initial condition.! Currently: " + v, 230304);
        }
        return;
    }
    static const char *names[] = {"4_nodes", "quadropole", "one_vortex", "two_votices", "three_vortices",
                                  "single_shear_layer", "tanh_shear_layer", "turbulence_gaussienne",
                                  "gaussian_blobs", "shielded_vortex", "two_cosine", "vortex_sheets"};
    for (int i = 0; i < 12; ++i)
        if (v == names[i]) { initial_condition_num = i; return; }
    if (v == "double_shear_layer") { initial_condition_num = 100; return; }
    if (v == "random_spectrum") { initial_condition_num = 101; return; }
    initial_condition = "zero";
    initial_condition_num = -1;
    fail("Ok, try it again please with an existing initial condition.! Currently: " + v, 230305);
}

void Settings::set_map_update_order(const std::string &v) {
    map_update_order = v;
    map_update_order_num = v == "2nd" ? 0 : v == "4th" ? 1 : v == "6th" ? 2 : -1;
}

void Settings::set_domain_bounds(const double *b) {
    for (int i = 0; i < 6; ++i) domain_bounds[i] = b[i];
    double lx = b[1] - b[0], ly = b[3] - b[2];
    if (!(lx > 1e-16 && ly > 1e-16))
        fail("Bonjour, there might be a problem in the engine room! We need your help! The domain bounds are not correct!",
             20230517);
}

Settings::Settings() {
    workspace = "./";
    simulation_type = "cmm_euler_2d";
    sim_name = "debug";
    grid_coarse = 512;
    grid_fine = 2048;
    grid_psi = 1024;
    grid_vort = grid_fine;
    initial_condition_num = 0;
    for (double &p : initial_params) p = 0;
    set_initial_condition("4_nodes");
    parse_array("{0}", initial_params, 10);
    initial_discrete = false;
    initial_discrete_grid = 2048;
    initial_discrete_location =
        "data/final-state-anticythere-1_gaussian_blobs_C1024_F8192_t3.07e+03_T500/Time_data/Time_270/Vorticity_W_4096.data";
    verbose = 3;
    double b[6] = {0, kTwoPi, 0, kTwoPi, 0, 0};
    set_domain_bounds(b);
    final_time = 3;
    set_dt_by_steps = true;
    factor_dt_by_grid = 1;
    steps_per_sec = 20;
    const std::string tmax = fmt16(kTMax);
    std::vector<std::string> cs = {
        "{is_instant=1,time_start=0,var=W-U,conv=1}",
        "{is_instant=1,time_start=" + tmax + ",var=W-U,conv=1}",
        "{is_instant=0,time_start=0,time_end=" + tmax + ",time_step=0.5,var ,conv=1}",
        "{is_instant=0,time_start=0,time_end=" + tmax + ",time_step=0.5,var=PartA_01-PartA_02-PartA_03-PartA_04,conv=0}"};
    for (auto &c : cs) { SaveComputational s; s.set_all(c); save_computational.push_back(s); }
    incomp_threshold = 1e-5;
    map_epsilon = 1e-3;
    skip_remapping = false;
    mem_RAM_CPU_remaps = 9000;
    save_map_stack = true;
    restart_time = 0;
    restart_location = "data/debug_4_nodes_C512_F2048_t20_T2.1/MapStack";
    lagrange_order = 0;
    set_time_integration("RK3");
    lagrange_override = -1;
    lagrange_init_higher_order = true;
    set_map_update_order("4th");
    molly_stencil = 0;
    freq_cut_psi = 512 / 4.0;  // preset rule grid_coarse/4 evaluated with the preset grid (128)
    { SaveSample s; s.set_all("{is_instant=0,time_start=0,time_end=" + tmax + ",time_step=0.5,var=W,grid=1024}"); save_sample.push_back(s); }
    scalar_name = "rectangle";
    scalar_discrete = false;
    scalar_discrete_grid = 2048;
    scalar_discrete_location = "I-am-not-implemented-yet!";
    forward_map = false;
    const std::string pi = fmt16(kPi), twopi = fmt16(kPi * 2.0);
    std::vector<std::string> ps = {
        "{num=100000,tau=0,seed=0,time_integration=RK3,init_name=sine_sheets,init_time=0,init_vel=0,init_param_1=" + pi +
            ",init_param_2=" + pi + ",init_param_3=" + twopi + ",init_param_4=" + twopi + "}",
        "{num=100000,tau=0.1,seed=0,time_integration=RK3,init_name=uniform,init_time=0,init_vel=0,init_param_1=" + pi +
            ",init_param_2=" + pi + ",init_param_3=" + twopi + ",init_param_4=" + twopi + "}"};
    for (auto &p : ps) { ParticlesAdvected a; a.set_all(p); particles_advected.push_back(a); }
    particles_steps = -1;
    seed = 0;
    map_precision = "fp64";
}

int Settings::set_variable(const std::string &line, const std::string &delim) {
    // Cuda-CMM/src/ui/settings.cu:347-436
    size_t pos = line.find(delim);
    if (pos == std::string::npos) return 0;
    const std::string c = line.substr(0, pos), v = line.substr(pos + delim.size());
    if (c == "workspace") workspace = v;
    else if (c == "simulation_type") simulation_type = v;
    else if (c == "sim_name") sim_name = v;
    else if (c == "grid_coarse") grid_coarse = std::stoi(v);
    else if (c == "grid_fine") grid_fine = std::stoi(v);
    else if (c == "grid_psi") grid_psi = std::stoi(v);
    else if (c == "grid_vort") grid_vort = std::stoi(v);
    else if (c == "domain_bounds") { double b[6] = {0, 0, 0, 0, 0, 0}; parse_array(v, b, 6); set_domain_bounds(b); }
    else if (c == "final_time") final_time = std::stod(v);
    else if (c == "factor_dt_by_grid") factor_dt_by_grid = std::stod(v);
    else if (c == "steps_per_sec") steps_per_sec = std::stoi(v);
    else if (c == "set_dt_by_steps") set_dt_by_steps = parse_bool(v);
    else if (c == "save_computational_num") { save_computational.assign(std::stoi(v), SaveComputational()); return 1; }
    else if (c == "mem_RAM_CPU_remaps") mem_RAM_CPU_remaps = std::stoi(v);
    else if (c == "save_map_stack") save_map_stack = parse_bool(v);
    else if (c == "restart_time") restart_time = std::stod(v);
    else if (c == "restart_location") restart_location = v;
    else if (c == "verbose") verbose = std::stoi(v);
    else if (c == "initial_condition") set_initial_condition(v);
    else if (c == "initial_params") parse_array(v, initial_params, 10);
    else if (c == "initial_discrete") initial_discrete = parse_bool(v);
    else if (c == "initial_discrete_grid") initial_discrete_grid = std::stoi(v);
    else if (c == "initial_discrete_location") initial_discrete_location = v;
    else if (c == "incomp_threshold") incomp_threshold = std::stod(v);
    else if (c == "map_epsilon") map_epsilon = std::stod(v);
    else if (c == "time_integration") set_time_integration(v);
    else if (c == "lagrange_override") lagrange_override = std::stoi(v);
    else if (c == "lagrange_init_higher_order") lagrange_init_higher_order = parse_bool(v);
    else if (c == "map_update_order") set_map_update_order(v);
    else if (c == "molly_stencil") molly_stencil = std::stoi(v);
    else if (c == "freq_cut_psi") freq_cut_psi = std::stod(v);
    else if (c == "skip_remapping") skip_remapping = parse_bool(v);
    else if (c == "save_sample_num") { save_sample.assign(std::stoi(v), SaveSample()); return 2; }
    else if (c == "scalar_name") scalar_name = v;
    else if (c == "scalar_discrete") scalar_discrete = parse_bool(v);
    else if (c == "scalar_discrete_grid") scalar_discrete_grid = std::stoi(v);
    else if (c == "scalar_discrete_location") scalar_discrete_location = v;
    else if (c == "save_zoom_num") { save_zoom.assign(std::stoi(v), SaveZoom()); return 3; }
    else if (c == "forward_map") forward_map = parse_bool(v);
    else if (c == "particles_forwarded_num") { particles_forwarded.assign(std::stoi(v), ParticlesForwarded()); return 5; }
    else if (c == "particles_advected_num") { particles_advected.assign(std::stoi(v), ParticlesAdvected()); return 4; }
    else if (c == "particles_steps") particles_steps = std::stoi(v);
    else if (c == "seed") seed = std::stoull(v);
    else if (c == "map_precision") map_precision = v;
    return 0;
}

void Settings::apply_commands(int argc, char **argv) {
    // COMMAND=VALUE from the command line; a block count takes the following "name=value" arguments
    for (int i = 0; i < argc; ++i) {
        std::string a = argv[i];
        size_t eq = a.find('=');
        if (eq == std::string::npos) continue;
        const std::string c = a.substr(0, eq), v = a.substr(eq + 1);
        if (c == "save_computational") { SaveComputational s; s.set_all(v); save_computational.push_back(s); continue; }
        if (c == "save_sample") { SaveSample s; s.set_all(v); save_sample.push_back(s); continue; }
        if (c == "save_zoom") { SaveZoom s; s.set_all(v); save_zoom.push_back(s); continue; }
        if (c == "particles_advected") { ParticlesAdvected s; s.set_all(v); particles_advected.push_back(s); continue; }
        if (c == "particles_forwarded") { ParticlesForwarded s; s.set_all(v); particles_forwarded.push_back(s); continue; }
        int kind = set_variable(a, "=");
        // on the command line the block entries follow as separate arguments
        if (kind) {
            if (kind == 1) save_computational.clear();
            else if (kind == 2) save_sample.clear();
            else if (kind == 3) save_zoom.clear();
            else if (kind == 4) particles_advected.clear();
            else particles_forwarded.clear();
        }
    }
}

static const std::string kHeader = "CMM Parameter file";
static const std::string kDelim = "\t=\t";

bool load_param_file(Settings &s, const std::string &path) {
    std::ifstream f(path);
    if (!f) {
        std::cout << "Unable to read parameter file.\n";
        return false;
    }
    std::string line;
    std::getline(f, line);
    if (line != kHeader) return false;
    std::getline(f, line);
    size_t pos = line.find(kDelim);
    if (pos == std::string::npos || line.substr(0, pos) != "param_version") return false;
    int version = std::stoi(line.substr(pos + kDelim.size()));
    std::getline(f, line);  // separator line
    while (std::getline(f, line)) {
        int kind = s.set_variable(line, kDelim);
        auto value_of = [](const std::string &l) {
            size_t p = l.find(kDelim);
            return p == std::string::npos ? std::string() : l.substr(p + kDelim.size());
        };
        if (kind == 1) for (auto &e : s.save_computational) { std::getline(f, line); if (line.find(kDelim) != std::string::npos) e.set_all(value_of(line)); }
        else if (kind == 2) for (auto &e : s.save_sample) { std::getline(f, line); if (line.find(kDelim) != std::string::npos) e.set_all(value_of(line)); }
        else if (kind == 3) for (auto &e : s.save_zoom) { std::getline(f, line); if (line.find(kDelim) != std::string::npos) e.set_all(value_of(line)); }
        else if (kind == 4) for (auto &e : s.particles_advected) { std::getline(f, line); if (line.find(kDelim) != std::string::npos) e.set_all(value_of(line)); }
        else if (kind == 5) for (auto &e : s.particles_forwarded) { std::getline(f, line); if (line.find(kDelim) != std::string::npos) e.set_all(value_of(line)); }
    }
    if (version < 2) std::cout << "Param-file version < 2: Save settings cannot be set." << std::endl;
    if (version < 3)
        std::cout << "Param-file version < 3: Initial condition parameters cannot be set."
                     "Check cmm-init.cu if the initial conditions contains parameters that need to be set"
                  << std::endl;
    return true;
}

void save_param_file(const Settings &s, const std::string &path) {
    std::ofstream f(path);
    if (!f) {
        std::cout << "Unable to create parameter file.\n";
        return;
    }
    auto w = [&](const std::string &k, const std::string &v) { f << k << kDelim << v << "\n"; };
    f << kHeader << "\n";
    w("param_version", "3");
    f << "\n";
    w("workspace", s.workspace);
    w("simulation_type", s.simulation_type);
    w("sim_name", s.sim_name);
    w("grid_coarse", s16(s.grid_coarse));
    w("grid_fine", s16(s.grid_fine));
    w("grid_psi", s16(s.grid_psi));
    w("grid_vort", s16(s.grid_vort));
    w("domain_bounds", array_string(s.domain_bounds, 6));
    w("final_time", s16(s.final_time));
    w("set_dt_by_steps", s16(s.set_dt_by_steps));
    if (s.set_dt_by_steps) w("steps_per_sec", s16(s.steps_per_sec));
    else w("factor_dt_by_grid", s16(s.factor_dt_by_grid));
    w("save_computational_num", s16(s.save_computational.size()));
    for (auto &e : s.save_computational) w("save_computational", e.str());
    w("mem_RAM_CPU_remaps", s16(s.mem_RAM_CPU_remaps));
    w("save_map_stack", s16(s.save_map_stack));
    w("restart_time", s16(s.restart_time));
    if (s.restart_time != 0) w("restart_location", s.restart_location);
    w("verbose", s16(s.verbose));
    w("initial_condition", s.initial_condition);
    w("initial_params", array_string(s.initial_params, 10));
    w("initial_discrete", s16(s.initial_discrete));
    if (s.initial_discrete) {
        w("initial_discrete_grid", s16(s.initial_discrete_grid));
        w("initial_discrete_location", s.initial_discrete_location);
    }
    w("incomp_threshold", s16(s.incomp_threshold));
    w("map_epsilon", s16(s.map_epsilon));
    w("time_integration", s.time_integration);
    w("lagrange_override", s16(s.lagrange_override));
    w("lagrange_init_higher_order", s16(s.lagrange_init_higher_order));
    w("map_update_order", s.map_update_order);
    w("molly_stencil", s16(s.molly_stencil));
    w("freq_cut_psi", s16(s.freq_cut_psi));
    w("skip_remapping", s16(s.skip_remapping));
    w("save_sample_num", s16(s.save_sample.size()));
    for (auto &e : s.save_sample) w("save_sample", e.str());
    w("scalar_name", s.scalar_name);
    w("scalar_discrete", s16(s.scalar_discrete));
    if (s.scalar_discrete) {
        w("scalar_discrete_grid", s16(s.scalar_discrete_grid));
        w("scalar_discrete_location", s.scalar_discrete_location);
    }
    w("save_zoom_num", s16(s.save_zoom.size()));
    for (auto &e : s.save_zoom) w("save_zoom", e.str());
    w("forward_map", s16(s.forward_map));
    if (s.forward_map) {
        w("particles_forwarded_num", s16(s.particles_forwarded.size()));
        for (auto &e : s.particles_forwarded) w("particles_forwarded", e.str());
    }
    w("particles_advected_num", s16(s.particles_advected.size()));
    for (auto &e : s.particles_advected) w("particles_advected", e.str());
    if (!s.particles_advected.empty()) w("particles_steps", s16(s.particles_steps));
}

}  // namespace cmmhost
