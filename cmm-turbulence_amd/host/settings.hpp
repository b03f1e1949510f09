// Settings of a CMM run: presets, COMMAND=VALUE parsing, block settings and the parameter file.
// The keys, value formats, presets and file format are those of the reference
// (Cuda-CMM/src/ui/settings.{h,cu}, Cuda-CMM/src/ui/cmm-param.cu) so existing parameter files and
// command lines work unchanged.
#pragma once

#include <string>
#include <vector>

namespace cmmhost {

constexpr double kPi = 3.141592653589793238462643383279502884197169399375105820974944592307816406286;
constexpr double kTwoPi = 6.283185307179586476925286766559005768394338798750211641949889184615632812572;
constexpr double kTMax = 10000;  // "final" save time, Cuda-CMM/src/ui/settings.h:26

// error(msg, code): print and exit with the code, Cuda-CMM/src/ui/globals.cu:8-12
[[noreturn]] void fail(const std::string &msg, int code);

bool parse_bool(const std::string &v);       // true / True / 1
std::string fmt16(double v);                  // 16 significant digits (str_t)
std::string fmt3(double v);                   // 3 significant digits (to_str of a double)

struct SaveComputational {
    bool is_instant = false;
    double time_start = 0, time_end = 0, time_step = 0;
    std::string var;
    bool conv = false;
    void set_all(const std::string &block);
    std::string str() const;
};

struct SaveSample {
    bool is_instant = false;
    double time_start = 0, time_end = 0, time_step = 0;
    std::string var;
    int grid = 0;
    void set_all(const std::string &block);
    std::string str() const;
};

struct SaveZoom {
    bool is_instant = false;
    double time_start = 0, time_end = 0, time_step = 0;
    std::string var;
    int grid = 0;
    double pos_x = 0, pos_y = 0, width_x = 0, width_y = 0;
    int rep = 0;
    double rep_fac = 0;
    void set_all(const std::string &block);
    std::string str() const;
};

struct ParticlesAdvected {
    int num = 0;
    double tau = 0;
    unsigned long long seed = 0;
    std::string time_integration;
    int time_integration_num = -1;
    std::string init_name;
    int init_num = -1;
    bool init_vel = false;
    double init_time = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0;
    void set_all(const std::string &block);
    std::string str() const;
};

struct ParticlesForwarded {
    int num = 0;
    unsigned long long seed = 0;
    std::string init_name;
    int init_num = -1;
    double init_time = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0;
    void set_all(const std::string &block);
    std::string str() const;
};

struct Settings {
    std::string workspace, simulation_type, sim_name, file_name;
    int grid_coarse, grid_fine, grid_psi, grid_vort;
    double domain_bounds[6];
    double final_time, factor_dt_by_grid;
    int steps_per_sec;
    bool set_dt_by_steps;
    std::vector<SaveComputational> save_computational;
    int mem_RAM_CPU_remaps;
    bool save_map_stack;
    double restart_time;
    std::string restart_location;
    int verbose;
    std::string initial_condition;
    int initial_condition_num;
    double initial_params[10];
    bool initial_discrete;
    int initial_discrete_grid;
    std::string initial_discrete_location;
    double incomp_threshold, map_epsilon;
    std::string time_integration;
    int time_integration_num;
    int lagrange_order, lagrange_override;
    bool lagrange_init_higher_order;
    std::string map_update_order;
    int map_update_order_num;
    int molly_stencil;
    double freq_cut_psi;
    bool skip_remapping;
    std::vector<SaveSample> save_sample;
    std::string scalar_name;
    bool scalar_discrete;
    int scalar_discrete_grid;
    std::string scalar_discrete_location;
    std::vector<SaveZoom> save_zoom;
    bool forward_map;
    std::vector<ParticlesForwarded> particles_forwarded;
    std::vector<ParticlesAdvected> particles_advected;
    int particles_steps;
    unsigned long long seed;  // seed of the random-spectrum initial condition (new key)
    std::string map_precision;  // "fp64" or "split": the engine's precision form (new key)

    Settings();                 // presets, Cuda-CMM/src/ui/settings.cu:18-327
    // one COMMAND<delim>VALUE; returns the block kind (1..5) when the next lines hold a block
    int set_variable(const std::string &line, const std::string &delim);
    void apply_commands(int argc, char **argv);

    void set_time_integration(const std::string &v);
    void set_initial_condition(const std::string &v);
    void set_map_update_order(const std::string &v);
    void set_domain_bounds(const double *b);
};

// parameter file, Cuda-CMM/src/ui/cmm-param.cu:39-229
bool load_param_file(Settings &s, const std::string &path);
void save_param_file(const Settings &s, const std::string &path);

}  // namespace cmmhost
