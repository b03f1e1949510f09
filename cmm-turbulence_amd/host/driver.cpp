// Euler time loop of a CMM run, driving libcmm.so through its C-ABI (include/cmm.h).
// Order of operations, monitoring files and output layout follow Cuda-CMM/src/simulation/cmm-euler2d.cu
// and Cuda-CMM/src/ui/cmm-io.cu.
#include "driver.hpp"

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <ctime>
#include <dirent.h>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cmm.h"

namespace cmmhost {

namespace {

using Clock = std::chrono::high_resolution_clock;

std::string to_s(int v) { return std::to_string(v); }
std::string zpad(int v, int w) {
    std::ostringstream os;
    os.width(w);
    os.fill('0');
    os << v;
    return os.str();
}

std::string now_string() {
    time_t t = time(nullptr);
    struct tm tm = *localtime(&t);
    char buf[80];
    strftime(buf, sizeof(buf), "%Y-%m-%d.%X", &tm);
    return buf;
}

std::string duration(double sec) {
    return fmt3(std::floor(sec / 3600.0)) + "h " + fmt3(std::floor(std::fmod(sec, 3600) / 60.0)) + "m " +
           fmt3(std::fmod(sec, 60)) + "s";
}

// log.txt plus stdout, Cuda-CMM/src/ui/cmm-io.cu:692-731
struct Logger {
    std::string path;
    explicit Logger(const Settings &s) {
        path = s.workspace + "data/" + s.file_name + "/log.txt";
        std::ofstream f(path, s.restart_time == 0 ? std::ios::out : std::ios::app);
        if (!f) {
            std::cerr << "Unable to open log file.. exitting" << std::endl;
            std::exit(0);
        }
        if (s.restart_time == 0) f << s.file_name << std::endl;
        else f << "\nContinuing simulation\n" << std::endl;
    }
    void push(const std::string &m) const {
        std::ofstream f(path, std::ios::app);
        if (f) f << "[" << now_string() << "]\t" << m << std::endl;
        std::cout << m << std::endl;
    }
};

void mkdir_p(const std::string &p) { mkdir(p.c_str(), 0777); }

void remove_tree(const std::string &p) {
    DIR *d = opendir(p.c_str());
    if (!d) return;
    while (struct dirent *e = readdir(d)) {
        std::string n = e->d_name;
        if (n == "." || n == "..") continue;
        std::string q = p + "/" + n;
        struct stat st;
        if (lstat(q.c_str(), &st) == 0 && S_ISDIR(st.st_mode)) remove_tree(q);
        else unlink(q.c_str());
    }
    closedir(d);
    rmdir(p.c_str());
}

// Cuda-CMM/src/ui/cmm-io.cu:23-64
void create_directories(const Settings &s) {
    const std::string base = s.workspace + "data/" + s.file_name;
    mkdir_p(s.workspace + "data");
    mkdir_p(base);
    if (s.restart_time == 0) remove_tree(base + "/Monitoring_data");
    mkdir_p(base + "/Monitoring_data");
    mkdir_p(base + "/Monitoring_data/Mesure");
    if (s.verbose >= 4) mkdir_p(base + "/Monitoring_data/Debug_globals");
    mkdir_p(base + "/Time_data");
    if (!s.save_zoom.empty()) mkdir_p(base + "/Zoom_data");
    if (!s.particles_advected.empty() || !s.particles_forwarded.empty()) mkdir_p(base + "/Particle_data");
}

// raw little-endian doubles, Cuda-CMM/src/ui/cmm-io.cu:73-103
void write_file(const Settings &s, const std::string &name, const double *v, size_t n, bool append) {
    std::string f = s.workspace + "data/" + s.file_name + name + ".data";
    std::ofstream o(f, std::ios::binary | (append ? std::ios::app : std::ios::out));
    if (!o) {
        std::cout << "Error saving file. Unable to open : " << f << std::endl;
        return;
    }
    o.write(reinterpret_cast<const char *>(v), sizeof(double) * n);
}

void append1(const Settings &s, const std::string &name, double v) { write_file(s, name, &v, 1, true); }

// readAllRealFromBinaryFile: exactly n doubles from a file (false when it is missing or short)
bool read_doubles(const std::string &path, double *v, size_t n) {
    std::ifstream f(path, std::ios::binary);
    return f && f.read(reinterpret_cast<char *>(v), sizeof(double) * n) && f.gcount() == (std::streamsize)(sizeof(double) * n);
}

// countFilesWithString (Cuda-CMM/src/simulation/cmm-simulation-host.cu:1234-1252): regular files
// in dir whose name contains key
int count_files_with(const std::string &dir, const std::string &key) {
    int count = 0;
    if (DIR *d = opendir(dir.c_str())) {
        while (dirent *ent = readdir(d))
            if (ent->d_type == DT_REG && std::string(ent->d_name).find(key) != std::string::npos) count++;
        closedir(d);
    } else {
        std::cerr << "Error: Could not open directory " << dir << std::endl;
    }
    return count;
}

// Debug statistics instead of arrays at verbose >= 4, Cuda-CMM/src/ui/cmm-io.cu:508-534 and
// Cuda-CMM/src/numerical/cmm-mesure.cu:220-250: on a grid "Avg" is max(0, max) h^2 and "L2" is
// sum h^2; on a plain array they are max(0, max)/N and sum/N.  The hash files are not written.
void debug_globals(const Settings &s, const std::string &name, const double *v, size_t n, double area) {
    double mn = v[0], mx = v[0], mxz = 0.0, sum = 0.0;
    for (size_t i = 0; i < n; ++i) {
        mn = std::min(mn, v[i]);
        mx = std::max(mx, v[i]);
        mxz = std::max(mxz, v[i]);
        sum += v[i];
    }
    const std::string p = "/Monitoring_data/Debug_globals/" + name;
    append1(s, p + "_Min", mn);
    append1(s, p + "_Max", mx);
    append1(s, p + "_Avg", area > 0 ? area * mxz : mxz / (double)n);
    append1(s, p + "_L2", area > 0 ? area * sum : sum / (double)n);
}

uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Particle initial positions, Cuda-CMM/src/simulation/cmm-init.cu:311-455.  Random sets draw
// u_j = mix64(mix64(seed) + j) >> 11 / 2^53 (j = 2i, 2i+1) in place of cuRAND.
std::vector<double> init_particles(const ParticlesAdvected &p, const double *b) {
    const int n = p.num;
    std::vector<double> pos(2 * (size_t)n);
    const double lx = b[1] - b[0], ly = b[3] - b[2];
    const uint64_t s0 = mix64(p.seed);
    auto u = [&](uint64_t j) { return (double)(mix64(s0 + j) >> 11) * (1.0 / 9007199254740992.0); };
    const int len = (int)std::ceil(std::sqrt((double)n));
    for (int i = 0; i < n; ++i) {
        double x = 0, y = 0;
        switch (p.init_num) {
        case 0:
            x = p.p1 + p.p3 * (u(2 * (uint64_t)i) - 0.5);
            y = p.p2 + p.p4 * (u(2 * (uint64_t)i + 1) - 0.5);
            break;
        case 1: {  // normal around 0.5 with unit deviation (Box-Muller), then the same rescale
            double u1 = std::max(u(2 * (uint64_t)i), 1e-300), u2 = u(2 * (uint64_t)i + 1);
            double r = std::sqrt(-2.0 * std::log(u1));
            x = p.p1 + p.p3 * (0.5 + r * std::cos(kTwoPi * u2) - 0.5);
            y = p.p2 + p.p4 * (0.5 + r * std::sin(kTwoPi * u2) - 0.5);
            break;
        }
        case 2:
            x = p.p1 + p.p3 * std::cos(i / (double)n * kTwoPi);
            y = p.p2 + p.p4 * std::sin(i / (double)n * kTwoPi);
            break;
        case 3:
            x = p.p1 + p.p3 * (-1 + 2 * (i % len) / (double)len);
            y = p.p2 + p.p4 * (-1 + 2 * (i / len) / (double)len);
            break;
        case 4:
            x = i / (double)n * kTwoPi + p.p1;
            y = std::sin(i / (double)n * kTwoPi) / 2.0 + p.p2;
            break;
        default:
            break;
        }
        pos[2 * i] = x - std::floor((x - b[0]) / lx) * lx;
        pos[2 * i + 1] = y - std::floor((y - b[2]) / ly) * ly;
    }
    return pos;
}

// forwarded particles start from the same distributions (init_particles with type 1)
std::vector<double> init_particles(const ParticlesForwarded &p, const double *b) {
    ParticlesAdvected a;
    a.num = p.num;
    a.seed = p.seed;
    a.init_num = p.init_num;
    a.p1 = p.p1;
    a.p2 = p.p2;
    a.p3 = p.p3;
    a.p4 = p.p4;
    return init_particles(a, b);
}

bool save_due(bool is_instant, double ts, double te, double tstep, double t_now, double dt_now, double dt) {
    // Cuda-CMM/src/ui/cmm-io.cu:300-312
    if (is_instant) return t_now - ts + dt * 1e-5 < dt_now && t_now - ts + dt * 1e-5 >= 0;
    return (std::fmod(t_now - ts + dt * 1e-5, tstep) < dt_now && t_now + dt * 1e-5 >= ts && t_now - dt * 1e-5 <= te) ||
           t_now == te;
}

struct Run {
    Settings &s;
    Logger log;
    cmm_euler *e = nullptr;
    cmm_grid gc, gf, gp;
    int L;
    std::vector<int> part_set;  // engine set per advected particle entry (-1 if not advected here)
    // forwarded particles: start positions and the stack index their maps start from (init_map of
    // Cuda-CMM/src/simulation/cmm-euler2d.cu:804-811; 0 until they start)
    std::vector<std::vector<double>> fpos;
    std::vector<int> finit_map;
    Run(Settings &st) : s(st), log(st) {}

    // apply_map_stack_points with direction 1 (Cuda-CMM/src/simulation/cmm-simulation-host.cu:213-291)
    // for forwarded set i: with an empty stack the current forward map; else the stacked maps from 0
    // (a set started at t = 0) or from max(1, init_map), and the current map once the set has started
    std::vector<double> forwarded(size_t i, const cmm_grid &gb) {
        const auto &pf = s.particles_forwarded[i];
        int steps = 0, ns = 0;
        double t = 0;
        cmm_euler_state(e, &steps, &t, &ns);
        int first = 0, cur = 1;
        if (ns > 0 && pf.init_time != 0) {
            first = std::max(1, finit_map[i]);
            cur = finit_map[i] != 0;
        }
        std::vector<double> out(fpos[i].size());
        check(cmm_euler_forward_points(e, gb, pf.num, fpos[i].data(), first, cur, out.data()), "cmm_euler_forward_points");
        return out;
    }

    void check(int rc, const char *what) {
        if (rc != CMM_OK) {
            log.push(std::string("Exitted early : ") + what + " failed: " + cmm_last_error());
            std::exit(0);
        }
    }

    std::vector<double> field(int f, size_t n) {
        std::vector<double> v(n);
        check(cmm_euler_get_field(e, f, v.data()), "cmm_euler_get_field");
        return v;
    }

    void write_var(const std::string &sub, const std::string &name, const double *v, size_t n, double area) {
        if (s.verbose < 4) write_file(s, sub + "/" + name, v, n, false);
        else debug_globals(s, name, v, n, area);
    }

    // writeTimeStep, Cuda-CMM/src/ui/cmm-io.cu:295-437
    std::string write_time_step(double t_now, double dt_now, double dt) {
        bool now = false;
        std::string var;
        for (auto &c : s.save_computational)
            if (save_due(c.is_instant, c.time_start, c.time_end, c.time_step, t_now, dt_now, dt)) {
                now = true;
                var += c.var;
            }
        if (!now) return "";
        std::string msg;
        std::string ts = std::abs(t_now - kTMax) < 1 ? "final" : fmt3(t_now);
        const bool all = var.find("All") != std::string::npos;
        auto has = [&](const char *k) { return var.find(k) != std::string::npos; };
        const long nc = (long)gc.nx * gc.ny, np = (long)gp.nx * gp.ny;
        const double ac = ((gc.x1 - gc.x0) / (float)gc.nx) * ((gc.y1 - gc.y0) / (float)gc.ny);
        const double ap = ((gp.x1 - gp.x0) / (float)gp.nx) * ((gp.y1 - gp.y0) / (float)gp.ny);
        const std::string sub = "/Time_data/Time_" + ts;
        bool saved_any = false;
        if (has("Vorticity") || has("W") || has("Stream") || has("Psi") || has("Map") || has("Chi") || has("Velocity") ||
            has("U") || has("Dist") || has("F") || all) {
            mkdir_p(s.workspace + "data/" + s.file_name + sub);
            if (s.verbose >= 4) append1(s, "/Monitoring_data/Debug_globals/Time_s", t_now);
            saved_any = true;
            msg = "Processed computational data";
        }
        if (has("Vorticity") || has("W") || all || has("Dist") || has("F")) {
            auto w = field(CMM_FIELD_W_COARSE, nc);
            if (has("Vorticity") || has("W") || all) write_var(sub, "Vorticity_W_coarse", w.data(), nc, ac);
            if (has("Dist") || has("F") || all) write_var(sub, "Distribution_coarse", w.data(), nc, ac);
        }
        if (has("Stream") || has("Psi") || has("Velocity") || has("U") || all) {
            auto psi = field(CMM_FIELD_PSI, 4 * np * L);
            if (has("Stream") || has("Psi") || all) write_var(sub, "Stream_function_Psi_psi", psi.data(), np, ap);
            if (has("Stream_H") || has("Psi_H") || all) write_var(sub, "Stream_function_Psi_H_psi", psi.data(), 4 * np, 0);
            if (has("Velocity") || has("U") || all) {
                write_var(sub, "Velocity_UX_psi", psi.data() + np, np, ap);
                write_var(sub, "Velocity_UY_psi", psi.data() + 2 * np, np, ap);
            }
        }
        if (has("Map_b") || has("Chi_b") || has("Map_H") || has("Chi_H") || all) {
            auto cx = field(CMM_FIELD_CHIX, 4 * nc), cy = field(CMM_FIELD_CHIY, 4 * nc);
            if (has("Map_b") || has("Chi_b") || all) {
                write_var(sub, "Map_ChiX_coarse", cx.data(), nc, ac);
                write_var(sub, "Map_ChiY_coarse", cy.data(), nc, ac);
            }
            if (has("Map_H") || has("Chi_H") || all) {
                write_var(sub, "Map_ChiX_H_coarse", cx.data(), 4 * nc, 0);
                write_var(sub, "Map_ChiY_H_coarse", cy.data(), 4 * nc, 0);
            }
        }
        // forward map on the coarse grid, Cuda-CMM/src/ui/cmm-io.cu:379-389
        if (s.forward_map && (has("Map_f") || has("Chi_f") || all)) {
            auto fx = field(CMM_FIELD_CHIX_F, 4 * nc), fy = field(CMM_FIELD_CHIY_F, 4 * nc);
            write_var(sub, "Map_ChiX_f_coarse", fx.data(), nc, ac);
            write_var(sub, "Map_ChiY_f_coarse", fy.data(), nc, ac);
            if (has("Map_f_H") || has("Chi_f_H") || all) {
                write_var(sub, "Map_ChiX_f_H_coarse", fx.data(), 4 * nc, 0);
                write_var(sub, "Map_ChiY_f_H_coarse", fy.data(), 4 * nc, 0);
            }
        }
        // particle positions, Cuda-CMM/src/ui/cmm-io.cu:396-434
        const std::string psub = "/Particle_data/Time_" + ts;
        if (has("PartA_") || all) {
            mkdir_p(s.workspace + "data/" + s.file_name + psub);
            if (!saved_any && s.verbose >= 4) append1(s, "/Monitoring_data/Debug_globals/Time_s", t_now);
            msg = "Processed particle data";
        }
        for (size_t i = 0; i < s.particles_advected.size(); ++i) {
            const std::string tag = zpad((int)i + 1, 2);
            if ((has(("PartA_" + tag).c_str()) || all) && part_set[i] >= 0) {
                std::vector<double> pos(2 * (size_t)s.particles_advected[i].num);
                check(cmm_euler_get_particles(e, part_set[i], pos.data()), "cmm_euler_get_particles");
                if (s.verbose < 4) write_file(s, psub + "/Particles_advected_pos_P" + tag, pos.data(), pos.size(), false);
                else debug_globals(s, "/Particles_advected_pos_P" + tag, pos.data(), pos.size(), 0);
            }
            // velocities: an inertial set's own, else psi's x and y derivatives at the positions
            // (Cuda-CMM/src/ui/cmm-io.cu:418-430)
            if ((has(("PartA_Vel_" + tag).c_str()) || all) && part_set[i] >= 0) {
                std::vector<double> vel(2 * (size_t)s.particles_advected[i].num);
                check(cmm_euler_get_particle_velocities(e, part_set[i], vel.data()), "cmm_euler_get_particle_velocities");
                if (s.verbose < 4) write_file(s, psub + "/Particles_advected_vel_U" + tag, vel.data(), vel.size(), false);
                else debug_globals(s, "/Particles_advected_vel_U" + tag, vel.data(), vel.size(), 0);
            }
        }
        return msg;
    }

    // compute_conservation_targets, Cuda-CMM/src/simulation/cmm-simulation-host.cu:615-698
    std::string conservation(double t_now, double dt_now, double dt) {
        bool now = false;
        for (auto &c : s.save_computational)
            if (c.conv && save_due(c.is_instant, c.time_start, c.time_end, c.time_step, t_now, dt_now, dt)) now = true;
        if (!now) return "";
        auto f8 = [](double v) {
            std::ostringstream os;
            os.precision(8);
            os << v;
            return os.str();
        };
        if (s.simulation_type == "cmm_vlasov_poisson_1d") {
            double m[9];
            check(cmm_euler_vp_diagnostics(e, m), "cmm_euler_vp_diagnostics");
            append1(s, "/Monitoring_data/Mesure/Time_s", t_now);
            append1(s, "/Monitoring_data/Mesure/Etot", m[0]);
            append1(s, "/Monitoring_data/Mesure/Ekin", m[1]);
            append1(s, "/Monitoring_data/Mesure/Epot", m[2]);
            append1(s, "/Monitoring_data/Mesure/Mass", m[3]);
            append1(s, "/Monitoring_data/Mesure/Momentum", m[4]);
            for (int k = 0; k < 4; ++k) append1(s, "/Monitoring_data/Mesure/Emode" + to_s(k + 1), m[5 + k]);
            return "Computed coarse Cons : Etot = " + f8(m[0]) + " \t Momenta = " + f8(m[4]) + " \t Ekin = " + f8(m[1]) +
                   " \t Epot = " + f8(m[2]) + " \t Mass = " + f8(m[3]);
        }
        double m[4];
        check(cmm_euler_diagnostics(e, m), "cmm_euler_diagnostics");
        append1(s, "/Monitoring_data/Mesure/Time_s", t_now);
        append1(s, "/Monitoring_data/Mesure/Energy", m[0]);
        append1(s, "/Monitoring_data/Mesure/Enstrophy", m[1]);
        append1(s, "/Monitoring_data/Mesure/Palinstrophy", m[2]);
        append1(s, "/Monitoring_data/Mesure/Max_vorticity", m[3]);
        return "Computed coarse Cons : Energ = " + f8(m[0]) + " \t Enstr = " + f8(m[1]) + " \t Palinstr = " + f8(m[2]) +
               " \t Wmax = " + f8(m[3]);
    }

    // sample_compute_and_write, Cuda-CMM/src/simulation/cmm-simulation-host.cu:706-893: fields
    // through the whole map stack on each due sample grid, written like the computational data,
    // with conservation values on that grid
    std::string sample(double t_now, double dt_now, double dt) {
        if (s.simulation_type == "cmm_vlasov_poisson_1d") return sample_vlasov(t_now, dt_now, dt);
        if (s.simulation_type != "cmm_euler_2d") return "";
        std::string msg;
        auto f8 = [](double v) {
            std::ostringstream os;
            os.precision(8);
            os << v;
            return os.str();
        };
        const double *b = s.domain_bounds;
        for (size_t i = 0; i < s.save_sample.size(); ++i) {
            const auto &c = s.save_sample[i];
            if (!save_due(c.is_instant, c.time_start, c.time_end, c.time_step, t_now, dt_now, dt)) continue;
            const int N = c.grid;
            const std::string tag = to_s(N);
            const cmm_grid gs{N, N, b[0], b[1], b[2], b[3]};
            const size_t n = (size_t)N * N;
            const double area = ((b[1] - b[0]) / (float)N) * ((b[3] - b[2]) / (float)N);
            const std::string &var = c.var;
            const bool all = var.find("All") != std::string::npos;
            auto has = [&](const char *k) { return var.find(k) != std::string::npos; };
            if (s.verbose >= 4) append1(s, "/Monitoring_data/Debug_globals/Time_s_" + tag, t_now);
            const bool scalar = has("Scalar") || has("Theta") || all;
            if (scalar && s.scalar_discrete) fail("a discrete passive scalar is not provided by this engine", 1);
            // scalar numbers, Cuda-CMM/src/ui/settings.h:431-434
            const int scalar_num = s.scalar_name == "rectangle" ? 0 : s.scalar_name == "gaussian" ? 1
                                 : s.scalar_name == "circular_ring" ? 2 : -1;
            const int scalar_ic = scalar && scalar_num >= 0 ? CMM_IC_SCALAR_RECTANGLE + scalar_num : -1;
            std::vector<double> out(CMM_SAMPLE_FIELDS * n);
            check(cmm_euler_sample(e, gs, scalar_ic, out.data()), "cmm_euler_sample");
            auto f = [&](int k) { return out.data() + (size_t)k * n; };
            std::string ts = std::abs(t_now - kTMax) < 1 ? "final" : fmt3(t_now);
            const std::string sub = "/Time_data/Time_" + ts;
            mkdir_p(s.workspace + "data/" + s.file_name + sub);
            // the forward map through the forward stack and the forwarded particles, :740-783
            if (s.forward_map && (has("Map_f") || has("Chi_f") || all || !s.particles_forwarded.empty())) {
                if (has("Map_f") || has("Chi_f") || all) {
                    std::vector<double> fxy(2 * n);
                    check(cmm_euler_forward_positions(e, gs, fxy.data()), "cmm_euler_forward_positions");
                    write_var(sub, "Map_ChiX_f_" + tag, fxy.data(), n, area);
                    write_var(sub, "Map_ChiY_f_" + tag, fxy.data() + n, n, area);
                }
                for (size_t k = 0; k < s.particles_forwarded.size(); ++k) {
                    const std::string ptag = zpad((int)k + 1, 2);
                    if (!(has(("PartF_" + ptag).c_str()) || all)) continue;
                    auto pos = forwarded(k, gs);
                    const std::string psub = "/Particle_data/Time_" + ts;
                    mkdir_p(s.workspace + "data/" + s.file_name + psub);
                    write_file(s, psub + "/Particles_forwarded_pos_P" + ptag, pos.data(), pos.size(), false);
                }
            }
            if (has("Map_b") || has("Chi_b") || all) {
                write_var(sub, "Map_ChiX_" + tag, f(CMM_SAMPLE_CHIX), n, area);
                write_var(sub, "Map_ChiY_" + tag, f(CMM_SAMPLE_CHIY), n, area);
            }
            if (scalar) {
                write_var(sub, "Scalar_Theta_" + tag, f(CMM_SAMPLE_THETA), n, area);
                // Compute_avg (Cuda-CMM/src/numerical/cmm-mesure.cu:225-228): h^2 max(0, max)
                double mx = 0.0;
                for (size_t k = 0; k < n; ++k) mx = std::max(mx, f(CMM_SAMPLE_THETA)[k]);
                append1(s, "/Monitoring_data/Mesure/Scalar_integral_" + tag, area * mx);
            }
            const double *w = f(CMM_SAMPLE_W);
            if (has("Vorticity") || has("W") || has("F") || all) write_var(sub, "Vorticity_W_" + tag, w, n, area);
            double ens = 0, pal = 0, eng = 0, wmax = w[0];
            for (size_t k = 0; k < n; ++k) {
                ens += w[k] * w[k];
                pal += f(CMM_SAMPLE_GRADX_W)[k] * f(CMM_SAMPLE_GRADX_W)[k] + f(CMM_SAMPLE_GRADY_W)[k] * f(CMM_SAMPLE_GRADY_W)[k];
                eng += f(CMM_SAMPLE_UX)[k] * f(CMM_SAMPLE_UX)[k] + f(CMM_SAMPLE_UY)[k] * f(CMM_SAMPLE_UY)[k];
                wmax = std::max(wmax, w[k]);
            }
            double m[4];
            m[0] = 0.5 * area * eng;
            m[1] = 0.5 * area * ens;
            m[2] = 0.5 * area * pal;
            // the reference takes the maximum twice (Compute_max for w_max and w_min), so this is |max w|
            m[3] = std::max(wmax, -wmax);
            if (has("Laplacian_W") || all) write_var(sub, "Laplacian_W_" + tag, f(CMM_SAMPLE_LAP_W), n, area);
            if (has("Grad_W") || all) {
                write_var(sub, "GradX_W_" + tag, f(CMM_SAMPLE_GRADX_W), n, area);
                write_var(sub, "GradY_W_" + tag, f(CMM_SAMPLE_GRADY_W), n, area);
            }
            if (has("Stream") || has("Psi") || all) write_var(sub, "Stream_function_Psi_" + tag, f(CMM_SAMPLE_PSI), n, area);
            if (has("Velocity") || has("U") || all) {
                write_var(sub, "Velocity_UX_" + tag, f(CMM_SAMPLE_UX), n, area);
                write_var(sub, "Velocity_UY_" + tag, f(CMM_SAMPLE_UY), n, area);
            }
            append1(s, "/Monitoring_data/Mesure/Time_s_" + tag, t_now);
            append1(s, "/Monitoring_data/Mesure/Energy_" + tag, m[0]);
            append1(s, "/Monitoring_data/Mesure/Enstrophy_" + tag, m[1]);
            append1(s, "/Monitoring_data/Mesure/Palinstrophy_" + tag, m[2]);
            append1(s, "/Monitoring_data/Mesure/Max_vorticity_" + tag, m[3]);
            msg += "Processed sample data " + to_s((int)i + 1) + " on grid " + tag + ", Cons : Energ = " + f8(m[0]) +
                   " \t Enstr = " + f8(m[1]) + " \t Palinstr = " + f8(m[2]) + " \t Wmax = " + f8(m[3]);
        }
        return msg;
    }

    // sample_compute_and_write_vlasov, Cuda-CMM/src/simulation/cmm-simulation-host.cu:898-1030: on
    // each due sample grid the map through the whole stack, the distribution there, the measures
    // of that sample (the potential energy from the sample's own phi) and the modes of psi.
    // Stream_function_Psi_<g>: the reference writes its reused sample buffer, whose first nx values
    // hold dphi/dx and whose tail holds f; here the first nx values are dphi/dx and the rest f
    // (the reference holds FFT scratch between nx and n/8).
    std::string sample_vlasov(double t_now, double dt_now, double dt) {
        std::string msg;
        auto f8 = [](double v) {
            std::ostringstream os;
            os.precision(8);
            os << v;
            return os.str();
        };
        const double *b = s.domain_bounds;
        for (size_t i = 0; i < s.save_sample.size(); ++i) {
            const auto &c = s.save_sample[i];
            if (!save_due(c.is_instant, c.time_start, c.time_end, c.time_step, t_now, dt_now, dt)) continue;
            const int N = c.grid;
            const std::string tag = to_s(N);
            const cmm_grid gs{N, N, b[0], b[1], b[2], b[3]};
            const size_t n = (size_t)N * N;
            const double area = ((b[1] - b[0]) / (float)N) * ((b[3] - b[2]) / (float)N);
            const std::string &var = c.var;
            const bool all = var.find("All") != std::string::npos;
            auto has = [&](const char *k) { return var.find(k) != std::string::npos; };
            if (s.verbose >= 4) append1(s, "/Monitoring_data/Debug_globals/Time_s_" + tag, t_now);
            std::vector<double> out(3 * n + N + 9);
            check(cmm_euler_vp_sample(e, gs, out.data()), "cmm_euler_vp_sample");
            std::string ts = std::abs(t_now - kTMax) < 1 ? "final" : fmt3(t_now);
            const std::string sub = "/Time_data/Time_" + ts;
            mkdir_p(s.workspace + "data/" + s.file_name + sub);
            if (s.forward_map && (has("Map_f") || has("Chi_f") || all)) {
                std::vector<double> fxy(2 * n);
                check(cmm_euler_forward_positions(e, gs, fxy.data()), "cmm_euler_forward_positions");
                write_var(sub, "Map_ChiX_f_" + tag, fxy.data(), n, area);
                write_var(sub, "Map_ChiY_f_" + tag, fxy.data() + n, n, area);
            }
            if (has("Map_b") || has("Chi_b") || all) {
                write_var(sub, "Map_ChiX_" + tag, out.data(), n, area);
                write_var(sub, "Map_ChiY_" + tag, out.data() + n, n, area);
            }
            const double *fv = out.data() + 2 * n;
            if (has("Dist") || has("F") || all) write_var(sub, "Distribution_F_" + tag, fv, n, area);
            if (has("Stream") || has("Psi") || all) {
                std::vector<double> buf(fv, fv + n);
                std::copy(out.begin() + 3 * n, out.begin() + 3 * n + N, buf.begin());
                write_var(sub, "Stream_function_Psi_" + tag, buf.data(), n, area);
            }
            const double *m = out.data() + 3 * n + N;
            append1(s, "/Monitoring_data/Mesure/Time_s_" + tag, t_now);
            append1(s, "/Monitoring_data/Mesure/Etot_" + tag, m[0]);
            append1(s, "/Monitoring_data/Mesure/Ekin_" + tag, m[1]);
            append1(s, "/Monitoring_data/Mesure/Epot_" + tag, m[2]);
            append1(s, "/Monitoring_data/Mesure/Mass_" + tag, m[3]);
            append1(s, "/Monitoring_data/Mesure/Momentum_" + tag, m[4]);
            for (int k = 0; k < 4; ++k) append1(s, "/Monitoring_data/Mesure/Emode" + to_s(k + 1) + "_" + tag, m[5 + k]);
            msg += "Saved sample data " + to_s((int)i + 1) + " on grid " + tag + ", Cons : Etot = " + f8(m[0]) +
                   " \t Momenta = " + f8(m[4]) + " \t Ekin = " + f8(m[1]) + " \t Epot = " + f8(m[2]) +
                   " \t Mass = " + f8(m[3]);
        }
        return msg;
    }

    // compute_zoom, Cuda-CMM/src/simulation/cmm-simulation-host.cu:1039-1227: for each due zoom, rep
    // windows around (pos_x, pos_y), the widths shrinking by rep_fac, each with the fields through
    // the whole map stack on a grid x grid window.  The statistics use the cell area of the full
    // domain on that grid, as the reference's zoom variable keeps the domain's grid.
    std::string zoom(double t_now, double dt_now, double dt) {
        std::string msg;
        const double *b = s.domain_bounds;
        const bool vp = s.simulation_type == "cmm_vlasov_poisson_1d";
        for (size_t i = 0; i < s.save_zoom.size(); ++i) {
            const auto &c = s.save_zoom[i];
            if (!save_due(c.is_instant, c.time_start, c.time_end, c.time_step, t_now, dt_now, dt)) continue;
            const int N = c.grid;
            const std::string tag = to_s(N), zname = "Zoom_" + to_s((int)i + 1);
            const size_t n = (size_t)N * N;
            const double area = ((b[1] - b[0]) / (float)N) * ((b[3] - b[2]) / (float)N);
            const std::string &var = c.var;
            const bool all = var.find("All") != std::string::npos;
            auto has = [&](const char *k) { return var.find(k) != std::string::npos; };
            msg += "Processed zoom data " + to_s((int)i + 1) + " on grid " + tag;
            if (s.verbose >= 4) append1(s, "/Monitoring_data/Debug_globals/Time_s_" + tag + "_" + zname, t_now);
            // the reference's k_h_sample_from_init switches: vorticity and distribution from the
            // initial condition number of the run (so a Vlasov run's "vorticity" is the vorticity
            // condition of the same number), the passive scalar from scalar_name
            const bool want_map = has("Map_b") || has("Chi_b") || all;
            const bool want_scalar = has("Scalar") || has("Theta") || all;
            const bool want_w = has("Vorticity") || has("W") || all;
            const bool want_f = has("Dist") || has("F") || all;
            const bool want_psi = has("Stream") || has("Psi") || all;
            if (want_scalar && s.scalar_discrete) fail("a discrete passive scalar is not provided by this engine", 1);
            const int scalar_num = s.scalar_name == "rectangle" ? 0 : s.scalar_name == "gaussian" ? 1
                                 : s.scalar_name == "circular_ring" ? 2 : -1;
            std::vector<int> ics;
            std::vector<std::string> names;
            if (want_scalar) {
                ics.push_back(scalar_num >= 0 ? CMM_IC_SCALAR_RECTANGLE + scalar_num : 399);
                names.push_back("Scalar_Theta_" + tag);
            }
            if (want_w) {
                ics.push_back(vp && !s.initial_discrete ? s.initial_condition_num : -1);
                names.push_back("Vorticity_W_" + tag);
            }
            if (want_f) {
                ics.push_back(vp || s.initial_discrete ? -1 : CMM_IC_LANDAU_DAMPING + s.initial_condition_num);
                names.push_back("Distribution_" + tag);
            }
            std::string ts = std::abs(t_now - kTMax) < 1 ? "final" : fmt3(t_now);
            double wx = c.width_x / 2.0, wy = c.width_y / 2.0;
            std::vector<double> out((3 + ics.size()) * n);
            for (int r = 0; r < c.rep; ++r) {
                const cmm_grid gz{N, N, c.pos_x - wx, c.pos_x + wx, c.pos_y - wy, c.pos_y + wy};
                check(cmm_euler_zoom(e, gz, ics.data(), (int)ics.size(), want_psi ? 1 : 0, out.data()), "cmm_euler_zoom");
                const std::string rep = zname + "_rep_" + to_s(r);
                const std::string sub = "/Zoom_data/Time_" + ts + "/" + rep;
                if (s.verbose < 4) mkdir_p(s.workspace + "data/" + s.file_name + sub);
                auto put = [&](const std::string &name, const double *v) {
                    if (s.verbose < 4) write_file(s, sub + "/" + name, v, n, false);
                    else debug_globals(s, name + "_" + rep, v, n, area);
                };
                // forward map and forwarded particles of the window, :1091-1140 (the particle files
                // carry the set's 0-based index, and only an explicit PartF_ entry saves them)
                if (s.forward_map && (has("Map_f") || has("Chi_f") || all || !s.particles_forwarded.empty())) {
                    if (has("Map_f") || has("Chi_f") || all) {
                        std::vector<double> fxy(2 * n);
                        check(cmm_euler_forward_positions(e, gz, fxy.data()), "cmm_euler_forward_positions");
                        put("Map_ChiX_f_" + tag, fxy.data());
                        put("Map_ChiY_f_" + tag, fxy.data() + n);
                    }
                    for (size_t k = 0; k < s.particles_forwarded.size(); ++k) {
                        if (!has(("PartF_" + zpad((int)k + 1, 2)).c_str())) continue;
                        auto pos = forwarded(k, gz);
                        mkdir_p(s.workspace + "data/" + s.file_name + sub);
                        write_file(s, sub + "/Particles_forwarded_pos_P" + zpad((int)k, 2), pos.data(), pos.size(), false);
                    }
                }
                if (want_map) {
                    put("Map_ChiX_" + tag, out.data());
                    put("Map_ChiY_" + tag, out.data() + n);
                }
                for (size_t k = 0; k < ics.size(); ++k) put(names[k], out.data() + (2 + k) * n);
                if (want_psi) put("Stream_function_Psi_" + tag, out.data() + (2 + ics.size()) * n);
                wx *= c.rep_fac;
                wy *= c.rep_fac;
            }
        }
        return msg;
    }

    // writeMapStack (Cuda-CMM/src/ui/cmm-io.cu:611-635): map m of the stack (m = -1: the active map,
    // file suffix "coarse") and psi level `level` written once per older level, as the reference
    // writes its newest level into every Stream_function_Psi_H_psi_i file
    // forward = true: the forward map, files Map_ChiX_f_H_*, and no psi
    void write_map_stack(int m, int level, bool forward = false) {
        mkdir_p(s.workspace + "data/" + s.file_name + "/MapStack");
        const size_t nc4 = 4 * (size_t)gc.nx * gc.ny, np4 = 4 * (size_t)gp.nx * gp.ny;
        std::vector<double> cx(nc4), cy(nc4);
        if (m < 0) {
            check(cmm_euler_get_field(e, forward ? CMM_FIELD_CHIX_F : CMM_FIELD_CHIX, cx.data()), "cmm_euler_get_field");
            check(cmm_euler_get_field(e, forward ? CMM_FIELD_CHIY_F : CMM_FIELD_CHIY, cy.data()), "cmm_euler_get_field");
        } else if (forward) {
            check(cmm_euler_get_forward_stack_map(e, m, cx.data(), cy.data()), "cmm_euler_get_forward_stack_map");
        } else {
            check(cmm_euler_get_stack_map(e, m, cx.data(), cy.data()), "cmm_euler_get_stack_map");
        }
        const std::string name = m < 0 ? "coarse" : to_s(m);
        const std::string f = forward ? "_f" : "";
        write_file(s, "/MapStack/Map_ChiX" + f + "_H_" + name, cx.data(), nc4, false);
        write_file(s, "/MapStack/Map_ChiY" + f + "_H_" + name, cy.data(), nc4, false);
        if (L > 1 && !forward) {
            std::vector<double> psi(np4 * L);
            check(cmm_euler_get_field(e, CMM_FIELD_PSI, psi.data()), "cmm_euler_get_field");
            for (int i = 1; i < L; ++i)
                write_file(s, "/MapStack/Stream_function_Psi_H_psi_" + to_s(i), psi.data() + np4 * level, np4, false);
        }
    }

    void save(double t_now, double dt_now, double dt) {
        std::string m = write_time_step(t_now, dt_now, dt);
        if (s.verbose >= 3 && !m.empty()) log.push(m);
        m = conservation(t_now, dt_now, dt);
        if (s.verbose >= 3 && !m.empty()) log.push(m);
        m = sample(t_now, dt_now, dt);
        if (s.verbose >= 3 && !m.empty()) log.push(m);
        m = zoom(t_now, dt_now, dt);
        if (s.verbose >= 3 && !m.empty()) log.push(m);
    }
};

}  // namespace

double compute_next_timestep(const Settings &s, double t, double dt) {
    double dt_now = dt;
    const double e = dt * 1e-5;
    for (auto &p : s.particles_advected)
        if (t + e - p.init_time < 0 && t + e + dt - p.init_time > 0) dt_now = std::fmin(dt_now, p.init_time - t);
    for (auto &p : s.particles_forwarded)
        if (t + e - p.init_time < 0 && t + e + dt - p.init_time > 0) dt_now = std::fmin(dt_now, p.init_time - t);
    auto targets = [&](bool inst, double ts, double te, double step) {
        if (inst && t + e - ts < 0 && t + e + dt - ts > 0) dt_now = std::fmin(dt_now, ts - t);
        if (!inst &&
            (std::fmod(t + e - ts, step) > std::fmod(t + dt + e - ts, step) ||
             (std::fmod(t + e - ts, step) < 0 && std::fmod(t + dt + e - ts, step) > 0)) &&
            t + dt + e >= ts && t - e <= te)
            dt_now = std::fmin(dt_now, step - std::fmod(t, step));
    };
    for (auto &c : s.save_computational) targets(c.is_instant, c.time_start, c.time_end, c.time_step);
    for (auto &c : s.save_sample) targets(c.is_instant, c.time_start, c.time_end, c.time_step);
    for (auto &c : s.save_zoom) targets(c.is_instant, c.time_start, c.time_end, c.time_step);
    return dt_now;
}

// readMapStack's folder: restart_location, or the run's own MapStack folder when it is empty
std::string restart_dir(const Settings &s) {
    return s.restart_location.empty() ? s.workspace + "data/" + s.file_name + "/MapStack" : s.restart_location;
}

int run_euler(Settings &s) {
    const auto begin = Clock::now();
    // the Vlasov-Poisson loop (Cuda-CMM/src/simulation/cmm-vlasov1d.cu) shares this driver: bounds
    // in units of pi (:64-75), f in place of omega, psi from the density, E_inc sanity abort (:838-840)
    const bool vp = s.simulation_type == "cmm_vlasov_poisson_1d";
    if (vp)
        for (int i = 0; i < 4; ++i) s.domain_bounds[i] *= kPi;
    auto ctime = [&]() { return std::chrono::duration_cast<std::chrono::microseconds>(Clock::now() - begin).count() / 1e6; };

    // dt and Lagrange order, Cuda-CMM/src/simulation/cmm-euler2d.cu:62-78
    const double dt = s.set_dt_by_steps ? 1.0 / s.steps_per_sec
                                        : 1.0 / std::max(s.grid_coarse, s.grid_coarse) / s.factor_dt_by_grid;
    for (auto &p : s.particles_advected) {
        int need = (p.time_integration_num == 20) ? 2 : (p.time_integration_num == 30 || p.time_integration_num == 31) ? 3
                 : (p.time_integration_num == 40 || p.time_integration_num == 41) ? 4 : 0;
        if (s.lagrange_order < need) s.lagrange_order = need;
    }
    if (s.lagrange_override != -1) s.lagrange_order = s.lagrange_override;
    const int iter_max = (int)(1.05 * std::ceil((s.final_time - s.restart_time) / dt));
    s.file_name = s.sim_name + "_" + s.initial_condition + "_C" + to_s(s.grid_coarse) + "_F" + to_s(s.grid_fine) + "_t" +
                  to_s((int)(1.0 / dt)) + "_T" + fmt3(s.final_time);
    create_directories(s);
    Run r(s);
    Logger &log = r.log;
    if (s.verbose >= 2) log.push("Starting memory initialization \t C-Time = " + duration(ctime()));
    if (s.verbose >= 3) {
        char buf[512];
        if (cmm_device_info(buf, sizeof(buf)) == CMM_OK) log.push(std::string("Device = ") + buf);
    }
    if (s.verbose >= 1) {
        log.push("Solving = " + s.simulation_type);
        log.push("Initial condition = " + s.initial_condition);
        log.push("Iter max = " + to_s(iter_max));
        log.push("Name of simulation = " + s.file_name);
    }
    // parts of the reference that are not on this engine's path stop here with a message
    if (!vp && s.initial_condition_num >= 7 && s.initial_condition_num <= 8)
        fail("initial condition " + s.initial_condition + " draws from cuRAND/std distributions and is not provided", 230305);


    const double *b = s.domain_bounds;
    r.gc = cmm_grid{s.grid_coarse, s.grid_coarse, b[0], b[1], b[2], b[3]};
    r.gf = cmm_grid{s.grid_fine, s.grid_fine, b[0], b[1], b[2], b[3]};
    r.gp = cmm_grid{s.grid_psi, s.grid_psi, b[0], b[1], b[2], b[3]};
    cmm_grid gv{s.grid_vort, s.grid_vort, b[0], b[1], b[2], b[3]};
    r.L = s.lagrange_order;

    cmm_euler_params p{};
    p.g_coarse = r.gc;
    p.g_fine = r.gf;
    p.g_psi = r.gp;
    p.g_vort = gv;
    p.dt = dt;
    p.time_integration_num = s.time_integration_num;
    p.map_update_order_num = s.map_update_order_num;
    p.l_order = s.lagrange_order;
    p.lagrange_init_higher_order = s.lagrange_init_higher_order;
    p.map_epsilon = s.map_epsilon;
    p.incomp_threshold = s.incomp_threshold;
    p.freq_cut_psi = s.freq_cut_psi;
    p.molly_stencil = s.molly_stencil;
    p.skip_remapping = s.skip_remapping;
    p.ic = vp ? CMM_IC_LANDAU_DAMPING + s.initial_condition_num : s.initial_condition_num;
    p.sim = vp ? CMM_SIM_VLASOV_POISSON_1D : CMM_SIM_EULER_2D;
    for (int i = 0; i < 10; ++i) p.ic_params[i] = s.initial_params[i];
    std::vector<double> disc;
    if (vp && s.initial_discrete) fail("a discrete initial condition is not provided for the Vlasov-Poisson flow", 1);
    if (vp && s.grid_vort > s.grid_psi) fail("vorticity grid should be smaller than phi_out grid", 1013);
    if (!vp && s.initial_condition_num == CMM_IC_RANDOM_SPECTRUM) {
        p.discrete = 1;
        p.g_disc = r.gf;
        p.seed = s.seed;
    } else if (s.initial_discrete) {
        // discrete initial condition read from file, Cuda-CMM/src/simulation/cmm-euler2d.cu:325-344
        p.discrete = 1;
        p.g_disc = cmm_grid{s.initial_discrete_grid, s.initial_discrete_grid, b[0], b[1], b[2], b[3]};
        disc.resize((size_t)s.initial_discrete_grid * s.initial_discrete_grid);
        std::ifstream f(s.initial_discrete_location, std::ios::binary);
        if (!f || !f.read(reinterpret_cast<char *>(disc.data()), sizeof(double) * disc.size())) {
            log.push("Unable to read discrete initial condition .. exitting");
            std::exit(0);
        }
        p.disc_real = disc.data();
        if (s.verbose >= 1) log.push("Using discrete initial condition");
    }
    // stack capacity from the memory budget, Cuda-CMM/src/simulation/cmm-euler2d.cu:366-379
    const double map_size = 8.0 * sizeof(double) * r.gc.nx * r.gc.ny / 1e6;
    int cap = (int)(double(s.mem_RAM_CPU_remaps) / map_size);
    if (s.forward_map) cap = (int)(cap / 2.0);  // the forward stack takes the other half, :368
    p.stack_capacity = cap;
    p.forward_map = s.forward_map ? 1 : 0;
    if (s.map_precision == "split") p.precision = CMM_PRECISION_SPLIT;
    else if (s.map_precision == "fp64") p.precision = CMM_PRECISION_FP64;
    else fail("map_precision must be fp64 or split", 1020);
    if (s.verbose >= 1) {
        log.push("Map size in MB = " + to_s((int)map_size));
        log.push("Map stack length on device = " + to_s(cap));
    }
    r.check(cmm_euler_create(&p, &r.e), "cmm_euler_create");

    // forwarded particles, Cuda-CMM/src/simulation/cmm-euler2d.cu:399-421: start positions that stay
    // fixed; the forward maps carry them at each save
    if (s.forward_map)
        for (size_t i = 0; i < s.particles_forwarded.size(); ++i) {
            r.fpos.push_back(init_particles(s.particles_forwarded[i], b));
            r.finit_map.push_back(0);
            if (s.verbose >= 1)
                log.push("Particles set F" + zpad((int)i + 1, 2) + " : Num = " + to_s(s.particles_forwarded[i].num));
        }

    // particles, Cuda-CMM/src/simulation/cmm-euler2d.cu:423-450: fluid (tau 0) and inertial sets, each
    // moving from its init_time on
    for (size_t i = 0; i < s.particles_advected.size(); ++i) {
        const auto &pa = s.particles_advected[i];
        r.part_set.push_back(-1);
        std::vector<double> pos, vel;
        if (s.restart_time != 0) {
            // readParticlesState, Cuda-CMM/src/ui/cmm-io.cu:659-688: positions, and velocities if tau != 0
            pos.resize(2 * (size_t)pa.num);
            const std::string f = restart_dir(s) + "/Particles_advected_pos_P" + zpad((int)i + 1, 2) + ".data";
            if (!read_doubles(f, pos.data(), pos.size())) {
                log.push("Failed to read in particles pos " + to_s((int)i) + " at location " + f);
                log.push("Exitting ..");
                std::exit(0);
            }
            if (pa.tau != 0) {
                vel.resize(2 * (size_t)pa.num);
                const std::string fv = restart_dir(s) + "/Particles_advected_vel_U" + zpad((int)i + 1, 2) + ".data";
                if (!read_doubles(fv, vel.data(), vel.size())) {
                    log.push("Failed to read in particles vel " + to_s((int)i) + " at location " + fv);
                    log.push("Exitting ..");
                    std::exit(0);
                }
            }
        } else {
            pos = init_particles(pa, b);
        }
        int set = -1;
        r.check(cmm_euler_add_particle_set(r.e, pa.num, pa.time_integration_num, pa.init_time, pa.tau, pa.init_vel ? 1 : 0,
                                           pos.data(), vel.empty() ? nullptr : vel.data(), &set),
                "cmm_euler_add_particle_set");
        r.part_set.back() = set;
        if (s.verbose >= 1) {
            std::ostringstream os;
            os.precision(8);
            os << "Particles set P" << zpad((int)i + 1, 2) << " : Num = " << pa.num << " \t Tau = " << pa.tau;
            log.push(os.str());
        }
    }
    if (s.verbose >= 1) log.push("Memory initialization finished \t C-Time = " + duration(ctime()));
    save_param_file(s, s.workspace + "data/" + s.file_name + "/params.txt");
    if (s.verbose >= 2) log.push("Starting simulation initialization \t C-Time = " + duration(ctime()));

    double warm[8];
    int nwarm = 0;
    if (s.restart_time != 0) {
        // restart from a saved map stack, Cuda-CMM/src/simulation/cmm-euler2d.cu:634-684 with readMapStack
        // (Cuda-CMM/src/simulation/cmm-simulation-host.cu:1255-1302)
        if (s.verbose >= 2) log.push("Initializing by loading previous state");
        const std::string dir = restart_dir(s);
        const size_t nc4 = 4 * (size_t)r.gc.nx * r.gc.ny, np4 = 4 * (size_t)r.gp.nx * r.gp.ny;
        const int n_maps = std::max(0, count_files_with(dir, "Map_ChiX_H") - 1);
        std::vector<double> stack(2 * nc4 * n_maps), cx(nc4), cy(nc4), psi_old(np4 * std::max(0, r.L - 1));
        std::string msg;
        for (int m = 0; m < n_maps; ++m) {
            for (int k = 0; k < 2; ++k) {
                const std::string f = dir + "/Map_Chi" + (k ? "Y" : "X") + "_H_" + to_s(m) + ".data";
                if (!read_doubles(f, stack.data() + (2 * m + k) * nc4, nc4))
                    msg += "Failed to read in map " + to_s(m) + " at location " + f + "\n";
            }
        }
        for (int k = 0; k < 2; ++k) {
            const std::string f = dir + "/Map_Chi" + (k ? "Y" : "X") + "_H_coarse.data";
            if (!read_doubles(f, (k ? cy : cx).data(), nc4)) msg += "Failed to read in active map at location " + f + "\n";
        }
        for (int l = 1; l < r.L; ++l) {
            const std::string f = dir + "/Stream_function_Psi_H_psi_" + to_s(l) + ".data";
            if (!read_doubles(f, psi_old.data() + (l - 1) * np4, np4))
                msg += "Failed to read in stream function" + to_s(l) + " at location " + f + "\n";
        }
        if (!msg.empty()) {
            log.push(msg);
            log.push("Exitting ..");
            std::exit(0);
        }
        r.check(cmm_euler_restart(r.e, s.restart_time, n_maps, stack.data(), cx.data(), cy.data(), psi_old.data()),
                "cmm_euler_restart");
        if (s.verbose >= 2) log.push("Loaded " + to_s(n_maps) + " maps to map stack");
        if (s.forward_map) {
            // the forward maps of the same folder (readMapStack with isForward, files Map_ChiX_f_H_*)
            for (int m = 0; m < n_maps; ++m)
                for (int k = 0; k < 2; ++k) {
                    const std::string f = dir + "/Map_Chi" + (k ? "Y" : "X") + "_f_H_" + to_s(m) + ".data";
                    if (!read_doubles(f, stack.data() + (2 * m + k) * nc4, nc4))
                        msg += "Failed to read in map " + to_s(m) + " at location " + f + "\n";
                }
            for (int k = 0; k < 2; ++k) {
                const std::string f = dir + "/Map_Chi" + (k ? "Y" : "X") + "_f_H_coarse.data";
                if (!read_doubles(f, (k ? cy : cx).data(), nc4)) msg += "Failed to read in active map at location " + f + "\n";
            }
            if (!msg.empty()) {
                log.push(msg);
                log.push("Exitting ..");
                std::exit(0);
            }
            r.check(cmm_euler_restart_forward(r.e, n_maps, stack.data(), cx.data(), cy.data()), "cmm_euler_restart_forward");
            if (s.verbose >= 2) log.push("Loaded " + to_s(n_maps) + " maps to forward map stack");
        }
        if (!s.particles_advected.empty() && s.verbose >= 2) log.push("Loaded data for particle sets");
    } else {
        r.check(cmm_euler_init(r.e, warm, &nwarm), "cmm_euler_init");
    }
    // warm-up log in the reference's format, Cuda-CMM/src/simulation/cmm-euler2d.cu:557-596
    if (s.restart_time == 0 && s.verbose >= 2 && r.L > 1) {
        int k = 0;
        for (int order = 1; order <= r.L; ++order) {
            int n = order - (order == r.L);
            if (!s.lagrange_init_higher_order && order != r.L) continue;
            const char *method = !s.lagrange_init_higher_order ? "EulerExp"
                                 : order == 1 ? "EulerExp" : order == 2 ? "RK2" : order == 3 ? "RK3" : "RK4";
            log.push("Init order = " + to_s(order) + "/" + to_s(r.L) + " \t Method = " + method);
            for (int st = 1; st <= n && k < nwarm; ++st, ++k)
                log.push("   Init step = " + to_s(st) + "/" + to_s(n) + " \t IncompErr = " + fmt3(warm[k]));
        }
    }
    r.save(s.restart_time, dt, dt);

    std::vector<double> tv(s.lagrange_order), dtv(s.lagrange_order);
    for (int i = 0; i < s.lagrange_order; ++i) { tv[i] = s.restart_time - (s.lagrange_order - 1 - i) * dt; dtv[i] = dt; }
    append1(s, "/Monitoring_data/Time_c", ctime());
    if (s.verbose >= 1) log.push("Simulation initialization finished \t C-Time = " + duration(ctime()));
    if (s.verbose >= 2) log.push("Starting simulation loop \t C-Time = " + duration(ctime()));

    int loop = 0, old = 0;
    bool cont = true;
    const int L = s.lagrange_order;
    while (s.final_time - tv[loop + L - 1] > dt * 1e-5 && loop < iter_max && cont) {
        const int lc = loop + L - 1;
        double dt_now = compute_next_timestep(s, tv[lc], dt);
        if (tv[lc] + dt_now > s.final_time) dt_now = s.final_time - tv[lc];
        // forwarded particles starting in this step remap both maps, :804-811
        if (s.forward_map) {
            const double t_next = tv[lc] + dt_now;
            bool start = false;
            int steps = 0, ns = 0;
            double t_cur = 0;
            cmm_euler_state(r.e, &steps, &t_cur, &ns);
            for (size_t i = 0; i < s.particles_forwarded.size(); ++i) {
                const double d = t_next - s.particles_forwarded[i].init_time + dt * 1e-5;
                if (d < dt_now && d >= 0) {
                    start = true;
                    r.finit_map[i] = ns + 1;
                }
            }
            if (start) r.check(cmm_euler_request_remap(r.e), "cmm_euler_request_remap");
        }
        cmm_step_info info;
        r.check(cmm_euler_step(r.e, dt_now, &info), "cmm_euler_step");
        tv.push_back(info.t);
        dtv.push_back(dt_now);
        if (info.saturated) {
            log.push("Stack Saturated : Exitting");
            cont = false;
        }
        if (vp && (info.incomp <= 0 || info.incomp > 1e15))
            fail("Dear captain there is an incompressibility error in coarse grid. Aborting.", 2345);
        if (info.remapped) {
            if (s.verbose >= 2)
                log.push("Refining Map : Step = " + to_s(loop) + " \t Maps = " + to_s(info.stack_size) + " \t Gap = " + to_s(loop - old));
            old = loop;
            // writeMapStack at the remap (Cuda-CMM/src/simulation/cmm-euler2d.cu:841): the pushed map and
            // the psi level that was newest during the step (level 1 once the step has run)
            if (s.save_map_stack) r.write_map_stack(info.stack_size - 1, 1);
            if (s.save_map_stack && s.forward_map) r.write_map_stack(info.stack_size - 1, 1, true);
        }
        const double gap = loop - old, counter = info.stack_size;
        r.save(info.t, dt_now, dt);
        loop++;
        append1(s, "/Monitoring_data/Time_s", info.t);
        append1(s, "/Monitoring_data/Error_incompressibility", info.incomp);
        if (s.forward_map) {
            append1(s, "/Monitoring_data/Error_incompressibility_forward", info.incomp_forward);
            append1(s, "/Monitoring_data/Error_invertibility", info.invertibility);
        }
        append1(s, "/Monitoring_data/Map_gaps", gap);
        append1(s, "/Monitoring_data/Map_counter", counter);
        const double c_time = ctime();
        if (s.verbose >= 2)
            log.push("Step = " + to_s(loop) + " \t S-Time = " + fmt3(info.t) + "/" + fmt3(s.final_time) + " \t E-inc = " +
                     fmt3(info.incomp) +
                     (s.forward_map ? " \t E-inc_f = " + fmt3(info.incomp_forward) + " \t E-inv = " + fmt3(info.invertibility)
                                    : std::string()) +
                     " \t C-Time = " + duration(c_time));
        append1(s, "/Monitoring_data/Time_c", c_time);
    }
    if (s.verbose >= 2) log.push("Simulation loop finished \t C-Time = " + duration(ctime()));
    r.save(kTMax, dt, dt);
    if (s.save_map_stack) {
        // Cuda-CMM/src/simulation/cmm-euler2d.cu:1033-1051: the active map, the newest psi level and
        // the particle state
        int steps = 0, n_stack = 0;
        double t_end = 0;
        cmm_euler_state(r.e, &steps, &t_end, &n_stack);
        if (s.verbose >= 2) {
            std::ostringstream os;
            os.precision(16);
            os << t_end;
            log.push("Saving last Map of MapStack : Maps = " + to_s(n_stack) + " \t Total size = " +
                     fmt3((n_stack + 1) * map_size) + "mb \t S-time = " + os.str());
        }
        r.write_map_stack(-1, 0);
        if (s.forward_map) {
            if (s.verbose >= 2)
                log.push("Saving last map of forward MapStack : Maps = " + to_s(n_stack) + " \t Total size = " +
                         fmt3((n_stack + 1) * map_size) + "mb");
            r.write_map_stack(-1, 0, true);
        }
        for (size_t i = 0; i < s.particles_advected.size(); ++i) {
            if (r.part_set[i] < 0) continue;
            std::vector<double> pos(2 * (size_t)s.particles_advected[i].num);
            r.check(cmm_euler_get_particles(r.e, r.part_set[i], pos.data()), "cmm_euler_get_particles");
            write_file(s, "/MapStack/Particles_advected_pos_P" + zpad((int)i + 1, 2), pos.data(), pos.size(), false);
            // writeParticlesState, Cuda-CMM/src/ui/cmm-io.cu:638-657: velocities of inertial sets
            if (s.particles_advected[i].tau != 0) {
                std::vector<double> vel(pos.size());
                r.check(cmm_euler_get_particle_velocities(r.e, r.part_set[i], vel.data()), "cmm_euler_get_particle_velocities");
                write_file(s, "/MapStack/Particles_advected_vel_U" + zpad((int)i + 1, 2), vel.data(), vel.size(), false);
            }
        }
    }
    cmm_euler_destroy(r.e);
    append1(s, "/Monitoring_data/Time_c", ctime());
    if (s.verbose >= 1) log.push("Finished simulation \t C-Time = " + duration(ctime()));
    return 0;
}

}  // namespace cmmhost
