// C-ABI entry points (include/cmm.h) and the Euler engine that runs the reference's time loop on
// the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <map>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cmm_internal.hpp"
#include "../../include/cmm.h"

using namespace cmm;

namespace {

hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

bool grid_ok(const cmm_grid &g) { return g.nx > 0 && g.ny > 0 && g.x1 > g.x0 && g.y1 > g.y0; }

PsiLevels levels(const double *const *p, int L) {
    PsiLevels P;
    for (int l = 0; l < 4; ++l) P.p[l] = p[l < L ? l : 0];
    return P;
}

// SplitMix64 finaliser (Steele, Lea, Flood 2014)
uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace

// ================================================================================================
extern "C" {

int cmm_version(void) { return 1; }

const char *cmm_last_error(void) { return cmm::last_error(); }

int cmm_device_info(char *buf, int len) {
    int dev;
    if (int rc = hip_check(hipGetDevice(&dev), "hipGetDevice")) return rc;
    hipDeviceProp_t p;
    if (int rc = hip_check(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties")) return rc;
    std::snprintf(buf, len, "%s, %s, CUs %d, LDS/block %zu, HBM %zu", p.name, p.gcnArchName, p.multiProcessorCount,
                  (size_t)p.sharedMemPerBlock, (size_t)p.totalGlobalMem);
    return CMM_OK;
}

double cmm_lagrange_coeff(const double *t, double t_next, int loop_ctr, int i_point, int l_order) {
    double coeff = 1;
    for (int i_p = 0; i_p < l_order; ++i_p)
        if (i_p != i_point) coeff *= (t_next - t[loop_ctr - i_p]) / (t[loop_ctr - i_point] - t[loop_ctr - i_p]);
    return coeff;
}

int cmm_advect_coeffs_fill(const double *t, const double *dt, int loop_ctr, int l_order, int tin, int muo, double eps,
                           int direction, cmm_advect_coeffs *c) {
    if (!t || !dt || !c || l_order < 1 || l_order > 4 || muo < 0 || muo > 2 || !(eps > 0)) {
        set_error("cmm_advect_coeffs_fill: bad argument");
        return CMM_ERR_ARG;
    }
    std::memset(c, 0, sizeof(*c));
    const int lc = loop_ctr + l_order - 1;
    for (int i = 0; i < l_order; ++i) {
        c->L1[i] = cmm_lagrange_coeff(t, t[lc + 1], lc, i, l_order);
        c->L12[i] = cmm_lagrange_coeff(t, t[lc] + dt[lc + 1] / 2.0, lc, i, l_order);
    }
    // stencil weights of the 2nd/4th/6th order map update with signs and 1/eps folded in,
    // Cuda-CMM/src/simulation/cmm-simulation-host.cu:75-90
    double h[3] = {0, 0, 0};
    if (muo == 2) { h[0] = 3.0 / 8.0; h[1] = -3.0 / 20.0; h[2] = 1.0 / 40.0; }
    else if (muo == 1) { h[0] = 1.0 / 3.0; h[1] = -1.0 / 12.0; }
    else { h[0] = 1.0 / 4.0; }
    for (int f = 0; f < 4 + 4 * muo; ++f) {
        const int ring = f / 4;
        c->c1[f] = h[ring];
        c->cx[f] = h[ring] * (1 - 2 * ((f / 2) % 2)) / eps / double(ring + 1);
        c->cy[f] = h[ring] * (1 - 2 * (((f + 1) / 2) % 2)) / eps / double(ring + 1);
        c->cxy[f] = h[ring] * (1 - 2 * (f % 2)) / eps / eps / double(ring + 1) / double(ring + 1);
    }
    c->dt = dt[lc + 1];
    c->eps = eps;
    c->time_integration_num = tin;
    c->map_update_order_num = muo;
    c->l_order = l_order;
    c->direction = direction;
    return CMM_OK;
}

int cmm_map_init_identity(cmm_grid g, double *cx, double *cy, void *stream) {
    if (!grid_ok(g) || !cx || !cy) { set_error("cmm_map_init_identity: bad argument"); return CMM_ERR_ARG; }
    return op_map_identity(g, cx, cy, S(stream));
}

size_t cmm_advect_workspace_bytes(cmm_grid gm, cmm_grid gp, const cmm_advect_coeffs *c) {
    if (!grid_ok(gm) || !grid_ok(gp) || !c) return 0;
    return op_advect_workspace(gm, gp, *c);
}

int cmm_map_advect(cmm_grid gm, cmm_grid gp, const double *cx, const double *cy, const double *const *psi,
                   const cmm_advect_coeffs *c, double *ncx, double *ncy, void *work, size_t work_bytes,
                   void *stream) {
    if (!grid_ok(gm) || !grid_ok(gp) || !cx || !cy || !psi || !c || !ncx || !ncy) {
        set_error("cmm_map_advect: bad argument");
        return CMM_ERR_ARG;
    }
    if (c->direction != -1 && c->direction != 1) { set_error("cmm_map_advect: direction must be -1 or 1"); return CMM_ERR_ARG; }
    if (c->l_order < 1 || c->l_order > 4 || c->map_update_order_num < 0 || c->map_update_order_num > 2) {
        set_error("cmm_map_advect: bad coefficients");
        return CMM_ERR_ARG;
    }
    if (ncx == cx || ncy == cy || ncx == cy || ncy == cx) { set_error("cmm_map_advect: output aliases input"); return CMM_ERR_ARG; }
    for (int l = 0; l < c->l_order; ++l)
        if (!psi[l]) { set_error("cmm_map_advect: null psi level"); return CMM_ERR_ARG; }
    return op_map_advect(gm, gp, cx, cy, levels(psi, c->l_order), *c, ncx, ncy, work, work_bytes, S(stream));
}

int cmm_incomp_error(cmm_grid gm, cmm_grid gc, const double *cx, const double *cy, double *out, void *stream) {
    if (!grid_ok(gm) || !grid_ok(gc) || !cx || !cy || !out) { set_error("cmm_incomp_error: bad argument"); return CMM_ERR_ARG; }
    return op_incomp(gm, gc, cx, cy, out, S(stream));
}

int cmm_invert_error(cmm_grid gb, cmm_grid gf, cmm_grid gc, const double *bx, const double *by, const double *fx,
                     const double *fy, double *out, void *stream) {
    if (!grid_ok(gb) || !grid_ok(gf) || !grid_ok(gc) || !bx || !by || !fx || !fy || !out) {
        set_error("cmm_invert_error: bad argument");
        return CMM_ERR_ARG;
    }
    return op_invert(gb, gf, gc, bx, by, fx, fy, out, S(stream));
}

int cmm_compose_positions(cmm_grid gm, cmm_grid gt, const double *cx, const double *cy, const double *stack,
                          int ns, double *xy, void *stream) {
    if (!grid_ok(gm) || !grid_ok(gt) || !cx || !cy || !xy || ns < 0 || (ns > 0 && !stack)) {
        set_error("cmm_compose_positions: bad argument");
        return CMM_ERR_ARG;
    }
    return op_compose_positions(gm, gt, cx, cy, stack, ns, xy, S(stream));
}

int cmm_stack_pack(cmm_grid gm, const double *stack, int n_maps, int format, void *packed, void *stream) {
    if (!grid_ok(gm) || n_maps < 0 || (n_maps > 0 && (!stack || !packed)) ||
        (format != CMM_STACK_RECORDS_FP64 && format != CMM_STACK_RECORDS_SPLIT)) {
        set_error("cmm_stack_pack: bad argument");
        return CMM_ERR_ARG;
    }
    const long N = (long)gm.nx * gm.ny;
    return op_pack_maps(gm, stack, stack + 4 * N, 8 * N, n_maps, format, packed, S(stream));
}

int cmm_compose_positions_packed(cmm_grid gm, cmm_grid gt, const double *cx, const double *cy, const void *packed,
                                 int ns, int format, double *xy, void *stream) {
    if (!grid_ok(gm) || !grid_ok(gt) || !cx || !cy || !xy || ns < 0 || (ns > 0 && !packed) ||
        (format != CMM_STACK_RECORDS_FP64 && format != CMM_STACK_RECORDS_SPLIT)) {
        set_error("cmm_compose_positions_packed: bad argument");
        return CMM_ERR_ARG;
    }
    return op_compose_positions(gm, gt, cx, cy, static_cast<const double *>(packed), ns, xy, S(stream), format);
}

int cmm_forward_positions(cmm_grid gm, cmm_grid gt, const double *cx, const double *cy, const double *stack,
                          int ns, double *xy, void *stream) {
    if (!grid_ok(gm) || !grid_ok(gt) || !cx || !cy || !xy || ns < 0 || (ns > 0 && !stack)) {
        set_error("cmm_forward_positions: bad argument");
        return CMM_ERR_ARG;
    }
    return op_forward_positions(gm, gt, cx, cy, stack, ns, xy, S(stream));
}

int cmm_forward_points(cmm_grid gm, cmm_grid gb, const double *cx, const double *cy, const double *stack, int ns,
                       int first_map, int apply_current, int n, const double *in, double *out, void *stream) {
    if (!grid_ok(gm) || !cx || !cy || ns < 0 || (ns > 0 && !stack) || n < 0 || (n > 0 && (!in || !out || in == out))) {
        set_error("cmm_forward_points: bad argument");
        return CMM_ERR_ARG;
    }
    return op_forward_points(gm, gb, stack, first_map, ns, cx, cy, apply_current, in, out, n, S(stream));
}

int cmm_sample_init(cmm_grid gm, cmm_grid gt, const double *cx, const double *cy, const double *stack, int ns, int ic,
                    const double *prm, int discrete, cmm_grid gd, const double *disc, double *out, void *stream) {
    if (!grid_ok(gm) || !grid_ok(gt) || !cx || !cy || !out || ns < 0 || (ns > 0 && !stack) ||
        (discrete && (!disc || !grid_ok(gd)))) {
        set_error("cmm_sample_init: bad argument");
        return CMM_ERR_ARG;
    }
    return op_sample_init(gm, gt, cx, cy, stack, ns, ic, prm, discrete, discrete ? gd : gt, disc, out, S(stream));
}

size_t cmm_translate_workspace_bytes(cmm_grid g) { return grid_ok(g) ? op_translate_workspace(g) : 0; }

int cmm_translate_ic(cmm_grid gm, cmm_grid gf, const double *cx, const double *cy, const double *stack, int ns, int ic,
                     const double *prm, int discrete, cmm_grid gd, const double *disc, double *W, void *work,
                     size_t wb, void *stream) {
    if (!grid_ok(gm) || !grid_ok(gf) || !cx || !cy || !W || ns < 0 || (ns > 0 && !stack) ||
        (discrete && (!disc || !grid_ok(gd)))) {
        set_error("cmm_translate_ic: bad argument");
        return CMM_ERR_ARG;
    }
    return op_translate_ic(gm, gf, cx, cy, stack, ns, ic, prm, discrete, discrete ? gd : gf, disc, W, work, wb, S(stream));
}

size_t cmm_hermite_workspace_bytes(cmm_grid g) { return op_hermite_workspace(g); }

int cmm_hermite_from_real(cmm_grid g, const double *real, double *herm, void *work, size_t wb, void *stream) {
    if (!real || !herm) { set_error("cmm_hermite_from_real: null pointer"); return CMM_ERR_ARG; }
    return op_hermite_from_real(g, real, herm, work, wb, S(stream));
}

int cmm_hermite_from_spectrum(cmm_grid g, const double *spec, double *herm, void *work, size_t wb, void *stream) {
    if (!spec || !herm) { set_error("cmm_hermite_from_spectrum: null pointer"); return CMM_ERR_ARG; }
    return op_hermite_from_spectrum(g, spec, herm, work, wb, S(stream));
}

size_t cmm_psi_workspace_bytes(cmm_grid gv, cmm_grid gp, double fcut) { return op_psi_workspace(gv, gp, fcut); }

size_t cmm_spectral_workspace_bytes(cmm_grid g) { return grid_ok(g) ? op_hermite_workspace(g) : 0; }

int cmm_spectral_from_real(cmm_grid g, const double *real, int op, double *herm, void *work, size_t wb, void *stream) {
    if (!grid_ok(g) || !real || !herm) { set_error("cmm_spectral_from_real: bad argument"); return CMM_ERR_ARG; }
    return op_spectral_from_real(g, real, op, herm, work, wb, S(stream));
}

size_t cmm_vp_psi_workspace_bytes(cmm_grid gv, cmm_grid gp) {
    if (!grid_ok(gv) || !grid_ok(gp)) return 0;
    return op_vp_psi_workspace(gv, gp);
}

int cmm_vp_psi_from_f(cmm_grid gm, cmm_grid gv, cmm_grid gf, cmm_grid gp, const double *cx, const double *cy,
                      const double *f, int molly, double *psi, void *work, size_t wb, void *stream) {
    if (!grid_ok(gm) || !grid_ok(gv) || !grid_ok(gf) || !grid_ok(gp) || !cx || !cy || !f || !psi ||
        (molly != 0 && molly != 4 && molly != 8)) {
        set_error("cmm_vp_psi_from_f: bad argument");
        return CMM_ERR_ARG;
    }
    return op_vp_psi_from_f(gm, gv, gf, gp, cx, cy, f, molly, psi, work, wb, S(stream));
}

int cmm_psi_from_vort(cmm_grid gm, cmm_grid gv, cmm_grid gf, cmm_grid gp, const double *cx, const double *cy,
                      const double *W, int molly, double fcut, double *psi, void *work, size_t wb, void *stream) {
    if (!grid_ok(gm) || !grid_ok(gf) || !cx || !cy || !W || !psi || (molly != 0 && molly != 4 && molly != 8)) {
        set_error("cmm_psi_from_vort: bad argument");
        return CMM_ERR_ARG;
    }
    return op_psi_from_vort(gm, gv, gf, gp, cx, cy, W, molly, fcut, psi, work, wb, S(stream));
}

int cmm_psi_from_vort_noalloc(cmm_grid gm, cmm_grid gv, cmm_grid gf, cmm_grid gp, const double *cx, const double *cy,
                              const double *W, int molly, double fcut, double *psi, void *stream) {
    if (!grid_ok(gm) || !grid_ok(gf) || !cx || !cy || !W || !psi || (molly != 0 && molly != 4 && molly != 8)) {
        set_error("cmm_psi_from_vort_noalloc: bad argument");
        return CMM_ERR_ARG;
    }
    return op_psi_from_vort_noalloc(gm, gv, gf, gp, cx, cy, W, molly, fcut, psi, S(stream));
}

int cmm_sample_vort(cmm_grid gm, cmm_grid gt, cmm_grid gf, const double *cx, const double *cy, const double *W,
                    int molly, double *out, void *stream) {
    if (!grid_ok(gm) || !grid_ok(gt) || !grid_ok(gf) || !cx || !cy || !W || !out) {
        set_error("cmm_sample_vort: bad argument");
        return CMM_ERR_ARG;
    }
    return op_sample_vort(gm, gt, gf, cx, cy, W, molly, out, S(stream));
}

int cmm_particles_advect(cmm_grid gp, const double *const *psi, int L, const double *Lp, double dt, int scheme, int n,
                         double *pos, void *stream) {
    if (!grid_ok(gp) || !psi || L < 1 || L > 4 || n < 0 || (n > 0 && !pos)) {
        set_error("cmm_particles_advect: bad argument");
        return CMM_ERR_ARG;
    }
    for (int l = 0; l < L; ++l)
        if (!psi[l]) { set_error("cmm_particles_advect: null psi level"); return CMM_ERR_ARG; }
    if (L < 2 && scheme != 10 && scheme != 20) {
        set_error("cmm_particles_advect: scheme needs at least two psi levels");
        return CMM_ERR_ARG;
    }
    return op_particles(gp, levels(psi, L), L, Lp, dt, scheme, n, pos, S(stream));
}

int cmm_map_to_split(cmm_grid g, const double *cx, const double *cy, float *M, void *stream) {
    if (!grid_ok(g) || !cx || !cy || !M) { set_error("cmm_map_to_split: bad argument"); return CMM_ERR_ARG; }
    return op_map_to32(g, cx, cy, M, S(stream));
}

int cmm_map_from_split(cmm_grid g, const float *M, double *cx, double *cy, void *stream) {
    if (!grid_ok(g) || !cx || !cy || !M) { set_error("cmm_map_from_split: bad argument"); return CMM_ERR_ARG; }
    return op_map_from32(g, M, cx, cy, S(stream));
}

int cmm_hermite_to_split(cmm_grid g, const double *H, float *H4, void *stream) {
    if (!grid_ok(g) || !H || !H4) { set_error("cmm_hermite_to_split: bad argument"); return CMM_ERR_ARG; }
    return op_interleave4_32(g, H, H4, S(stream));
}

size_t cmm_map_advect_split_workspace_bytes(cmm_grid gp, const cmm_advect_coeffs *c) {
    if (!grid_ok(gp) || !c) return 0;
    return op_advect32_workspace(gp, *c);
}

int cmm_map_advect_split(cmm_grid gm, cmm_grid gp, const float *M_in, const double *const *psi,
                         const cmm_advect_coeffs *c, float *M_out, void *work, size_t work_bytes, void *stream) {
    if (!grid_ok(gm) || !grid_ok(gp) || !M_in || !M_out || !psi || !c || M_in == M_out) {
        set_error("cmm_map_advect_split: bad argument");
        return CMM_ERR_ARG;
    }
    if (c->l_order < 2 || c->l_order > 4 || c->map_update_order_num < 0 || c->map_update_order_num > 2) {
        set_error("cmm_map_advect_split: bad coefficients");
        return CMM_ERR_ARG;
    }
    for (int l = 0; l < c->l_order; ++l)
        if (!psi[l]) { set_error("cmm_map_advect_split: null psi level"); return CMM_ERR_ARG; }
    return op_map_advect32(gm, gp, levels(psi, c->l_order), *c, M_in, M_out, work, work_bytes, S(stream));
}

int cmm_incomp_error_split(cmm_grid gm, cmm_grid gc, const float *M, double *out, void *stream) {
    if (!grid_ok(gm) || !grid_ok(gc) || !M || !out) { set_error("cmm_incomp_error_split: bad argument"); return CMM_ERR_ARG; }
    return op_incomp32(gm, gc, M, out, S(stream));
}

int cmm_psi_from_vort_split(cmm_grid gm, cmm_grid gv, cmm_grid gf, cmm_grid gp, const float *M, const float *W4,
                            int molly, double fcut, double *psi, void *work, size_t wb, void *stream) {
    if (!grid_ok(gm) || !grid_ok(gf) || !M || !W4 || !psi || (molly != 0 && molly != 4 && molly != 8)) {
        set_error("cmm_psi_from_vort_split: bad argument");
        return CMM_ERR_ARG;
    }
    return op_psi_from_vort32(gm, gv, gf, gp, M, W4, molly, fcut, psi, work, wb, S(stream));
}

}  // extern "C"

// ================================================================================================
// Engine

// Events that only order streams on the device (never waited on by the host) need only a
// device-scope release.
constexpr unsigned kDeviceEvent = hipEventDisableTiming | hipEventReleaseToDevice;

struct cmm_euler {
    cmm_euler_params p;
    hipStream_t st = nullptr;
    long nc = 0, nf = 0, np = 0, nd = 0;
    double *cx = nullptr, *cy = nullptr, *ncx = nullptr, *ncy = nullptr;
    double *psi_buf[4] = {nullptr, nullptr, nullptr, nullptr};
    double *lvl[4] = {nullptr, nullptr, nullptr, nullptr};  // level -> buffer, newest first
    double *W = nullptr, *disc = nullptr, *stack = nullptr, *tmp = nullptr, *tmp4 = nullptr;
    double *W4 = nullptr;  // W node-interleaved (Euler: read by the psi sampler)
    // split form: the stack again in the split stack format (FP64 values, FP32 derivatives), kept
    // beside the FP64 stack and read by every composition; the FP64 stack serves saves and readback
    double *stack_s = nullptr;
    double *m8[2] = {nullptr, nullptr};  // the map node-interleaved (ping-pong), see engine_advect
    double *sbuf = nullptr;              // sample-grid buffers of cmm_euler_sample
    size_t sbuf_n = 0;
    int m8_cur = 0;
    bool m8_valid = false;
    // split-precision form: the map as displacement records (12 floats per node; ping-pong) and the fine
    // vorticity as FP32 node records.  cx/cy and m32[m32_cur] hold the same map when both are
    // fresh; a consumer brings the form it reads up to date, a writer marks the other stale.
    float *m32[2] = {nullptr, nullptr};
    float *W4f = nullptr;
    int m32_cur = 0;
    bool m32_fresh = false, planes_fresh = true;
    // dry: a step's host state advances without device work (the work is replayed from a captured
    // graph); capturing: the step's work is being captured (no event records)
    bool dry = false, capturing = false;
    // graph steps (captured or replayed): the advection reads its Lagrange weights from lw_dev
    // (device), which the batch fills from lw_host (pinned) before each graph launch, since the
    // weights follow the whole time history and a graph's kernel arguments are frozen
    bool lw_graph = false;
    double *lw_dev = nullptr, *lw_host = nullptr;
    // Vlasov-Poisson: H(v) and H'(v) of the bump function (they depend on the grid only)
    double *vpH = nullptr;
    IncompTune tune;  // the check's tile and grid, from cmm_euler_autotune
    // the step's check runs inside the psi sampler's pass (the check grid is the vorticity grid);
    // sampled: the current step's sampling is queued, its spectral part not yet
    bool fuse_check = false, sampled = false;
    // the next step's combined psi fields (cpre, three fields), written by the psi solve's last
    // transform for the weights of a next step as long as this one (pre); the next advection reads
    // them in place of combining when its own weights, levels and scheme are those
    void *cpre = nullptr;
    PsiPre pre{};
    int pre_tin = 0;
    bool pre_valid = false, pre_enabled = true;
    // grouped ensemble steps: while set, the step records the arguments of its launches here instead
    // of launching them (the batch launches every engine's work together, op_group_vp_step)
    GroupInstance *grp = nullptr;
    // corrections (device Hermite fields on g_psi) added to the Lagrange extrapolations of psi
    // to t_{n+1} and t_n + dt/2 in the next steps (cmm_euler_set_psi_correction)
    const double *corr1 = nullptr, *corr12 = nullptr;
    int nstack = 0;
    // forward map (forward_map only): current, next, and its stack, pushed with the backward one
    double *fx = nullptr, *fy = nullptr, *nfx = nullptr, *nfy = nullptr, *fstack = nullptr;
    bool force_remap = false;  // set by cmm_euler_request_remap for the next step
    void *work = nullptr;
    size_t work_bytes = 0;
    unsigned long long *slot = nullptr;
    unsigned long long *h_slot = nullptr;  // pinned host copy of slot
    hipEvent_t ev_inc = nullptr;           // set after the E_inc copy of a step
    // particles run on a second stream after the psi solve and overlap the next step's advection;
    // the next psi solve (which overwrites the oldest level) waits for them
    hipStream_t st2 = nullptr;
    hipEvent_t ev_psi = nullptr, ev_part = nullptr;
    // single-engine steps (cmm_euler_step) run the check on a third stream beside the psi solve
    // (both only read the new map); ensemble steps keep it on the engine stream
    hipStream_t st3 = nullptr;
    hipEvent_t ev_fork = nullptr;
    bool part_pending = false;
    std::vector<double *> part;
    std::vector<int> part_n, part_scheme;
    // per set: start time (the set moves once t_{n+1} + 1e-5 dt reaches it), tau (non-zero:
    // inertial particles with velocities part_vel), the start-velocity rule and whether the caller
    // gave the velocities
    std::vector<double *> part_vel;
    std::vector<double> part_t0, part_tau;
    std::vector<int> part_init_vel, part_vel_given;
    std::vector<double> tv, dtv;
    int loop = 0;
    bool initialised = false;
    // per-kernel event timing: [advect, incomp, psi_row, psi_col, psi_c2r, particles, remap, psi_combine]
    bool prof = false;     // events in the current step
    int prof_every = 0;    // events on every prof_every-th step (0: none)
    // event sets of the last kProfSets steps; a set's intervals are read when the set comes round
    // again (its events long complete), so profiling never makes the host wait inside a step
    static constexpr int kProfSets = 4;
    hipEvent_t evs[kProfSets][12] = {};
    hipEvent_t *ev = evs[0];
    int cur = 0;  // index of ev in evs
    std::vector<std::array<int, 3>> pend[kProfSets];  // (slot, from, to) still to be read
    double ms[CMM_PROF_SLOTS] = {};
    int cnt[CMM_PROF_SLOTS] = {};
};

namespace {

int dalloc(double **p, size_t n) {
    if (hipMalloc(p, sizeof(double) * n) != hipSuccess) {
        *p = nullptr;
        set_error("hipMalloc failed (" + std::to_string(sizeof(double) * n) + " bytes)");
        return CMM_ERR_ALLOC;
    }
    return CMM_OK;
}

int falloc(float **p, size_t n) {
    if (hipMalloc(p, sizeof(float) * n) != hipSuccess) {
        *p = nullptr;
        set_error("hipMalloc failed (" + std::to_string(sizeof(float) * n) + " bytes)");
        return CMM_ERR_ALLOC;
    }
    return CMM_OK;
}

bool split(const cmm_euler *e) { return e->p.precision == CMM_PRECISION_SPLIT; }

// the stack the compositions read, and its format
const double *comp_stack(const cmm_euler *e) { return e->stack_s ? e->stack_s : e->stack; }
int comp_fmt(const cmm_euler *e) { return e->stack_s ? kStackRec48 : kStackRec64; }

int sync_planes(cmm_euler *e) {
    if (e->planes_fresh) return CMM_OK;
    if (!e->dry)
        if (int rc = op_map_from32(e->p.g_coarse, e->m32[e->m32_cur], e->cx, e->cy, e->st)) return rc;
    e->planes_fresh = true;
    return CMM_OK;
}

int sync_m32(cmm_euler *e) {
    if (e->m32_fresh) return CMM_OK;
    if (!e->dry)
        if (int rc = op_map_to32(e->p.g_coarse, e->cx, e->cy, e->m32[e->m32_cur], e->st)) return rc;
    e->m32_fresh = true;
    return CMM_OK;
}

// after a write to cx/cy (identity, restart, an FP64 advection)
void planes_written(cmm_euler *e) {
    e->planes_fresh = true;
    e->m32_fresh = false;
    e->m8_valid = false;
}

int engine_translate(cmm_euler *e) {
    if (int rc = sync_planes(e)) return rc;
    if (int rc = op_translate_ic(e->p.g_coarse, e->p.g_fine, e->cx, e->cy, comp_stack(e), e->nstack, e->p.ic,
                                 e->p.ic_params, e->p.discrete, e->p.discrete ? e->p.g_disc : e->p.g_fine, e->disc,
                                 e->W, e->work, e->work_bytes, e->st, comp_fmt(e)))
        return rc;
    // the psi sampler reads the fine vorticity node-interleaved (one 32-byte record per corner,
    // 16 bytes in the split form)
    if (e->W4f) return op_interleave4_32(e->p.g_fine, e->W, e->W4f, e->st);
    return e->W4 ? op_interleave4(e->p.g_fine, e->W, e->W4, e->st) : CMM_OK;
}

int engine_psi(cmm_euler *e, double *out, hipEvent_t *ev = nullptr, const PsiPre *pre = nullptr) {
    if (e->grp) {
        if (e->p.sim != CMM_SIM_VLASOV_POISSON_1D || split(e)) {
            set_error("grouped step: psi of an engine that is not an FP64 Vlasov-Poisson engine");
            return CMM_ERR_UNSUPPORTED;
        }
        GroupInstance &g = *e->grp;
        g.gv = e->p.g_vort;
        g.gf = e->p.g_fine;
        g.pcx = e->cx;
        g.pcy = e->cy;
        g.W4 = e->W4;
        g.psi = out;
        g.psi_work = e->work;
        g.psi_work_bytes = e->work_bytes;
        g.parts |= 4;
        return CMM_OK;
    }
    if (e->dry) return split(e) ? (e->p.sim == CMM_SIM_VLASOV_POISSON_1D ? sync_planes(e) : sync_m32(e)) : CMM_OK;
    if (e->p.sim == CMM_SIM_VLASOV_POISSON_1D) {
        // one operator; its whole time goes to the first psi profile slot
        if (int rc = sync_planes(e)) return rc;
        if (ev) (void)hipEventRecord(ev[0], e->st);
        int rc = op_vp_psi_from_f(e->p.g_coarse, e->p.g_vort, e->p.g_fine, e->p.g_psi, e->cx, e->cy, e->W,
                                  e->p.molly_stencil, out, e->work, e->work_bytes, e->st, e->W4, e->vpH);
        if (ev)
            for (int i = 1; i < 4; ++i) (void)hipEventRecord(ev[i], e->st);
        return rc;
    }
    if (split(e)) {
        if (int rc = sync_m32(e)) return rc;
        return op_psi_from_vort32(e->p.g_coarse, e->p.g_vort, e->p.g_fine, e->p.g_psi, e->m32[e->m32_cur], e->W4f,
                                  e->p.molly_stencil, e->p.freq_cut_psi, out, e->work, e->work_bytes, e->st, ev, pre);
    }
    return op_psi_from_vort(e->p.g_coarse, e->p.g_vort, e->p.g_fine, e->p.g_psi, e->cx, e->cy, e->W,
                            e->p.molly_stencil, e->p.freq_cut_psi, out, e->work, e->work_bytes, e->st, ev, e->W4, pre);
}

// the next step's combined fields to be written by this step's psi solve, for a next step as long
// as this one (null where the engine state does not allow them)
const PsiPre *prepare_pre(cmm_euler *e) {
    const cmm_euler_params &p = e->p;
    e->pre_valid = false;
    if (!e->pre_enabled || !e->cpre || e->dry || e->capturing || e->lw_graph || e->corr1 || e->corr12 ||
        p.sim != CMM_SIM_EULER_2D || !e->initialised)
        return nullptr;
    const int L = p.l_order, lc = e->loop + L - 1;
    if ((int)e->tv.size() < lc + 2) return nullptr;
    std::vector<double> t2(e->tv.begin(), e->tv.begin() + lc + 2), d2(e->dtv.begin(), e->dtv.begin() + lc + 2);
    t2.push_back(t2[lc + 1] + d2[lc + 1]);
    d2.push_back(d2[lc + 1]);
    cmm_advect_coeffs c;
    if (cmm_advect_coeffs_fill(t2.data(), d2.data(), e->loop + 1, L, p.time_integration_num, p.map_update_order_num,
                               p.map_epsilon, -1, &c))
        return nullptr;
    const int f = op_advect_fields(c);
    if (!f) return nullptr;
    const bool sp = split(e) && op_advect32_supported(c);
    if (!sp && op_advect_workspace(p.g_coarse, p.g_psi, c) == 0) return nullptr;
    const size_t field = (sp ? sizeof(float) : sizeof(double)) * 4 * (size_t)e->np;
    PsiPre &q = e->pre;
    q = PsiPre{};
    for (int l = 1; l < L; ++l) q.P.p[l] = e->lvl[l];
    q.L = L;
    q.form = sp ? 1 : 0;
    std::memcpy(q.w1, c.L1, sizeof(q.w1));
    std::memcpy(q.w12, c.L12, sizeof(q.w12));
    char *base = static_cast<char *>(e->cpre);
    q.C1 = base;
    q.C12 = (f & 2) ? base + field : nullptr;
    q.C0 = (f & 4) ? base + 2 * field : nullptr;
    e->pre_tin = c.time_integration_num;
    return &e->pre;
}

// the precombined fields for this advection's coefficients, or null
const PsiPre *take_pre(cmm_euler *e, const cmm_advect_coeffs &c, int form) {
    const bool ok = e->pre_valid && !e->lw_graph && !(e->initialised && (e->corr1 || e->corr12)) &&
                    e->pre.form == form && e->pre.L == c.l_order && e->pre_tin == c.time_integration_num &&
                    std::memcmp(e->pre.w1, c.L1, sizeof(double) * c.l_order) == 0 &&
                    std::memcmp(e->pre.w12, c.L12, sizeof(double) * c.l_order) == 0 &&
                    (op_advect_fields(c) & ~(1 | (e->pre.C12 ? 2 : 0) | (e->pre.C0 ? 4 : 0))) == 0;
    bool levels = true;
    for (int l = 1; l < c.l_order; ++l) levels = levels && e->pre.P.p[l] == e->lvl[l];
    e->pre_valid = false;
    return ok && levels ? &e->pre : nullptr;
}

bool check_fusable(const cmm_euler *e) {
    return e->p.sim == CMM_SIM_EULER_2D && !e->p.forward_map && op_check_fusable(e->p.g_coarse, e->p.g_vort, e->p.g_fine);
}

// the psi solve's sampling with the step's check in the same pass (fuse_check), and its spectral
// part into out
int engine_psi_sample(cmm_euler *e, hipEvent_t *ev) {
    if (split(e)) {
        if (int rc = sync_m32(e)) return rc;
        if (e->dry) return CMM_OK;
        return op_psi_sample32(e->p.g_coarse, e->p.g_vort, e->p.g_fine, e->p.g_psi, e->m32[e->m32_cur], e->W4f,
                               e->p.molly_stencil, e->p.freq_cut_psi, e->work, e->work_bytes, e->st, ev, e->slot);
    }
    if (e->dry) return CMM_OK;
    return op_psi_sample(e->p.g_coarse, e->p.g_vort, e->p.g_fine, e->p.g_psi, e->cx, e->cy, e->W, e->p.molly_stencil,
                         e->p.freq_cut_psi, e->work, e->work_bytes, e->st, ev, e->W4, e->slot);
}

int engine_psi_spectral(cmm_euler *e, double *out, hipEvent_t *ev, const PsiPre *pre = nullptr) {
    if (e->dry) return CMM_OK;
    if (split(e))
        return op_psi_spectral32(e->p.g_vort, e->p.g_psi, e->p.freq_cut_psi, out, e->work, e->work_bytes, e->st, ev, pre);
    return op_psi_spectral(e->p.g_vort, e->p.g_psi, e->p.freq_cut_psi, out, e->work, e->work_bytes, e->st, ev, pre);
}

int engine_advect(cmm_euler *e, const double *t, const double *dt, int loop_ctr, int L, int tin) {
    cmm_advect_coeffs c;
    if (int rc = cmm_advect_coeffs_fill(t, dt, loop_ctr, L, tin, e->p.map_update_order_num, e->p.map_epsilon, -1, &c))
        return rc;
    PsiLevels P;
    for (int l = 0; l < 4; ++l) P.p[l] = e->lvl[l < L ? l : 0];
    if (e->lw_graph) {
        std::memcpy(e->lw_host, c.L1, sizeof(c.L1));
        std::memcpy(e->lw_host + 4, c.L12, sizeof(c.L12));
    }
    const double *lw = e->lw_graph ? e->lw_dev : nullptr;
    // the combined fields the previous step's psi solve wrote for these coefficients, if any
    const PsiPre *pre_in = take_pre(e, c, split(e) && op_advect32_supported(c) ? 1 : 0);
    if (split(e) && op_advect32_supported(c)) {
        if (int rc = sync_m32(e)) return rc;
        // the Vlasov-Poisson flow: the velocity from the separable psi levels, as the FP64 form
        const bool sep = e->p.sim == CMM_SIM_VLASOV_POISSON_1D && e->vpH && !pre_in &&
                         op_vp_advect_supported(e->p.g_coarse, e->p.g_psi, c) &&
                         !(e->initialised && (e->corr1 || e->corr12));
        if (sep && !e->dry) {
            if (int rc = op_map_advect32_vp(e->p.g_coarse, e->p.g_psi, P, c, e->m32[e->m32_cur], e->m32[1 - e->m32_cur],
                                            e->vpH, e->work, e->work_bytes, e->st, e->prof ? e->ev[10] : nullptr, lw))
                return rc;
        } else if (!e->dry)
            // the Vlasov-Poisson stream function -H(v) + phi(x) has a large smooth part in v: its
            // combined fields stay FP64 records; Euler's are packed FP32 records keeping the values'
            // low parts
            if (int rc = op_map_advect32(e->p.g_coarse, e->p.g_psi, P, c, e->m32[e->m32_cur], e->m32[1 - e->m32_cur],
                                         e->work, e->work_bytes, e->st, e->prof ? e->ev[10] : nullptr,
                                         e->p.sim == CMM_SIM_VLASOV_POISSON_1D ? PsiRec64 : PsiRec32Pk,
                                         e->initialised ? e->corr1 : nullptr,
                                         e->initialised ? e->corr12 : nullptr, lw, pre_in))
                return rc;
        e->m32_cur = 1 - e->m32_cur;
        e->m32_fresh = true;
        e->planes_fresh = false;
        return CMM_OK;
    }
    // the split form runs the schemes that read psi levels on their own (and the one-level
    // warm-up steps) on the FP64 planes
    if (int rc = sync_planes(e)) return rc;
    if (e->grp && !(e->p.sim == CMM_SIM_VLASOV_POISSON_1D && e->m8[0] && e->vpH &&
                    op_vp_advect_supported(e->p.g_coarse, e->p.g_psi, c) && !(e->initialised && (e->corr1 || e->corr12)))) {
        set_error("grouped step: an advection that is not the separable Vlasov-Poisson one");
        return CMM_ERR_UNSUPPORTED;
    }
    // the Vlasov-Poisson flow: the velocities from the separable psi levels (1D tables of the
    // combined phi and of the bump function instead of 2D gathers)
    if (e->p.sim == CMM_SIM_VLASOV_POISSON_1D && e->m8[0] && e->vpH &&
        op_vp_advect_supported(e->p.g_coarse, e->p.g_psi, c) && !(e->initialised && (e->corr1 || e->corr12))) {
        const double *m8_in = e->m8_valid ? e->m8[e->m8_cur] : nullptr;
        double *m8_out = e->m8[1 - e->m8_cur];
        if (e->grp) {
            GroupInstance &g = *e->grp;
            g.gm = e->p.g_coarse;
            g.gp = e->p.g_psi;
            g.P = P;
            g.c = c;
            g.m8_in = m8_in;
            g.m8_out = m8_out;
            g.ncx = e->ncx;
            g.ncy = e->ncy;
            g.H = e->vpH;
            g.adv_work = e->work;
            g.adv_work_bytes = e->work_bytes;
            g.parts |= 1;
        } else if (!e->dry)
            if (int rc = op_map_advect_vp(e->p.g_coarse, e->p.g_psi, P, c, e->cx, e->cy, m8_in, m8_out, e->vpH, e->ncx,
                                          e->ncy, e->work, e->work_bytes, e->st, e->prof ? e->ev[10] : nullptr, lw))
                return rc;
        std::swap(e->cx, e->ncx);
        std::swap(e->cy, e->ncy);
        e->m8_valid = true;
        e->m8_cur = 1 - e->m8_cur;
        e->m32_fresh = false;
        return CMM_OK;
    }
    // the workspace is shared with the psi solve, which runs later on the same stream
    // event 10 separates the psi combination from the advection kernel in the profile
    // m8: the map node-interleaved, written by the previous advection (valid until the map is
    // reset); with a workspace the kernel reads it and writes the next one
    const bool comb = op_advect_workspace(e->p.g_coarse, e->p.g_psi, c) > 0;
    const double *m8_in = comb && e->m8_valid ? e->m8[e->m8_cur] : nullptr;
    double *m8_out = comb ? e->m8[1 - e->m8_cur] : nullptr;
    if (!e->dry)
        if (int rc = op_map_advect(e->p.g_coarse, e->p.g_psi, e->cx, e->cy, P, c, e->ncx, e->ncy, e->work,
                                   e->work_bytes, e->st, e->prof ? e->ev[10] : nullptr, m8_in, m8_out,
                                   e->initialised ? e->corr1 : nullptr, e->initialised ? e->corr12 : nullptr, lw,
                                   comb ? pre_in : nullptr))
            return rc;
    std::swap(e->cx, e->ncx);
    std::swap(e->cy, e->ncy);
    e->m8_valid = comb;
    if (comb) e->m8_cur = 1 - e->m8_cur;
    e->m32_fresh = false;
    return CMM_OK;
}

// the forward map with the same psi levels and times, Cuda-CMM/src/simulation/cmm-euler2d.cu:780-785
int engine_advect_forward(cmm_euler *e, const double *t, const double *dt, int loop_ctr, int L, int tin) {
    cmm_advect_coeffs c;
    if (int rc = cmm_advect_coeffs_fill(t, dt, loop_ctr, L, tin, e->p.map_update_order_num, e->p.map_epsilon, 1, &c))
        return rc;
    PsiLevels P;
    for (int l = 0; l < 4; ++l) P.p[l] = e->lvl[l < L ? l : 0];
    if (int rc = op_map_advect(e->p.g_coarse, e->p.g_psi, e->fx, e->fy, P, c, e->nfx, e->nfy, nullptr, 0, e->st, nullptr,
                               nullptr, nullptr))
        return rc;
    std::swap(e->fx, e->nfx);
    std::swap(e->fy, e->nfy);
    return CMM_OK;
}

// the check on whichever form of the map is current, on stream s (a conversion to the FP64
// planes, when needed, runs on the engine stream before)
int launch_incomp(cmm_euler *e, hipStream_t s) {
    if (split(e) && e->m32_fresh) {
        if (e->grp) { set_error("grouped step: a split-form check"); return CMM_ERR_UNSUPPORTED; }
        return e->dry ? CMM_OK : op_incomp32_launch(e->p.g_coarse, e->p.g_fine, e->m32[e->m32_cur], e->slot, s, e->tune);
    }
    if (int rc = sync_planes(e)) return rc;
    if (e->grp) {
        GroupInstance &g = *e->grp;
        g.gc = e->p.g_fine;
        g.ccx = e->cx;
        g.ccy = e->cy;
        g.tune = e->tune;
        g.parts |= 2;
        return CMM_OK;
    }
    return e->dry ? CMM_OK : op_incomp_launch(e->p.g_coarse, e->p.g_fine, e->cx, e->cy, e->slot, s, e->tune);
}

int engine_incomp(cmm_euler *e, double *out) {
    if (int rc = launch_incomp(e, e->st)) return rc;
    unsigned long long h = 0;
    if (int rc = hip_check(hipMemcpyAsync(&h, e->slot, sizeof(h), hipMemcpyDeviceToHost, e->st), "hipMemcpyAsync")) return rc;
    if (int rc = hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize")) return rc;
    std::memcpy(out, &h, sizeof(double));
    return CMM_OK;
}

int copy_level(cmm_euler *e, int dst, int src) {
    return hip_check(hipMemcpyAsync(e->lvl[dst], e->lvl[src], sizeof(double) * 4 * e->np, hipMemcpyDeviceToDevice, e->st),
                     "hipMemcpyAsync psi");
}

void prof_mark(cmm_euler *e, int idx) {
    if (e->prof) (void)hipEventRecord(e->ev[idx], e->st);
}

// queue the interval [ev a, ev b] of the current step for slot `phase`
void prof_acc(cmm_euler *e, int phase, int a, int b) {
    if (!e->prof) return;
    e->pend[e->cur].push_back({phase, a, b});
}

// read the queued intervals of event set s
void prof_drain(cmm_euler *e, int s) {
    for (const auto &q : e->pend[s]) {
        float m = 0;
        if (hipEventSynchronize(e->evs[s][q[2]]) == hipSuccess &&
            hipEventElapsedTime(&m, e->evs[s][q[1]], e->evs[s][q[2]]) == hipSuccess) {
            e->ms[q[0]] += m;
            e->cnt[q[0]] += 1;
        }
    }
    e->pend[s].clear();
}

// Random-spectrum field (no reference counterpart), as a column-major half spectrum:
// omega_hat = A sqrt(k E(k)/pi) exp(i theta), E(k) = k^4 exp(-2 (k/6)^2), theta = 2 pi U(seed,kx,ky),
// theta(0,-ky) = -theta(0,ky), |kx|,|ky| <= min(64, n/2-1), A such that 0.5 h^2 sum omega^2 = 1.
void random_spectrum_columns(int n, uint64_t seed, std::vector<std::complex<double>> &cols) {
    const int nh = n / 2 + 1;
    const int km = std::min(64, n / 2 - 1);
    const double pi = kPi;
    cols.assign((size_t)nh * n, std::complex<double>(0, 0));
    double s2 = 0;
    for (int kx = 0; kx <= km; ++kx)
        for (int ky = -km; ky <= km; ++ky) {
            if (kx == 0 && ky <= 0) continue;
            double k = std::sqrt((double)kx * kx + (double)ky * ky);
            double E = k * k * k * k * std::exp(-2.0 * (k / 6.0) * (k / 6.0));
            double a = std::sqrt(k * E / pi);
            uint64_t key = ((uint64_t)(kx + 32768) << 32) | (uint64_t)(ky + 32768);
            double u = (double)(mix64(mix64(seed) + key) >> 11) * (1.0 / 9007199254740992.0);
            double th = 2.0 * pi * u;
            std::complex<double> v(a * std::cos(th), a * std::sin(th));
            int row = ky < 0 ? ky + n : ky;
            cols[(size_t)kx * n + row] = v;
            if (kx == 0) cols[(size_t)(n - ky)] = std::conj(v);
            s2 += 2.0 * std::norm(v);
        }
    const double A = 1.0 / std::sqrt(2.0 * pi * pi * s2);
    for (auto &v : cols) v *= A;
}

int engine_setup_discrete(cmm_euler *e) {
    const cmm_grid &gd = e->p.g_disc;
    e->nd = (long)gd.nx * gd.ny;
    if (int rc = dalloc(&e->disc, 4 * e->nd)) return rc;
    double *real = nullptr;
    if (int rc = dalloc(&real, e->nd)) return rc;
    int rc = CMM_OK;
    if (e->p.ic == CMM_IC_RANDOM_SPECTRUM) {
        std::vector<std::complex<double>> cols;
        random_spectrum_columns(gd.nx, e->p.seed, cols);
        double *dspec = nullptr;
        rc = dalloc(&dspec, 2 * cols.size());
        if (!rc) rc = hip_check(hipMemcpyAsync(dspec, cols.data(), sizeof(double) * 2 * cols.size(), hipMemcpyHostToDevice, e->st), "upload spectrum");
        // real field = C2R of the spectrum (plane 0 of its Hermite form)
        if (!rc) rc = op_hermite_from_spectrum(gd, dspec, e->disc, e->work, e->work_bytes, e->st);
        if (!rc) rc = hip_check(hipMemcpyAsync(real, e->disc, sizeof(double) * e->nd, hipMemcpyDeviceToDevice, e->st), "copy field");
        if (!rc) rc = hip_check(hipStreamSynchronize(e->st), "sync");
        if (dspec) (void)hipFree(dspec);
    } else {
        if (!e->p.disc_real) { set_error("discrete initial condition without data"); rc = CMM_ERR_ARG; }
        if (!rc) rc = hip_check(hipMemcpyAsync(real, e->p.disc_real, sizeof(double) * e->nd, hipMemcpyHostToDevice, e->st), "upload field");
    }
    // Hermite form of the discrete field, Cuda-CMM/src/simulation/cmm-euler2d.cu:325-344
    if (!rc) rc = op_hermite_from_real(gd, real, e->disc, e->work, e->work_bytes, e->st);
    if (!rc) rc = hip_check(hipStreamSynchronize(e->st), "sync");
    (void)hipFree(real);
    return rc;
}

bool pow2(int n) { return n >= 8 && n <= 4096 && (n & (n - 1)) == 0; }

}  // namespace

extern "C" {

int cmm_euler_create(const cmm_euler_params *p, cmm_euler **out) {
    if (!p || !out) { set_error("cmm_euler_create: null argument"); return CMM_ERR_ARG; }
    *out = nullptr;
    if (!grid_ok(p->g_coarse) || !grid_ok(p->g_fine) || !grid_ok(p->g_psi) || !grid_ok(p->g_vort) ||
        !pow2(p->g_fine.nx) || !pow2(p->g_fine.ny) || !pow2(p->g_psi.nx) || !pow2(p->g_vort.nx) ||
        p->g_psi.nx != p->g_psi.ny || p->g_vort.nx != p->g_vort.ny) {
        set_error("cmm_euler_create: grids must be square powers of two between 8 and 4096 (fine, psi, vort)");
        return CMM_ERR_ARG;
    }
    if (p->sim != CMM_SIM_EULER_2D && p->sim != CMM_SIM_VLASOV_POISSON_1D) {
        set_error("cmm_euler_create: unknown simulation type");
        return CMM_ERR_ARG;
    }
    if (p->sim == CMM_SIM_VLASOV_POISSON_1D && (p->g_vort.nx > p->g_psi.nx || p->g_vort.ny > p->g_psi.ny)) {
        set_error("cmm_euler_create: vorticity grid should be smaller than phi_out grid (Vlasov-Poisson needs grid_vort <= grid_psi)");
        return CMM_ERR_ARG;
    }
    if (p->l_order < 1 || p->l_order > 4 || !(p->dt > 0) || p->stack_capacity < 0 || p->map_update_order_num < 0 ||
        p->map_update_order_num > 2) {
        set_error("cmm_euler_create: bad time-stepping parameters");
        return CMM_ERR_ARG;
    }
    if (p->precision != CMM_PRECISION_FP64 && p->precision != CMM_PRECISION_SPLIT) {
        set_error("cmm_euler_create: precision must be CMM_PRECISION_FP64 or CMM_PRECISION_SPLIT");
        return CMM_ERR_ARG;
    }
    if (p->precision == CMM_PRECISION_SPLIT && p->forward_map) {
        set_error("cmm_euler_create: the split-precision form runs the backward map only");
        return CMM_ERR_ARG;
    }
    if (p->forward_map != 0 && p->forward_map != 1) {
        set_error("cmm_euler_create: forward_map must be 0 or 1");
        return CMM_ERR_ARG;
    }
    if (p->discrete && (!grid_ok(p->g_disc) || !pow2(p->g_disc.nx) || p->g_disc.nx != p->g_disc.ny)) {
        set_error("cmm_euler_create: discrete grid must be a square power of two");
        return CMM_ERR_ARG;
    }
    cmm_euler *e = new cmm_euler();
    e->p = *p;
    e->nc = (long)p->g_coarse.nx * p->g_coarse.ny;
    e->nf = (long)p->g_fine.nx * p->g_fine.ny;
    e->np = (long)p->g_psi.nx * p->g_psi.ny;
    int rc = hip_check(hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking), "hipStreamCreate");
    if (!rc) rc = dalloc(&e->cx, 4 * e->nc);
    if (!rc) rc = dalloc(&e->cy, 4 * e->nc);
    if (!rc) rc = dalloc(&e->ncx, 4 * e->nc);
    if (!rc) rc = dalloc(&e->ncy, 4 * e->nc);
    for (int l = 0; l < p->l_order && !rc; ++l) {
        rc = dalloc(&e->psi_buf[l], 4 * e->np);
        e->lvl[l] = e->psi_buf[l];
    }
    if (!rc) rc = dalloc(&e->W, 4 * e->nf);
    const bool sp = p->precision == CMM_PRECISION_SPLIT;
    if (!rc && p->l_order >= 2 && !sp) rc = dalloc(&e->m8[0], 8 * e->nc);
    if (!rc && p->l_order >= 2 && !sp) rc = dalloc(&e->m8[1], 8 * e->nc);
    // the fine field node-interleaved for the column samplers (the Vlasov-Poisson flow samples FP64
    // records in both forms)
    if (!rc && (!sp || p->sim == CMM_SIM_VLASOV_POISSON_1D)) rc = dalloc(&e->W4, 4 * e->nf);
    if (!rc && p->sim == CMM_SIM_VLASOV_POISSON_1D) {
        rc = dalloc(&e->vpH, 3 * (size_t)p->g_psi.nx);
        if (!rc) rc = op_vp_bump(p->g_vort, p->g_psi, e->vpH, e->st);
    }
    if (!rc && sp) rc = falloc(&e->m32[0], 12 * e->nc);
    if (!rc && sp) rc = falloc(&e->m32[1], 12 * e->nc);
    if (!rc && sp && p->sim == CMM_SIM_EULER_2D) rc = falloc(&e->W4f, 4 * e->nf);
    // the next step's combined psi fields written by the psi solve (three fields of FP64 records,
    // or of packed FP32 records in the split form)
    if (!rc && p->sim == CMM_SIM_EULER_2D)
        rc = hip_check(hipMalloc(&e->cpre, 3 * (sp ? sizeof(float) : sizeof(double)) * 4 * (size_t)e->np),
                       "hipMalloc (combined psi fields)");
    if (!rc && p->stack_capacity > 0) rc = dalloc(&e->stack, (size_t)8 * e->nc * p->stack_capacity);
    if (!rc && sp && p->stack_capacity > 0) rc = dalloc(&e->stack_s, (size_t)6 * e->nc * p->stack_capacity);
    if (p->forward_map) {
        if (!rc) rc = dalloc(&e->fx, 4 * e->nc);
        if (!rc) rc = dalloc(&e->fy, 4 * e->nc);
        if (!rc) rc = dalloc(&e->nfx, 4 * e->nc);
        if (!rc) rc = dalloc(&e->nfy, 4 * e->nc);
        if (!rc && p->stack_capacity > 0) rc = dalloc(&e->fstack, (size_t)8 * e->nc * p->stack_capacity);
    }
    if (!rc) {
        size_t w = std::max(op_psi_workspace(p->g_vort, p->g_psi, p->freq_cut_psi), op_translate_workspace(p->g_fine));
        if (p->discrete) w = std::max(w, op_hermite_workspace(p->g_disc));
        // Lagrange-weighted psi combinations of the map advection (up to three Hermite fields on
        // g_psi) and the interleaved map
        if (p->l_order >= 2) w = std::max(w, 12 * sizeof(double) * (size_t)e->np + 8 * sizeof(double) * (size_t)e->nc);
        if (p->sim == CMM_SIM_VLASOV_POISSON_1D) w = std::max(w, op_vp_psi_workspace(p->g_vort, p->g_psi));
        e->work_bytes = w;
        if (hipMalloc(&e->work, w) != hipSuccess) { set_error("hipMalloc workspace"); rc = CMM_ERR_ALLOC; }
    }
    if (!rc && hipMalloc(&e->slot, 256) != hipSuccess) { set_error("hipMalloc slot"); rc = CMM_ERR_ALLOC; }
    if (!rc && hipHostMalloc((void **)&e->h_slot, 256, hipHostMallocDefault) != hipSuccess) {
        e->h_slot = nullptr;
        set_error("hipHostMalloc slot");
        rc = CMM_ERR_ALLOC;
    }
    if (!rc) {  // bytes 64..127 of the slots: the Lagrange weights of graph steps
        e->lw_dev = reinterpret_cast<double *>(e->slot + 8);
        e->lw_host = reinterpret_cast<double *>(e->h_slot + 8);
    }
    if (!rc) rc = hip_check(hipEventCreateWithFlags(&e->ev_inc, hipEventDisableTiming), "hipEventCreate");
    if (!rc) rc = hip_check(hipEventCreateWithFlags(&e->ev_psi, kDeviceEvent), "hipEventCreate");
    if (!rc) rc = hip_check(hipEventCreateWithFlags(&e->ev_part, kDeviceEvent), "hipEventCreate");
    if (!rc) rc = hip_check(hipStreamCreateWithFlags(&e->st2, hipStreamNonBlocking), "hipStreamCreate");
    if (!rc) rc = hip_check(hipStreamCreateWithFlags(&e->st3, hipStreamNonBlocking), "hipStreamCreate");
    if (!rc) rc = hip_check(hipEventCreateWithFlags(&e->ev_fork, kDeviceEvent), "hipEventCreate");
    for (int s = 0; s < cmm_euler::kProfSets && !rc; ++s)
        for (int i = 0; i < 12 && !rc; ++i) rc = hip_check(hipEventCreate(&e->evs[s][i]), "hipEventCreate");
    if (!rc && p->discrete) rc = engine_setup_discrete(e);
    if (rc) {
        cmm_euler_destroy(e);
        return rc;
    }
    // the split form's sampler keeps fewer registers without the check, and its check overlaps
    // the sampler from the third stream: measured faster unfused
    e->fuse_check = check_fusable(e) && !split(e);
    *out = e;
    return CMM_OK;
}

int cmm_euler_set_check_fusion(cmm_euler *e, int on) {
    if (!e) { set_error("cmm_euler_set_check_fusion: null engine"); return CMM_ERR_ARG; }
    if (on && !check_fusable(e)) {
        set_error("cmm_euler_set_check_fusion: the check grid must be the vorticity grid, refining the map grid by a "
                  "multiple of 4 over the same domain, in an Euler run without a forward map");
        return CMM_ERR_ARG;
    }
    e->fuse_check = on != 0;
    return CMM_OK;
}

int cmm_euler_set_psi_precombine(cmm_euler *e, int on) {
    if (!e) { set_error("cmm_euler_set_psi_precombine: null engine"); return CMM_ERR_ARG; }
    e->pre_enabled = on != 0;
    e->pre_valid = false;
    return CMM_OK;
}

int cmm_euler_get_psi_precombine(cmm_euler *e, int *on) {
    if (!e || !on) { set_error("cmm_euler_get_psi_precombine: bad argument"); return CMM_ERR_ARG; }
    *on = e->pre_enabled && e->cpre ? 1 : 0;
    return CMM_OK;
}

int cmm_euler_get_check_fusion(cmm_euler *e, int *on) {
    if (!e || !on) { set_error("cmm_euler_get_check_fusion: bad argument"); return CMM_ERR_ARG; }
    *on = e->fuse_check ? 1 : 0;
    return CMM_OK;
}

int cmm_euler_destroy(cmm_euler *e) {
    if (!e) return CMM_OK;
    if (e->st) (void)hipStreamSynchronize(e->st);
    double *bufs[] = {e->cx, e->cy, e->ncx, e->ncy, e->W, e->W4, e->vpH, e->m8[0], e->m8[1], e->sbuf, e->disc, e->stack, e->tmp, e->tmp4,
                      e->psi_buf[0], e->psi_buf[1], e->psi_buf[2], e->psi_buf[3], e->fx, e->fy, e->nfx, e->nfy, e->fstack, e->stack_s};
    for (double *b : bufs)
        if (b) (void)hipFree(b);
    if (e->cpre) (void)hipFree(e->cpre);
    for (double *b : e->part)
        if (b) (void)hipFree(b);
    for (double *b : e->part_vel)
        if (b) (void)hipFree(b);
    for (float *b : {e->m32[0], e->m32[1], e->W4f})
        if (b) (void)hipFree(b);
    if (e->work) (void)hipFree(e->work);
    if (e->slot) (void)hipFree(e->slot);
    if (e->st2) (void)hipStreamSynchronize(e->st2);
    if (e->h_slot) (void)hipHostFree(e->h_slot);
    if (e->ev_inc) (void)hipEventDestroy(e->ev_inc);
    if (e->ev_psi) (void)hipEventDestroy(e->ev_psi);
    if (e->ev_part) (void)hipEventDestroy(e->ev_part);
    if (e->st2) (void)hipStreamDestroy(e->st2);
    if (e->st3) {
        (void)hipStreamSynchronize(e->st3);
        (void)hipStreamDestroy(e->st3);
    }
    if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
    for (int s = 0; s < cmm_euler::kProfSets; ++s)
        for (int i = 0; i < 12; ++i)
            if (e->evs[s][i]) (void)hipEventDestroy(e->evs[s][i]);
    if (e->st) (void)hipStreamDestroy(e->st);
    delete e;
    return CMM_OK;
}

void *cmm_euler_stream(cmm_euler *e) { return e ? (void *)e->st : nullptr; }

int cmm_euler_add_particle_set(cmm_euler *e, int n, int scheme, double init_time, double tau, int init_vel,
                               const double *pos, const double *vel, int *set) {
    if (!e || n < 0 || (n > 0 && !pos) || !std::isfinite(init_time) || !std::isfinite(tau)) {
        set_error("cmm_euler_add_particle_set: bad argument");
        return CMM_ERR_ARG;
    }
    if (scheme != 10 && scheme != 20 && scheme != 30 && scheme != 31 && scheme != 40 && scheme != 41) {
        set_error("cmm_euler_add_particle_set: unknown scheme");
        return CMM_ERR_ARG;
    }
    double *d = nullptr, *v = nullptr;
    if (n > 0) {
        if (int rc = dalloc(&d, 2 * (size_t)n)) return rc;
        if (int rc = hip_check(hipMemcpy(d, pos, sizeof(double) * 2 * n, hipMemcpyHostToDevice), "upload particles")) {
            (void)hipFree(d);
            return rc;
        }
        if (tau != 0) {
            if (int rc = dalloc(&v, 2 * (size_t)n)) { (void)hipFree(d); return rc; }
            int rc = vel ? hip_check(hipMemcpy(v, vel, sizeof(double) * 2 * n, hipMemcpyHostToDevice), "upload velocities")
                         : hip_check(hipMemset(v, 0, sizeof(double) * 2 * n), "clear velocities");
            // a set added after the warm-up starts at the current psi, as cmm_euler_init would
            if (!rc && !vel && e->initialised && init_vel)
                rc = op_particles_vel_init(e->p.g_psi, e->lvl[0], n, 1, d, v, e->st);
            if (!rc) rc = hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
            if (rc) { (void)hipFree(d); (void)hipFree(v); return rc; }
        }
    }
    e->part.push_back(d);
    e->part_n.push_back(n);
    e->part_scheme.push_back(scheme);
    e->part_vel.push_back(v);
    e->part_t0.push_back(init_time);
    e->part_tau.push_back(tau);
    e->part_init_vel.push_back(init_vel ? 1 : 0);
    e->part_vel_given.push_back(vel ? 1 : 0);
    if (set) *set = (int)e->part.size() - 1;
    return CMM_OK;
}

int cmm_euler_add_particles(cmm_euler *e, int n, int scheme, const double *pos, int *set) {
    return cmm_euler_add_particle_set(e, n, scheme, 0.0, 0.0, 0, pos, nullptr, set);
}

int cmm_euler_init(cmm_euler *e, double *warm, int *nwarm) {
    if (!e) { set_error("cmm_euler_init: null engine"); return CMM_ERR_ARG; }
    const cmm_euler_params &p = e->p;
    const int L = p.l_order;
    int nw = 0;
    e->pre_valid = false;
    // Cuda-CMM/src/simulation/cmm-euler2d.cu:495-512
    if (int rc = op_map_identity(p.g_coarse, e->cx, e->cy, e->st)) return rc;
    if (p.forward_map)
        if (int rc = op_map_identity(p.g_coarse, e->fx, e->fy, e->st)) return rc;
    planes_written(e);
    e->nstack = 0;
    e->force_remap = false;
    if (int rc = engine_translate(e)) return rc;
    if (int rc = engine_psi(e, e->lvl[0])) return rc;
    // Lagrange warm-up, Cuda-CMM/src/simulation/cmm-euler2d.cu:520-619
    if (L > 1) {
        const double dtn = p.dt * std::pow(-1.0, L);
        double t_init[5] = {0, dtn, 2 * dtn, 3 * dtn, 4 * dtn};
        double dt_init[5] = {dtn, dtn, dtn, dtn, dtn};
        for (int order = 1; order <= L; ++order) {
            const int nsteps = order - (order == L);
            if (!p.lagrange_init_higher_order && order != L) {
                for (int i = 0; i < 5; ++i) { dt_init[i] = -dt_init[i]; t_init[i] = -t_init[i]; }
                continue;
            }
            int tin = 10, lo = 1;
            if (p.lagrange_init_higher_order) {
                lo = order;
                tin = order == 1 ? 10 : order == 2 ? 21 : order == 3 ? 30 : 40;
            }
            for (int s = 1; s <= nsteps; ++s) {
                if (int rc = engine_advect(e, t_init, dt_init, 0, lo, tin)) return rc;
                double einc;
                if (int rc = engine_incomp(e, &einc)) return rc;
                if (nw < 8 && warm) warm[nw] = einc;
                nw++;
                for (int il = 1; il <= nsteps; ++il)
                    if (int rc = copy_level(e, nsteps - il + 1, nsteps - il)) return rc;
                if (int rc = engine_psi(e, e->lvl[0])) return rc;
            }
            for (int i = 0; i < 5; ++i) { dt_init[i] = -dt_init[i]; t_init[i] = -t_init[i]; }
            std::swap(e->lvl[0], e->lvl[nsteps]);
            if (nsteps >= 3) std::swap(e->lvl[1], e->lvl[2]);
            if (int rc = op_map_identity(p.g_coarse, e->cx, e->cy, e->st)) return rc;
            planes_written(e);
        }
    }
    if (nwarm) *nwarm = std::min(nw, 8);
    // start velocities of inertial particles from psi at t = 0, Cuda-CMM/src/simulation/cmm-euler2d.cu:620-632
    for (size_t s = 0; s < e->part.size(); ++s)
        if (e->part_vel[s] && !e->part_vel_given[s])
            if (int rc = op_particles_vel_init(p.g_psi, e->lvl[0], e->part_n[s], e->part_init_vel[s], e->part[s],
                                               e->part_vel[s], e->st))
                return rc;
    // time vectors, Cuda-CMM/src/simulation/cmm-euler2d.cu:709-715
    e->tv.assign(L, 0.0);
    e->dtv.assign(L, p.dt);
    for (int i = 0; i < L; ++i) e->tv[i] = -(L - 1 - i) * p.dt;
    e->loop = 0;
    e->initialised = true;
    return hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
}

int cmm_euler_restart(cmm_euler *e, double t_restart, int n_maps, const double *stack_host, const double *chix_host,
                      const double *chiy_host, const double *psi_old_host) {
    if (!e || !chix_host || !chiy_host || n_maps < 0 || (n_maps > 0 && !stack_host) ||
        (e->p.l_order > 1 && !psi_old_host)) {
        set_error("cmm_euler_restart: bad argument");
        return CMM_ERR_ARG;
    }
    const cmm_euler_params &p = e->p;
    const int L = p.l_order;
    e->pre_valid = false;
    if (n_maps > p.stack_capacity) {
        set_error("cmm_euler_restart: more maps than the stack holds");
        return CMM_ERR_ARG;
    }
    auto put = [&](double *dst, const double *src, size_t cnt) {
        return hip_check(hipMemcpyAsync(dst, src, sizeof(double) * cnt, hipMemcpyHostToDevice, e->st), "put");
    };
    // readMapStack (Cuda-CMM/src/simulation/cmm-simulation-host.cu:1255-1302): the stacked maps, the
    // initial condition translated through them with the active map still the identity, then the
    // active map and the older psi levels
    // each map goes through a device buffer into the stack's node records (and the split records)
    int rc = CMM_OK;
    if (n_maps > 0) {
        double *tmp = nullptr;
        rc = dalloc(&tmp, 8 * (size_t)e->nc);
        for (int m = 0; m < n_maps && !rc; ++m) {
            rc = put(tmp, stack_host + (size_t)8 * e->nc * m, (size_t)8 * e->nc);
            if (!rc) rc = op_pack_maps(p.g_coarse, tmp, tmp + 4 * e->nc, 0, 1, kStackRec64, e->stack + (size_t)8 * e->nc * m, e->st);
            if (!rc && e->stack_s)
                rc = op_pack_maps(p.g_coarse, tmp, tmp + 4 * e->nc, 0, 1, kStackRec48, e->stack_s + (size_t)6 * e->nc * m, e->st);
            // the buffer is reused for the next map after these kernels read it
            if (!rc) rc = hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
        }
        if (tmp) (void)hipFree(tmp);
    }
    e->nstack = n_maps;
    if (!rc) rc = op_map_identity(p.g_coarse, e->cx, e->cy, e->st);
    if (!rc && p.forward_map) rc = op_map_identity(p.g_coarse, e->fx, e->fy, e->st);
    planes_written(e);
    e->force_remap = false;
    if (!rc) rc = engine_translate(e);
    if (!rc) rc = put(e->cx, chix_host, (size_t)4 * e->nc);
    if (!rc) rc = put(e->cy, chiy_host, (size_t)4 * e->nc);
    planes_written(e);
    for (int l = 1; l < L && !rc; ++l) rc = put(e->lvl[l], psi_old_host + (size_t)4 * e->np * (l - 1), (size_t)4 * e->np);
    // the newest psi from the loaded state (Cuda-CMM/src/simulation/cmm-euler2d.cu:664-666); no warm-up
    if (!rc) rc = engine_psi(e, e->lvl[0]);
    if (rc) return rc;
    // time vectors from the restart time, Cuda-CMM/src/simulation/cmm-euler2d.cu:709-715
    e->tv.assign(L, 0.0);
    e->dtv.assign(L, p.dt);
    for (int i = 0; i < L; ++i) e->tv[i] = t_restart - (L - 1 - i) * p.dt;
    e->loop = 0;
    e->initialised = true;
    return hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
}

int cmm_euler_get_stack_map(cmm_euler *e, int m, double *chix_host, double *chiy_host) {
    if (!e || m < 0 || m >= e->nstack || !chix_host || !chiy_host) {
        set_error("cmm_euler_get_stack_map: bad argument");
        return CMM_ERR_ARG;
    }
    // the stack holds node records (8 doubles per node): copied out, then split into the planes
    const double *slot = e->stack + (size_t)8 * e->nc * m;
    std::vector<double> rec(8 * (size_t)e->nc);
    int rc = hip_check(hipMemcpyAsync(rec.data(), slot, sizeof(double) * 8 * e->nc, hipMemcpyDeviceToHost, e->st), "get");
    if (!rc) rc = hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
    if (rc) return rc;
    for (long i = 0; i < e->nc; ++i)
        for (int q = 0; q < 4; ++q) {
            chix_host[q * e->nc + i] = rec[8 * i + q];
            chiy_host[q * e->nc + i] = rec[8 * i + 4 + q];
        }
    return CMM_OK;
}

int cmm_euler_restart_forward(cmm_euler *e, int n_maps, const double *stack_host, const double *chix_host,
                              const double *chiy_host) {
    if (!e || !e->fx || !chix_host || !chiy_host || n_maps != e->nstack || (n_maps > 0 && !stack_host)) {
        set_error("cmm_euler_restart_forward: bad argument (forward map engine, restarted, with as many forward maps as backward maps)");
        return CMM_ERR_ARG;
    }
    auto put = [&](double *dst, const double *src, size_t cnt) {
        return hip_check(hipMemcpyAsync(dst, src, sizeof(double) * cnt, hipMemcpyHostToDevice, e->st), "put");
    };
    int rc = n_maps > 0 ? put(e->fstack, stack_host, (size_t)8 * e->nc * n_maps) : CMM_OK;
    if (!rc) rc = put(e->fx, chix_host, (size_t)4 * e->nc);
    if (!rc) rc = put(e->fy, chiy_host, (size_t)4 * e->nc);
    if (rc) return rc;
    return hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
}

int cmm_euler_get_forward_stack_map(cmm_euler *e, int m, double *chix_host, double *chiy_host) {
    if (!e || !e->fstack || m < 0 || m >= e->nstack || !chix_host || !chiy_host) {
        set_error("cmm_euler_get_forward_stack_map: bad argument");
        return CMM_ERR_ARG;
    }
    const double *slot = e->fstack + (size_t)8 * e->nc * m;
    int rc = hip_check(hipMemcpyAsync(chix_host, slot, sizeof(double) * 4 * e->nc, hipMemcpyDeviceToHost, e->st), "get");
    if (!rc)
        rc = hip_check(hipMemcpyAsync(chiy_host, slot + 4 * e->nc, sizeof(double) * 4 * e->nc, hipMemcpyDeviceToHost, e->st),
                       "get");
    if (rc) return rc;
    return hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
}

int cmm_euler_request_remap(cmm_euler *e) {
    if (!e) { set_error("cmm_euler_request_remap: null engine"); return CMM_ERR_ARG; }
    e->force_remap = true;
    return CMM_OK;
}

namespace {

// The device work of a step up to the psi solve, queued without a host wait: advection, check,
// E_inc copied to pinned memory, psi history rotated and the new psi.  With e->dry only the host
// state advances (the work comes from a captured graph replay).  overlap: the check and the E_inc
// copy run on the third stream beside the psi solve (single-engine steps only).
// The step's first part: time vectors, advection, the check and the copy of E_inc to pinned memory.
// *forked: the check runs on the third stream (overlap, single-engine steps; the caller waits for
// ev_inc before it queues more work for the engine).
int step_queue_a(cmm_euler *e, double dt_now, bool overlap, bool *forked) {
    const cmm_euler_params &p = e->p;
    const int L = p.l_order;
    const int lc = e->loop + L - 1;
    e->tv.resize(lc + 2);
    e->dtv.resize(lc + 2);
    e->tv[lc + 1] = e->tv[lc] + dt_now;
    e->dtv[lc + 1] = dt_now;
    e->prof = e->prof_every > 0 && e->loop % e->prof_every == 0;
    if (e->prof) {
        const int s = e->loop % cmm_euler::kProfSets;
        prof_drain(e, s);
        e->cur = s;
        e->ev = e->evs[s];
    }
    // advection, Cuda-CMM/src/simulation/cmm-euler2d.cu:774-778
    prof_mark(e, 0);
    if (int rc = engine_advect(e, e->tv.data(), e->dtv.data(), e->loop, L, p.time_integration_num)) return rc;
    if (p.forward_map)
        if (int rc = engine_advect_forward(e, e->tv.data(), e->dtv.data(), e->loop, L, p.time_integration_num)) return rc;
    // incompressibility on the fine grid, :796: the check and the copy of E_inc to pinned host
    // memory are queued, and the host waits for that copy only after the psi solve is queued too.
    // With overlap they run on the third stream beside the psi solve: both only read the new map,
    // which the next advection does not overwrite (it writes the other buffer, after the host has
    // waited for E_inc); a conversion of the map that the check needs runs first, on the engine
    // stream.
    // With fuse_check the check runs in the pass of the psi solve's sampler over the same grid,
    // which is queued here; the step's second part queues the spectral rest of the psi solve.
    const bool fuse = e->fuse_check;
    const bool fork = overlap && e->st3 && !e->dry && !e->capturing && !fuse;
    if (fork && !(split(e) && e->m32_fresh))
        if (int rc = sync_planes(e)) return rc;
    const hipStream_t sc = fork ? e->st3 : e->st;
    if (fork) {
        if (int rc = hip_check(hipEventRecord(e->ev_fork, e->st), "hipEventRecord")) return rc;
        if (int rc = hip_check(hipStreamWaitEvent(e->st3, e->ev_fork, 0), "hipStreamWaitEvent")) return rc;
    }
    if (e->prof) (void)hipEventRecord(e->ev[1], sc);
    if (fuse) {
        // the check has no time of its own: its slot closes at once, the sampling counts to the
        // psi row slot as without the fusion
        if (e->prof) (void)hipEventRecord(e->ev[2], sc);
        if (int rc = engine_psi_sample(e, e->prof ? &e->ev[5] : nullptr)) return rc;
        e->sampled = true;
    } else if (int rc = launch_incomp(e, sc)) {
        return rc;
    }
    // forward map: invertibility and E_inc on the coarse grid, :798-803
    if (p.forward_map) {
        if (int rc = op_invert_launch(p.g_coarse, p.g_coarse, p.g_coarse, e->cx, e->cy, e->fx, e->fy, e->slot + 2, sc))
            return rc;
        if (int rc = op_incomp_launch(p.g_coarse, p.g_coarse, e->fx, e->fy, e->slot + 1, sc)) return rc;
    }
    if (!e->dry && !e->grp)
        if (int rc = hip_check(hipMemcpyAsync(e->h_slot, e->slot, sizeof(unsigned long long) * (p.forward_map ? 3 : 1),
                                              hipMemcpyDeviceToHost, sc),
                               "hipMemcpyAsync"))
            return rc;
    if (!e->dry && !e->capturing && !e->grp)
        if (int rc = hip_check(hipEventRecord(e->ev_inc, sc), "hipEventRecord")) return rc;
    if (e->prof && !fuse) (void)hipEventRecord(e->ev[2], sc);
    *forked = fork;
    return CMM_OK;
}

// The step's second part: the psi history rotated and the new psi (forked: part a put the check
// on the third stream).
int step_queue_b(cmm_euler *e, bool fork) {
    const int L = e->p.l_order;
    // psi history: the oldest level becomes the newest (same contents as the reference's
    // copies at :865-868), then the new psi, :871.  It is queued before E_inc is known: without a
    // remap (the common case) it is the step's psi; a remap below recomputes it afterwards in
    // stream order, from the reset map and the new fine vorticity, as the reference orders it.
    double *oldest = e->lvl[L - 1];
    for (int l = L - 1; l > 0; --l) e->lvl[l] = e->lvl[l - 1];
    e->lvl[0] = oldest;
    if (e->part_pending && !e->grp)  // the previous step's particles still read the level about to be rewritten
        if (int rc = hip_check(hipStreamWaitEvent(e->st, e->ev_part, 0), "hipStreamWaitEvent")) return rc;
    const PsiPre *pre = prepare_pre(e);
    if (e->sampled) {
        e->sampled = false;
        if (int rc = engine_psi_spectral(e, e->lvl[0], e->prof ? &e->ev[5] : nullptr, pre)) return rc;
    } else if (int rc = engine_psi(e, e->lvl[0], e->prof ? &e->ev[5] : nullptr, pre)) {
        return rc;
    }
    e->pre_valid = pre != nullptr;
    // no rejoin of the check's stream: the host waits for ev_inc, recorded there after the check,
    // before it queues anything else for this engine (step_finish), so all later work on the
    // engine stream follows the check; a stream wait would only add a barrier packet (about 12 us
    // idle) in front of the next step
    (void)fork;
    return CMM_OK;
}

int step_queue(cmm_euler *e, double dt_now, bool overlap = false) {
    bool fork = false;
    if (int rc = step_queue_a(e, dt_now, overlap, &fork)) return rc;
    return step_queue_b(e, fork);
}

// The rest of a step: E_inc (waited for unless the caller has synchronised), remap, particles,
// bookkeeping.
int step_finish(cmm_euler *e, cmm_step_info *info, bool waited) {
    const cmm_euler_params &p = e->p;
    const int L = p.l_order;
    const int lc = e->loop + L - 1;
    if (!waited)
        if (int rc = hip_check(hipEventSynchronize(e->ev_inc), "hipEventSynchronize")) return rc;
    double einc, einc_f = 0, einv = 0;
    std::memcpy(&einc, e->h_slot, sizeof(double));
    if (p.forward_map) {
        std::memcpy(&einc_f, e->h_slot + 1, sizeof(double));
        std::memcpy(&einv, e->h_slot + 2, sizeof(double));
    }
    int remapped = 0, saturated = 0;
    const bool forced = e->force_remap;
    e->force_remap = false;
    // remap, :814-853 (translate before push; the map is reset afterwards)
    if ((einc > p.incomp_threshold && !p.skip_remapping) || forced) {
        if (e->nstack >= p.stack_capacity) {
            saturated = 1;
        } else {
            prof_mark(e, 3);
            if (int rc = engine_translate(e)) return rc;
            // the pushed map as node records (and split records for the split form's compositions)
            if (int rc = op_pack_maps(p.g_coarse, e->cx, e->cy, 0, 1, kStackRec64, e->stack + (size_t)8 * e->nc * e->nstack, e->st))
                return rc;
            if (e->stack_s)
                if (int rc = op_pack_maps(p.g_coarse, e->cx, e->cy, 0, 1, kStackRec48, e->stack_s + (size_t)6 * e->nc * e->nstack, e->st))
                    return rc;
            if (p.forward_map) {  // :847-851
                double *fslot = e->fstack + (size_t)8 * e->nc * e->nstack;
                if (int rc = hip_check(hipMemcpyAsync(fslot, e->fx, sizeof(double) * 4 * e->nc, hipMemcpyDeviceToDevice, e->st), "push forward map")) return rc;
                if (int rc = hip_check(hipMemcpyAsync(fslot + 4 * e->nc, e->fy, sizeof(double) * 4 * e->nc, hipMemcpyDeviceToDevice, e->st), "push forward map")) return rc;
                if (int rc = op_map_identity(p.g_coarse, e->fx, e->fy, e->st)) return rc;
            }
            e->nstack++;
            if (int rc = op_map_identity(p.g_coarse, e->cx, e->cy, e->st)) return rc;
            planes_written(e);
            // the new psi also rewrites the next step's combined fields
            const PsiPre *pre = prepare_pre(e);
            if (int rc = engine_psi(e, e->lvl[0], nullptr, pre)) return rc;
            e->pre_valid = pre != nullptr;
            prof_mark(e, 4);
            remapped = 1;
        }
    }
    // particles, :881 with Cuda-CMM/src/numerical/cmm-particles.cu:63-87, on the second stream: a set
    // moves once t_{n+1} + 1e-5 dt reaches its start time; an inertial set's velocity is set again
    // at the step that reaches it (:883-895)
    bool moved = false;
    if (!e->part.empty()) {
        double Lp[4] = {0, 0, 0, 0};
        for (int i = 0; i < L; ++i) Lp[i] = cmm_lagrange_coeff(e->tv.data(), e->tv[lc] + e->dtv[lc + 1] / 2.0, lc + 1, i, L);
        PsiLevels P;
        for (int l = 0; l < 4; ++l) P.p[l] = e->lvl[l < L ? l : 0];
        const double t1 = e->tv[lc + 1], dt1 = e->dtv[lc + 1];
        for (size_t s = 0; s < e->part.size(); ++s) {
            if (e->part_n[s] <= 0) continue;
            if (L < 2 && e->part_scheme[s] != 10 && e->part_scheme[s] != 20) continue;
            const bool move = t1 + 1e-5 * dt1 >= e->part_t0[s];
            const bool start = e->part_vel[s] && e->tv[lc] - e->part_t0[s] + p.dt * 1e-5 < 0 &&
                               t1 - e->part_t0[s] + p.dt * 1e-5 > 0;
            if (!move && !start) continue;
            if (!moved) {
                if (int rc = hip_check(hipEventRecord(e->ev_psi, e->st), "hipEventRecord")) return rc;
                if (int rc = hip_check(hipStreamWaitEvent(e->st2, e->ev_psi, 0), "hipStreamWaitEvent")) return rc;
                moved = true;
            }
            if (move && !e->part_vel[s])
                if (int rc = op_particles(p.g_psi, P, L, Lp, dt1, e->part_scheme[s], e->part_n[s], e->part[s], e->st2))
                    return rc;
            if (move && e->part_vel[s])
                if (int rc = op_particles_inertial(p.g_psi, P, L, Lp, dt1, e->part_scheme[s], e->part_tau[s],
                                                   e->part_n[s], e->part[s], e->part_vel[s], e->st2))
                    return rc;
            if (start)
                if (int rc = op_particles_vel_init(p.g_psi, e->lvl[0], e->part_n[s], e->part_init_vel[s], e->part[s],
                                                   e->part_vel[s], e->st2))
                    return rc;
        }
    }
    if (moved) {
        if (int rc = hip_check(hipEventRecord(e->ev_part, e->st2), "hipEventRecord")) return rc;
        e->part_pending = true;
    }
    if (e->prof) (void)hipEventRecord(e->ev[9], moved ? e->st2 : e->st);
    prof_acc(e, 7, 0, 10);
    prof_acc(e, 0, 10, 1);
    prof_acc(e, 1, 1, 2);
    if (remapped) prof_acc(e, 6, 3, 4);
    prof_acc(e, 2, 5, 6);
    prof_acc(e, 3, 6, 7);
    prof_acc(e, 4, 7, 8);
    prof_acc(e, 5, 8, 9);
    e->loop++;
    if (info) {
        info->t = e->tv[lc + 1];
        info->incomp = einc;
        info->remapped = remapped;
        info->stack_size = e->nstack;
        info->saturated = saturated;
        info->incomp_forward = einc_f;
        info->invertibility = einv;
    }
    return CMM_OK;
}

}  // namespace

int cmm_euler_step(cmm_euler *e, double dt_now, cmm_step_info *info) {
    if (!e || !e->initialised) { set_error("cmm_euler_step: engine not initialised"); return CMM_ERR_ARG; }
    if (!(dt_now > 0)) { set_error("cmm_euler_step: dt must be positive"); return CMM_ERR_ARG; }
    if (int rc = step_queue(e, dt_now, true)) return rc;
    return step_finish(e, info, false);
}

// ------------------------------------------------------------------------------------------------
// Ensembles: several engines stepped together, their queued step work in one captured graph

struct cmm_batch {
    std::vector<cmm_euler *> e;
    bool graphs = true, graphs_allowed = true;
    // A graph step is one captured graph launched on `master`: it forks to every engine stream
    // (one branch per engine) and joins them back.  A branch holds the engine's whole step: it
    // waits for tail[i] (the engine stream's work before the step: the previous step, a remap),
    // queues the advection, the check and the E_inc copy, records ev_ak[i] (external event
    // node, which the host waits for), queues the psi history and psi solve and records ev_bk[i],
    // which the engine stream waits for.  So each engine's kernels follow each other as in an
    // eager step, the host learns E_inc engine by engine while psi solves still run, and an
    // engine's next step starts when its own psi solve is done.  One graph per (step phase, state
    // flags of every engine): the psi ring and the map ping-pong rotate device pointers with
    // periods dividing 12, so a graph replays exactly the launches of the step it was captured on.
    hipStream_t master = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    std::map<std::string, hipGraphExec_t> cache;
    std::vector<hipEvent_t> tail, ev_ak, ev_bk;
    int captured = 0, launched = 0, eager = 0;
    // every engine's Lagrange weights (8 doubles each) in one pinned and one device array, so
    // that a graph step copies them with one transfer; the engines point into them while batched
    double *lw_host = nullptr, *lw_dev = nullptr;
    // grouped steps (op_group_vp_step): the recorded launches of every engine, the job tables
    // (pinned staging, device copy), E_inc of every engine (device, pinned) and the step's end
    bool grouping = false;
    int grouped = 0;
    int g_blocks = 0;  // grid cap of the grouped launches (0: one workgroup per block of work)
    // the engines split into lanes of about equal work, each lane's grouped launches on its own
    // stream (lane 0: master), so that the lanes' kernels overlap
    struct Lane {
        std::vector<int> members;
        std::vector<GroupInstance> inst;
        hipStream_t st = nullptr;
        bool own_stream = false;
        void *host = nullptr, *dev = nullptr;
        unsigned long long *hslots = nullptr, *dslots = nullptr;
        hipEvent_t done = nullptr;
    };
    int g_lanes = 1;
    std::vector<Lane> lanes;
    std::vector<std::pair<int, int>> lane_of;  // engine -> (lane, position in it)
};

namespace {

std::string batch_key(const cmm_batch *b) {
    std::string k;
    for (const cmm_euler *e : b->e) {
        k += std::to_string(e->loop % 12);
        k += e->m8_valid ? 'v' : '-';
        k += e->m32_fresh ? 'm' : '-';
        k += e->planes_fresh ? 'p' : '-';
        k += ';';
    }
    return k;
}

// Whether an engine's next step of length dt can run in a grouped step: an initialised FP64
// Vlasov-Poisson engine in its steady state (interleaved map current, bump function cached, the
// column sampler's node-interleaved field) whose scheme takes the separable advection.
bool group_ready(const cmm_euler *e, double dt) {
    const cmm_euler_params &p = e->p;
    if (p.sim != CMM_SIM_VLASOV_POISSON_1D || split(e) || p.forward_map || !e->initialised || !e->m8[0] ||
        !e->m8_valid || !e->vpH || !e->W4 || p.molly_stencil != 0 || e->corr1 || e->corr12 || e->prof_every)
        return false;
    const int L = p.l_order, lc = e->loop + L - 1;
    if ((int)e->tv.size() < lc + 1) return false;
    std::vector<double> t2(e->tv.begin(), e->tv.begin() + lc + 1), d2(e->dtv.begin(), e->dtv.begin() + lc + 1);
    t2.push_back(t2[lc] + dt);
    d2.push_back(dt);
    cmm_advect_coeffs c;
    if (cmm_advect_coeffs_fill(t2.data(), d2.data(), e->loop, L, p.time_integration_num, p.map_update_order_num,
                               p.map_epsilon, -1, &c))
        return false;
    return op_vp_advect_supported(p.g_coarse, p.g_psi, c);
}

int batch_normalise(cmm_euler *e);

// whether the batch's engines can share grouped launches (op_group_vp_alike on what the engines'
// parameters fix), checked before any engine records a step
bool batch_group_alike(const cmm_batch *b);

// One grouped step of every engine: each engine's step is run with its launches recorded
// (step_queue_a/b with grp set), then op_group_vp_step launches all of them together on the master
// stream, ordered after every engine stream's earlier work; the engines' streams are ordered after
// it, and each engine finishes its step (E_inc, remap, bookkeeping) as after its own.
void batch_free_lanes(cmm_batch *b) {
    for (cmm_batch::Lane &l : b->lanes) {
        if (l.st) (void)hipStreamSynchronize(l.st);
        if (l.dev) (void)hipFree(l.dev);
        if (l.dslots) (void)hipFree(l.dslots);
        if (l.host) (void)hipHostFree(l.host);
        if (l.hslots) (void)hipHostFree(l.hslots);
        if (l.done) (void)hipEventDestroy(l.done);
        if (l.own_stream && l.st) (void)hipStreamDestroy(l.st);
    }
    b->lanes.clear();
}

// b->g_lanes lanes (fewer where there are fewer engines): the engines dealt largest first to the
// lane with the least work so far (work: fine-grid nodes), each lane with its job tables and slots
int batch_lanes(cmm_batch *b) {
    const int n = (int)b->e.size(), nl = std::max(1, std::min(b->g_lanes, n));
    if ((int)b->lanes.size() == nl) return CMM_OK;
    batch_free_lanes(b);
    b->lanes.resize(nl);
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    auto work = [&](int i) { return (double)b->e[i]->nf; };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work(x) > work(y); });
    std::vector<double> load(nl, 0.0);
    b->lane_of.assign(n, {0, 0});
    for (int i : order) {
        const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        b->lane_of[i] = {l, (int)b->lanes[l].members.size()};
        b->lanes[l].members.push_back(i);
        load[l] += work(i);
    }
    for (int l = 0; l < nl; ++l) {
        cmm_batch::Lane &ln = b->lanes[l];
        const int m = (int)ln.members.size();
        const size_t jb = op_group_vp_job_bytes(m);
        if (hipMalloc(&ln.dev, jb) != hipSuccess || hipHostMalloc(&ln.host, jb, hipHostMallocDefault) != hipSuccess ||
            hipMalloc((void **)&ln.dslots, sizeof(unsigned long long) * m) != hipSuccess ||
            hipHostMalloc((void **)&ln.hslots, sizeof(unsigned long long) * m, hipHostMallocDefault) != hipSuccess) {
            set_error("grouped step: allocation of the job tables");
            return CMM_ERR_ALLOC;
        }
        if (int rc = hip_check(hipEventCreateWithFlags(&ln.done, hipEventDisableTiming), "hipEventCreate")) return rc;
        if (l == 0) {
            ln.st = b->master;
        } else {
            if (int rc = hip_check(hipStreamCreateWithFlags(&ln.st, hipStreamNonBlocking), "hipStreamCreate")) return rc;
            ln.own_stream = true;
        }
    }
    return CMM_OK;
}

bool batch_group_alike(const cmm_batch *b) {
    std::vector<GroupInstance> g(b->e.size());
    for (size_t i = 0; i < b->e.size(); ++i) {
        const cmm_euler *e = b->e[i];
        g[i].gm = e->p.g_coarse;
        g[i].gp = e->p.g_psi;
        g[i].gv = e->p.g_vort;
        g[i].gc = e->p.g_fine;
        g[i].W4 = e->W4;
        g[i].tune = e->tune;
        g[i].c.time_integration_num = e->p.time_integration_num;
        g[i].c.l_order = e->p.l_order;
    }
    return op_group_vp_alike(g.data(), (int)g.size());
}

int batch_step_grouped(cmm_batch *b, double dt, cmm_step_info *infos) {
    const int n = (int)b->e.size();
    if (int rc = batch_lanes(b)) return rc;
    for (int i = 0; i < n; ++i) {
        cmm_batch::Lane &ln = b->lanes[b->lane_of[i].first];
        if (int rc = hip_check(hipEventRecord(b->tail[i], b->e[i]->st), "hipEventRecord (tail)")) return rc;
        if (int rc = hip_check(hipStreamWaitEvent(ln.st, b->tail[i], 0), "hipStreamWaitEvent (tail)")) return rc;
    }
    for (cmm_batch::Lane &ln : b->lanes) ln.inst.assign(ln.members.size(), GroupInstance{});
    for (int i = 0; i < n; ++i) {
        cmm_euler *e = b->e[i];
        e->grp = &b->lanes[b->lane_of[i].first].inst[b->lane_of[i].second];
        bool fork = false;
        int rc = step_queue_a(e, dt, false, &fork);
        if (!rc) rc = step_queue_b(e, false);
        e->grp = nullptr;
        if (rc) return rc;
    }
    for (cmm_batch::Lane &ln : b->lanes) {
        if (int rc = op_group_vp_step(ln.inst.data(), (int)ln.inst.size(), ln.dev, ln.host, ln.dslots, ln.hslots, ln.st,
                                      b->g_blocks))
            return rc;
        if (int rc = hip_check(hipEventRecord(ln.done, ln.st), "hipEventRecord")) return rc;
    }
    for (cmm_batch::Lane &ln : b->lanes)
        if (int rc = hip_check(hipEventSynchronize(ln.done), "hipEventSynchronize")) return rc;
    for (int i = 0; i < n; ++i) {
        cmm_euler *e = b->e[i];
        const cmm_batch::Lane &ln = b->lanes[b->lane_of[i].first];
        std::memcpy(e->h_slot, ln.hslots + b->lane_of[i].second, sizeof(unsigned long long));
        if (int rc = hip_check(hipStreamWaitEvent(e->st, ln.done, 0), "hipStreamWaitEvent")) return rc;
        // the interleaved map is brought back after a remap whether or not the caller takes infos
        cmm_step_info own{};
        cmm_step_info &info = infos ? infos[i] : own;
        if (int rc = step_finish(e, &info, true)) return rc;
        if (info.remapped)
            if (int rc = batch_normalise(e)) return rc;
    }
    b->grouped++;
    return CMM_OK;
}

// after a remap, bring the forms of the map a replayed step reads up to date, so that the next
// step's launches are those of the captured steady state
int batch_normalise(cmm_euler *e) {
    if (split(e)) return sync_m32(e);
    if (e->m8[0] && !e->m8_valid) {
        if (int rc = op_interleave_map(e->p.g_coarse, e->cx, e->cy, e->m8[e->m8_cur], e->st)) return rc;
        e->m8_valid = true;
    }
    return CMM_OK;
}

// An event record (record) or wait node on the capture of stream st, depending on the stream's
// current capture leaves and becoming its only leaf: the engines' branches of the two phase graphs
// are chained by such nodes (they record and wait in launch order, across graph launches)
int capture_event_node(hipStream_t st, hipEvent_t ev, bool record) {
    hipStreamCaptureStatus status = hipStreamCaptureStatusNone;
    unsigned long long id = 0;
    hipGraph_t graph = nullptr;
    const hipGraphNode_t *deps = nullptr;
    size_t n_deps = 0;
    if (int rc = hip_check(hipStreamGetCaptureInfo_v2(st, &status, &id, &graph, &deps, &n_deps), "hipStreamGetCaptureInfo"))
        return rc;
    if (status != hipStreamCaptureStatusActive || !graph) {
        set_error("batch capture: engine stream not capturing");
        return CMM_ERR_HIP;
    }
    std::vector<hipGraphNode_t> d(deps, deps + n_deps);
    hipGraphNode_t node = nullptr;
    if (record) {
        if (int rc = hip_check(hipGraphAddEventRecordNode(&node, graph, d.data(), d.size(), ev), "hipGraphAddEventRecordNode"))
            return rc;
    } else {
        if (int rc = hip_check(hipGraphAddEventWaitNode(&node, graph, d.data(), d.size(), ev), "hipGraphAddEventWaitNode"))
            return rc;
    }
    return hip_check(hipStreamUpdateCaptureDependencies(st, &node, 1, hipStreamSetCaptureDependencies),
                     "hipStreamUpdateCaptureDependencies");
}

// The step of every engine captured into one graph from the master stream: a branch per engine
// stream, opened by a wait for the engine's tail event, with a record node after the E_inc copy
// and one at the end (external event nodes), joined back into the master stream.
int batch_capture(cmm_batch *b, double dt, hipGraphExec_t *out) {
    int rc = hip_check(hipStreamBeginCapture(b->master, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    if (rc) return rc;
    rc = hip_check(hipEventRecord(b->fork, b->master), "hipEventRecord (fork)");
    for (size_t i = 0; i < b->e.size() && !rc; ++i) {
        cmm_euler *e = b->e[i];
        rc = hip_check(hipStreamWaitEvent(e->st, b->fork, 0), "hipStreamWaitEvent (fork)");
        if (!rc) rc = capture_event_node(e->st, b->tail[i], false);
        e->capturing = true;
        bool fork = false;
        if (!rc) rc = step_queue_a(e, dt, false, &fork);
        e->lw_graph = false;
        if (!rc) rc = capture_event_node(e->st, b->ev_ak[i], true);
        if (!rc) rc = step_queue_b(e, false);
        e->capturing = false;
        if (!rc) rc = capture_event_node(e->st, b->ev_bk[i], true);
        if (!rc) rc = hip_check(hipEventRecord(b->join[i], e->st), "hipEventRecord (join)");
        if (!rc) rc = hip_check(hipStreamWaitEvent(b->master, b->join[i], 0), "hipStreamWaitEvent (join)");
    }
    // the capture always ends; a failure inside it keeps its own message
    hipGraph_t g = nullptr;
    const hipError_t er = hipStreamEndCapture(b->master, &g);
    if (!rc) rc = hip_check(er, "hipStreamEndCapture");
    if (!rc) rc = hip_check(hipGraphInstantiate(out, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (g) (void)hipGraphDestroy(g);
    return rc;
}

// the graph of this step's state, cached or captured; with a cached graph the engines only advance
// their host state (dry pass) and the graph carries the launches
int batch_graph(cmm_batch *b, double dt, hipGraphExec_t *exec) {
    const std::string key = batch_key(b);
    auto it = b->cache.find(key);
    for (cmm_euler *e : b->e) e->lw_graph = true;
    int rc = CMM_OK;
    if (it == b->cache.end()) {
        rc = batch_capture(b, dt, exec);
        if (!rc) {
            b->cache[key] = *exec;
            b->captured++;
        }
    } else {
        *exec = it->second;
        for (cmm_euler *e : b->e) {
            e->dry = true;
            bool fork = false;
            rc = step_queue_a(e, dt, false, &fork);
            e->lw_graph = false;
            if (!rc) rc = step_queue_b(e, false);
            e->dry = false;
            if (rc) break;
        }
    }
    for (cmm_euler *e : b->e) e->lw_graph = false;
    return rc;
}

}  // namespace

int cmm_batch_create(cmm_euler *const *engines, int n, int use_graphs, cmm_batch **out) {
    if (!engines || n <= 0 || !out) { set_error("cmm_batch_create: bad argument"); return CMM_ERR_ARG; }
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        const cmm_euler *e = engines[i];
        if (!e || !e->initialised || e->p.forward_map || !e->part.empty()) {
            set_error("cmm_batch_create: engines must be initialised, without a forward map or particles");
            return CMM_ERR_ARG;
        }
        for (int j = 0; j < i; ++j)
            if (engines[j] == e) { set_error("cmm_batch_create: an engine appears twice"); return CMM_ERR_ARG; }
    }
    cmm_batch *b = new cmm_batch();
    b->e.assign(engines, engines + n);
    b->graphs = b->graphs_allowed = use_graphs != 0;
    int rc = hip_check(hipStreamCreateWithFlags(&b->master, hipStreamNonBlocking), "hipStreamCreate");
    if (!rc) rc = hip_check(hipEventCreateWithFlags(&b->fork, kDeviceEvent), "hipEventCreate");
    for (auto *v : {&b->join, &b->tail, &b->ev_bk}) {
        v->assign(n, nullptr);
        for (int i = 0; i < n && !rc; ++i) rc = hip_check(hipEventCreateWithFlags(&(*v)[i], kDeviceEvent), "hipEventCreate");
    }
    // the host waits for these after the E_inc copies to pinned memory: system-scope release
    b->ev_ak.assign(n, nullptr);
    for (int i = 0; i < n && !rc; ++i) rc = hip_check(hipEventCreateWithFlags(&b->ev_ak[i], hipEventDisableTiming), "hipEventCreate");
    if (!rc && hipHostMalloc((void **)&b->lw_host, sizeof(double) * 8 * n, hipHostMallocDefault) != hipSuccess) {
        b->lw_host = nullptr;
        set_error("hipHostMalloc batch weights");
        rc = CMM_ERR_ALLOC;
    }
    if (!rc && hipMalloc(&b->lw_dev, sizeof(double) * 8 * n) != hipSuccess) {
        b->lw_dev = nullptr;
        set_error("hipMalloc batch weights");
        rc = CMM_ERR_ALLOC;
    }
    if (rc) { cmm_batch_destroy(b); return rc; }
    for (int i = 0; i < n; ++i) {
        b->e[i]->lw_host = b->lw_host + 8 * i;
        b->e[i]->lw_dev = b->lw_dev + 8 * i;
    }
    *out = b;
    return CMM_OK;
}

int cmm_batch_destroy(cmm_batch *b) {
    if (!b) return CMM_OK;
    if (b->master) (void)hipStreamSynchronize(b->master);
    // the engines' own weight slots again (bytes 64..127 of their slots)
    if (b->lw_host)
        for (cmm_euler *e : b->e) {
            e->lw_dev = reinterpret_cast<double *>(e->slot + 8);
            e->lw_host = reinterpret_cast<double *>(e->h_slot + 8);
        }
    for (auto &kv : b->cache) (void)hipGraphExecDestroy(kv.second);
    batch_free_lanes(b);
    for (auto *v : {&b->join, &b->tail, &b->ev_ak, &b->ev_bk})
        for (hipEvent_t ev : *v)
            if (ev) (void)hipEventDestroy(ev);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->lw_dev) (void)hipFree(b->lw_dev);
    if (b->lw_host) (void)hipHostFree(b->lw_host);
    if (b->master) (void)hipStreamDestroy(b->master);
    delete b;
    return CMM_OK;
}

int cmm_batch_step(cmm_batch *b, double dt, cmm_step_info *infos) {
    if (!b || !(dt > 0)) { set_error("cmm_batch_step: bad argument"); return CMM_ERR_ARG; }
    for (cmm_euler *e : b->e)
        if (e->prof_every || !e->part.empty()) { set_error("cmm_batch_step: profiling and particles are not batched"); return CMM_ERR_ARG; }
    // grouped launches where every engine is ready for them
    bool grouped = b->grouping && (int)b->e.size() <= kGroupMax;
    for (cmm_euler *e : b->e) grouped = grouped && group_ready(e, dt);
    grouped = grouped && batch_group_alike(b);
    if (grouped) return batch_step_grouped(b, dt, infos);
    // a graph replays the steady state: the nominal step of every engine, after its first step
    bool graph = b->graphs;
    for (cmm_euler *e : b->e) graph = graph && e->loop >= 1 && dt == e->p.dt;
    if (graph) {
        // each engine's stream work so far (the previous step, a remap): the engine's branch
        // waits for it
        for (size_t i = 0; i < b->e.size(); ++i)
            if (int rc = hip_check(hipEventRecord(b->tail[i], b->e[i]->st), "hipEventRecord (tail)")) return rc;
        hipGraphExec_t g = nullptr;
        if (int rc = batch_graph(b, dt, &g)) return rc;
        // this step's Lagrange weights, in stream order before the graph reads them (their last
        // readers, the previous step's advections, have completed: the host waited for ev_ak)
        if (int rc = hip_check(hipMemcpyAsync(b->lw_dev, b->lw_host, sizeof(double) * 8 * b->e.size(),
                                              hipMemcpyHostToDevice, b->master),
                               "hipMemcpyAsync weights"))
            return rc;
        if (int rc = hip_check(hipGraphLaunch(g, b->master), "hipGraphLaunch")) return rc;
        // every engine stream's later work (a remap, field reads, the next eager step) after its
        // own branch
        for (size_t i = 0; i < b->e.size(); ++i)
            if (int rc = hip_check(hipStreamWaitEvent(b->e[i]->st, b->ev_bk[i], 0), "hipStreamWaitEvent")) return rc;
        b->launched++;
    } else {
        for (cmm_euler *e : b->e)
            if (int rc = step_queue(e, dt)) return rc;
        b->eager++;
    }
    for (size_t i = 0; i < b->e.size(); ++i) {
        cmm_euler *e = b->e[i];
        // E_inc of the engine is in pinned memory once its branch has passed ev_ak
        if (graph)
            if (int rc = hip_check(hipEventSynchronize(b->ev_ak[i]), "hipEventSynchronize")) return rc;
        cmm_step_info own{};
        cmm_step_info &info = infos ? infos[i] : own;
        if (int rc = step_finish(e, &info, graph)) return rc;
        if (info.remapped)
            if (int rc = batch_normalise(e)) return rc;
    }
    return CMM_OK;
}

int cmm_batch_set_grouping(cmm_batch *b, int on) {
    if (!b) { set_error("cmm_batch_set_grouping: null batch"); return CMM_ERR_ARG; }
    b->grouping = on != 0;
    return CMM_OK;
}

namespace {

// every engine's streams and the batch's stream drained
int batch_sync(cmm_batch *b) {
    if (int rc = hip_check(hipStreamSynchronize(b->master), "hipStreamSynchronize")) return rc;
    for (cmm_euler *e : b->e) {
        if (int rc = hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize")) return rc;
        if (int rc = hip_check(hipStreamSynchronize(e->st2), "hipStreamSynchronize")) return rc;
    }
    return CMM_OK;
}

void batch_set_mode(cmm_batch *b, int mode) {
    b->grouping = mode == CMM_BATCH_GROUPED || mode == CMM_BATCH_GROUPED_PERSISTENT || mode == CMM_BATCH_GROUPED_LANES;
    b->g_blocks = mode == CMM_BATCH_GROUPED_PERSISTENT ? 2048 : 0;
    b->g_lanes = mode == CMM_BATCH_GROUPED_LANES ? 4 : 1;
    b->graphs = mode == CMM_BATCH_GRAPHS;
}

}  // namespace

int cmm_batch_set_mode(cmm_batch *b, int mode) {
    if (!b || mode < CMM_BATCH_EAGER || mode > CMM_BATCH_GROUPED_LANES || (mode == CMM_BATCH_GRAPHS && !b->graphs_allowed)) {
        set_error("cmm_batch_set_mode: bad argument");
        return CMM_ERR_ARG;
    }
    batch_set_mode(b, mode);
    return CMM_OK;
}

// The batch's launch form chosen by timing: each candidate (grouped launches with one workgroup per
// block of work, persistent grouped launches of 2048 workgroups, grouped launches in 4 lanes of
// engines on their own streams, captured graphs, eager queues)
// runs steps_per_mode real steps of length dt back to back (the run advances; captured graphs
// first take the 12 steps that capture them, untimed), between full synchronisations; the fastest
// stays.  All forms give the same results bit for bit.
int cmm_batch_autotune(cmm_batch *b, double dt, int steps_per_mode, int *mode_out, double *ms_out) {
    if (!b || !(dt > 0) || steps_per_mode < 1) { set_error("cmm_batch_autotune: bad argument"); return CMM_ERR_ARG; }
    // the first step of fresh engines interleaves their maps (grouped steps start after it)
    bool ready = (int)b->e.size() <= kGroupMax;
    for (cmm_euler *e : b->e) ready = ready && group_ready(e, dt);
    if (!ready) {
        batch_set_mode(b, CMM_BATCH_EAGER);
        if (int rc = cmm_batch_step(b, dt, nullptr)) return rc;
        ready = (int)b->e.size() <= kGroupMax;
        for (cmm_euler *e : b->e) ready = ready && group_ready(e, dt);
    }
    ready = ready && batch_group_alike(b);
    std::vector<int> cand;
    if (ready) {
        cand.push_back(CMM_BATCH_GROUPED);
        cand.push_back(CMM_BATCH_GROUPED_PERSISTENT);
        if (b->e.size() > 1) cand.push_back(CMM_BATCH_GROUPED_LANES);
    }
    if (b->graphs_allowed) cand.push_back(CMM_BATCH_GRAPHS);
    cand.push_back(CMM_BATCH_EAGER);
    int best = CMM_BATCH_EAGER;
    double best_ms = 1e300;
    for (int mode : cand) {
        batch_set_mode(b, mode);
        if (mode == CMM_BATCH_GRAPHS)
            for (int i = 0; i < 12; ++i)
                if (int rc = cmm_batch_step(b, dt, nullptr)) return rc;
        if (int rc = batch_sync(b)) return rc;
        // each step's time (a step returns once every engine's E_inc is in and its remap queued);
        // the median keeps a step that remaps from deciding the choice
        std::vector<double> t(steps_per_mode);
        auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < steps_per_mode; ++i) {
            if (int rc = cmm_batch_step(b, dt, nullptr)) return rc;
            if (i == steps_per_mode - 1)
                if (int rc = batch_sync(b)) return rc;
            const auto t1 = std::chrono::steady_clock::now();
            t[i] = std::chrono::duration<double, std::milli>(t1 - t0).count();
            t0 = t1;
        }
        std::sort(t.begin(), t.end());
        const int m = steps_per_mode / 2;
        const double ms = steps_per_mode % 2 ? t[m] : 0.5 * (t[m - 1] + t[m]);
        if (ms < best_ms) {
            best_ms = ms;
            best = mode;
        }
    }
    batch_set_mode(b, best);
    if (mode_out) *mode_out = best;
    if (ms_out) *ms_out = best_ms;
    return CMM_OK;
}

int cmm_batch_grouped_steps(cmm_batch *b, int *grouped_steps) {
    if (!b || !grouped_steps) { set_error("cmm_batch_grouped_steps: bad argument"); return CMM_ERR_ARG; }
    *grouped_steps = b->grouped;
    return CMM_OK;
}

int cmm_batch_stats(cmm_batch *b, int *graphs_captured, int *graph_launches, int *eager_steps) {
    if (!b) { set_error("cmm_batch_stats: null batch"); return CMM_ERR_ARG; }
    if (graphs_captured) *graphs_captured = b->captured;
    if (graph_launches) *graph_launches = b->launched;
    if (eager_steps) *eager_steps = b->eager;
    return CMM_OK;
}

int cmm_euler_set_check_tile(cmm_euler *e, int tile, int max_blocks) {
    if (!e || tile < 0 || tile > 4 || max_blocks < 1) {
        set_error("cmm_euler_set_check_tile: bad argument (tile 0 to 4; max_blocks >= 1)");
        return CMM_ERR_ARG;
    }
    e->tune.tile = tile;
    e->tune.max_blocks = max_blocks;
    return CMM_OK;
}

// The check's tile edge (4, 2, 1 where it divides the refinement ratio) and block cap, timed on the
// engine's current map with events (the fastest of 5 runs per candidate after one untimed run)
int cmm_euler_autotune(cmm_euler *e, int *tile_out, int *max_blocks_out, double *ms_out) {
    if (!e || !e->initialised) { set_error("cmm_euler_autotune: engine not initialised"); return CMM_ERR_ARG; }
    const int r = e->p.g_fine.nx / e->p.g_coarse.nx;
    hipEvent_t a = nullptr, z = nullptr;
    int rc = hip_check(hipEventCreate(&a), "hipEventCreate");
    if (!rc) rc = hip_check(hipEventCreate(&z), "hipEventCreate");
    IncompTune best = e->tune;
    float best_ms = 1e30f;
    for (int tile : {4, 2, 1}) {
        if (r <= 0 || r % tile) continue;
        for (int cap : {256, 1024, 4096, 16384}) {
            if (rc) break;
            e->tune.tile = tile;
            e->tune.max_blocks = cap;
            float m_min = 1e30f;
            for (int rep = 0; rep < 6 && !rc; ++rep) {
                rc = hip_check(hipEventRecord(a, e->st), "hipEventRecord");
                if (!rc) rc = launch_incomp(e, e->st);
                if (!rc) rc = hip_check(hipEventRecord(z, e->st), "hipEventRecord");
                if (!rc) rc = hip_check(hipEventSynchronize(z), "hipEventSynchronize");
                float m = 0;
                if (!rc) rc = hip_check(hipEventElapsedTime(&m, a, z), "hipEventElapsedTime");
                if (rep > 0) m_min = std::min(m_min, m);
            }
            if (!rc && m_min < best_ms) {
                best_ms = m_min;
                best = e->tune;
            }
        }
    }
    e->tune = best;
    if (a) (void)hipEventDestroy(a);
    if (z) (void)hipEventDestroy(z);
    if (rc) return rc;
    if (tile_out) *tile_out = best.tile;
    if (max_blocks_out) *max_blocks_out = best.max_blocks;
    if (ms_out) *ms_out = best_ms;
    return CMM_OK;
}

int cmm_euler_psi_level(cmm_euler *e, int level, const double **dev_ptr) {
    if (!e || !dev_ptr || level < 0 || level >= e->p.l_order) { set_error("cmm_euler_psi_level: bad argument"); return CMM_ERR_ARG; }
    *dev_ptr = e->lvl[level];
    return CMM_OK;
}

int cmm_euler_copy_psi_level(cmm_euler *e, int level, int plane, double *dst) {
    if (!e || !dst || level < 0 || level >= e->p.l_order || plane < 0 || plane > 3) {
        set_error("cmm_euler_copy_psi_level: bad argument");
        return CMM_ERR_ARG;
    }
    return hip_check(hipMemcpyAsync(dst, e->lvl[level] + (size_t)plane * e->np, sizeof(double) * e->np,
                                    hipMemcpyDeviceToDevice, e->st),
                     "hipMemcpyAsync");
}

int cmm_euler_set_psi_correction(cmm_euler *e, const double *d1, const double *d12) {
    if (!e) { set_error("cmm_euler_set_psi_correction: null engine"); return CMM_ERR_ARG; }
    if ((d1 || d12) && (e->p.l_order < 2 || e->p.forward_map)) {
        set_error("cmm_euler_set_psi_correction: needs two or more psi levels and no forward map");
        return CMM_ERR_ARG;
    }
    e->corr1 = d1;
    e->corr12 = d12;
    e->pre_valid = false;
    return CMM_OK;
}

int cmm_euler_get_field(cmm_euler *e, int field, double *out) {
    if (!e || !out) { set_error("cmm_euler_get_field: bad argument"); return CMM_ERR_ARG; }
    int rc = sync_planes(e);
    if (rc) return rc;
    switch (field) {
    case CMM_FIELD_CHIX: rc = hip_check(hipMemcpyAsync(out, e->cx, sizeof(double) * 4 * e->nc, hipMemcpyDeviceToHost, e->st), "get"); break;
    case CMM_FIELD_CHIY: rc = hip_check(hipMemcpyAsync(out, e->cy, sizeof(double) * 4 * e->nc, hipMemcpyDeviceToHost, e->st), "get"); break;
    case CMM_FIELD_PSI:
        for (int l = 0; l < e->p.l_order && !rc; ++l)
            rc = hip_check(hipMemcpyAsync(out + (size_t)4 * e->np * l, e->lvl[l], sizeof(double) * 4 * e->np, hipMemcpyDeviceToHost, e->st), "get");
        break;
    case CMM_FIELD_W_FINE: rc = hip_check(hipMemcpyAsync(out, e->W, sizeof(double) * 4 * e->nf, hipMemcpyDeviceToHost, e->st), "get"); break;
    case CMM_FIELD_W_COARSE: {
        if (!e->tmp) rc = dalloc(&e->tmp, e->nc);
        if (!rc) rc = op_sample_vort(e->p.g_coarse, e->p.g_coarse, e->p.g_fine, e->cx, e->cy, e->W, 0, e->tmp, e->st);
        if (!rc) rc = hip_check(hipMemcpyAsync(out, e->tmp, sizeof(double) * e->nc, hipMemcpyDeviceToHost, e->st), "get");
        break;
    }
    case CMM_FIELD_CHIX_F:
    case CMM_FIELD_CHIY_F:
        if (!e->fx) { set_error("cmm_euler_get_field: the engine has no forward map"); return CMM_ERR_ARG; }
        rc = hip_check(hipMemcpyAsync(out, field == CMM_FIELD_CHIX_F ? e->fx : e->fy, sizeof(double) * 4 * e->nc,
                                      hipMemcpyDeviceToHost, e->st), "get");
        break;
    case CMM_FIELD_DISC:
        if (!e->disc) { set_error("no discrete initial condition"); return CMM_ERR_ARG; }
        rc = hip_check(hipMemcpyAsync(out, e->disc, sizeof(double) * 4 * e->nd, hipMemcpyDeviceToHost, e->st), "get");
        break;
    default: set_error("cmm_euler_get_field: unknown field"); return CMM_ERR_ARG;
    }
    if (rc) return rc;
    return hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
}

int cmm_euler_get_particles(cmm_euler *e, int set, double *out) {
    if (!e || set < 0 || set >= (int)e->part.size() || !out) { set_error("cmm_euler_get_particles: bad argument"); return CMM_ERR_ARG; }
    if (e->part_n[set] <= 0) return CMM_OK;
    if (int rc = hip_check(hipMemcpyAsync(out, e->part[set], sizeof(double) * 2 * e->part_n[set], hipMemcpyDeviceToHost, e->st2), "get")) return rc;
    return hip_check(hipStreamSynchronize(e->st2), "hipStreamSynchronize");
}

int cmm_euler_get_particle_velocities(cmm_euler *e, int set, double *out) {
    if (!e || set < 0 || set >= (int)e->part.size() || !out) {
        set_error("cmm_euler_get_particle_velocities: bad argument");
        return CMM_ERR_ARG;
    }
    const int n = e->part_n[set];
    if (n <= 0) return CMM_OK;
    if (e->part_vel[set]) {
        if (int rc = hip_check(hipMemcpyAsync(out, e->part_vel[set], sizeof(double) * 2 * n, hipMemcpyDeviceToHost, e->st2), "get"))
            return rc;
        return hip_check(hipStreamSynchronize(e->st2), "hipStreamSynchronize");
    }
    // fluid particles: psi's x and y derivatives at the positions (the engine stream's psi, after the
    // second stream's particle step)
    double *d = nullptr;
    if (int rc = dalloc(&d, 2 * (size_t)n)) return rc;
    int rc = hip_check(hipStreamSynchronize(e->st2), "hipStreamSynchronize");
    if (!rc) rc = op_points_dxdy(e->p.g_psi, e->lvl[0], n, e->part[set], d, e->st);
    if (!rc) rc = hip_check(hipMemcpyAsync(out, d, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, e->st), "get");
    if (!rc) rc = hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
    (void)hipFree(d);
    return rc;
}

int cmm_euler_diagnostics(cmm_euler *e, double *out) {
    if (!e || !out) { set_error("cmm_euler_diagnostics: bad argument"); return CMM_ERR_ARG; }
    const cmm_grid &gc = e->p.g_coarse, &gp = e->p.g_psi;
    int rc = sync_planes(e);
    if (!e->tmp) rc = dalloc(&e->tmp, e->nc);
    if (!rc && !e->tmp4) rc = dalloc(&e->tmp4, 4 * e->nc);
    // coarse vorticity and its spectral gradient (Hermite planes 1, 2 are grad_x, grad_y of
    // Cuda-CMM/src/numerical/cmm-fft.cu:457-474)
    if (!rc) rc = op_sample_vort(gc, gc, e->p.g_fine, e->cx, e->cy, e->W, 0, e->tmp, e->st);
    const bool fft_ok = gc.nx == gc.ny && gc.nx >= 8 && gc.nx <= 4096 && (gc.nx & (gc.nx - 1)) == 0;
    if (!rc && fft_ok) rc = op_hermite_from_real(gc, e->tmp, e->tmp4, nullptr, 0, e->st);
    std::vector<double> w(fft_ok ? 4 * e->nc : e->nc), ps(2 * e->np);
    if (!rc) rc = hip_check(hipMemcpyAsync(w.data(), fft_ok ? e->tmp4 : e->tmp, sizeof(double) * w.size(), hipMemcpyDeviceToHost, e->st), "get");
    if (!rc) rc = hip_check(hipMemcpyAsync(ps.data(), e->lvl[0] + e->np, sizeof(double) * 2 * e->np, hipMemcpyDeviceToHost, e->st), "get");
    if (!rc) rc = hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
    if (rc) return rc;
    // Cuda-CMM/src/numerical/cmm-mesure.cu:30-89 and cmm-simulation-host.cu:667-673
    const DGrid dc = make_dgrid(gc), dp = make_dgrid(gp);
    double se = 0, sz = 0, sp = 0, wmax = 0, wmin = 0;
    for (long i = 0; i < 2 * e->np; ++i) se += ps[i] * ps[i];
    for (long i = 0; i < e->nc; ++i) {
        sz += w[i] * w[i];
        wmax = std::max(wmax, w[i]);
        wmin = std::min(wmin, w[i]);
    }
    double spx = 0, spy = 0;
    if (fft_ok)
        for (long i = 0; i < e->nc; ++i) { spx += w[e->nc + i] * w[e->nc + i]; spy += w[2 * e->nc + i] * w[2 * e->nc + i]; }
    sp = 0.5 * dc.hx * dc.hy * spx + 0.5 * dc.hx * dc.hy * spy;
    out[0] = 0.5 * dp.hx * dp.hy * se;
    out[1] = 0.5 * dc.hx * dc.hy * sz;
    out[2] = fft_ok ? sp : NAN;
    out[3] = std::max(wmax, -wmin);
    return CMM_OK;
}

int cmm_euler_vp_diagnostics(cmm_euler *e, double *out) {
    if (!e || !out) { set_error("cmm_euler_vp_diagnostics: bad argument"); return CMM_ERR_ARG; }
    const cmm_grid &gc = e->p.g_coarse, &gp = e->p.g_psi;
    int rc = sync_planes(e);
    if (!e->tmp) rc = dalloc(&e->tmp, e->nc);
    if (!rc) rc = op_sample_vort(gc, gc, e->p.g_fine, e->cx, e->cy, e->W, 0, e->tmp, e->st);
    std::vector<double> f(e->nc), ex(gp.nx);
    if (!rc) rc = hip_check(hipMemcpyAsync(f.data(), e->tmp, sizeof(double) * e->nc, hipMemcpyDeviceToHost, e->st), "get");
    if (!rc) rc = hip_check(hipMemcpyAsync(ex.data(), e->lvl[0] + e->np, sizeof(double) * gp.nx, hipMemcpyDeviceToHost, e->st), "get");
    if (!rc) rc = hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
    if (rc) return rc;
    const DGrid dc = make_dgrid(gc), dp = make_dgrid(gp);
    double mass = 0, kin = 0, mom = 0, pot = 0;
    for (int iy = 0; iy < gc.ny; ++iy) {
        const double v = gc.y0 + iy * dc.hy;
        for (int ix = 0; ix < gc.nx; ++ix) {
            const double fv = f[(size_t)iy * gc.nx + ix];
            mass += fv;
            kin += v * v * fv;
            mom += v * fv;
        }
    }
    for (int ix = 0; ix < gp.nx; ++ix) pot += ex[ix] * ex[ix];
    out[1] = 0.5 * dc.hy * dc.hx * kin;
    out[2] = 0.5 * dp.hx * pot;
    out[0] = out[1] + out[2];
    out[3] = dc.hy * dc.hx * mass;
    out[4] = dc.hy * dc.hx * mom;
    // Compute_kth_Mode: |R2C(dpsi/dx row)[k]| / NX for k = 1..4
    for (int k = 1; k <= 4; ++k) {
        double re = 0, im = 0;
        for (int ix = 0; ix < gp.nx; ++ix) {
            const double a = -2.0 * 3.141592653589793238462643383279502884 * (double)((long)k * ix % gp.nx) / gp.nx;
            re += ex[ix] * std::cos(a);
            im += ex[ix] * std::sin(a);
        }
        out[4 + k] = k <= gp.nx / 2 ? std::sqrt(re * re + im * im) / gp.nx : 0.0;
    }
    return CMM_OK;
}

int cmm_euler_sample(cmm_euler *e, cmm_grid gs, int scalar_ic, double *out) {
    if (!e || !out || !grid_ok(gs)) { set_error("cmm_euler_sample: bad argument"); return CMM_ERR_ARG; }
    const cmm_grid &gc = e->p.g_coarse;
    if (gs.x0 != gc.x0 || gs.x1 != gc.x1 || gs.y0 != gc.y0 || gs.y1 != gc.y1) {
        set_error("cmm_euler_sample: the sample grid must have the map's bounds");
        return CMM_ERR_ARG;
    }
    const size_t n = (size_t)gs.nx * gs.ny;
    // device buffer: positions (2n), W (n), theta (n), one Hermite result (4n), kept for the next call
    if (e->sbuf_n < n) {
        if (e->sbuf) (void)hipFree(e->sbuf);
        e->sbuf = nullptr;
        e->sbuf_n = 0;
        if (int rc = dalloc(&e->sbuf, 8 * n)) return rc;
        e->sbuf_n = n;
    }
    if (int rc = sync_planes(e)) return rc;
    double *xy = e->sbuf, *w = xy + 2 * n, *th = w + n, *h = th + n;
    std::vector<double> hxy(2 * n);
    auto get = [&](double *dst, const double *src, size_t cnt) {
        return hip_check(hipMemcpyAsync(dst, src, sizeof(double) * cnt, hipMemcpyDeviceToHost, e->st), "get");
    };
    int rc = op_compose_positions(gc, gs, e->cx, e->cy, comp_stack(e), e->nstack, xy, e->st, comp_fmt(e));
    if (!rc) rc = get(hxy.data(), xy, 2 * n);
    if (!rc) rc = op_sample_init(gc, gs, e->cx, e->cy, comp_stack(e), e->nstack, e->p.ic, e->p.ic_params, e->p.discrete,
                                 e->p.discrete ? e->p.g_disc : e->p.g_fine, e->disc, w, e->st, comp_fmt(e));
    if (!rc) rc = get(out + CMM_SAMPLE_W * n, w, n);
    if (!rc && scalar_ic >= 0)
        rc = op_sample_init(gc, gs, e->cx, e->cy, comp_stack(e), e->nstack, scalar_ic, e->p.ic_params, 0, gs, nullptr, th,
                            e->st, comp_fmt(e));
    if (!rc && scalar_ic >= 0) rc = get(out + CMM_SAMPLE_THETA * n, th, n);
    // spectral derivatives of the sampled field (the psi-solve workspace is free between steps)
    if (!rc) rc = op_spectral_from_real(gs, w, 2, h, nullptr, 0, e->st);
    if (!rc) rc = get(out + CMM_SAMPLE_LAP_W * n, h, n);
    if (!rc) rc = op_spectral_from_real(gs, w, 0, h, nullptr, 0, e->st);
    if (!rc) rc = get(out + CMM_SAMPLE_GRADX_W * n, h + n, 2 * n);
    if (!rc) rc = op_spectral_from_real(gs, w, 1, h, nullptr, 0, e->st);
    if (!rc) rc = get(out + CMM_SAMPLE_PSI * n, h, 3 * n);
    if (!rc) rc = hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
    if (rc) return rc;
    for (size_t i = 0; i < n; ++i) {
        out[CMM_SAMPLE_CHIX * n + i] = hxy[2 * i];
        out[CMM_SAMPLE_CHIY * n + i] = hxy[2 * i + 1];
    }
    if (scalar_ic < 0) std::fill(out + CMM_SAMPLE_THETA * n, out + (CMM_SAMPLE_THETA + 1) * n, 0.0);
    return CMM_OK;
}

int cmm_euler_vp_sample(cmm_euler *e, cmm_grid gs, double *out) {
    if (!e || !out || !grid_ok(gs)) { set_error("cmm_euler_vp_sample: bad argument"); return CMM_ERR_ARG; }
    if (e->p.sim != CMM_SIM_VLASOV_POISSON_1D) {
        set_error("cmm_euler_vp_sample: not a Vlasov-Poisson engine");
        return CMM_ERR_ARG;
    }
    const cmm_grid &gc = e->p.g_coarse;
    if (gs.x0 != gc.x0 || gs.x1 != gc.x1 || gs.y0 != gc.y0 || gs.y1 != gc.y1) {
        set_error("cmm_euler_vp_sample: the sample grid must have the map's bounds");
        return CMM_ERR_ARG;
    }
    // the per-x Fourier solve wants a square power-of-two grid
    if (gs.nx != gs.ny) { set_error("cmm_euler_vp_sample: the sample grid must be square"); return CMM_ERR_ARG; }
    const size_t n = (size_t)gs.nx * gs.ny, nx = gs.nx;
    const size_t wbytes = op_vp_efield_workspace(gs);
    // device buffer: positions (2n), f (n), phi / phi' (3 nx), the field solve's workspace
    const size_t need = 3 * n + 3 * nx + (wbytes + sizeof(double) - 1) / sizeof(double);
    if (e->sbuf_n * 8 < need) {  // the buffer holds 8 sbuf_n doubles
        if (e->sbuf) (void)hipFree(e->sbuf);
        e->sbuf = nullptr;
        e->sbuf_n = 0;
        if (int rc = dalloc(&e->sbuf, 8 * ((need + 7) / 8))) return rc;
        e->sbuf_n = (need + 7) / 8;
    }
    if (int rc = sync_planes(e)) return rc;
    double *xy = e->sbuf, *f = xy + 2 * n, *phi3 = f + n, *work = phi3 + 3 * nx;
    std::vector<double> hxy(2 * n);
    auto get = [&](double *dst, const double *src, size_t cnt) {
        return hip_check(hipMemcpyAsync(dst, src, sizeof(double) * cnt, hipMemcpyDeviceToHost, e->st), "get");
    };
    int rc = op_compose_positions(gc, gs, e->cx, e->cy, comp_stack(e), e->nstack, xy, e->st, comp_fmt(e));
    if (!rc) rc = get(hxy.data(), xy, 2 * n);
    if (!rc) rc = op_sample_init(gc, gs, e->cx, e->cy, comp_stack(e), e->nstack, e->p.ic, e->p.ic_params, e->p.discrete,
                                 e->p.discrete ? e->p.g_disc : e->p.g_fine, e->disc, f, e->st, comp_fmt(e));
    if (!rc) rc = get(out + 2 * n, f, n);
    if (!rc) rc = op_vp_efield(gs, f, phi3, work, wbytes, e->st);
    if (!rc) rc = get(out + 3 * n, phi3 + 2 * nx, nx);
    if (!rc) rc = hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
    if (rc) return rc;
    for (size_t i = 0; i < n; ++i) {
        out[i] = hxy[2 * i];
        out[n + i] = hxy[2 * i + 1];
    }
    // Compute_Mass, Compute_Kinetic_Energy, Compute_Momentum, Compute_Potential_Energy on the sample
    const DGrid d = make_dgrid(gs);
    const double *fv = out + 2 * n, *E = out + 3 * n;
    double mass = 0, kin = 0, mom = 0, pot = 0;
    for (int iy = 0; iy < gs.ny; ++iy) {
        const double v = gs.y0 + iy * d.hy;
        for (int ix = 0; ix < gs.nx; ++ix) {
            const double a = fv[(size_t)iy * nx + ix];
            mass += a;
            kin += v * v * a;
            mom += v * a;
        }
    }
    for (size_t ix = 0; ix < nx; ++ix) pot += E[ix] * E[ix];
    double *m = out + 3 * n + nx;
    m[1] = 0.5 * d.hy * d.hx * kin;
    m[2] = 0.5 * d.hx * pot;
    m[0] = m[1] + m[2];
    m[3] = d.hy * d.hx * mass;
    m[4] = d.hy * d.hx * mom;
    // the modes come from the engine's psi, as in the conservation diagnostics
    double diag[9];
    if (int rc2 = cmm_euler_vp_diagnostics(e, diag)) return rc2;
    for (int k = 0; k < 4; ++k) m[5 + k] = diag[5 + k];
    return CMM_OK;
}

int cmm_euler_zoom(cmm_euler *e, cmm_grid gz, const int *ics, int n_ics, int want_psi, double *out) {
    if (!e || !out || gz.nx < 1 || gz.ny < 1 || n_ics < 0 || (n_ics > 0 && !ics) || !(gz.x1 > gz.x0) ||
        !(gz.y1 > gz.y0)) {
        set_error("cmm_euler_zoom: bad argument");
        return CMM_ERR_ARG;
    }
    const cmm_grid &gc = e->p.g_coarse;
    const size_t n = (size_t)gz.nx * gz.ny;
    if (e->sbuf_n < n) {
        if (e->sbuf) (void)hipFree(e->sbuf);
        e->sbuf = nullptr;
        e->sbuf_n = 0;
        if (int rc = dalloc(&e->sbuf, 8 * n)) return rc;
        e->sbuf_n = n;
    }
    if (int rc = sync_planes(e)) return rc;
    double *xy = e->sbuf, *f = xy + 2 * n;
    std::vector<double> hxy(2 * n);
    auto get = [&](double *dst, const double *src, size_t cnt) {
        if (int rc = hip_check(hipMemcpyAsync(dst, src, sizeof(double) * cnt, hipMemcpyDeviceToHost, e->st), "get"))
            return rc;
        return hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
    };
    int rc = op_compose_positions(gc, gz, e->cx, e->cy, comp_stack(e), e->nstack, xy, e->st, comp_fmt(e));
    if (!rc) rc = get(hxy.data(), xy, 2 * n);
    for (int k = 0; !rc && k < n_ics; ++k) {
        const bool own = ics[k] < 0;
        const int disc = own ? e->p.discrete : 0;
        rc = op_sample_init(gc, gz, e->cx, e->cy, comp_stack(e), e->nstack, own ? e->p.ic : ics[k], e->p.ic_params, disc,
                            disc ? e->p.g_disc : gz, disc ? e->disc : nullptr, f, e->st, comp_fmt(e));
        if (!rc) rc = get(out + (2 + k) * n, f, n);
    }
    if (!rc && want_psi) {
        rc = op_sample_hermite(e->p.g_psi, e->lvl[0], gz, f, e->st);
        if (!rc) rc = get(out + (2 + n_ics) * n, f, n);
    }
    if (rc) return rc;
    for (size_t i = 0; i < n; ++i) {
        out[i] = hxy[2 * i];
        out[n + i] = hxy[2 * i + 1];
    }
    return CMM_OK;
}

int cmm_euler_forward_positions(cmm_euler *e, cmm_grid gt, double *out) {
    if (!e || !out || !e->fx || gt.nx < 1 || gt.ny < 1 || !(gt.x1 > gt.x0) || !(gt.y1 > gt.y0)) {
        set_error("cmm_euler_forward_positions: bad argument (needs a forward map engine)");
        return CMM_ERR_ARG;
    }
    const size_t n = (size_t)gt.nx * gt.ny;
    if (e->sbuf_n < n) {
        if (e->sbuf) (void)hipFree(e->sbuf);
        e->sbuf = nullptr;
        e->sbuf_n = 0;
        if (int rc = dalloc(&e->sbuf, 8 * n)) return rc;
        e->sbuf_n = n;
    }
    std::vector<double> hxy(2 * n);
    int rc = op_forward_positions(e->p.g_coarse, gt, e->fx, e->fy, e->fstack, e->nstack, e->sbuf, e->st);
    if (!rc) rc = hip_check(hipMemcpyAsync(hxy.data(), e->sbuf, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, e->st), "get");
    if (!rc) rc = hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
    if (rc) return rc;
    for (size_t i = 0; i < n; ++i) {
        out[i] = hxy[2 * i];
        out[n + i] = hxy[2 * i + 1];
    }
    return CMM_OK;
}

int cmm_euler_forward_points(cmm_euler *e, cmm_grid gb, int n, const double *pos, int first_map, int apply_current,
                             double *out) {
    if (!e || !e->fx || n < 0 || (n > 0 && (!pos || !out))) {
        set_error("cmm_euler_forward_points: bad argument (needs a forward map engine)");
        return CMM_ERR_ARG;
    }
    if (n == 0) return CMM_OK;
    double *d = nullptr;
    if (int rc = dalloc(&d, 4 * (size_t)n)) return rc;
    int rc = hip_check(hipMemcpyAsync(d, pos, sizeof(double) * 2 * n, hipMemcpyHostToDevice, e->st), "put");
    if (!rc)
        rc = op_forward_points(e->p.g_coarse, gb, e->fstack, first_map, e->nstack, e->fx, e->fy, apply_current, d,
                               d + 2 * (size_t)n, n, e->st);
    if (!rc) rc = hip_check(hipMemcpyAsync(out, d + 2 * (size_t)n, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, e->st), "get");
    if (!rc) rc = hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
    (void)hipFree(d);
    return rc;
}

int cmm_euler_state(cmm_euler *e, int *steps, double *t, int *stack_size) {
    if (!e) { set_error("cmm_euler_state: null engine"); return CMM_ERR_ARG; }
    if (steps) *steps = e->loop;
    if (t) *t = e->tv.empty() ? 0.0 : e->tv[e->loop + e->p.l_order - 1];
    if (stack_size) *stack_size = e->nstack;
    return CMM_OK;
}

int cmm_euler_profile(cmm_euler *e, int enable) {
    if (!e) { set_error("cmm_euler_profile: null engine"); return CMM_ERR_ARG; }
    if (enable < 0) { set_error("cmm_euler_profile: negative sampling period"); return CMM_ERR_ARG; }
    for (int s = 0; s < cmm_euler::kProfSets; ++s) prof_drain(e, s);
    e->prof_every = enable;
    // the next step decides whether it records (step_queue_a); until then calls outside a step
    // record as the setting says
    e->prof = enable != 0;
    for (int i = 0; i < CMM_PROF_SLOTS; ++i) { e->ms[i] = 0; e->cnt[i] = 0; }
    return CMM_OK;
}

int cmm_euler_profile_read(cmm_euler *e, double *ms, int *cnt) {
    if (!e) { set_error("cmm_euler_profile_read: null engine"); return CMM_ERR_ARG; }
    for (int s = 0; s < cmm_euler::kProfSets; ++s) prof_drain(e, s);
    for (int i = 0; i < CMM_PROF_SLOTS; ++i) {
        if (ms) ms[i] = e->ms[i];
        if (cnt) cnt[i] = e->cnt[i];
    }
    return CMM_OK;
}

}  // extern "C"
