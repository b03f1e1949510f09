/*
 * C-ABI of the MI355X CMM engine (libcmm.so).
 *
 * Plain C: raw device pointers (allocated by the caller, e.g. hipMalloc or torch), sizes, and a
 * hipStream_t passed as void*.  Every call returns 0 on success or a negative CMM_ERR_* code;
 * cmm_last_error() gives the message of the calling thread's last failure.  Calls that need
 * scratch memory take (work, work_bytes); with work == NULL the library uses a per-device buffer
 * it allocates and keeps (calls on one device then share it and must not overlap).
 *
 * Each entry point replaces a host operator of the reference (Cuda-CMM/src/...) as cited.
 * Layouts: include/cmm_types.h.
 */
#ifndef CMM_H
#define CMM_H

#include <stddef.h>
#include <stdint.h>
#include "cmm_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- library ----------------------------------------------------------------------------- */
int cmm_version(void);
const char *cmm_last_error(void);
/* Writes "name, arch, CUs, LDS per block, HBM bytes" of the current device into buf. */
int cmm_device_info(char *buf, int len);

/* ---- coefficients (host only) --------------------------------------------------------------
 * Lagrange-in-time weights and footpoint stencil weights of one advection step; replaces the
 * host half of advect_using_stream_hermite (Cuda-CMM/src/simulation/cmm-simulation-host.cu:64-90)
 * and get_L_coefficient (Cuda-CMM/src/numerical/cmm-timestep.cu:474-484).  t and dt are the
 * time vectors of the driver, loop_ctr the step counter (Cuda-CMM/src/simulation/cmm-euler2d.cu:709-767). */
double cmm_lagrange_coeff(const double *t, double t_next, int loop_ctr, int i_point, int l_order);
int cmm_advect_coeffs_fill(const double *t, const double *dt, int loop_ctr, int l_order, int time_integration_num,
                           int map_update_order_num, double eps, int direction, cmm_advect_coeffs *c);

/* ---- map ----------------------------------------------------------------------------------- */
/* identity map; replaces k_init_diffeo (Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:27-41) */
int cmm_map_init_identity(cmm_grid g, double *chix, double *chiy, void *stream);

/* Advection of the Hermite map through psi; replaces advect_using_stream_hermite +
 * k_advect_using_stream_hermite (Cuda-CMM/src/simulation/cmm-simulation-host.cu:64-104,
 * Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:207-298).  c->direction -1 advects a backward
 * map (trace back, then the old map); +1 a forward map (the old map, then forward in time, with
 * the forward schemes of Cuda-CMM/src/numerical/cmm-timestep.cu; the workspace is not used).
 * psi_levels is a host array of
 * c->l_order device pointers (newest first, 4*Npsi doubles each).  chix_new/chiy_new must not
 * alias chix/chiy.
 * work may be NULL: every footpoint evaluation then interpolates each psi level and weights the
 * velocities (the reference's order of operations).  With work_bytes >=
 * cmm_advect_workspace_bytes(g_map, g_psi, c) (> 0 for l_order >= 2 and schemes 10, 21, 30, 40, 41)
 * the call first forms the Lagrange-weighted psi combinations (and copies of level 0 and of the
 * map) in work, once, node-interleaved, and interpolates those: one gather per evaluation instead
 * of one per level, in 16-byte loads; equal up to rounding. */
size_t cmm_advect_workspace_bytes(cmm_grid g_map, cmm_grid g_psi, const cmm_advect_coeffs *c);
int cmm_map_advect(cmm_grid g_map, cmm_grid g_psi, const double *chix, const double *chiy,
                   const double *const *psi_levels, const cmm_advect_coeffs *c, double *chix_new,
                   double *chiy_new, void *work, size_t work_bytes, void *stream);

/* max |1 - det grad X| over the cell centres of g_check, returned to the host (synchronises the
 * stream); replaces incompressibility_check (Cuda-CMM/src/simulation/cmm-simulation-host.cu:44-50). */
int cmm_incomp_error(cmm_grid g_map, cmm_grid g_check, const double *chix, const double *chiy, double *out,
                     void *stream);
/* Invertibility of the backward and forward maps: max over the cell centres of g_check of
 * |x - X_b(X_f(x))|; replaces invertibility_check (Cuda-CMM/src/simulation/cmm-simulation-host.cu:52-60)
 * with k_invertibility_check (Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:177-198). */
int cmm_invert_error(cmm_grid g_back, cmm_grid g_fwd, cmm_grid g_check, const double *chix_b, const double *chiy_b,
                     const double *chix_f, const double *chiy_f, double *out, void *stream);

/* ---- stack composition and initial condition ---------------------------------------------------
 * The stack is device resident: map m at stack + 8*Nc*m (ChiX planes then ChiY planes), m = 0
 * oldest.  Positions X_1(...X_M(X_cur(x))) on g_target; replaces apply_map_stack
 * (Cuda-CMM/src/simulation/cmm-simulation-host.cu:112-152). xy: 2*Nt interleaved. */
int cmm_compose_positions(cmm_grid g_map, cmm_grid g_target, const double *chix, const double *chiy,
                          const double *stack, int n_stack, double *xy, void *stream);
/* Stacked maps repacked as node records, the form the engine keeps its stack in: a map evaluation
 * then gathers four records (two to four cache lines) instead of 32 values from eight planes.  No
 * counterpart in the reference (its MapStack is plane-major FP64, the layout above).
 *   CMM_STACK_RECORDS_FP64:  64 bytes per node, ChiX planes then ChiY planes; map m at byte 64*Nc*m.
 *   CMM_STACK_RECORDS_SPLIT: 48 bytes per node, X and Y values in FP64, then dX/dx, dX/dy, d2X/dxdy,
 *     dY/dx, dY/dy, d2Y/dxdy in FP32 and 8 unused bytes; map m at byte 48*Nc*m (split-precision
 *     engines).  Positions differ from the FP64 stack's by the FP32 rounding of the derivative
 *     planes: per layer at most h/4 (|dX/dx| + |dX/dy|) 2^-24 + h^2/16 |d2X/dxdy| 2^-24 (likewise Y),
 *     carried through the later layers by their gradients. */
#define CMM_STACK_RECORDS_FP64 1
#define CMM_STACK_RECORDS_SPLIT 2
int cmm_stack_pack(cmm_grid g_map, const double *stack, int n_maps, int format, void *packed, void *stream);
/* cmm_compose_positions through a stack packed by cmm_stack_pack; with CMM_STACK_RECORDS_FP64 the
 * positions equal cmm_compose_positions' bit for bit. */
int cmm_compose_positions_packed(cmm_grid g_map, cmm_grid g_target, const double *chix, const double *chiy,
                                 const void *packed, int n_stack, int format, double *xy, void *stream);
/* Forward positions X_cur(X_{M-1}(...X_0(x))) on g_target through a stack of forward maps (same
 * layout), apply_map_stack with direction 1 (Cuda-CMM/src/simulation/cmm-simulation-host.cu:128-148).
 * xy: 2*Nt interleaved. */
int cmm_forward_positions(cmm_grid g_map, cmm_grid g_target, const double *chix, const double *chiy,
                          const double *stack, int n_stack, double *xy, void *stream);
/* Points through forward stacked maps first_map..n_stack-1 and then, if apply_current, the current
 * forward map: k_h_sample_points_map (Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:97-118) as
 * apply_map_stack_points applies it (Cuda-CMM/src/simulation/cmm-simulation-host.cu:213-291).  A point
 * strictly inside g_bounds' box is mapped and wrapped (x - floor(x / Lx) Lx), any other becomes NaN.
 * in/out: 2*n interleaved device arrays (distinct). */
int cmm_forward_points(cmm_grid g_map, cmm_grid g_bounds, const double *chix, const double *chiy, const double *stack,
                       int n_stack, int first_map, int apply_current, int n, const double *in, double *out,
                       void *stream);
/* Initial condition at the composed positions; k_h_sample_from_init
 * (Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:342-373).  discrete != 0 samples the Hermite
 * field disc_herm on g_disc instead of the analytic condition ic. */
int cmm_sample_init(cmm_grid g_map, cmm_grid g_target, const double *chix, const double *chiy, const double *stack,
                    int n_stack, int ic, const double *ic_params, int discrete, cmm_grid g_disc,
                    const double *disc_herm, double *out, void *stream);
/* New fine vorticity Hermite after a remap: the initial condition at the composed positions
 * (one sampling pass into the workspace), then R2C + /N + fourier_hermite; replaces
 * translate_initial_condition_through_map_stack (Cuda-CMM/src/simulation/cmm-simulation-host.cu:299-327).
 * The workspace holds the spectral buffers and 8 bytes per fine point. */
size_t cmm_translate_workspace_bytes(cmm_grid g_fine);
int cmm_translate_ic(cmm_grid g_map, cmm_grid g_fine, const double *chix, const double *chiy, const double *stack,
                     int n_stack, int ic, const double *ic_params, int discrete, cmm_grid g_disc,
                     const double *disc_herm, double *w_herm, void *work, size_t work_bytes, void *stream);

/* ---- spectral ------------------------------------------------------------------------------ */
/* Hermite planes of a real periodic field (R2C, /N, fourier_hermite); replaces the
 * cufftExecD2Z + k_normalize_h + fourier_hermite sequence (Cuda-CMM/src/numerical/cmm-fft.cu:398-429,
 * Cuda-CMM/src/simulation/cmm-euler2d.cu:335-339). Powers of two, 8..4096. */
size_t cmm_hermite_workspace_bytes(cmm_grid g);
int cmm_hermite_from_real(cmm_grid g, const double *real, double *herm, void *work, size_t work_bytes,
                          void *stream);
/* Hermite planes from a half spectrum given column-major (spec[kx*ny + ky], kx <= nx/2, complex
 * interleaved), no scaling: the C2R half of fourier_hermite. */
int cmm_hermite_from_spectrum(cmm_grid g, const double *spec_cols, double *herm, void *work, size_t work_bytes,
                              void *stream);
/* Hermite planes [g, g_x, g_y, g_xy] of g = F (op 0), of the inverse Laplacian -F/|k|^2 with the
 * mean removed (op 1) or of the Laplacian -|k|^2 F (op 2) of a real periodic field F; replaces
 * fourier_hermite, i_laplacian, laplacian, grad_x and grad_y on a sample grid
 * (Cuda-CMM/src/numerical/cmm-fft.cu:398-474).  Powers of two, 8..4096. */
size_t cmm_spectral_workspace_bytes(cmm_grid g);
int cmm_spectral_from_real(cmm_grid g, const double *real, int op, double *herm, void *work, size_t work_bytes,
                           void *stream);
/* psi Hermite from the map and the fine vorticity: sample omega(X(x)) on g_vort (mollifier 0/4/8),
 * R2C, /N, move to the psi spectrum, circular cut-off at freq_cut, -1/|k|^2, fourier_hermite;
 * replaces evaluate_stream_hermite (Cuda-CMM/src/simulation/cmm-simulation-host.cu:333-375).
 * Sign convention of the reference: Laplace(psi) = omega, velocity (dpsi/dy, -dpsi/dx).
 * The workspace holds the transposed half spectrum, two psi-grid column buffers and the sampled
 * vorticity field (8 bytes per g_vort point). */
size_t cmm_psi_workspace_bytes(cmm_grid g_vort, cmm_grid g_psi, double freq_cut);
int cmm_psi_from_vort(cmm_grid g_map, cmm_grid g_vort, cmm_grid g_fine, cmm_grid g_psi, const double *chix,
                      const double *chiy, const double *w_herm, int molly_stencil, double freq_cut, double *psi_out,
                      void *work, size_t work_bytes, void *stream);
/* cmm_psi_from_vort without any allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same kernels and results as cmm_psi_from_vort.  Its workspace and the FFT
 * twiddle tables live in a static device array the library holds on every GPU it runs on until the
 * process ends (192 MiB); a call whose workspace, cmm_psi_workspace_bytes(g_vort, g_psi, freq_cut),
 * exceeds CMM_PSI_NOALLOC_MAX_WORKSPACE returns CMM_ERR_ARG without launching anything (the
 * 4096^2 vorticity grid fits for freq_cut up to 511).  The first call on a device looks the
 * array's address up and copies the twiddle tables with a synchronous copy: make it outside graph
 * capture.  All calls on one GPU share the array: calls that overlap on one GPU (on different
 * streams or host threads) race. */
#define CMM_PSI_NOALLOC_MAX_WORKSPACE ((size_t)201195520)
int cmm_psi_from_vort_noalloc(cmm_grid g_map, cmm_grid g_vort, cmm_grid g_fine, cmm_grid g_psi, const double *chix,
                              const double *chiy, const double *w_herm, int molly_stencil, double freq_cut,
                              double *psi_out, void *stream);
/* omega(X(x)) on g_target without transform (coarse vorticity output); k_apply_map_and_sample_from_hermite
 * with padd_inline_fft = false (Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:402-461). */
int cmm_sample_vort(cmm_grid g_map, cmm_grid g_target, cmm_grid g_fine, const double *chix, const double *chiy,
                    const double *w_herm, int molly_stencil, double *out, void *stream);

/* psi(x, v) Hermite of the 1D-1V Vlasov-Poisson flow from the distribution function's fine
 * Hermite field: sample f(X(x)) on g_vort (mollifier 0/4/8), integrate over v, solve the 1-D
 * Poisson equation spectrally, zero-pad to g_psi, add the periodified -v^2/2 part, Hermite planes;
 * replaces evaluate_potential_from_density_hermite with calc_phi, integral_v, solve_1D_laplace,
 * zero_padding_1D, compute_periodification_function and k_assemble_psi_periodified
 * (Cuda-CMM/src/simulation/cmm-simulation-host.cu:385-471, :478-494, :534-571,
 * Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:504-512, :570-586, Cuda-CMM/src/numerical/cmm-fft.cu:367-391).
 * g_vort must not be finer than g_psi and g_psi must be square (the reference's errors 1013 and
 * 20230904); x and v bounds are g_psi's (x0, x1) and (y0, y1). */
size_t cmm_vp_psi_workspace_bytes(cmm_grid g_vort, cmm_grid g_psi);
int cmm_vp_psi_from_f(cmm_grid g_map, cmm_grid g_vort, cmm_grid g_fine, cmm_grid g_psi, const double *chix,
                      const double *chiy, const double *f_herm, int molly_stencil, double *psi_out, void *work,
                      size_t work_bytes, void *stream);

/* ---- particles ----------------------------------------------------------------------------- */
/* Fluid particles, one step; replaces particles_advect + Particle_advect
 * (Cuda-CMM/src/numerical/cmm-particles.cu:63-87, :119-276).  psi_levels: host array of l_order
 * device pointers, level 0 at t_{n+1}, level 1 at t_n; Lp12: Lagrange weights at t_n + dt/2. */
int cmm_particles_advect(cmm_grid g_psi, const double *const *psi_levels, int l_order, const double *Lp12, double dt,
                         int scheme, int n, double *pos, void *stream);

/* ---- split-precision form (no counterpart in the reference) ----------------------------------
 * The map is stored as the identity plus a displacement D = X - identity, one 48-byte record of
 * 12 floats per node: (Dx hi, Dx lo, Dy hi, Dy lo), the value planes as FP32 pairs hi + lo;
 * (dX/dx - 1, dX/dy, dY/dx, dY/dy - 1); (d2X/dxdy, d2Y/dxdy, 0, 0).  The gathered Hermite fields
 * (the Lagrange-combined psi fields, the fine vorticity) are FP32 node records
 * (f, fx, fy, fxy); the combined psi fields keep their values to 2^-31 relative or better, since the
 * value of a stream function of order one rounded to FP32 would move the velocity by
 * 2^-24 |psi| 6/h: an 8-bit code of f - (float)f takes the low byte of fxy, which keeps 15
 * explicit mantissa bits.  Arithmetic, the footpoint integration, the Jacobian check, the psi levels,
 * the FFTs and the Poisson divide stay FP64.  These operators are the split forms of
 * cmm_map_advect, cmm_incomp_error and cmm_psi_from_vort. */
enum { CMM_PRECISION_FP64 = 0, CMM_PRECISION_SPLIT = 1 };
/* plane-major FP64 map (4*N doubles per component) <-> split map (12*N floats) */
int cmm_map_to_split(cmm_grid g, const double *chix, const double *chiy, float *map_split, void *stream);
int cmm_map_from_split(cmm_grid g, const float *map_split, double *chix, double *chiy, void *stream);
/* FP32 node records (4*N floats) of a plane-major FP64 Hermite field */
int cmm_hermite_to_split(cmm_grid g, const double *herm, float *herm_split, void *stream);
/* Backward advection of the split map (schemes 10, 21, 30, 40, 41 with l_order >= 2); the
 * workspace (cmm_map_advect_split_workspace_bytes) holds the combined psi fields (one 16-byte
 * record per psi node each). */
size_t cmm_map_advect_split_workspace_bytes(cmm_grid g_psi, const cmm_advect_coeffs *c);
int cmm_map_advect_split(cmm_grid g_map, cmm_grid g_psi, const float *map_in, const double *const *psi_levels,
                         const cmm_advect_coeffs *c, float *map_out, void *work, size_t work_bytes, void *stream);
int cmm_incomp_error_split(cmm_grid g_map, cmm_grid g_check, const float *map_split, double *out, void *stream);
/* workspace as cmm_psi_workspace_bytes; vort_split: FP32 records of the fine vorticity */
int cmm_psi_from_vort_split(cmm_grid g_map, cmm_grid g_vort, cmm_grid g_fine, cmm_grid g_psi, const float *map_split,
                            const float *vort_split, int molly_stencil, double freq_cut, double *psi_out, void *work,
                            size_t work_bytes, void *stream);

/* ---- engine: the Euler time loop on the device ------------------------------------------------
 * Mirrors cuda_euler_2d (Cuda-CMM/src/simulation/cmm-euler2d.cu:45-960) for the backward map:
 * initial sampling, Lagrange warm-up, and per step advection, incompressibility check, remap with
 * a device-resident stack, psi update and particles. */
typedef struct cmm_euler cmm_euler;

typedef struct cmm_euler_params {
    cmm_grid g_coarse, g_fine, g_psi, g_vort;
    double dt;                      /* nominal step */
    int time_integration_num, map_update_order_num, l_order, lagrange_init_higher_order;
    double map_epsilon, incomp_threshold, freq_cut_psi;
    int molly_stencil, skip_remapping;
    int ic;                         /* CMM_IC_* */
    double ic_params[10];
    int discrete;                   /* initial condition from a discrete field */
    cmm_grid g_disc;
    const double *disc_real;        /* host real field on g_disc (NULL with CMM_IC_RANDOM_SPECTRUM) */
    uint64_t seed;                  /* CMM_IC_RANDOM_SPECTRUM seed */
    int stack_capacity;             /* maps the device stack holds */
    int sim;                        /* CMM_SIM_EULER_2D or CMM_SIM_VLASOV_POISSON_1D */
    int forward_map;                /* 1: also advect the forward map on g_coarse, with its own stack
                                       of stack_capacity maps (Cuda-CMM/src/simulation/cmm-euler2d.cu:156-159,
                                       386-397, 781-811) */
    int precision;                  /* CMM_PRECISION_FP64, or CMM_PRECISION_SPLIT: the split-precision
                                       form of the map, the combined psi fields and (Euler) the fine
                                       vorticity (backward map only; no counterpart in the reference) */
} cmm_euler_params;

/* The 1D-1V Vlasov-Poisson loop (cuda_vlasov_1d, Cuda-CMM/src/simulation/cmm-vlasov1d.cu:48-1148)
 * is the same loop with f in place of omega: its initial conditions are CMM_IC_LANDAU_DAMPING..
 * and its psi comes from cmm_vp_psi_from_f; the map lives on (x, v) = (x0..x1, y0..y1). */
enum { CMM_SIM_EULER_2D = 0, CMM_SIM_VLASOV_POISSON_1D = 1 };

typedef struct cmm_step_info {
    double t;                       /* time after the step */
    double incomp;                  /* E_inc after advection (before a remap) */
    int remapped;                   /* 1 if the map was pushed and reset */
    int stack_size;                 /* maps on the stack after the step */
    int saturated;                  /* 1 if a remap was needed but the stack was full */
    double incomp_forward;          /* forward map: E_inc on g_coarse (0 without a forward map) */
    double invertibility;           /* forward map: max |X_f(X_b(x)) - x| on g_coarse (0 without) */
} cmm_step_info;

enum {
    CMM_FIELD_CHIX = 0, CMM_FIELD_CHIY = 1, /* 4*Nc, current map */
    CMM_FIELD_PSI = 2,                      /* l_order*4*Npsi, newest level first */
    CMM_FIELD_W_FINE = 3,                   /* 4*Nf fine vorticity Hermite */
    CMM_FIELD_W_COARSE = 4,                 /* Nc, omega(X(x)) sampled on the coarse grid */
    CMM_FIELD_DISC = 5,                     /* 4*Ndisc discrete initial Hermite */
    CMM_FIELD_CHIX_F = 6, CMM_FIELD_CHIY_F = 7 /* 4*Nc, current forward map (forward_map only) */
};

int cmm_euler_create(const cmm_euler_params *p, cmm_euler **out);
int cmm_euler_destroy(cmm_euler *e);
/* Stream the engine runs on (hipStream_t as void*).  Particle sets advance on a second engine
 * stream, ordered against this one with events; cmm_euler_get_particles waits for them. */
void *cmm_euler_stream(cmm_euler *e);
/* Adds a fluid-particle set (positions copied from host) that moves from the first step; returns
 * its index in *set.  Same as cmm_euler_add_particle_set with init_time 0 and tau 0. */
int cmm_euler_add_particles(cmm_euler *e, int n, int scheme, const double *pos_host, int *set);
/* Adds a particle set with the reference's per-set settings (ParticlesAdvected: tau, init_time,
 * init_vel, time_integration; Cuda-CMM/src/ui/settings.h, parsed at Cuda-CMM/src/ui/settings.cu):
 * - the set moves in every step with t_{n+1} + 1e-5 dt >= init_time
 *   (particles_advect, Cuda-CMM/src/numerical/cmm-particles.cu:63-87);
 * - tau == 0: fluid particles (Particle_advect, :119-276);
 * - tau != 0: inertial particles x' = v, v' = -(v - u)/tau (Particle_advect_inertia, :280-551).
 *   Their velocity starts as the fluid velocity of psi after the warm-up (init_vel != 0) or zero,
 *   and is set again by that rule in the step that reaches init_time
 *   (Cuda-CMM/src/simulation/cmm-euler2d.cu:620-632, :883-895); vel_host (2n doubles, may be NULL)
 *   gives the velocities instead, as a restart reads them (readParticlesState,
 *   Cuda-CMM/src/ui/cmm-io.cu:659-688). */
int cmm_euler_add_particle_set(cmm_euler *e, int n, int scheme, double init_time, double tau, int init_vel,
                               const double *pos_host, const double *vel_host, int *set);
/* Initial state and Lagrange warm-up; warmup_incomp[8] receives the logged errors. */
int cmm_euler_init(cmm_euler *e, double *warmup_incomp, int *n_warmup);
/* Restart from a saved state instead of cmm_euler_init: readMapStack and the restart branch of the
 * time loop (Cuda-CMM/src/simulation/cmm-simulation-host.cu:1255-1302,
 * Cuda-CMM/src/simulation/cmm-euler2d.cu:634-684).  stack_host holds n_maps maps (8*Nc doubles each:
 * ChiX planes then ChiY planes, oldest first); the fine vorticity is the initial condition
 * translated through them; chix_host/chiy_host (4*Nc) become the active map; psi_old_host holds
 * l_order-1 psi levels (4*Npsi each) for levels 1.. ; level 0 is computed.  Times restart at
 * t_restart; there is no Lagrange warm-up. */
int cmm_euler_restart(cmm_euler *e, double t_restart, int n_maps, const double *stack_host, const double *chix_host,
                      const double *chiy_host, const double *psi_old_host);
/* Map m of the stack (0 = oldest) to host memory, 4*Nc doubles each (writeMapStack,
 * Cuda-CMM/src/ui/cmm-io.cu:611-635). */
int cmm_euler_get_stack_map(cmm_euler *e, int m, double *chix_host, double *chiy_host);
/* Forward-map part of a restart (forward_map only), after cmm_euler_restart: n_maps forward maps
 * (8*Nc doubles each, oldest first) and the active forward map (readMapStack with isForward,
 * Cuda-CMM/src/simulation/cmm-simulation-host.cu:1262-1291).  n_maps equals the backward count. */
int cmm_euler_restart_forward(cmm_euler *e, int n_maps, const double *stack_host, const double *chix_host,
                              const double *chiy_host);
/* Forward map m of the forward stack (0 = oldest), as cmm_euler_get_stack_map. */
int cmm_euler_get_forward_stack_map(cmm_euler *e, int m, double *chix_host, double *chiy_host);
/* Makes the next cmm_euler_step remap whatever E_inc is (the forwarded-particle start of
 * Cuda-CMM/src/simulation/cmm-euler2d.cu:804-814, which pushes both maps at the particles' start
 * time).  The stack limit still applies. */
int cmm_euler_request_remap(cmm_euler *e);
/* One step of length dt_now (the driver clips dt to save targets and the final time). */
int cmm_euler_step(cmm_euler *e, double dt_now, cmm_step_info *info);
/* Copies a field to host memory (size per CMM_FIELD_*). */
int cmm_euler_get_field(cmm_euler *e, int field, double *host_out);
int cmm_euler_get_particles(cmm_euler *e, int set, double *host_out);
/* The velocity output of a particle set (writeParticles' PartA_Vel_, Cuda-CMM/src/ui/cmm-io.cu:418-430):
 * an inertial set's velocities; for fluid particles (d psi/dx, d psi/dy) of the newest psi at the
 * positions, NaN outside the psi grid's bounds (k_h_sample_points_dxdy,
 * Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:138-155).  2n doubles. */
int cmm_euler_get_particle_velocities(cmm_euler *e, int set, double *host_out);
/* Conservation diagnostics of compute_conservation_targets
 * (Cuda-CMM/src/simulation/cmm-simulation-host.cu:615-698): out[0] energy 0.5 h^2 sum |grad psi|^2,
 * out[1] enstrophy and out[2] palinstrophy of the coarse vorticity, out[3] max |omega|. */
int cmm_euler_diagnostics(cmm_euler *e, double *out);
/* Vlasov-Poisson diagnostics of compute_conservation_targets (Cuda-CMM/src/simulation/cmm-simulation-host.cu:642-661,
 * Cuda-CMM/src/numerical/cmm-mesure.cu:97-202) on the coarse f sample and psi level 0: out[0] Etot,
 * out[1] Ekin = 1/2 hx hv sum f v^2, out[2] Epot = 1/2 hx sum (dpsi/dx)^2 over the first row,
 * out[3] Mass = hx hv sum f, out[4] Momentum = hx hv sum f v, out[5..8] |E_k|, k = 1..4, of the
 * first row of dpsi/dx (R2C / NX). */
int cmm_euler_vp_diagnostics(cmm_euler *e, double *out);
/* Fields through the whole map stack on a sample grid, the computational part of
 * sample_compute_and_write (Cuda-CMM/src/simulation/cmm-simulation-host.cu:706-893): host_out
 * receives 10 arrays of g_sample.nx*g_sample.ny doubles: [CMM_SAMPLE_CHIX, CMM_SAMPLE_CHIY] the
 * composed positions, CMM_SAMPLE_W the initial condition there, CMM_SAMPLE_THETA the passive
 * scalar scalar_ic there (zeros when scalar_ic < 0), CMM_SAMPLE_LAP_W, CMM_SAMPLE_GRADX_W,
 * CMM_SAMPLE_GRADY_W, CMM_SAMPLE_PSI (inverse Laplacian of W), CMM_SAMPLE_UX, CMM_SAMPLE_UY (its
 * x and y gradients).  g_sample: powers of two, 8..4096, same bounds as the map. */
enum {
    CMM_SAMPLE_CHIX = 0, CMM_SAMPLE_CHIY = 1, CMM_SAMPLE_W = 2, CMM_SAMPLE_THETA = 3, CMM_SAMPLE_LAP_W = 4,
    CMM_SAMPLE_GRADX_W = 5, CMM_SAMPLE_GRADY_W = 6, CMM_SAMPLE_PSI = 7, CMM_SAMPLE_UX = 8, CMM_SAMPLE_UY = 9,
    CMM_SAMPLE_FIELDS = 10
};
int cmm_euler_sample(cmm_euler *e, cmm_grid g_sample, int scalar_ic, double *host_out);
/* The Vlasov-Poisson sample outputs of sample_compute_and_write_vlasov
 * (Cuda-CMM/src/simulation/cmm-simulation-host.cu:898-1030) on g_sample (n = nx*ny points, nx ==
 * ny a power of two, 8..4096, the map's bounds): host_out receives [0, n) and [n, 2n) the positions
 * X and V through the current map and the whole stack, [2n, 3n) the distribution f0 at those
 * positions, [3n, 3n + nx) E(x) = dphi/dx of the sample's own potential (phi_xx = hv sum_v f - 1
 * in Fourier space, Cuda-CMM/src/simulation/cmm-simulation-host.cu:436-471; the column integral
 * starts from zero), then 9 values: Etot, Ekin = 1/2 hx hv sum f v^2, Epot = 1/2 hx sum E^2,
 * Mass = hx hv sum f, Momentum = hx hv sum f v (Cuda-CMM/src/numerical/cmm-mesure.cu:97-183), and
 * |E_k|, k = 1..4, of psi level 0 (Compute_kth_Mode, as cmm_euler_vp_diagnostics).  Vlasov-Poisson
 * engines only. */
int cmm_euler_vp_sample(cmm_euler *e, cmm_grid g_sample, double *host_out);
/* A zoom window of compute_zoom (Cuda-CMM/src/simulation/cmm-simulation-host.cu:1039-1227) on
 * g_zoom (any size >= 1, any bounds): host_out receives, each g_zoom.nx*g_zoom.ny doubles, the
 * positions X and Y through the current map and the whole stack (wrapped into the domain after
 * each layer), then for each k < n_ics the initial condition ics[k] at those positions (ics[k] < 0:
 * the engine's own initial condition, discrete or not; else a CMM_IC_* value), then, when
 * want_psi, the newest psi level evaluated at the zoom grid points themselves (k_h_sample). */
int cmm_euler_zoom(cmm_euler *e, cmm_grid g_zoom, const int *ics, int n_ics, int want_psi, double *host_out);
/* The forward map through the forward stack on the points of g_target (apply_map_stack with
 * direction 1, Cuda-CMM/src/simulation/cmm-simulation-host.cu:128-148): the oldest stacked forward
 * map first, then the newer ones, then the current forward map, each result wrapped into the
 * domain.  host_out: X positions (n) then Y positions (n), n = g_target.nx * g_target.ny, row-major. */
int cmm_euler_forward_positions(cmm_euler *e, cmm_grid g_target, double *host_out);
/* Forwarded particles (apply_map_stack_points, direction 1, Cuda-CMM/src/simulation/cmm-simulation-host.cu:213-291):
 * n points (x, y interleaved) through the forward stacked maps first_map .. stack_size-1 and then,
 * if apply_current, the current forward map; with nothing to apply the points are copied.  Each
 * application takes a point strictly inside the bounds of g_bounds through the map and wraps the
 * result by the domain lengths (x - floor(x / Lx) Lx); any other point becomes NaN. */
int cmm_euler_forward_points(cmm_euler *e, cmm_grid g_bounds, int n, const double *pos_host, int first_map,
                             int apply_current, double *host_out);
/* Steps taken, current time, stack size. */
int cmm_euler_state(cmm_euler *e, int *steps, double *t, int *stack_size);
/* Per-kernel timing with events on the engine stream: enable (clears the counters; 0 turns the
 * events off, n >= 1 records them on every n-th step, since each event adds a packet to the
 * stream), then read the accumulated milliseconds and launch counts of CMM_PROF_SLOTS slots:
 * [map advection, incompressibility check, psi sample+row R2C, psi column stage, psi row C2R,
 *  particles, remap (translate + push + the psi recomputed after it), Lagrange-weighted psi
 *  combination before the advection]. */
#define CMM_PROF_SLOTS 8
int cmm_euler_profile(cmm_euler *e, int enable);
int cmm_euler_profile_read(cmm_euler *e, double *ms_out, int *count_out);


/* ---- ensembles (no counterpart in the reference) ------------------------------------------------
 * Engines of any grid sizes stepped together: each step queues every engine's advection, check,
 * E_inc copy and psi solve, then waits once and finishes each engine (remap when E_inc exceeds its
 * threshold, in stream order as cmm_euler_step does).  With use_graphs, a step is one captured
 * graph (one branch per engine stream, holding that engine's whole step), kept per step phase
 * and replayed; the host waits for each engine's E_inc copy inside its branch, while the psi
 * solves may still run, and every later use of an engine (its calls, the next step) is ordered
 * after its own branch on the device.  Steps of another length than the engines' nominal dt (a
 * clipped last step) and each engine's first step run without a
 * graph.  Engines must be initialised, without particles, a forward map or profiling; results equal
 * cmm_euler_step's. */
typedef struct cmm_batch cmm_batch;
int cmm_batch_create(cmm_euler *const *engines, int n, int use_graphs, cmm_batch **out);
int cmm_batch_destroy(cmm_batch *b);
/* one step of length dt for every engine; infos: n entries (or NULL) */
int cmm_batch_step(cmm_batch *b, double dt, cmm_step_info *infos);
int cmm_batch_stats(cmm_batch *b, int *graphs_captured, int *graph_launches, int *eager_steps);
/* Grouped steps (on by default): when every engine of the batch is an FP64 Vlasov-Poisson engine in
 * its steady state whose scheme takes the separable advection (the ensembles of
 * Cuda-CMM/src/simulation/cmm-vlasov1d.cu:48, one instance per engine), a step runs each kind of
 * launch of all engines as one launch over the concatenated blocks of every instance, with the same
 * per-instance results as their own launches; other steps run as above.  Off by default;
 * cmm_batch_autotune chooses. */
int cmm_batch_set_grouping(cmm_batch *b, int on);
int cmm_batch_grouped_steps(cmm_batch *b, int *grouped_steps);
/* Autotuner of the batch's launch form: grouped launches (one workgroup per block of work,
 * persistent ones of 2048 workgroups, or grouped launches per lane of engines, 4 lanes of about
 * equal work on their own streams), captured graphs (when the batch was created with them) and
 * eager queues each run steps_per_mode real steps of length dt (the run advances; graphs first
 * capture theirs in 12 untimed steps); the form with the shortest median step stays (ms_out: that
 * median).  Every form gives the same results. */
enum { CMM_BATCH_EAGER = 0, CMM_BATCH_GRAPHS = 1, CMM_BATCH_GROUPED = 2, CMM_BATCH_GROUPED_PERSISTENT = 3,
       CMM_BATCH_GROUPED_LANES = 4 };
int cmm_batch_autotune(cmm_batch *b, double dt, int steps_per_mode, int *mode_out, double *ms_out);
/* the launch form set directly (CMM_BATCH_*; graphs only for a batch created with them) */
int cmm_batch_set_mode(cmm_batch *b, int mode);
/* Autotuner: times the incompressibility check's tile edge (4, 2, 1 dividing the fine/coarse
 * ratio) and block cap on the engine's current map and keeps the fastest for its later steps. */
int cmm_euler_autotune(cmm_euler *e, int *tile_out, int *max_blocks_out, double *ms_out);
/* Sets the check's tile edge (1 to 4; 0 or one that does not divide the fine/coarse ratio: the
 * largest of 4, 3, 2, 1 that does) and block cap, as cmm_euler_autotune chooses them (to replay a
 * tuned launch shape). E_inc does not depend on the cap; tile edges round the determinant near 1
 * differently (about 1e-14). Grouped batch steps need the same tile edge in every engine; a batch
 * whose engines differ steps in its other forms. */
int cmm_euler_set_check_tile(cmm_euler *e, int tile, int max_blocks);
/* Whether a step's incompressibility check (Cuda-CMM/src/simulation/cmm-euler2d.cu:796) runs in the
 * pass of the psi solve's vorticity sampler instead of a kernel of its own: possible when the check
 * grid is the vorticity grid, refining the map grid by a multiple of 4 over the same domain, in an
 * Euler run without a forward map, and then on by default in the FP64 form (off in the split form,
 * whose sampler runs faster without it).  The check points, their arithmetic and the remap decision
 * are the same; the map is read once for both. */
int cmm_euler_set_check_fusion(cmm_euler *e, int on);
int cmm_euler_get_check_fusion(cmm_euler *e, int *on);
/* Whether a step's psi solve also writes the next step's Lagrange-combined psi fields (the
 * combination inside advect_using_stream_hermite, Cuda-CMM/src/simulation/cmm-simulation-host.cu:64-104)
 * from its last transform, for a next step of the same length: the next advection then reads them
 * instead of combining the levels again, when its weights, levels and scheme are those.  The sums
 * are the same.  On by default (Euler runs). */
int cmm_euler_set_psi_precombine(cmm_euler *e, int on);
int cmm_euler_get_psi_precombine(cmm_euler *e, int *on);


/* ---- learned psi extrapolation (no counterpart in the reference) --------------------------------
 * A 3x3 convolutional network with 32 channels per layer (the 3 input and 2 output channels
 * zero-padded to 32), BF16 weights and activations in NHWC layout (uint16_t bit patterns), FP32
 * accumulation on the matrix cores, periodic padding; widths are multiples of 32.  Weights of a
 * layer: W[co][tap][ci], 32 x 9 x 32, tap = 3 (dy + 1) + (dx + 1); bias: 32 floats.
 * mode 0: ReLU, bf16 out; 1: linear, bf16 out; 2: linear, float out (32 channels per pixel). */
int cmm_cnn_forward(int n_img, int H, int W, const uint16_t *in, const uint16_t *w, const float *bias, int mode,
                    void *out, void *stream);
/* gradient with respect to the layer input: wt holds the transposed, tap-flipped weights
 * (cmm_cnn_weights_bf16) */
int cmm_cnn_backward_data(int n_img, int H, int W, const uint16_t *g, const uint16_t *wt, uint16_t *g_in, void *stream);
/* cmm_cnn_backward_data with the ReLU derivative of the layer below in the same pass: g_in is zero
 * where act (the layer input, a ReLU output, same layout) is not positive, as cmm_cnn_relu_mask
 * after cmm_cnn_backward_data leaves it */
int cmm_cnn_backward_data_relu(int n_img, int H, int W, const uint16_t *g, const uint16_t *wt, const uint16_t *act,
                               uint16_t *g_in, void *stream);
/* dw (32 x 9 x 32 floats) and db (32 floats) accumulate the weight and bias gradients */
int cmm_cnn_backward_weights(int n_img, int H, int W, const uint16_t *g, const uint16_t *in, float *dw, float *db,
                             void *stream);
/* g = 0 where the layer output act is not positive */
int cmm_cnn_relu_mask(long n, const uint16_t *act, uint16_t *g, void *stream);
int cmm_cnn_fill(long n, float *p, float value, void *stream);
/* bf16 weights and their transposed, tap-flipped copies from FP32 master weights of n_layers layers */
int cmm_cnn_weights_bf16(int n_layers, const float *w32, uint16_t *w, uint16_t *wt, void *stream);
/* AdamW step (decoupled weight decay) on n FP32 parameters, gradient scaled by gscale; step >= 1 */
int cmm_cnn_adamw(long n, float *p, const float *g, float *m, float *v, float lr, float beta1, float beta2, float eps,
                  float weight_decay, float gscale, int step, void *stream);
/* network input from three psi value planes (newest first): psi0/s0, (psi0-psi1)/s1,
 * (psi0-2psi1+psi2)/s2 (scales: 3 floats in device memory, read by the kernel in stream order) */
int cmm_cnn_input(long n, const double *psi0, const double *psi1, const double *psi2, const float *scales,
                  uint16_t *out, void *stream);
/* training targets (2 floats per pixel): (psi_next - sum L1 psi_l)/r1, (psi_half - sum L12 psi_l)/r12
 * (L1, L12: 3 doubles each in host memory; rscales = {r1, r12}: 2 floats in device memory) */
int cmm_cnn_target(long n, const double *psi0, const double *psi1, const double *psi2, const double *psi_next,
                   const double *psi_half, const double *L1, const double *L12, const float *rscales, float *target,
                   void *stream);
/* squared-error loss (added to *loss) and its gradient pred - target in channels 0, 1 of g */
int cmm_cnn_mse(long n, const float *pred, const float *target, uint16_t *g, double *loss, void *stream);
/* correction value fields d1 = r1 pred[.][0], d12 = r12 pred[.][1] (rscales = {r1, r12}: 2 floats
 * in device memory) */
int cmm_cnn_correction(long n, const float *pred, const float *rscales, double *d1, double *d12, void *stream);
/* out = w0 psi0 + w1 psi1 + w2 psi2 (+ corr) (w: 3 doubles in host memory) */
int cmm_extrapolate(long n, const double *psi0, const double *psi1, const double *psi2, const double *w,
                    const double *corr, double *out, void *stream);
/* out2[0] = sum (a - b)^2, out2[1] = sum b^2 (device doubles) */
int cmm_dev_norms(long n, const double *a, const double *b, double *out2, void *stream);
/* plane `plane` of psi level `level` (Npsi doubles) copied to device memory dst on the engine stream */
int cmm_euler_copy_psi_level(cmm_euler *e, int level, int plane, double *dst);
/* device pointer of psi level `level` (0 newest; 4*Npsi doubles) of an engine */
int cmm_euler_psi_level(cmm_euler *e, int level, const double **dev_ptr);
/* corrections (device Hermite fields on g_psi, 4*Npsi doubles; NULL: none) added to the Lagrange
 * extrapolations of psi to t_{n+1} and to t_n + dt/2 in the following steps' map advection
 * (schemes that combine the levels: 10, 21, 30, 40, 41); the caller keeps them alive */
int cmm_euler_set_psi_correction(cmm_euler *e, const double *d1, const double *d12);

#ifdef __cplusplus
}
#endif
#endif
