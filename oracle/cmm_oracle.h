/*
 * CMM ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A CPU (C++/OpenMP, FP64) restatement of the reference's map-advection path, used as the
 * checker for the HIP library in tests/, in __graft_entry__.smoke() and as the CPU baseline
 * of bench.py.  Nothing in the product (cmm-turbulence_amd/) links, loads or calls it.
 *
 * Every function takes host pointers in the layouts documented in include/cmm_types.h and
 * cites the reference file:line it restates.  Pinning status of each function is listed in
 * the header comment of cmm_oracle.cpp.
 */
#ifndef CMM_ORACLE_H
#define CMM_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/cmm_types.h"

#ifdef __cplusplus
extern "C" {
#endif

double cmmo_grid_h(int n, double a, double b);

/* Lagrange coefficient, Cuda-CMM/src/numerical/cmm-timestep.cu:474-484 */
double cmmo_lagrange_coeff(const double *t, double t_next, int loop_ctr, int i_point, int l_order);
/* Coefficients of one advection step, Cuda-CMM/src/simulation/cmm-simulation-host.cu:64-90 */
int cmmo_advect_coeffs_fill(const double *t, const double *dt, int loop_ctr, int l_order,
                            int time_integration_num, int map_update_order_num, double eps,
                            int direction, cmm_advect_coeffs *c);

/* Hermite evaluation, Cuda-CMM/src/numerical/cmm-hermite.cu:323-331 / :396-420 / :446-460 / :479-503 */
double cmmo_hermite_eval(cmm_grid g, const double *H, double x, double y);
void cmmo_hermite_grad(cmm_grid g, const double *H, int n_levels, double x, double y, double *u);
void cmmo_diffeo_eval(cmm_grid g, const double *Hx, const double *Hy, double x, double y, double *xo, double *yo);
double cmmo_diffeo_det(cmm_grid g, const double *Hx, const double *Hy, double x, double y);

/* Identity map, Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:27-41 */
int cmmo_map_init_identity(cmm_grid g, double *chix, double *chiy);
/* Backward map advection, Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:207-298 */
int cmmo_map_advect(cmm_grid gm, cmm_grid gpsi, const double *chix, const double *chiy,
                    const double *psi, const cmm_advect_coeffs *c, double *chix_new, double *chiy_new);
/* One footpoint of the backward (c->direction -1) or forward (+1) time integrator at (x, y),
 * Cuda-CMM/src/numerical/cmm-timestep.cu:28-470; out = (x_f, y_f) */
int cmmo_footpoint(cmm_grid gpsi, const double *psi, const cmm_advect_coeffs *c, double x, double y, double *out);
/* max |1 - det grad X| on cell centres of gcheck, Cuda-CMM/src/simulation/cmm-simulation-host.cu:44-50 */
double cmmo_incomp_error(cmm_grid gm, cmm_grid gcheck, const double *chix, const double *chiy);
double cmmo_invert_error(cmm_grid g_back, cmm_grid g_fwd, cmm_grid gcheck, const double *bx, const double *by,
                         const double *fx, const double *fy);
/* Positions X_1(...X_M(X_cur(x))) on gt, Cuda-CMM/src/simulation/cmm-simulation-host.cu:112-152 */
int cmmo_compose_positions(cmm_grid gm, cmm_grid gt, const double *chix, const double *chiy,
                           const double *stack_x, const double *stack_y, int n_stack, double *xy);
/* Forward positions X_cur(X_{M-1}(...X_0(x))) on gt, Cuda-CMM/src/simulation/cmm-simulation-host.cu:128-148 */
int cmmo_forward_positions(cmm_grid gm, cmm_grid gt, const double *chix, const double *chiy,
                           const double *stack_x, const double *stack_y, int n_stack, double *xy);
/* Forwarded particles through stacked maps first..n_stack-1 then (if apply_current) the current map,
 * Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:97-118 with cmm-simulation-host.cu:213-291;
 * bounds = (x0, x1, y0, y1) of the target grid */
int cmmo_forward_points(cmm_grid gm, const double *bounds, const double *chix, const double *chiy,
                        const double *stack_x, const double *stack_y, int n_stack, int first_map, int apply_current,
                        int n, const double *in, double *out);
/* Initial condition at positions, Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:342-373 */
int cmmo_sample_init(cmm_grid gt, const double *xy, int ic, const double *ic_params,
                     const double *disc_herm, cmm_grid gdisc, int discrete, double *out);
double cmmo_init_vorticity(double x, double y, int ic, const double *ic_params);
/* omega(X(x)) on gtarget, Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:402-461 */
int cmmo_sample_vort(cmm_grid gm, cmm_grid gtarget, cmm_grid gfine, const double *chix, const double *chiy,
                     const double *w_herm, int molly, double *out);

/* FFTs (unnormalised).  Spectrum layout: ny rows of (nx/2+1) complex (re,im) pairs. */
int cmmo_rfft2(int nx, int ny, const double *in, double *out);
int cmmo_irfft2(int nx, int ny, const double *in, double *out);
/* real field -> Hermite (R2C, /N, fourier_hermite), Cuda-CMM/src/simulation/cmm-simulation-host.cu:318-325 */
int cmmo_hermite_from_real(cmm_grid g, const double *real, double *herm);
/* psi Hermite from map and fine vorticity, Cuda-CMM/src/simulation/cmm-simulation-host.cu:333-375 */
int cmmo_psi_from_vort(cmm_grid gm, cmm_grid gvort, cmm_grid gfine, cmm_grid gpsi,
                       const double *chix, const double *chiy, const double *w_herm,
                       int molly, double freq_cut, double *psi_out);
/* Hermite planes of F (op 0), i_laplacian(F) (op 1) or laplacian(F) (op 2),
 * Cuda-CMM/src/numerical/cmm-fft.cu:28-78, :398-474 */
int cmmo_spectral_from_real(cmm_grid g, const double *real, int op, double *herm);
/* Vlasov-Poisson psi(x, v) from the distribution function's Hermite field,
 * Cuda-CMM/src/simulation/cmm-simulation-host.cu:385-471, :478-494, :534-571 (see cmm_oracle.cpp) */
int cmmo_vp_psi_from_f(cmm_grid gm, cmm_grid gvort, cmm_grid gfine, cmm_grid gpsi, const double *chix,
                       const double *chiy, const double *f_herm, int molly, double *psi_out);
/* Fluid particles, Cuda-CMM/src/numerical/cmm-particles.cu:119-276 */
int cmmo_particles_advect(cmm_grid gpsi, const double *psi, int l_order, const double *Lp12,
                          double dt, int scheme, int n, double *pos);
/* Inertial particles, Cuda-CMM/src/numerical/cmm-particles.cu:280-551 (vel in/out, 2n), and their
 * start velocity (fluid != 0: velocity of psi0, else zero), :49-59 */
int cmmo_particles_inertial(cmm_grid gpsi, const double *psi, int l_order, const double *Lp, double dt, int scheme,
                            double tau, int n, double *pos, double *vel);
int cmmo_particles_vel_init(cmm_grid gpsi, const double *psi0, int n, int fluid, const double *pos, double *vel);
/* Random-spectrum initial field on an n*n grid (see cmm_oracle.cpp) */
int cmmo_ic_random_spectrum(int n, uint64_t seed, double *out);

/* One Euler simulation in the order of Cuda-CMM/src/simulation/cmm-euler2d.cu:45-960. */
typedef struct cmmo_euler_params {
    cmm_grid g_coarse, g_fine, g_psi, g_vort;
    double dt, final_time;
    int time_integration_num, map_update_order_num, l_order;
    int lagrange_init_higher_order;
    double map_epsilon, incomp_threshold, freq_cut_psi;
    int molly_stencil, skip_remapping;
    int ic;                    /* CMM_IC_* */
    double ic_params[10];
    int discrete;              /* sample the initial condition from a Hermite field (disc_herm) */
    cmm_grid g_disc;
    const double *disc_herm;   /* 4*N Hermite of the discrete initial condition (host) */
    int stack_capacity;        /* maps the stack can hold */
    int n_particles, particle_scheme;
    double *particles;         /* in/out, 2*n_particles */
    int max_steps;             /* stop after this many steps (<=0: run to final_time) */
    int sim;                   /* 0: 2D Euler (cmm-euler2d.cu), 1: 1D-1V Vlasov-Poisson (cmm-vlasov1d.cu) */
    int forward_map;           /* also advect the forward map (cmm-euler2d.cu:780-785) */
    double particle_init_time; /* the particle set moves once t_{n+1} + 1e-5 dt reaches it */
    double particle_tau;       /* != 0: inertial particles with velocities particle_vel */
    int particle_init_vel;     /* inertial start velocity: psi's velocity (1) or zero (0) */
    int particle_vel_given;    /* particle_vel holds start velocities (no start rule after the warm-up) */
    double *particle_vel;      /* in/out, 2*n_particles (inertial only) */
} cmmo_euler_params;

typedef struct cmmo_euler_result {
    int steps;                 /* steps taken */
    double warmup_incomp[8];   /* incompressibility error logged by each warm-up sub-step */
    int n_warmup;
    double *time;              /* [steps] t after each step (optional) */
    double *incomp;            /* [steps] E_inc after each step (optional) */
    double *map_counter;       /* [steps] stack size after each step (optional) */
    double *chix, *chiy;       /* final map (optional, 4*Nc each) */
    double *psi;               /* final psi history (optional, l_order*4*Npsi) */
    double *w_herm;            /* final fine vorticity Hermite (optional, 4*Nf) */
    double *vort_coarse;       /* final coarse vorticity sample (optional, Nc) */
    double *incomp_fwd;        /* [steps] E_inc of the forward map on the coarse grid (optional) */
    double *invert;            /* [steps] invertibility error (optional) */
    double *fchix, *fchiy;     /* final forward map (optional, 4*Nc each) */
} cmmo_euler_result;

int cmmo_euler_run(const cmmo_euler_params *p, cmmo_euler_result *r);

#ifdef __cplusplus
}
#endif
#endif
