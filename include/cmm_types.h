/*
 * Plain-C types shared by the CMM C-ABI (include/cmm.h) and its test oracle.
 *
 * Layout contract (same as the reference, Cuda-CMM/src/grid/cmm-grid2d.h:25-39 and
 * Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:27-41):
 *   - grids are nx*ny, periodic, row-major with x fastest: index = iy*nx + ix;
 *   - the spacing is h = (x1-x0)/(float)nx, i.e. the divisor is the float cast of the count
 *     (Cuda-CMM/src/grid/cmm-grid2d.cu:27-28);
 *   - a Hermite field is four planes of nx*ny doubles: [f, df/dx, df/dy, d2f/dxdy];
 *   - a map is two Hermite fields, ChiX and ChiY;
 *   - psi history is l_order Hermite fields (newest first), each 4*nx*ny doubles;
 *   - particles are interleaved (x0,y0,x1,y1,...).
 */
#ifndef CMM_TYPES_H
#define CMM_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

/* Periodic grid: nx*ny points on [x0,x1)x[y0,y1). */
typedef struct cmm_grid {
    int nx, ny;
    double x0, x1, y0, y1;
} cmm_grid;

/*
 * Per-step coefficients for backward map advection.  The host fills them with
 * cmm_advect_coeffs_fill(), which follows Cuda-CMM/src/simulation/cmm-simulation-host.cu:64-90.
 */
typedef struct cmm_advect_coeffs {
    double L1[4];              /* Lagrange weights at t_{n+1} over psi levels 0..l_order-1 */
    double L12[4];             /* Lagrange weights at t_n + dt/2 */
    double c1[12], cx[12], cy[12], cxy[12]; /* footpoint stencil weights (sign and 1/eps folded in) */
    double dt;                 /* step size of this step */
    double eps;                /* footpoint distance (map_epsilon) */
    int time_integration_num;  /* 10 EulerExp, 20 AB2, 21 RK2/Heun, 30 RK3, 31 RK3Mod, 40 RK4, 41 RK4Mod */
    int map_update_order_num;  /* 0: 2nd (4 footpoints), 1: 4th (8), 2: 6th (12) */
    int l_order;               /* number of psi levels (1..4) */
    int direction;             /* -1: backward map, +1: forward map */
} cmm_advect_coeffs;

/* Initial-condition identifiers.  Vorticity: 0..11 follow Cuda-CMM/src/ui/settings.h:313-324;
 * 100 and 101 are conditions the reference does not have.  Distribution functions of the 1D-1V
 * Vlasov-Poisson flow (x, v): 200 + the reference's number (Cuda-CMM/src/ui/settings.h:300-303,
 * Cuda-CMM/src/simulation/cmm-init.cu:457-493). */
enum {
    CMM_IC_4_NODES = 0,
    CMM_IC_QUADROPOLE = 1,
    CMM_IC_ONE_VORTEX = 2,
    CMM_IC_TWO_VORTICES = 3,
    CMM_IC_THREE_VORTICES = 4,
    CMM_IC_SINGLE_SHEAR_LAYER = 5,
    CMM_IC_TANH_SHEAR_LAYER = 6,
    CMM_IC_SHIELDED_VORTEX = 9,
    CMM_IC_TWO_COSINE = 10,
    CMM_IC_VORTEX_SHEETS = 11,
    CMM_IC_DOUBLE_SHEAR_LAYER = 100,   /* delta^-1 sech^2((y-pi/2)/delta) - delta^-1 sech^2((y-3pi/2)/delta) + 0.05 cos x */
    CMM_IC_RANDOM_SPECTRUM = 101,       /* discrete: seeded phases, E(k) ~ k^4 exp(-2 (k/6)^2), enstrophy 1 */
    CMM_IC_LANDAU_DAMPING = 200,        /* (1 + 0.05 cos 0.5x) exp(-v^2/2)/sqrt(2 pi) */
    CMM_IC_TWO_STREAM = 201,            /* (1 + 0.05 cos 0.2x)(exp(-(v-3)^2/2) + exp(-(v+3)^2/2))/(2 sqrt(2 pi)) */
    CMM_IC_NONLIN_LANDAU_DAMPING = 202, /* (1 + 0.5 cos 0.5x) exp(-v^2/2)/sqrt(2 pi) */
    /* ensemble distribution functions (no reference counterpart), amplitude a = ic_params[0]:
       (1 + a cos 0.5x) exp(-v^2/2)/sqrt(2 pi) and (1 + a cos 0.5x) v^2 exp(-v^2/2)/sqrt(2 pi) */
    CMM_IC_LANDAU_AMP = 203,
    CMM_IC_TWO_STREAM_V2 = 204,
    /* passive scalars, 300 + the reference's number (Cuda-CMM/src/ui/settings.h:431-434,
     * d_init_scalar, Cuda-CMM/src/simulation/cmm-init.cu:249-301) */
    CMM_IC_SCALAR_RECTANGLE = 300,
    CMM_IC_SCALAR_GAUSSIAN = 301,
    CMM_IC_SCALAR_CIRCULAR_RING = 302
};

/* Error codes returned by every C-ABI call (0 = success). */
enum {
    CMM_OK = 0,
    CMM_ERR_ARG = -1,        /* invalid argument (size, null pointer, unsupported option) */
    CMM_ERR_HIP = -2,        /* a HIP runtime call failed */
    CMM_ERR_ALLOC = -3,      /* device allocation failed */
    CMM_ERR_UNSUPPORTED = -4 /* option recognised but not on this path */
};

#ifdef __cplusplus
}
#endif
#endif
