/*
 * CMM ORACLE — TEST INFRASTRUCTURE ONLY (see cmm_oracle.h).
 *
 * CPU FP64 restatement of the reference's map-advection path.  Written from the formulas of
 * the reference sources cited at each function; the reference is CUDA-only (nvcc, cuFFT,
 * Thrust, cuRAND) and is not built or run here.  The FFT is this file's own radix-2
 * transform with the row-column C2R semantics stated at cmmo_irfft2().
 *
 * Pinning (what checks each function, all in tests/):
 *   hermite eval/grad/diffeo/det ... bicubic-polynomial exactness, identity map, brute-force
 *                                    tensor-product formula on tiny grids (test_oracle_math.py)
 *   lagrange / stencil coefficients  closed forms (3,-3,1), (1.875,-1.25,0.375), weights sum
 *   map advection .................. psi=0 identity, linear psi translation, fixture
 *                                    euler2D_one_step map statistics after warm-up + 1 step
 *   time schemes (all backward map   uniform velocity polynomial in time: exact quadrature of
 *   schemes, particle schemes) ..... each scheme (uneven level spacing where the scheme allows
 *                                    it); local-error order 2^(p+1) per halving of dt on a
 *                                    time-dependent bicubic flow; forward schemes: order in a
 *                                    steady flow, RK4Mod = the backward formula
 *                                    (test_oracle_schemes.py)
 *   footpoint stencils 2nd/4th/6th . exact on polynomials of total degree 1/3/5 (6th: every
 *                                    bicubic), closed-form residuals x^2 + eps^2 and
 *                                    x^2 y^2 - 4 eps^4 beyond (test_oracle_schemes.py)
 *   incompressibility .............. identity gives 0, fixture E_inc one-step and 11b series
 *   FFT / fourier_hermite / psi .... numpy.fft golden vectors, 4_nodes closed form, fixture
 *                                    psi / velocity statistics
 *   composition .................... translations compose, empty stack = IC, 11b remap step
 *                                    and post-remap energy/enstrophy
 *   particles ...................... linear psi gives straight lines, RK3 order test
 *   random-spectrum / double-shear . enstrophy = 1, spectrum support, periodicity (no reference
 *                                    implementation exists: these initial conditions are new)
 */
#include "cmm_oracle.h"

#include <cmath>
#include <limits>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <complex>
#include <algorithm>
#include <random>
#include <omp.h>

namespace {

const double kPi = 3.141592653589793238462643383279502884197169399375105820974944592307816406286;
const double kTwoPi = 6.283185307179586476925286766559005768394338798750211641949889184615632812572;

struct Grid {
    int nx, ny;
    long n;
    double hx, hy, x0, x1, y0, y1;
};

// h = (b-a)/(float)N, Cuda-CMM/src/grid/cmm-grid2d.cu:27-28
Grid mk(const cmm_grid &g) {
    Grid r;
    r.nx = g.nx; r.ny = g.ny; r.n = (long)g.nx * g.ny;
    r.x0 = g.x0; r.x1 = g.x1; r.y0 = g.y0; r.y1 = g.y1;
    r.hx = (g.x1 - g.x0) / (float)g.nx;
    r.hy = (g.y1 - g.y0) / (float)g.ny;
    return r;
}

// periodic cell lookup, Cuda-CMM/src/numerical/cmm-hermite.cu:152-164 and :191-208.
// The wrap count is a true floor division; the reference's (I/N - (I<0)) differs only for
// exact negative multiples of N, where it indexes one period outside the array.
struct Cell {
    long I[4];      // I00, I10, I01, I11
    int wx0, wx1, wy0, wy1;
    double dx, dy;
};

inline void wrap_index(long i, int n, int &r, int &w) {
    long q = i % n;
    if (q < 0) q += n;
    r = (int)q;
    w = (int)((i - q) / n);
}

inline void cell(const Grid &g, double x, double y, Cell &c) {
    double fx = (x - g.x0) / g.hx, fy = (y - g.y0) / g.hy;
    double flx = std::floor(fx), fly = std::floor(fy);
    long ix = (long)flx, iy = (long)fly;
    if (!(flx > -2147483648.0 && flx < 2147483647.0)) ix = 0;
    if (!(fly > -2147483648.0 && fly < 2147483647.0)) iy = 0;
    c.dx = fx - (double)ix;
    c.dy = fy - (double)iy;
    int rx0, rx1, ry0, ry1;
    wrap_index(ix, g.nx, rx0, c.wx0);
    wrap_index(ix + 1, g.nx, rx1, c.wx1);
    wrap_index(iy, g.ny, ry0, c.wy0);
    wrap_index(iy + 1, g.ny, ry1, c.wy1);
    c.I[0] = (long)ry0 * g.nx + rx0;
    c.I[1] = (long)ry0 * g.nx + rx1;
    c.I[2] = (long)ry1 * g.nx + rx0;
    c.I[3] = (long)ry1 * g.nx + rx1;
}

// cubic Hermite basis in the reference's f3/f4 form, Cuda-CMM/src/numerical/cmm-hermite.h:27-30 and
// Cuda-CMM/src/numerical/cmm-hermite.cu:240-277
inline double f3(double x) { return (3 - 2 * x) * x * x; }
inline double f3x(double x) { return (6 - 6 * x) * x; }
inline double f4(double x) { return (x - 1) * x * x; }
inline double f4x(double x) { return (3 * x - 2) * x; }

inline void basis(double d, double b[4]) { b[0] = f3(1 - d); b[1] = f3(d); b[2] = -f4(1 - d); b[3] = f4(d); }
inline void basis_d(double d, double b[4]) { b[0] = -f3x(1 - d); b[1] = f3x(d); b[2] = f4x(1 - d); b[3] = f4x(d); }

// tensor-product sum over 4 planes, Cuda-CMM/src/numerical/cmm-hermite.cu:31-45
inline double hmult(const double *H, long N, const double bX[4], const double bY[4], const long I[4],
                    double hx, double hy) {
    double out = 0.0;
    for (int o = 0; o < 4; ++o) {
        int ox = 2 * (o & 1), oy = 2 * ((o >> 1) & 1);
        double t = 0.0;
        for (int p = 0; p < 4; ++p) t += bY[(p >> 1) + oy] * bX[(p & 1) + ox] * H[o * N + I[p]];
        if (o & 1) t *= hx;
        if (o & 2) t *= hy;
        out += t;
    }
    return out;
}

// value plane gets + I_w*L on warped corners, Cuda-CMM/src/numerical/cmm-hermite.cu:67-90
inline double hmult_warp(const double *H, long N, const double bX[4], const double bY[4], const long I[4],
                         const int W[4], double L, double hx, double hy) {
    double out = 0.0;
    for (int o = 0; o < 4; ++o) {
        int ox = 2 * (o & 1), oy = 2 * ((o >> 1) & 1);
        double t = 0.0;
        for (int p = 0; p < 4; ++p) {
            double v = (o == 0) ? (H[I[p]] + W[p] * L) : H[o * N + I[p]];
            t += bY[(p >> 1) + oy] * bX[(p & 1) + ox] * v;
        }
        if (o & 1) t *= hx;
        if (o & 2) t *= hy;
        out += t;
    }
    return out;
}

double herm_eval(const Grid &g, const double *H, double x, double y) {
    Cell c; cell(g, x, y, c);
    double bX[4], bY[4]; basis(c.dx, bX); basis(c.dy, bY);
    return hmult(H, g.n, bX, bY, c.I, g.hx, g.hy);
}

// u = (d psi/dy, -d psi/dx) per level, Cuda-CMM/src/numerical/cmm-hermite.cu:396-420
void herm_grad(const Grid &g, const double *H, int nl, double x, double y, double *u) {
    Cell c; cell(g, x, y, c);
    double bX[4], bY[4];
    basis_d(c.dx, bX); basis(c.dy, bY);
    for (int l = 0; l < nl; ++l) u[2 * l + 1] = -hmult(H + 4 * g.n * l, g.n, bX, bY, c.I, g.hx, g.hy) / g.hx;
    basis(c.dx, bX); basis_d(c.dy, bY);
    for (int l = 0; l < nl; ++l) u[2 * l] = hmult(H + 4 * g.n * l, g.n, bX, bY, c.I, g.hx, g.hy) / g.hy;
}

// Cuda-CMM/src/numerical/cmm-hermite.cu:446-460
void diffeo(const Grid &g, const double *Hx, const double *Hy, double x, double y, double &xo, double &yo) {
    Cell c; cell(g, x, y, c);
    double bX[4], bY[4]; basis(c.dx, bX); basis(c.dy, bY);
    double Lx = g.nx * g.hx, Ly = g.ny * g.hy;
    int wx[4] = {c.wx0, c.wx1, c.wx0, c.wx1};
    int wy[4] = {c.wy0, c.wy0, c.wy1, c.wy1};
    xo = hmult_warp(Hx, g.n, bX, bY, c.I, wx, Lx, g.hx, g.hy);
    yo = hmult_warp(Hy, g.n, bX, bY, c.I, wy, Ly, g.hx, g.hy);
}

// det grad X, Cuda-CMM/src/numerical/cmm-hermite.cu:479-503
double diffeo_det(const Grid &g, const double *Hx, const double *Hy, double x, double y) {
    Cell c; cell(g, x, y, c);
    double Lx = g.nx * g.hx, Ly = g.ny * g.hy;
    int wx[4] = {c.wx0, c.wx1, c.wx0, c.wx1};
    int wy[4] = {c.wy0, c.wy0, c.wy1, c.wy1};
    double bX[4], bY[4];
    basis_d(c.dx, bX); basis(c.dy, bY);
    double Xx = hmult_warp(Hx, g.n, bX, bY, c.I, wx, Lx, g.hx, g.hy) / g.hx;
    double Yx = hmult_warp(Hy, g.n, bX, bY, c.I, wy, Ly, g.hx, g.hy) / g.hy;
    basis(c.dx, bX); basis_d(c.dy, bY);
    double Xy = hmult_warp(Hx, g.n, bX, bY, c.I, wx, Lx, g.hx, g.hy) / g.hx;
    double Yy = hmult_warp(Hy, g.n, bX, bY, c.I, wy, Ly, g.hx, g.hy) / g.hy;
    return Xx * Yy - Xy * Yx;
}

// Lagrange-weighted velocity, Cuda-CMM/src/numerical/cmm-timestep.cu:175-207 (loop shape)
inline void lsum(const double *w, const double *u, int nl, double &kx, double &ky) {
    kx = w[0] * u[0]; ky = w[0] * u[1];
    for (int l = 1; l < nl; ++l) { kx += w[l] * u[2 * l]; ky += w[l] * u[2 * l + 1]; }
}

// backward time integrators, Cuda-CMM/src/numerical/cmm-timestep.cu:35-45, 68-85, 115-137,
// 175-207, 261-304, 357-386, 441-470
void step_backward(const Grid &g, const double *psi, const cmm_advect_coeffs &c, const double xin[2], double xout[2]) {
    const int L = c.l_order;
    const double dt = c.dt;
    const long N4 = 4 * g.n;
    double u[8];
    switch (c.time_integration_num) {
    case 10: {
        herm_grad(g, psi, L, xin[0], xin[1], u);
        double kx, ky; lsum(c.L1, u, L, kx, ky);
        xout[0] = xin[0] - dt * kx; xout[1] = xin[1] - dt * ky;
        break;
    }
    case 20: {
        double uu[4], xp[2] = {xin[0], xin[1]};
        xout[0] = xin[0]; xout[1] = xin[1];
        for (int it = 0; it < 10; ++it) {
            herm_grad(g, psi + N4, 1, xp[0], xp[1], uu + 2);
            xp[0] = xout[0] - dt * uu[2]; xp[1] = xout[1] - dt * uu[3];
            herm_grad(g, psi + N4, 1, xp[0], xp[1], uu + 2);
            herm_grad(g, psi, 1, xout[0], xout[1], uu);
            xout[0] = xin[0] - dt * (1.5 * uu[0] - 0.5 * uu[2]);
            xout[1] = xin[1] - dt * (1.5 * uu[1] - 0.5 * uu[3]);
        }
        break;
    }
    case 21: {
        herm_grad(g, psi, L, xin[0], xin[1], u);
        double k1x, k1y; lsum(c.L1, u, L, k1x, k1y);
        double p[2] = {xin[0] - dt * k1x, xin[1] - dt * k1y};
        herm_grad(g, psi, 1, p[0], p[1], u);
        xout[0] = xin[0] - dt * (k1x + u[0]) / 2;
        xout[1] = xin[1] - dt * (k1y + u[1]) / 2;
        break;
    }
    case 30: {
        double k[6];
        herm_grad(g, psi, L, xin[0], xin[1], u);
        lsum(c.L1, u, L, k[0], k[1]);
        double p[2] = {xin[0] - dt * k[0] / 2.0, xin[1] - dt * k[1] / 2.0};
        herm_grad(g, psi, L, p[0], p[1], u);
        lsum(c.L12, u, L, k[2], k[3]);
        p[0] = xin[0] + dt * k[0] - 2 * dt * k[2];
        p[1] = xin[1] + dt * k[1] - 2 * dt * k[3];
        herm_grad(g, psi, 1, p[0], p[1], k + 4);
        xout[0] = xin[0] - dt * (k[0] + 4 * k[2] + k[4]) / 6.0;
        xout[1] = xin[1] - dt * (k[1] + 4 * k[3] + k[5]) / 6.0;
        break;
    }
    case 31: {
        double k[6];
        herm_grad(g, psi, L, xin[0], xin[1], u);
        lsum(c.L1, u, L, k[0], k[1]);
        double p[2] = {xin[0] - dt * k[0], xin[1] - dt * k[1]};
        herm_grad(g, psi, 1, p[0], p[1], k + 2);
        p[0] = xin[0] + 2 * dt * (-2 * k[0] + k[2]);
        p[1] = xin[1] + 2 * dt * (-2 * k[1] + k[3]);
        herm_grad(g, psi + N4, 1, p[0], p[1], k + 4);
        xout[0] = xin[0] - dt * (5 * k[0] + 8 * k[2] - k[4]) / 12.0;
        xout[1] = xin[1] - dt * (5 * k[1] + 8 * k[3] - k[5]) / 12.0;
        break;
    }
    case 40: {
        double k[8];
        herm_grad(g, psi, L, xin[0], xin[1], u);
        lsum(c.L1, u, L, k[0], k[1]);
        double p[2] = {xin[0] - dt * k[0] / 2.0, xin[1] - dt * k[1] / 2.0};
        herm_grad(g, psi, L, p[0], p[1], u);
        lsum(c.L12, u, L, k[2], k[3]);
        p[0] = xin[0] - dt * k[2] / 2.0; p[1] = xin[1] - dt * k[3] / 2.0;
        herm_grad(g, psi, L, p[0], p[1], u);
        lsum(c.L12, u, L, k[4], k[5]);
        p[0] = xin[0] - dt * k[4]; p[1] = xin[1] - dt * k[5];
        herm_grad(g, psi, 1, p[0], p[1], k + 6);
        xout[0] = xin[0] - dt * (k[0] + 2 * k[2] + 2 * k[4] + k[6]) / 6.0;
        xout[1] = xin[1] - dt * (k[1] + 2 * k[3] + 2 * k[5] + k[7]) / 6.0;
        break;
    }
    case 41: {
        double k[8];
        herm_grad(g, psi, L, xin[0], xin[1], u);
        lsum(c.L1, u, L, k[0], k[1]);
        double p[2] = {xin[0] - dt * k[0], xin[1] - dt * k[1]};
        herm_grad(g, psi, 1, p[0], p[1], k + 2);
        p[0] = xin[0] - dt * (3 * k[0] + k[2]) / 8.0;
        p[1] = xin[1] - dt * (3 * k[1] + k[3]) / 8.0;
        herm_grad(g, psi, L, p[0], p[1], u);
        lsum(c.L12, u, L, k[4], k[5]);
        p[0] = xin[0] + dt * (k[2] - 4 * k[4]) / 3.0;
        p[1] = xin[1] + dt * (k[3] - 4 * k[5]) / 3.0;
        herm_grad(g, psi, 1, p[0], p[1], k + 6);
        xout[0] = xin[0] - dt * (2 * k[0] - k[2] + 8 * k[4] + 3 * k[6]) / 12.0;
        xout[1] = xin[1] - dt * (2 * k[1] - k[3] + 8 * k[5] + 3 * k[7]) / 12.0;
        break;
    }
    default:
        xout[0] = xout[1] = 0;
    }
}

// forward time integrators of the forward map, Cuda-CMM/src/numerical/cmm-timestep.cu:28-33 (EulerExp),
// :48-65 (Heun), :86-109 (AB2), :143-174 (RK3), :218-260 (RK4), :314-337 (RK3Mod), :384-430 (RK4Mod).
// As in the reference: RK4 weights its second stage with L1 and its fourth with L12, and RK4Mod is
// the backward formula.
void step_forward(const Grid &g, const double *psi, const cmm_advect_coeffs &c, const double xin[2], double xout[2]) {
    const int L = c.l_order;
    const double dt = c.dt;
    const long N4 = 4 * g.n;
    double u[8];
    switch (c.time_integration_num) {
    case 10: {
        herm_grad(g, psi, 1, xin[0], xin[1], u);
        xout[0] = xin[0] + dt * u[0]; xout[1] = xin[1] + dt * u[1];
        break;
    }
    case 20: {
        double uu[4], xp[2] = {xin[0], xin[1]};
        xout[0] = xin[0]; xout[1] = xin[1];
        for (int it = 0; it < 10; ++it) {
            herm_grad(g, psi + N4, 1, xp[0], xp[1], uu + 2);
            xp[0] = xout[0] + dt * uu[2]; xp[1] = xout[1] + dt * uu[3];
            herm_grad(g, psi + N4, 1, xp[0], xp[1], uu + 2);
            herm_grad(g, psi, 1, xout[0], xout[1], uu);
            xout[0] = xin[0] + dt * (1.5 * uu[0] - 0.5 * uu[2]);
            xout[1] = xin[1] + dt * (1.5 * uu[1] - 0.5 * uu[3]);
        }
        break;
    }
    case 21: {
        double k[4];
        herm_grad(g, psi, 1, xin[0], xin[1], k);
        double p[2] = {xin[0] + dt * k[0], xin[1] + dt * k[1]};
        herm_grad(g, psi, L, p[0], p[1], u);
        lsum(c.L1, u, L, k[2], k[3]);
        xout[0] = xin[0] + dt * (k[0] + k[2]) / 2;
        xout[1] = xin[1] + dt * (k[1] + k[3]) / 2;
        break;
    }
    case 30: {
        double k[6];
        herm_grad(g, psi, 1, xin[0], xin[1], k);
        double p[2] = {xin[0] + dt * k[0] / 2.0, xin[1] + dt * k[1] / 2.0};
        herm_grad(g, psi, L, p[0], p[1], u);
        lsum(c.L12, u, L, k[2], k[3]);
        p[0] = xin[0] - dt * k[0] + 2 * dt * k[2];
        p[1] = xin[1] - dt * k[1] + 2 * dt * k[3];
        herm_grad(g, psi, L, p[0], p[1], u);
        lsum(c.L1, u, L, k[4], k[5]);
        xout[0] = xin[0] + dt * (k[0] + 4 * k[2] + k[4]) / 6.0;
        xout[1] = xin[1] + dt * (k[1] + 4 * k[3] + k[5]) / 6.0;
        break;
    }
    case 31: {
        double k[6];
        herm_grad(g, psi, 1, xin[0], xin[1], k);
        double p[2] = {xin[0] - dt * k[0], xin[1] - dt * k[1]};
        herm_grad(g, psi + N4, 1, p[0], p[1], k + 2);
        p[0] = xin[0] - dt * (8 * k[0] + 2 * k[2]) / 5.0;
        p[1] = xin[1] - dt * (8 * k[1] + 2 * k[3]) / 5.0;
        herm_grad(g, psi + 2 * N4, 1, p[0], p[1], k + 4);
        xout[0] = xin[0] + dt * (23 * k[0] - 16 * k[2] + 5 * k[4]) / 12.0;
        xout[1] = xin[1] + dt * (23 * k[1] - 16 * k[3] + 5 * k[5]) / 12.0;
        break;
    }
    case 40: {
        double k[8];
        herm_grad(g, psi, 1, xin[0], xin[1], k);
        double p[2] = {xin[0] + dt * k[0] / 2.0, xin[1] + dt * k[1] / 2.0};
        herm_grad(g, psi, L, p[0], p[1], u);
        lsum(c.L1, u, L, k[2], k[3]);
        p[0] = xin[0] + dt * k[2] / 2.0; p[1] = xin[1] + dt * k[3] / 2.0;
        herm_grad(g, psi, L, p[0], p[1], u);
        lsum(c.L12, u, L, k[4], k[5]);
        p[0] = xin[0] + dt * k[4]; p[1] = xin[1] + dt * k[5];
        herm_grad(g, psi, L, p[0], p[1], u);
        lsum(c.L12, u, L, k[6], k[7]);
        xout[0] = xin[0] + dt * (k[0] + 2 * k[2] + 2 * k[4] + k[6]) / 6.0;
        xout[1] = xin[1] + dt * (k[1] + 2 * k[3] + 2 * k[5] + k[7]) / 6.0;
        break;
    }
    case 41: {
        double k[8];
        herm_grad(g, psi, L, xin[0], xin[1], u);
        lsum(c.L1, u, L, k[0], k[1]);
        double p[2] = {xin[0] - dt * k[0], xin[1] - dt * k[1]};
        herm_grad(g, psi, 1, p[0], p[1], k + 2);
        p[0] = xin[0] - dt * (3 * k[0] + k[2]) / 8.0;
        p[1] = xin[1] - dt * (3 * k[1] + k[3]) / 8.0;
        herm_grad(g, psi, L, p[0], p[1], u);
        lsum(c.L12, u, L, k[4], k[5]);
        p[0] = xin[0] + dt * (k[2] - 4 * k[4]) / 3.0;
        p[1] = xin[1] + dt * (k[3] - 4 * k[5]) / 3.0;
        herm_grad(g, psi, 1, p[0], p[1], k + 6);
        xout[0] = xin[0] - dt * (2 * k[0] - k[2] + 8 * k[4] + 3 * k[6]) / 12.0;
        xout[1] = xin[1] - dt * (2 * k[1] - k[3] + 8 * k[5] + 3 * k[7]) / 12.0;
        break;
    }
    default:
        xout[0] = xout[1] = 0;
    }
}

// ----------------------------------------------------------------------------------------------
// FFT: iterative radix-2 complex transform, twiddles from long double.
typedef std::complex<double> cplx;

struct FFTPlan {
    int n, logn;
    std::vector<cplx> w;     // e^{-2 pi i k/n}, k < n/2
    std::vector<int> rev;
};

FFTPlan make_plan(int n) {
    FFTPlan p;
    p.n = n;
    p.logn = 0;
    while ((1 << p.logn) < n) ++p.logn;
    p.w.resize(n / 2 > 0 ? n / 2 : 1);
    for (int k = 0; k < n / 2; ++k) {
        long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
        p.w[k] = cplx((double)cosl(a), (double)sinl(a));
    }
    p.rev.resize(n);
    for (int i = 0; i < n; ++i) {
        int r = 0;
        for (int b = 0; b < p.logn; ++b) if (i & (1 << b)) r |= 1 << (p.logn - 1 - b);
        p.rev[i] = r;
    }
    return p;
}

// sign = -1: forward (e^{-i}), +1: inverse without normalisation
void fft1(const FFTPlan &p, cplx *a, int sign) {
    const int n = p.n;
    for (int i = 0; i < n; ++i) if (i < p.rev[i]) std::swap(a[i], a[p.rev[i]]);
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int j = 0; j < half; ++j) {
                cplx w = p.w[j * step];
                if (sign > 0) w = std::conj(w);
                cplx u = a[i + j], v = a[i + j + half] * w;
                a[i + j] = u + v;
                a[i + j + half] = u - v;
            }
        }
    }
}

bool pow2(int n) { return n >= 2 && (n & (n - 1)) == 0; }

// R2C: rows (x) then columns (y); out[y][kx], kx <= nx/2
void rfft2(int nx, int ny, const double *in, cplx *out) {
    const int nh = nx / 2 + 1;
    FFTPlan px = make_plan(nx), py = make_plan(ny);
#pragma omp parallel
    {
        std::vector<cplx> row(nx);
#pragma omp for
        for (int y = 0; y < ny; ++y) {
            for (int x = 0; x < nx; ++x) row[x] = cplx(in[(long)y * nx + x], 0.0);
            fft1(px, row.data(), -1);
            for (int k = 0; k < nh; ++k) out[(long)y * nh + k] = row[k];
        }
        std::vector<cplx> col(ny);
#pragma omp for
        for (int k = 0; k < nh; ++k) {
            for (int y = 0; y < ny; ++y) col[y] = out[(long)y * nh + k];
            fft1(py, col.data(), -1);
            for (int y = 0; y < ny; ++y) out[(long)y * nh + k] = col[y];
        }
    }
}

// C2R, unnormalised, row-column semantics: inverse complex transform along y of every kx column,
// then real inverse along x with the imaginary parts of the kx=0 and kx=nx/2 bins dropped.
// (Pinned against numpy.fft.irfft2 * nx*ny, which states the same semantics.)
void irfft2(int nx, int ny, const cplx *in, double *out) {
    const int nh = nx / 2 + 1;
    FFTPlan px = make_plan(nx), py = make_plan(ny);
    std::vector<cplx> tmp((size_t)nh * ny);
#pragma omp parallel
    {
        std::vector<cplx> col(ny);
#pragma omp for
        for (int k = 0; k < nh; ++k) {
            for (int y = 0; y < ny; ++y) col[y] = in[(long)y * nh + k];
            fft1(py, col.data(), +1);
            for (int y = 0; y < ny; ++y) tmp[(long)y * nh + k] = col[y];
        }
        std::vector<cplx> row(nx);
#pragma omp for
        for (int y = 0; y < ny; ++y) {
            const cplx *r = &tmp[(long)y * nh];
            row[0] = cplx(r[0].real(), 0.0);
            row[nx / 2] = cplx(r[nx / 2].real(), 0.0);
            for (int k = 1; k < nx / 2; ++k) { row[k] = r[k]; row[nx - k] = std::conj(r[k]); }
            fft1(px, row.data(), +1);
            for (int x = 0; x < nx; ++x) out[(long)y * nx + x] = row[x].real();
        }
    }
}

// signed wavenumber with the Nyquist index positive, Cuda-CMM/src/numerical/cmm-fft.cu:72-73
inline double kfac(const Grid &g, int axis) {
    return axis == 0 ? kTwoPi / (g.hx * g.nx) : kTwoPi / (g.hy * g.ny);
}

// fourier_hermite, Cuda-CMM/src/numerical/cmm-fft.cu:398-429: planes f, dx, dy, dxdy
void fourier_hermite(const Grid &g, const cplx *F, double *out) {
    const int nx = g.nx, ny = g.ny, nh = nx / 2 + 1;
    const long nf = (long)nh * ny;
    std::vector<cplx> a(nf), b(nf);
    const double fx = kfac(g, 0), fy = kfac(g, 1);
    // dy (k_fft_dy_h :169-187) and dxdy = dx(dy) (k_fft_dx_h :147-165)
#pragma omp parallel for
    for (int iy = 0; iy < ny; ++iy) {
        double ky = fy * (iy - (iy > ny / 2) * ny);
        for (int ix = 0; ix < nh; ++ix) {
            double kx = fx * (ix - (ix > nx / 2) * nx);
            long i = (long)iy * nh + ix;
            cplx t = F[i];
            cplx d(-t.imag() * ky, t.real() * ky);
            a[i] = d;
            b[i] = cplx(-d.imag() * kx, d.real() * kx);
        }
    }
    irfft2(nx, ny, b.data(), out + 3 * g.n);
    irfft2(nx, ny, a.data(), out + 2 * g.n);
#pragma omp parallel for
    for (int iy = 0; iy < ny; ++iy) {
        for (int ix = 0; ix < nh; ++ix) {
            double kx = fx * (ix - (ix > nx / 2) * nx);
            long i = (long)iy * nh + ix;
            cplx t = F[i];
            a[i] = cplx(-t.imag() * kx, t.real() * kx);
        }
    }
    irfft2(nx, ny, a.data(), out + 1 * g.n);
    irfft2(nx, ny, F, out);
}

// ----------------------------------------------------------------------------------------------
// Initial conditions, Cuda-CMM/src/simulation/cmm-init.cu:30-241 (and new ones)
double init_vorticity(double x, double y, int ic, const double *prm) {
    switch (ic) {
    // distribution functions f(x, v), d_init_distirbution_function (Cuda-CMM/src/simulation/cmm-init.cu:457-493)
    case CMM_IC_LANDAU_DAMPING:
        return (1 + 0.05 * std::cos(0.5 * x)) * std::exp(-y * y * 0.5) / std::sqrt(2 * kPi);
    case CMM_IC_TWO_STREAM:
        return 1 / (2 * std::sqrt(2 * kPi)) * (1 + 0.05 * std::cos(0.2 * x)) *
               (std::exp(-(y - 3) * (y - 3) * 0.5) + std::exp(-(y + 3) * (y + 3) * 0.5));
    case CMM_IC_NONLIN_LANDAU_DAMPING:
        return (1 + 0.5 * std::cos(0.5 * x)) * std::exp(-y * y * 0.5) / std::sqrt(2 * kPi);
    // ensemble distribution functions with amplitude prm[0] (no reference counterpart)
    case CMM_IC_LANDAU_AMP:
        return (1 + prm[0] * std::cos(0.5 * x)) * std::exp(-y * y * 0.5) / std::sqrt(2 * kPi);
    case CMM_IC_TWO_STREAM_V2:
        return (1 + prm[0] * std::cos(0.5 * x)) * y * y * std::exp(-y * y * 0.5) / std::sqrt(2 * kPi);
    // passive scalars, d_init_scalar (Cuda-CMM/src/simulation/cmm-init.cu:249-301)
    case CMM_IC_SCALAR_RECTANGLE: {
        x -= std::floor(x / kTwoPi) * kTwoPi;
        y -= std::floor(y / kTwoPi) * kTwoPi;
        const double cx0 = kPi, cy0 = kPi / 2.0, wx = 2 * kPi, wy = kPi;
        return (x > cx0 - wx / 2.0 && x < cx0 + wx / 2.0 && y > cy0 - wy / 2.0 && y < cy0 + wy / 2.0) ? 1.0 : 0.0;
    }
    case CMM_IC_SCALAR_GAUSSIAN: {
        const double mu = kPi, var = kPi / 4.0;
        return 1.0 / (kTwoPi * var * var) *
               std::exp(-((x - mu) * (x - mu) / (2 * var * var) + (y - mu) * (y - mu) / (2 * var * var)));
    }
    case CMM_IC_SCALAR_CIRCULAR_RING: {
        const double r = kPi / 2.0, rw = r / 8.0, s1 = 10.0, s2 = 10.0;
        const double d2 = (x - kPi) * (x - kPi) + (y - kPi) * (y - kPi);
        return (std::tanh(s1 * (d2 - (r + rw) * (r + rw))) - std::tanh(s2 * (d2 - (r - rw) * (r - rw)))) / 2.0;
    }
    case CMM_IC_4_NODES: {
        x -= std::floor(x / kTwoPi) * kTwoPi;
        y -= std::floor(y / kTwoPi) * kTwoPi;
        return std::cos(x) + std::cos(y) + 0.6 * std::cos(2 * x) + 0.2 * std::cos(3 * x);
    }
    case CMM_IC_QUADROPOLE: {
        double ret = 0;
        for (int iy = -2; iy <= 2; iy++)
            for (int ix = -2; ix <= 2; ix++) {
                double dx = x - kPi / 2 + ix * 2 * kPi, dy = y - kPi / 2 + iy * 2 * kPi;
                double A = 0.6258473, s = 0.5;
                double B = A / (s * s * s * s) * (dx * dy) * (dx * dx + dy * dy - 6 * s * s);
                double D = (dx * dx + dy * dy) / (2 * s * s);
                ret += B * std::exp(-D);
            }
        return ret;
    }
    case CMM_IC_ONE_VORTEX: {
        double sigma = kTwoPi * 0.125;
        return 1.0 / (kTwoPi * sigma * sigma) * std::exp(-((x - kPi) * (x - kPi) + (y - kPi) * (y - kPi)) / (2 * sigma * sigma));
    }
    case CMM_IC_SINGLE_SHEAR_LAYER:
        return (1 + 0.01 * std::cos(2 * x)) * std::exp(-50.0 * (y - kPi) * (y - kPi));
    case CMM_IC_TANH_SHEAR_LAYER: {
        double fac = 5, inst_strength = 0.01, inst_freq = 2;
        double d = 14 * inst_freq / kTwoPi;
        double e = std::exp(-d * (y - kPi)) + std::exp(d * (y - kPi));
        return fac * (1 + inst_strength * std::cos(inst_freq * x)) * 4 / e / e;
    }
    case CMM_IC_SHIELDED_VORTEX: {
        double nu = 2e-1, nu_fac = 1 / (2 * nu * nu * nu), nu_c = 4 * nu, nu_s = 4 * nu;
        double xr = kPi - x, yr = kPi - y;
        return nu_fac * (nu_c - xr * xr - yr * yr) * std::exp(-(xr * xr + yr * yr) / nu_s);
    }
    case CMM_IC_TWO_COSINE: {
        x -= std::floor(x / kTwoPi) * kTwoPi;
        y -= std::floor(y / kTwoPi) * kTwoPi;
        return std::cos(x) * std::cos(y);
    }
    case CMM_IC_VORTEX_SHEETS: {
        double d = 1 / std::sqrt(prm ? prm[0] : 1.0);
        double s = y - kPi - std::sin(x) / 2.0;
        return std::exp(-s * s / (2 * d * d)) / std::sqrt(2 * kPi) / d;
    }
    case CMM_IC_DOUBLE_SHEAR_LAYER: {
        // periodic images of the two layers summed; terms beyond |n|=2 are below 1e-40
        const double delta = kPi / 15.0;
        double ret = 0.05 * std::cos(x);
        for (int im = -2; im <= 2; ++im) {
            double a = (y + im * kTwoPi - kPi / 2) / delta;
            double b = (y + im * kTwoPi - 3 * kPi / 2) / delta;
            double sa = 1.0 / std::cosh(a), sb = 1.0 / std::cosh(b);
            ret += (sa * sa - sb * sb) / delta;
        }
        return ret;
    }
    default:
        return 0.0;
    }
}

void sample_init(const Grid &gt, const double *xy, int ic, const double *prm, const double *disc, const Grid &gd,
                 int discrete, double *out) {
#pragma omp parallel for
    for (long i = 0; i < gt.n; ++i) {
        double x = xy[2 * i], y = xy[2 * i + 1];
        out[i] = discrete ? herm_eval(gd, disc, x, y) : init_vorticity(x, y, ic, prm);
    }
}

// composition, Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:302-340
void compose(const Grid &gm, const Grid &gt, const double *cx, const double *cy, const double *sx, const double *sy,
             int ns, double *xy) {
    const double LX = gm.x1 - gm.x0, LY = gm.y1 - gm.y0;
#pragma omp parallel for
    for (int iy = 0; iy < gt.ny; ++iy) {
        for (int ix = 0; ix < gt.nx; ++ix) {
            long i = (long)iy * gt.nx + ix;
            double x = gt.x0 + ix * gt.hx, y = gt.y0 + iy * gt.hy;
            double px, py;
            diffeo(gm, cx, cy, x, y, px, py);
            x = px - std::floor((px - gm.x0) / LX) * LX;
            y = py - std::floor((py - gm.y0) / LY) * LY;
            for (int m = ns - 1; m >= 0; --m) {
                diffeo(gm, sx + 4 * gm.n * (long)m, sy + 4 * gm.n * (long)m, x, y, px, py);
                x = px - std::floor((px - gm.x0) / LX) * LX;
                y = py - std::floor((py - gm.y0) / LY) * LY;
            }
            xy[2 * i] = x; xy[2 * i + 1] = y;
        }
    }
}

// Forward composition, apply_map_stack with direction 1 (Cuda-CMM/src/simulation/cmm-simulation-host.cu:128-148):
// the oldest stacked forward map first (k_h_sample_map_compact), the newer ones after it and the
// current forward map last (k_apply_map_compact, Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:302-340),
// each result wrapped into the domain; with an empty stack the current map alone.
void compose_forward(const Grid &gm, const Grid &gt, const double *cx, const double *cy, const double *sx,
                     const double *sy, int ns, double *xy) {
    const double LX = gm.x1 - gm.x0, LY = gm.y1 - gm.y0;
#pragma omp parallel for
    for (int iy = 0; iy < gt.ny; ++iy) {
        for (int ix = 0; ix < gt.nx; ++ix) {
            long i = (long)iy * gt.nx + ix;
            double x = gt.x0 + ix * gt.hx, y = gt.y0 + iy * gt.hy;
            double px, py;
            for (int m = 0; m <= ns; ++m) {
                const double *mx = m < ns ? sx + 4 * gm.n * (long)m : cx;
                const double *my = m < ns ? sy + 4 * gm.n * (long)m : cy;
                diffeo(gm, mx, my, x, y, px, py);
                x = px - std::floor((px - gm.x0) / LX) * LX;
                y = py - std::floor((py - gm.y0) / LY) * LY;
            }
            xy[2 * i] = x; xy[2 * i + 1] = y;
        }
    }
}

// Forwarded particles: k_h_sample_points_map (Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:97-118)
// applied with the stacked forward maps first..ns-1 and then, if cur, the current forward map
// (the selection of Cuda-CMM/src/simulation/cmm-simulation-host.cu:213-291).  A point strictly inside
// the bounds b = (x0, x1, y0, y1) goes through the map and is wrapped by x - floor(x / Lx) Lx, with
// Lx the map domain's length; any other point becomes NaN.
void forward_points(const Grid &gm, const double *b, const double *cx, const double *cy, const double *sx,
                    const double *sy, int ns, int first, int cur, int n, const double *in, double *out) {
    const double LX = gm.x1 - gm.x0, LY = gm.y1 - gm.y0;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (int i = 0; i < n; ++i) {
        double x = in[2 * i], y = in[2 * i + 1];
        auto apply = [&](const double *mx, const double *my) {
            if (x > b[0] && x < b[1] && y > b[2] && y < b[3]) {
                double px, py;
                diffeo(gm, mx, my, x, y, px, py);
                x = px - std::floor(px / LX) * LX;
                y = py - std::floor(py / LY) * LY;
            } else {
                x = y = nan;
            }
        };
        for (int m = first < 0 ? 0 : first; m < ns; ++m) apply(sx + 4 * gm.n * (long)m, sy + 4 * gm.n * (long)m);
        if (cur) apply(cx, cy);
        out[2 * i] = x; out[2 * i + 1] = y;
    }
}

// Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:402-461; positions use the map grid's origin
void sample_vort(const Grid &gm, const Grid &gv, const Grid &gf, const double *cx, const double *cy, const double *H,
                 int molly, double *out) {
#pragma omp parallel for
    for (int iy = 0; iy < gv.ny; ++iy) {
        for (int ix = 0; ix < gv.nx; ++ix) {
            double x = gm.x0 + ix * gv.hx, y = gm.y0 + iy * gv.hy;
            double x2, y2, v;
            if (molly == 4) {
                diffeo(gm, cx, cy, x, y, x2, y2);
                v = herm_eval(gf, H, x2, y2) / 3.0;
                for (int m = 0; m < 4; ++m) {
                    // reference: (i_molly>>1+1)&1 parses as (i_molly>>2)&1
                    x2 = x + gv.hx / 2.0 * ((m >> 2) & 1) * (-1 + 2 * (m & 1));
                    y2 = y + gv.hy / 2.0 * ((m >> 1) & 1) * (-1 + 2 * (m & 1));
                    diffeo(gm, cx, cy, x2, y2, x2, y2);
                    v += herm_eval(gf, H, x2, y2) / 6.0;
                }
            } else if (molly == 8) {
                v = 0;
                for (int m = 0; m < 9; ++m) {
                    x2 = x + gv.hx * (-1 + m % 3) / 2.0;
                    y2 = y + gv.hy * (-1 + m / 3) / 2.0;
                    diffeo(gm, cx, cy, x2, y2, x2, y2);
                    // reference: (1 + (m%3)&1) parses as ((1 + m%3) & 1)
                    v += ((1 + (m % 3)) & 1) * ((1 + (m / 3)) & 1) * herm_eval(gf, H, x2, y2) / 16.0;
                }
            } else {
                diffeo(gm, cx, cy, x, y, x2, y2);
                v = herm_eval(gf, H, x2, y2);
            }
            out[(long)iy * gv.nx + ix] = v;
        }
    }
}

// psi solve, Cuda-CMM/src/simulation/cmm-simulation-host.cu:333-375 with
// k_fft_grid_move (cmm-fft.cu:294-323), k_fft_cut_off_scale_h (:267-286), k_fft_iLap_h (:55-78)
void psi_from_vort(const Grid &gm, const Grid &gv, const Grid &gf, const Grid &gp, const double *cx, const double *cy,
                   const double *H, int molly, double fcut, double *psi) {
    std::vector<double> w(gv.n);
    sample_vort(gm, gv, gf, cx, cy, H, molly, w.data());
    const int nhv = gv.nx / 2 + 1, nhp = gp.nx / 2 + 1;
    std::vector<cplx> W((size_t)nhv * gv.ny), P((size_t)nhp * gp.ny);
    rfft2(gv.nx, gv.ny, w.data(), W.data());
    for (auto &v : W) v = cplx(v.real() / (double)gv.n, v.imag() / (double)gv.n);
    if (gv.nx != gp.nx || gv.ny != gp.ny) {
        for (int iyc = 0; iyc < gp.ny; ++iyc)
            for (int ixc = 0; ixc < nhp; ++ixc) {
                long o = (long)iyc * nhp + ixc;
                int mx = ixc + (ixc > gp.nx / 2.0) * (gp.nx - 2 * ixc);
                int my = iyc + (iyc > gp.ny / 2.0) * (gp.ny - 2 * iyc);
                if (mx > gv.nx / 2.0 || my > gv.ny / 2.0) { P[o] = 0; continue; }
                int xs = ixc + (ixc > gp.nx / 2.0) * (gv.nx - gp.nx);
                int ys = iyc + (iyc > gp.ny / 2.0) * (gv.ny - gp.ny);
                P[o] = W[(long)ys * nhv + xs];
            }
    } else {
        P = W;
    }
    for (int iy = 0; iy < gp.ny; ++iy)
        for (int ix = 0; ix < nhp; ++ix) {
            long i = (long)iy * nhp + ix;
            int mx = ix + (ix > gp.nx / 2.0) * (gp.nx - 2 * ix);
            int my = iy + (iy > gp.ny / 2.0) * (gp.ny - 2 * iy);
            if ((mx * mx + my * my) > fcut * fcut || i == 0) P[i] = 0;
        }
    const double fx = kfac(gp, 0), fy = kfac(gp, 1);
    for (int iy = 0; iy < gp.ny; ++iy)
        for (int ix = 0; ix < nhp; ++ix) {
            long i = (long)iy * nhp + ix;
            if (i == 0) { P[0] = 0; continue; }
            double kx = fx * (ix - (ix > gp.nx / 2) * gp.nx);
            double ky = fy * (iy - (iy > gp.ny / 2) * gp.ny);
            double k2 = kx * kx + ky * ky;
            P[i] = cplx(-P[i].real() / k2, -P[i].imag() / k2);
        }
    fourier_hermite(gp, P.data(), psi);
}

// Cuda-CMM/src/numerical/cmm-particles.cu:119-276 (fluid particles; psi level 0 = t_{n+1}, 1 = t_n)
// 1-D transforms of length n (power of two), unnormalised.  irfft1 has the C2R semantics of the
// 2-D version: the imaginary parts of bins 0 and n/2 are dropped.
void rfft1(int n, const double *in, cplx *out) {
    FFTPlan pl = make_plan(n);
    std::vector<cplx> a(n);
    for (int i = 0; i < n; ++i) a[i] = cplx(in[i], 0.0);
    fft1(pl, a.data(), -1);
    for (int k = 0; k <= n / 2; ++k) out[k] = a[k];
}

void irfft1(int n, const cplx *in, double *out) {
    FFTPlan pl = make_plan(n);
    std::vector<cplx> a(n);
    a[0] = cplx(in[0].real(), 0.0);
    a[n / 2] = cplx(in[n / 2].real(), 0.0);
    for (int k = 1; k < n / 2; ++k) { a[k] = in[k]; a[n - k] = std::conj(in[k]); }
    fft1(pl, a.data(), +1);
    for (int i = 0; i < n; ++i) out[i] = a[i].real();
}

// Vlasov-Poisson psi (x, v) from the distribution function,
// evaluate_potential_from_density_hermite (Cuda-CMM/src/simulation/cmm-simulation-host.cu:385-415):
//  1. f(X(x)) on the vort grid (k_apply_map_and_sample_from_hermite, as for the vorticity);
//  2. integral_v (Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:570-586): I(x) = hv sum_v f - 1.
//     The reference's output row aliases the input's first row (both are the temporary array),
//     so the row v0 enters twice: I(x) = hv (f(x, v0) + sum_iv f(x, v_iv)) - 1;
//  3. solve_1D_laplace (:417-432): R2C, /NX, then division by +kx^2 (k_fft_iLap_hx_1D,
//     Cuda-CMM/src/numerical/cmm-fft.cu), DC set to 0;
//  4. calc_phi (:436-471): zero_padding_1D (Cuda-CMM/src/numerical/cmm-fft.cu:367-391) keeps the bins
//     k <= NX_vort/2 on the psi grid, then C2R -> phi(x); a vort grid finer than the psi grid is
//     rejected (reference error 1013);
//  5. compute_periodification_function (:534-571) with compute_bump_function (:478-494) on the psi
//     grid (NX == NY required, reference error 20230904): sigma(v) = 1/2 - 1/2 tanh(2 pi (|v| - 0.4 Lv)
//     / (0.1 Lv)), minus its mean, divided by its maximum; R2C, /NY, division by +ky^2, DC 0, C2R
//     -> H(v);
//  6. k_assemble_psi_periodified (Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:504-512):
//     psi(x, v) = phi(x) - H(v), then R2C, /N and fourier_hermite.
int vp_psi_from_f(const Grid &gm, const Grid &gv, const Grid &gf, const Grid &gp, const double *cx, const double *cy,
                  const double *Hf, int molly, double *psi) {
    if (gv.nx > gp.nx || gv.ny > gp.ny || gp.nx != gp.ny) return CMM_ERR_ARG;
    std::vector<double> f(gv.n);
    sample_vort(gm, gv, gf, cx, cy, Hf, molly, f.data());
    const int nv = gv.nx, np = gp.nx;
    std::vector<double> I(nv);
    for (int ix = 0; ix < nv; ++ix) {
        double a = f[ix];
        for (int iy = 0; iy < gv.ny; ++iy) a += f[(long)iy * nv + ix];
        a *= gv.hy;
        a += -1.0;
        I[ix] = a;
    }
    std::vector<cplx> F(nv / 2 + 1), P(np / 2 + 1, cplx(0, 0));
    rfft1(nv, I.data(), F.data());
    const double fkx = kTwoPi / (gv.hx * gv.nx);
    for (int k = 0; k <= nv / 2; ++k) {
        cplx v(F[k].real() / (double)nv, F[k].imag() / (double)nv);
        if (k == 0) { P[0] = 0; continue; }
        const double kx = fkx * k, k2 = kx * kx;
        P[k] = cplx(v.real() / k2, v.imag() / k2);
    }
    std::vector<double> phi(np);
    irfft1(np, P.data(), phi.data());
    // bump function and its double integral along v
    const int ny = gp.ny;
    const double Lv = gp.y1 - gp.y0;
    std::vector<double> sg(ny);
    double mean = 0;
    for (int iy = 0; iy < ny; ++iy) {
        const double v = gp.y0 + iy * gp.hy;
        sg[iy] = 0.5 - 0.5 * std::tanh(2 * kPi * (std::fabs(v) - 0.4 * Lv) / (0.1 * Lv));
        mean += sg[iy];
    }
    mean /= (double)ny;
    double mx = -1e300;
    for (int iy = 0; iy < ny; ++iy) { sg[iy] -= mean; mx = std::max(mx, sg[iy]); }
    for (int iy = 0; iy < ny; ++iy) sg[iy] /= mx;
    std::vector<cplx> S(ny / 2 + 1);
    rfft1(ny, sg.data(), S.data());
    const double fky = kTwoPi / (gp.hy * gp.ny);
    for (int k = 0; k <= ny / 2; ++k) {
        cplx v(S[k].real() / (double)ny, S[k].imag() / (double)ny);
        if (k == 0) { S[0] = 0; continue; }
        const double ky = fky * k, k2 = ky * ky;
        S[k] = cplx(v.real() / k2, v.imag() / k2);
    }
    std::vector<double> H(ny);
    irfft1(ny, S.data(), H.data());
    std::vector<double> real(gp.n);
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < np; ++ix) real[(long)iy * np + ix] = phi[ix] - H[iy];
    const int nh = gp.nx / 2 + 1;
    std::vector<cplx> R((size_t)nh * gp.ny);
    rfft2(gp.nx, gp.ny, real.data(), R.data());
    for (auto &v : R) v = cplx(v.real() / (double)gp.n, v.imag() / (double)gp.n);
    fourier_hermite(gp, R.data(), psi);
    return CMM_OK;
}

void particle_step(const Grid &g, const double *psi, int L, const double *Lp, double dt, int scheme, double p[2]) {
    const long N4 = 4 * g.n;
    double u[8], k[8];
    switch (scheme) {
    case 10:
        herm_grad(g, psi + N4, 1, p[0], p[1], u);
        p[0] += dt * u[0]; p[1] += dt * u[1];
        break;
    case 20:
        herm_grad(g, psi + N4, 1, p[0], p[1], k);
        k[2] = p[0] + dt * k[0]; k[3] = p[1] + dt * k[1];
        herm_grad(g, psi, 1, k[2], k[3], k + 2);
        p[0] += dt * (k[0] + k[2]) / 2; p[1] += dt * (k[1] + k[3]) / 2;
        break;
    case 30: {
        herm_grad(g, psi + N4, 1, p[0], p[1], k);
        double q[2] = {p[0] + dt * k[0] / 2.0, p[1] + dt * k[1] / 2.0};
        herm_grad(g, psi, L, q[0], q[1], u);
        lsum(Lp, u, L, k[2], k[3]);
        q[0] = p[0] - dt * k[0] + 2 * dt * k[2]; q[1] = p[1] - dt * k[1] + 2 * dt * k[3];
        herm_grad(g, psi, 1, q[0], q[1], k + 4);
        p[0] += dt * (k[0] + 4 * k[2] + k[4]) / 6.0;
        p[1] += dt * (k[1] + 4 * k[3] + k[5]) / 6.0;
        break;
    }
    case 40: {
        herm_grad(g, psi + N4, 1, p[0], p[1], k);
        double q[2] = {p[0] + dt * k[0] / 2.0, p[1] + dt * k[1] / 2.0};
        herm_grad(g, psi, L, q[0], q[1], u);
        lsum(Lp, u, L, k[2], k[3]);
        q[0] = p[0] + dt * k[2] / 2.0; q[1] = p[1] + dt * k[3] / 2.0;
        herm_grad(g, psi, L, q[0], q[1], u);
        lsum(Lp, u, L, k[4], k[5]);
        q[0] = p[0] + dt * k[4]; q[1] = p[1] + dt * k[5];
        herm_grad(g, psi, 1, q[0], q[1], k + 6);
        p[0] += dt * (k[0] + 2 * k[2] + 2 * k[4] + k[6]) / 6.0;
        p[1] += dt * (k[1] + 2 * k[3] + 2 * k[5] + k[7]) / 6.0;
        break;
    }
    case 31: {
        herm_grad(g, psi + N4, 1, p[0], p[1], k);
        double q[2] = {p[0] + dt * k[0], p[1] + dt * k[1]};
        herm_grad(g, psi, 1, q[0], q[1], k + 2);
        q[0] = p[0] + dt * (k[0] - 2 * k[2]); q[1] = p[1] + dt * (k[1] - 2 * k[3]);
        herm_grad(g, psi + 2 * N4, 1, q[0], q[1], k + 4);
        p[0] += dt * (8 * k[0] + 5 * k[2] - k[4]) / 12.0;
        p[1] += dt * (8 * k[1] + 5 * k[3] - k[5]) / 12.0;
        break;
    }
    case 41: {
        herm_grad(g, psi + N4, 1, p[0], p[1], k);
        double q[2] = {p[0] + dt * k[0], p[1] + dt * k[1]};
        herm_grad(g, psi, 1, q[0], q[1], k + 2);
        q[0] = p[0] + dt / 8.0 * (3 * k[0] + k[2]); q[1] = p[1] + dt / 8.0 * (3 * k[1] + k[3]);
        herm_grad(g, psi, L, q[0], q[1], u);
        lsum(Lp, u, L, k[4], k[5]);
        q[0] = p[0] + dt / 3.0 * (-k[2] + 4 * k[4]); q[1] = p[1] + dt / 3.0 * (-k[3] + 4 * k[5]);
        herm_grad(g, psi, 1, q[0], q[1], k + 6);
        p[0] += dt * (2 * k[0] - k[2] + 8 * k[4] + 3 * k[6]) / 12.0;
        p[1] += dt * (2 * k[1] - k[3] + 8 * k[5] + 3 * k[7]) / 12.0;
        break;
    }
    default:
        p[0] = p[1] = 0;
    }
    const double LX = g.x1 - g.x0, LY = g.y1 - g.y0;
    p[0] = p[0] - std::floor((p[0] - g.x0) / LX) * LX;
    p[1] = p[1] - std::floor((p[1] - g.y0) / LY) * LY;
}

// Inertial particles, Particle_advect_inertia (Cuda-CMM/src/numerical/cmm-particles.cu:280-551), written
// here as explicit Runge-Kutta tableaus over the state (x, y, vx, vy) with
//   F(state) = (vx, vy, -(vx - ux)/tau, -(vy - uy)/tau),
// where each stage reads the fluid velocity u from its own source: psi level 1 (t_n), level 0
// (t_{n+1}), level 2 (t_n - dt) or the Lagrange combination Lp of the levels to t_n + dt/2.
// Tableaus (a = stage offsets, b = weights, src per stage):
//   10 EulerExp  src (1)          b (1)
//   20 Heun      src (1, 0)       a21 = 1                         b (1, 1)/2
//   30 RK3       src (1, h, 0)    a21 = 1/2; a31 = -1, a32 = 2    b (1, 4, 1)/6
//   40 RK4       src (1, h, h, 0) a21 = a32 = 1/2, a43 = 1        b (1, 2, 2, 1)/6
//   31 RK3Mod    src (1, 0, 2)    a21 = 1; a31 = 1, a32 = -2      b (8, 5, -1)/12
//   41 RK4Mod    src (1, 0, h, 0) a21 = 1; a31 = 3/8, a32 = 1/8; a42 = -1/3, a43 = 4/3
//                                                                 b (2, -1, 8, 3)/12
void particle_step_inertial(const Grid &g, const double *psi, int L, const double *Lp, double dt, int scheme, double tau,
                            double p[2], double v[2]) {
    const long N4 = 4 * g.n;
    enum { HALF = -1 };
    int ns = 0;
    int src[4] = {1, 1, 1, 1};
    double a[4][4] = {{0}}, b[4] = {0};
    switch (scheme) {
    case 10: ns = 1; b[0] = 1; break;
    case 20: ns = 2; src[1] = 0; a[1][0] = 1; b[0] = b[1] = 0.5; break;
    case 30:
        ns = 3; src[1] = HALF; src[2] = 0;
        a[1][0] = 0.5; a[2][0] = -1; a[2][1] = 2;
        b[0] = 1.0 / 6; b[1] = 4.0 / 6; b[2] = 1.0 / 6;
        break;
    case 40:
        ns = 4; src[1] = src[2] = HALF; src[3] = 0;
        a[1][0] = 0.5; a[2][1] = 0.5; a[3][2] = 1;
        b[0] = b[3] = 1.0 / 6; b[1] = b[2] = 2.0 / 6;
        break;
    case 31:
        ns = 3; src[1] = 0; src[2] = 2;
        a[1][0] = 1; a[2][0] = 1; a[2][1] = -2;
        b[0] = 8.0 / 12; b[1] = 5.0 / 12; b[2] = -1.0 / 12;
        break;
    case 41:
        ns = 4; src[1] = 0; src[2] = HALF; src[3] = 0;
        a[1][0] = 1; a[2][0] = 3.0 / 8; a[2][1] = 1.0 / 8; a[3][1] = -1.0 / 3; a[3][2] = 4.0 / 3;
        b[0] = 2.0 / 12; b[1] = -1.0 / 12; b[2] = 8.0 / 12; b[3] = 3.0 / 12;
        break;
    default:
        p[0] = p[1] = v[0] = v[1] = 0;
        break;
    }
    if (ns > 0) {
        double K[4][4];  // stage derivatives of (x, y, vx, vy)
        for (int st = 0; st < ns; ++st) {
            double z[4] = {p[0], p[1], v[0], v[1]};
            for (int j = 0; j < st; ++j)
                for (int q = 0; q < 4; ++q) z[q] += dt * a[st][j] * K[j][q];
            double u[8], ux, uy;
            if (src[st] == HALF) {
                herm_grad(g, psi, L, z[0], z[1], u);
                lsum(Lp, u, L, ux, uy);
            } else {
                herm_grad(g, psi + src[st] * N4, 1, z[0], z[1], u);
                ux = u[0];
                uy = u[1];
            }
            K[st][0] = z[2];
            K[st][1] = z[3];
            K[st][2] = -(z[2] - ux) / tau;
            K[st][3] = -(z[3] - uy) / tau;
        }
        double dz[4] = {0, 0, 0, 0};
        for (int st = 0; st < ns; ++st)
            for (int q = 0; q < 4; ++q) dz[q] += b[st] * K[st][q];
        p[0] += dt * dz[0];
        p[1] += dt * dz[1];
        v[0] += dt * dz[2];
        v[1] += dt * dz[3];
    }
    const double LX = g.x1 - g.x0, LY = g.y1 - g.y0;
    p[0] = p[0] - std::floor((p[0] - g.x0) / LX) * LX;
    p[1] = p[1] - std::floor((p[1] - g.y0) / LY) * LY;
}

// SplitMix64 finaliser (Steele, Lea, Flood 2014)
inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
inline double unit_from(uint64_t seed, uint64_t key) {
    return (double)(mix64(mix64(seed) + key) >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace

// ================================================================================================
extern "C" {

double cmmo_grid_h(int n, double a, double b) { return (b - a) / (float)n; }

double cmmo_lagrange_coeff(const double *t, double t_next, int loop_ctr, int i_point, int l_order) {
    double coeff = 1;
    for (int i_p = 0; i_p < l_order; ++i_p)
        if (i_p != i_point) coeff *= (t_next - t[loop_ctr - i_p]) / (t[loop_ctr - i_point] - t[loop_ctr - i_p]);
    return coeff;
}

int cmmo_advect_coeffs_fill(const double *t, const double *dt, int loop_ctr, int l_order, int tin, int muo,
                            double eps, int direction, cmm_advect_coeffs *c) {
    if (!t || !dt || !c || l_order < 1 || l_order > 4 || muo < 0 || muo > 2) return CMM_ERR_ARG;
    std::memset(c, 0, sizeof(*c));
    int lc = loop_ctr + l_order - 1;
    for (int i = 0; i < l_order; ++i) {
        c->L1[i] = cmmo_lagrange_coeff(t, t[lc + 1], lc, i, l_order);
        c->L12[i] = cmmo_lagrange_coeff(t, t[lc] + dt[lc + 1] / 2.0, lc, i, l_order);
    }
    double h[3] = {0, 0, 0};
    if (muo == 2) { h[0] = 3.0 / 8.0; h[1] = -3.0 / 20.0; h[2] = 1.0 / 40.0; }
    else if (muo == 1) { h[0] = 1.0 / 3.0; h[1] = -1.0 / 12.0; }
    else { h[0] = 1.0 / 4.0; }
    for (int f = 0; f < 4 + 4 * muo; ++f) {
        c->c1[f] = h[f / 4];
        c->cx[f] = h[f / 4] * (1 - 2 * ((f / 2) % 2)) / eps / double(f / 4 + 1);
        c->cy[f] = h[f / 4] * (1 - 2 * (((f + 1) / 2) % 2)) / eps / double(f / 4 + 1);
        c->cxy[f] = h[f / 4] * (1 - 2 * (f % 2)) / eps / eps / double(f / 4 + 1) / double(f / 4 + 1);
    }
    c->dt = dt[lc + 1];
    c->eps = eps;
    c->time_integration_num = tin;
    c->map_update_order_num = muo;
    c->l_order = l_order;
    c->direction = direction;
    return CMM_OK;
}

double cmmo_hermite_eval(cmm_grid g, const double *H, double x, double y) { return herm_eval(mk(g), H, x, y); }
void cmmo_hermite_grad(cmm_grid g, const double *H, int nl, double x, double y, double *u) { herm_grad(mk(g), H, nl, x, y, u); }
void cmmo_diffeo_eval(cmm_grid g, const double *Hx, const double *Hy, double x, double y, double *xo, double *yo) {
    diffeo(mk(g), Hx, Hy, x, y, *xo, *yo);
}
double cmmo_diffeo_det(cmm_grid g, const double *Hx, const double *Hy, double x, double y) { return diffeo_det(mk(g), Hx, Hy, x, y); }
double cmmo_init_vorticity(double x, double y, int ic, const double *prm) { return init_vorticity(x, y, ic, prm); }

int cmmo_map_init_identity(cmm_grid gg, double *cx, double *cy) {
    Grid g = mk(gg);
#pragma omp parallel for
    for (int iy = 0; iy < g.ny; ++iy)
        for (int ix = 0; ix < g.nx; ++ix) {
            long i = (long)iy * g.nx + ix;
            cx[i] = g.x0 + ix * g.hx; cy[i] = g.y0 + iy * g.hy;
            cx[g.n + i] = 1; cy[2 * g.n + i] = 1;
            cx[2 * g.n + i] = 0; cy[g.n + i] = 0; cx[3 * g.n + i] = 0; cy[3 * g.n + i] = 0;
        }
    return CMM_OK;
}

int cmmo_map_advect(cmm_grid ggm, cmm_grid ggp, const double *cx, const double *cy, const double *psi,
                    const cmm_advect_coeffs *c, double *nx_, double *ny_) {
    if (!c || (c->direction != -1 && c->direction != 1)) return CMM_ERR_UNSUPPORTED;
    Grid gm = mk(ggm), gp = mk(ggp);
    const int nfoot = 4 + 4 * c->map_update_order_num;
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < gm.ny; ++iy) {
        for (int ix = 0; ix < gm.nx; ++ix) {
            double ax[4] = {0, 0, 0, 0}, ay[4] = {0, 0, 0, 0};
            for (int k = 0; k < nfoot; ++k) {
                double xe[2], xf[2];
                xe[0] = gm.x0 + ix * gm.hx + ((k >> 2) + 1) * c->eps * (1 - 2 * ((k >> 1) & 1));
                xe[1] = gm.y0 + iy * gm.hy + ((k >> 2) + 1) * c->eps * (1 - 2 * (((k + 1) >> 1) & 1));
                double px, py;
                if (c->direction == -1) {
                    // backward: trace back, then the old map, Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:232-252
                    step_backward(gp, psi, *c, xe, xf);
                    diffeo(gm, cx, cy, xf[0], xf[1], px, py);
                } else {
                    // forward: the old map first, then forward in time, :254-274
                    diffeo(gm, cx, cy, xe[0], xe[1], xf[0], xf[1]);
                    double xo[2];
                    step_forward(gp, psi, *c, xf, xo);
                    px = xo[0];
                    py = xo[1];
                }
                ax[0] += px * c->c1[k]; ay[0] += py * c->c1[k];
                ax[1] += px * c->cx[k]; ay[1] += py * c->cx[k];
                ax[2] += px * c->cy[k]; ay[2] += py * c->cy[k];
                ax[3] += px * c->cxy[k]; ay[3] += py * c->cxy[k];
            }
            long i = (long)iy * gm.nx + ix;
            for (int p = 0; p < 4; ++p) { nx_[p * gm.n + i] = ax[p]; ny_[p * gm.n + i] = ay[p]; }
        }
    }
    return CMM_OK;
}

// One footpoint of the map advection: the backward (direction -1) or forward (+1) time
// integrator applied to the point (x, y), Cuda-CMM/src/numerical/cmm-timestep.cu:28-470, without
// the old map or the stencil (what k_advect_using_stream_hermite feeds them,
// Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:232-274)
int cmmo_footpoint(cmm_grid ggp, const double *psi, const cmm_advect_coeffs *c, double x, double y, double *out) {
    if (!c || !psi || !out || (c->direction != -1 && c->direction != 1)) return CMM_ERR_ARG;
    Grid gp = mk(ggp);
    const double xin[2] = {x, y};
    if (c->direction == -1)
        step_backward(gp, psi, *c, xin, out);
    else
        step_forward(gp, psi, *c, xin, out);
    return CMM_OK;
}

// max over cell centres of the check grid of |x - X_b(X_f(x))|, k_invertibility_check and
// invertibility_check (Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:177-198,
// Cuda-CMM/src/simulation/cmm-simulation-host.cu:52-60)
double cmmo_invert_error(cmm_grid ggb, cmm_grid ggf, cmm_grid ggc, const double *bx, const double *by, const double *fx,
                         const double *fy) {
    Grid gb = mk(ggb), gf = mk(ggf), gc = mk(ggc);
    double mx = 0.0;
#pragma omp parallel for reduction(max : mx)
    for (int iy = 0; iy < gc.ny; ++iy)
        for (int ix = 0; ix < gc.nx; ++ix) {
            double x = gc.x0 + ix * gc.hx + 0.5 * gc.hx;
            double y = gc.y0 + iy * gc.hy + 0.5 * gc.hy;
            double xi, yi;
            diffeo(gf, fx, fy, x, y, xi, yi);
            diffeo(gb, bx, by, xi, yi, xi, yi);
            double d = std::hypot(x - xi, y - yi);
            if (d > mx) mx = d;
        }
    return mx;
}

double cmmo_incomp_error(cmm_grid ggm, cmm_grid ggc, const double *cx, const double *cy) {
    Grid gm = mk(ggm), gc = mk(ggc);
    double mx = 0.0;
#pragma omp parallel for reduction(max : mx)
    for (int iy = 0; iy < gc.ny; ++iy)
        for (int ix = 0; ix < gc.nx; ++ix) {
            double x = gm.x0 + ix * gc.hx + 0.5 * gc.hx;
            double y = gm.y0 + iy * gc.hy + 0.5 * gc.hy;
            double d = std::fabs(1 - diffeo_det(gm, cx, cy, x, y));
            if (d > mx) mx = d;
        }
    return mx;
}

int cmmo_compose_positions(cmm_grid gm, cmm_grid gt, const double *cx, const double *cy, const double *sx,
                           const double *sy, int ns, double *xy) {
    compose(mk(gm), mk(gt), cx, cy, sx, sy, ns, xy);
    return CMM_OK;
}

int cmmo_forward_positions(cmm_grid gm, cmm_grid gt, const double *cx, const double *cy, const double *sx,
                           const double *sy, int ns, double *xy) {
    compose_forward(mk(gm), mk(gt), cx, cy, sx, sy, ns, xy);
    return CMM_OK;
}

int cmmo_forward_points(cmm_grid gm, const double *bounds, const double *cx, const double *cy, const double *sx,
                        const double *sy, int ns, int first, int cur, int n, const double *in, double *out) {
    forward_points(mk(gm), bounds, cx, cy, sx, sy, ns, first, cur, n, in, out);
    return CMM_OK;
}

int cmmo_sample_init(cmm_grid gt, const double *xy, int ic, const double *prm, const double *disc, cmm_grid gd,
                     int discrete, double *out) {
    sample_init(mk(gt), xy, ic, prm, disc, mk(gd), discrete, out);
    return CMM_OK;
}

int cmmo_sample_vort(cmm_grid gm, cmm_grid gt, cmm_grid gf, const double *cx, const double *cy, const double *H,
                     int molly, double *out) {
    sample_vort(mk(gm), mk(gt), mk(gf), cx, cy, H, molly, out);
    return CMM_OK;
}

int cmmo_rfft2(int nx, int ny, const double *in, double *out) {
    if (!pow2(nx) || !pow2(ny)) return CMM_ERR_ARG;
    rfft2(nx, ny, in, reinterpret_cast<cplx *>(out));
    return CMM_OK;
}

int cmmo_irfft2(int nx, int ny, const double *in, double *out) {
    if (!pow2(nx) || !pow2(ny)) return CMM_ERR_ARG;
    irfft2(nx, ny, reinterpret_cast<const cplx *>(in), out);
    return CMM_OK;
}

int cmmo_hermite_from_real(cmm_grid gg, const double *real, double *herm) {
    Grid g = mk(gg);
    if (!pow2(g.nx) || !pow2(g.ny)) return CMM_ERR_ARG;
    std::vector<cplx> F((size_t)(g.nx / 2 + 1) * g.ny);
    rfft2(g.nx, g.ny, real, F.data());
    for (auto &v : F) v = cplx(v.real() / (double)g.n, v.imag() / (double)g.n);
    fourier_hermite(g, F.data(), herm);
    return CMM_OK;
}

int cmmo_psi_from_vort(cmm_grid gm, cmm_grid gv, cmm_grid gf, cmm_grid gp, const double *cx, const double *cy,
                       const double *H, int molly, double fcut, double *psi) {
    if (!pow2(gv.nx) || !pow2(gp.nx) || gv.nx != gv.ny || gp.nx != gp.ny) return CMM_ERR_ARG;
    psi_from_vort(mk(gm), mk(gv), mk(gf), mk(gp), cx, cy, H, molly, fcut, psi);
    return CMM_OK;
}

// Hermite planes of F, of i_laplacian(F) (-F/|k|^2, DC 0) or of laplacian(F) (-|k|^2 F),
// Cuda-CMM/src/numerical/cmm-fft.cu:28-78, :398-474
int cmmo_spectral_from_real(cmm_grid gg, const double *real, int op, double *herm) {
    const Grid g = mk(gg);
    if (!pow2(g.nx) || !pow2(g.ny) || op < 0 || op > 2) return CMM_ERR_ARG;
    const int nh = g.nx / 2 + 1;
    std::vector<cplx> F((size_t)nh * g.ny);
    rfft2(g.nx, g.ny, real, F.data());
    const double fx = kfac(g, 0), fy = kfac(g, 1);
    for (int iy = 0; iy < g.ny; ++iy)
        for (int ix = 0; ix < nh; ++ix) {
            const long i = (long)iy * nh + ix;
            cplx v(F[i].real() / (double)g.n, F[i].imag() / (double)g.n);
            const double kx = fx * (ix - (ix > g.nx / 2) * g.nx), ky = fy * (iy - (iy > g.ny / 2) * g.ny);
            const double k2 = kx * kx + ky * ky;
            if (op == 1) v = i == 0 ? cplx(0, 0) : cplx(-v.real() / k2, -v.imag() / k2);
            if (op == 2) v = cplx(-v.real() * k2, -v.imag() * k2);
            F[i] = v;
        }
    fourier_hermite(g, F.data(), herm);
    return CMM_OK;
}

int cmmo_vp_psi_from_f(cmm_grid gm, cmm_grid gv, cmm_grid gf, cmm_grid gp, const double *cx, const double *cy,
                       const double *f_herm, int molly, double *psi_out) {
    return vp_psi_from_f(mk(gm), mk(gv), mk(gf), mk(gp), cx, cy, f_herm, molly, psi_out);
}

int cmmo_particles_advect(cmm_grid gg, const double *psi, int L, const double *Lp, double dt, int scheme, int n,
                          double *pos) {
    Grid g = mk(gg);
#pragma omp parallel for
    for (int i = 0; i < n; ++i) particle_step(g, psi, L, Lp, dt, scheme, pos + 2 * i);
    return CMM_OK;
}

int cmmo_particles_inertial(cmm_grid gg, const double *psi, int L, const double *Lp, double dt, int scheme, double tau,
                            int n, double *pos, double *vel) {
    Grid g = mk(gg);
#pragma omp parallel for
    for (int i = 0; i < n; ++i) particle_step_inertial(g, psi, L, Lp, dt, scheme, tau, pos + 2 * i, vel + 2 * i);
    return CMM_OK;
}

// start velocity of inertial particles: the velocity of psi (level 0) or zero,
// Cuda-CMM/src/numerical/cmm-particles.cu:49-59 and Cuda-CMM/src/simulation/cmm-euler2d.cu:620-632
int cmmo_particles_vel_init(cmm_grid gg, const double *psi0, int n, int fluid, const double *pos, double *vel) {
    Grid g = mk(gg);
    for (int i = 0; i < n; ++i) {
        double u[2] = {0, 0};
        if (fluid) herm_grad(g, psi0, 1, pos[2 * i], pos[2 * i + 1], u);
        vel[2 * i] = u[0];
        vel[2 * i + 1] = u[1];
    }
    return CMM_OK;
}

/*
 * Random-spectrum field on an n*n grid over [0,2pi)^2 (no reference counterpart):
 *   omega_hat(kx,ky) = A sqrt(k E(k)/pi) exp(i theta),  E(k) = k^4 exp(-2 (k/6)^2),
 *   theta = 2 pi U(seed, kx, ky) for kx > 0; theta(0,-ky) = -theta(0,ky) so the kx = 0 column is
 *   Hermitian; modes with |kx|,|ky| <= min(64, n/2-1) only (E < 1e-140 beyond); A sets the
 *   enstrophy 0.5 h^2 sum omega^2 = 2 pi^2 sum_full |omega_hat|^2 to 1.
 * U(seed, kx, ky) = top 53 bits of mix64(mix64(seed) + ((kx+32768) << 32 | (ky+32768))) / 2^53.
 */
int cmmo_ic_random_spectrum(int n, uint64_t seed, double *out) {
    if (!pow2(n) || n < 8) return CMM_ERR_ARG;
    const int nh = n / 2 + 1;
    const int km = std::min(64, n / 2 - 1);
    std::vector<cplx> F((size_t)nh * n, cplx(0, 0));
    double s2 = 0;
    for (int kx = 0; kx <= km; ++kx)
        for (int ky = -km; ky <= km; ++ky) {
            if (kx == 0 && ky <= 0) continue;
            double k = std::sqrt((double)kx * kx + (double)ky * ky);
            double E = k * k * k * k * std::exp(-2.0 * (k / 6.0) * (k / 6.0));
            double a = std::sqrt(k * E / kPi);
            uint64_t key = ((uint64_t)(kx + 32768) << 32) | (uint64_t)(ky + 32768);
            double th = kTwoPi * unit_from(seed, key);
            cplx v(a * std::cos(th), a * std::sin(th));
            int row = ky < 0 ? ky + n : ky;
            F[(size_t)row * nh + kx] = v;
            if (kx == 0) F[(size_t)(n - ky) * nh] = std::conj(v);
            s2 += (kx == 0 ? 2.0 : 2.0) * std::norm(v);
        }
    // kx = 0 pairs were both stored (2 |v|^2); kx > 0 modes count twice in the full spectrum
    double A = 1.0 / std::sqrt(2.0 * kPi * kPi * s2);
    for (auto &v : F) v *= A;
    irfft2(n, n, F.data(), out);
    return CMM_OK;
}

// ------------------------------------------------------------------------------------------------
// Full Euler run in the order of Cuda-CMM/src/simulation/cmm-euler2d.cu:495-960
int cmmo_euler_run(const cmmo_euler_params *p, cmmo_euler_result *r) {
    if (!p || !r) return CMM_ERR_ARG;
    if (p->sim == 1 && (p->g_vort.nx > p->g_psi.nx || p->g_psi.nx != p->g_psi.ny)) return CMM_ERR_ARG;
    const Grid gc = mk(p->g_coarse), gf = mk(p->g_fine), gp = mk(p->g_psi);
    const int L = p->l_order;
    const long Nc = gc.n, Nf = gf.n, Np = gp.n;
    std::vector<double> cx(4 * Nc), cy(4 * Nc), ncx(4 * Nc), ncy(4 * Nc);
    std::vector<double> W(4 * Nf), psi((size_t)4 * Np * L, 0.0), xy(2 * Nf), samp(Nf);
    std::vector<double> stx, sty;
    // forward map and its stack (only with p->forward_map)
    std::vector<double> fx(4 * Nc), fy(4 * Nc), nfx(4 * Nc), nfy(4 * Nc), fstx, fsty;
    int nstack = 0;
    r->n_warmup = 0;

    auto translate = [&]() {
        compose(gc, gf, cx.data(), cy.data(), stx.data(), sty.data(), nstack, xy.data());
        sample_init(gf, xy.data(), p->ic, p->ic_params, p->disc_herm, mk(p->g_disc), p->discrete, samp.data());
        cmmo_hermite_from_real(p->g_fine, samp.data(), W.data());
    };
    auto eval_psi = [&]() {
        if (p->sim == 1)
            vp_psi_from_f(gc, mk(p->g_vort), gf, gp, cx.data(), cy.data(), W.data(), p->molly_stencil, psi.data());
        else
            psi_from_vort(gc, mk(p->g_vort), gf, gp, cx.data(), cy.data(), W.data(), p->molly_stencil,
                          p->freq_cut_psi, psi.data());
    };
    auto advect = [&](const double *t, const double *dt, int loop_ctr, int lo, int tin) {
        cmm_advect_coeffs c;
        cmmo_advect_coeffs_fill(t, dt, loop_ctr, lo, tin, p->map_update_order_num, p->map_epsilon, -1, &c);
        cmmo_map_advect(p->g_coarse, p->g_psi, cx.data(), cy.data(), psi.data(), &c, ncx.data(), ncy.data());
        std::swap(cx, ncx); std::swap(cy, ncy);
    };

    cmmo_map_init_identity(p->g_coarse, cx.data(), cy.data());
    cmmo_map_init_identity(p->g_coarse, fx.data(), fy.data());
    translate();
    eval_psi();

    // Lagrange warm-up, Cuda-CMM/src/simulation/cmm-euler2d.cu:520-619
    if (L > 1) {
        double dtn = p->dt * std::pow(-1.0, L);
        double t_init[5] = {0, dtn, 2 * dtn, 3 * dtn, 4 * dtn};
        double dt_init[5] = {dtn, dtn, dtn, dtn, dtn};
        for (int order = 1; order <= L; ++order) {
            int nsteps = order - (order == L);
            if (!p->lagrange_init_higher_order && order != L) {
                for (int i = 0; i < 5; ++i) { dt_init[i] = -dt_init[i]; t_init[i] = -t_init[i]; }
                continue;
            }
            int tin = 10, lo = 1;
            if (p->lagrange_init_higher_order) {
                lo = order;
                tin = order == 1 ? 10 : order == 2 ? 21 : order == 3 ? 30 : 40;
            }
            for (int s = 1; s <= nsteps; ++s) {
                advect(t_init, dt_init, 0, lo, tin);
                double e = cmmo_incomp_error(p->g_coarse, p->g_fine, cx.data(), cy.data());
                if (r->n_warmup < 8) r->warmup_incomp[r->n_warmup++] = e;
                for (int il = 1; il <= nsteps; ++il)
                    std::memcpy(&psi[(size_t)4 * Np * (nsteps - il + 1)], &psi[(size_t)4 * Np * (nsteps - il)],
                                sizeof(double) * 4 * Np);
                eval_psi();
            }
            for (int i = 0; i < 5; ++i) { dt_init[i] = -dt_init[i]; t_init[i] = -t_init[i]; }
            // k_swap_h of the outer (and for order 4 the inner) history slots
            std::swap_ranges(psi.begin(), psi.begin() + 4 * Np, psi.begin() + (size_t)4 * Np * nsteps);
            if (nsteps >= 3)
                std::swap_ranges(psi.begin() + 4 * Np, psi.begin() + 8 * Np, psi.begin() + 8 * Np);
            cmmo_map_init_identity(p->g_coarse, cx.data(), cy.data());
        }
    }

    // start velocities of inertial particles, Cuda-CMM/src/simulation/cmm-euler2d.cu:620-632
    const bool inertial = p->n_particles > 0 && p->particles && p->particle_tau != 0 && p->particle_vel;
    if (inertial && !p->particle_vel_given)
        cmmo_particles_vel_init(p->g_psi, psi.data(), p->n_particles, p->particle_init_vel, p->particles,
                                p->particle_vel);

    // main loop, Cuda-CMM/src/simulation/cmm-euler2d.cu:709-960
    const int iter_max = (int)(1.05 * std::ceil(p->final_time / p->dt));
    std::vector<double> tv(iter_max + L + 1), dtv(iter_max + L + 1);
    for (int i = 0; i < L; ++i) { tv[i] = -(L - 1 - i) * p->dt; dtv[i] = p->dt; }
    int loop = 0;
    bool cont = true;
    Grid gpart = gp;
    double dummy_t = 0;
    (void)dummy_t;
    while (p->final_time - tv[loop + L - 1] > p->dt * 1e-5 && loop < iter_max && cont) {
        if (p->max_steps != 0 && loop >= p->max_steps) break;  // max_steps < 0: warm-up only
        int lc = loop + L - 1;
        double dtn = p->dt;
        if (tv[lc] + dtn > p->final_time) dtn = p->final_time - tv[lc];
        tv[lc + 1] = tv[lc] + dtn;
        dtv[lc + 1] = dtn;
        advect(tv.data(), dtv.data(), loop, L, p->time_integration_num);
        double einc_f = 0, einv = 0;
        if (p->forward_map) {
            // the forward map with the same psi history and coefficients, :780-785
            cmm_advect_coeffs c;
            cmmo_advect_coeffs_fill(tv.data(), dtv.data(), loop, L, p->time_integration_num, p->map_update_order_num,
                                    p->map_epsilon, 1, &c);
            cmmo_map_advect(p->g_coarse, p->g_psi, fx.data(), fy.data(), psi.data(), &c, nfx.data(), nfy.data());
            std::swap(fx, nfx); std::swap(fy, nfy);
        }
        double einc = cmmo_incomp_error(p->g_coarse, p->g_fine, cx.data(), cy.data());
        if (p->forward_map) {
            // :799-802, both on the coarse grid
            einv = cmmo_invert_error(p->g_coarse, p->g_coarse, p->g_coarse, cx.data(), cy.data(), fx.data(), fy.data());
            einc_f = cmmo_incomp_error(p->g_coarse, p->g_coarse, fx.data(), fy.data());
        }
        if (einc > p->incomp_threshold && !p->skip_remapping) {
            if (nstack >= p->stack_capacity) {
                cont = false;
            } else {
                translate();
                stx.insert(stx.end(), cx.begin(), cx.end());
                sty.insert(sty.end(), cy.begin(), cy.end());
                nstack++;
                cmmo_map_init_identity(p->g_coarse, cx.data(), cy.data());
                if (p->forward_map) {
                    fstx.insert(fstx.end(), fx.begin(), fx.end());
                    fsty.insert(fsty.end(), fy.begin(), fy.end());
                    cmmo_map_init_identity(p->g_coarse, fx.data(), fy.data());
                }
            }
        }
        if (r->incomp_fwd) r->incomp_fwd[loop] = einc_f;
        if (r->invert) r->invert[loop] = einv;
        for (int il = 1; il < L; ++il)
            std::memcpy(&psi[(size_t)4 * Np * (L - il)], &psi[(size_t)4 * Np * (L - il - 1)], sizeof(double) * 4 * Np);
        eval_psi();
        // particles from their start time on (Cuda-CMM/src/numerical/cmm-particles.cu:63-87); inertial
        // ones get their start velocity again in the step reaching it (cmm-euler2d.cu:883-895)
        if (p->n_particles > 0 && p->particles) {
            double Lp[4] = {0, 0, 0, 0};
            for (int i = 0; i < L; ++i) Lp[i] = cmmo_lagrange_coeff(tv.data(), tv[lc] + dtv[lc + 1] / 2.0, lc + 1, i, L);
            const double t0 = p->particle_init_time;
            if (tv[lc + 1] + 1e-5 * dtv[lc + 1] >= t0) {
                if (inertial)
                    cmmo_particles_inertial(p->g_psi, psi.data(), L, Lp, dtv[lc + 1], p->particle_scheme, p->particle_tau,
                                            p->n_particles, p->particles, p->particle_vel);
                else
                    cmmo_particles_advect(p->g_psi, psi.data(), L, Lp, dtv[lc + 1], p->particle_scheme,
                                          p->n_particles, p->particles);
            }
            if (inertial && tv[lc] - t0 + p->dt * 1e-5 < 0 && tv[lc + 1] - t0 + p->dt * 1e-5 > 0)
                cmmo_particles_vel_init(p->g_psi, psi.data(), p->n_particles, p->particle_init_vel, p->particles,
                                        p->particle_vel);
        }
        if (r->time) r->time[loop] = tv[lc + 1];
        if (r->incomp) r->incomp[loop] = einc;
        if (r->map_counter) r->map_counter[loop] = nstack;
        loop++;
    }
    (void)gpart;
    r->steps = loop;
    if (r->chix) std::memcpy(r->chix, cx.data(), sizeof(double) * 4 * Nc);
    if (r->chiy) std::memcpy(r->chiy, cy.data(), sizeof(double) * 4 * Nc);
    if (r->psi) std::memcpy(r->psi, psi.data(), sizeof(double) * 4 * Np * L);
    if (r->w_herm) std::memcpy(r->w_herm, W.data(), sizeof(double) * 4 * Nf);
    if (r->vort_coarse) sample_vort(gc, gc, gf, cx.data(), cy.data(), W.data(), 0, r->vort_coarse);
    if (r->fchix) std::memcpy(r->fchix, fx.data(), sizeof(double) * 4 * Nc);
    if (r->fchiy) std::memcpy(r->fchiy, fy.data(), sizeof(double) * 4 * Nc);
    return CMM_OK;
}

}  // extern "C"
