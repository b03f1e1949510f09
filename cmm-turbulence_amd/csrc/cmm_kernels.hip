// HIP kernels of the CMM map-advection path for gfx950, and the C-ABI launchers declared in
// include/cmm.h.  One thread per grid point for the gather kernels; one workgroup per row or
// column for the LDS-resident FFT kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/cmm.h"
#include "cmm_core.hpp"
#include "cmm_internal.hpp"
#include "../../include/cmm.h"

using namespace cmm;

namespace cmm {

// ------------------------------------------------------------------------------------------------
// error state
static thread_local std::string g_last_error;

void set_error(const std::string &s) { g_last_error = s; }
const char *last_error() { return g_last_error.c_str(); }

int hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return CMM_ERR_HIP;
    }
    return CMM_OK;
}

// XCD-aware block order: the hardware deals consecutive workgroups round-robin over the 8 XCDs;
// this returns the logical index so that each XCD works on one contiguous range (shared L2 lines).
__device__ __forceinline__ int xcd_block(int bid, int nblocks) {
    const int xcd = bid & 7, slot = bid >> 3;
    const int per = nblocks >> 3, extra = nblocks & 7;
    return xcd * per + (xcd < extra ? xcd : extra) + slot;
}

// The tiles [t_begin, t_end) of this block when ntiles tiles are split into gridDim.x contiguous
// ranges (whole multiples of 256) and the ranges are dealt to the XCDs in contiguous runs
// (xcd_block): an XCD then reads one band of the map rows, so the rows shared by neighbouring
// tiles are fetched into one L2 rather than into several.
__device__ __forceinline__ void tile_range_of(long ntiles, int lb, int nblocks, long &t_begin, long &t_end) {
    const long per = ((ntiles + nblocks - 1) / nblocks + 255) / 256 * 256;
    t_begin = lb * per;
    t_end = min(ntiles, t_begin + per);
}

__device__ __forceinline__ void tile_range(long ntiles, long &t_begin, long &t_end) {
    tile_range_of(ntiles, xcd_block(blockIdx.x, gridDim.x), gridDim.x, t_begin, t_end);
}

// ------------------------------------------------------------------------------------------------
// Map kernels

// identity map, Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:27-41
__global__ void k_map_identity(DGrid g, double *__restrict__ cx, double *__restrict__ cy) {
    int ix = blockIdx.x * blockDim.x + threadIdx.x;
    int iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= g.nx || iy >= g.ny) return;
    int i = iy * g.nx + ix;
    cx[i] = g.x0 + ix * g.hx;
    cy[i] = g.y0 + iy * g.hy;
    cx[g.n + i] = 1.0;
    cy[2 * g.n + i] = 1.0;
    cx[2 * g.n + i] = 0.0;
    cy[g.n + i] = 0.0;
    cx[3 * g.n + i] = 0.0;
    cy[3 * g.n + i] = 0.0;
}

__device__ __forceinline__ void lsum(const double *w, const double *u, int nl, double &kx, double &ky) {
    kx = w[0] * u[0];
    ky = w[0] * u[1];
#pragma unroll
    for (int l = 1; l < 4; ++l) {  // at most 4 levels; constant indices keep u in registers
        if (l >= nl) break;
        kx += w[l] * u[2 * l];
        ky += w[l] * u[2 * l + 1];
    }
}

// Velocity sources for the footpoint integration.  w1 / w12: velocity extrapolated in time with the
// Lagrange weights L1 (to t_{n+1}) / L12 (to t_n + dt/2); lev(l): velocity of psi level l.
//
// VelLevels evaluates every level at the point and weights the velocities, the reference's order
// (Cuda-CMM/src/numerical/cmm-timestep.cu:175-207 with cmm-hermite.cu:396-420).
struct VelLevels {
    DGrid g;
    PsiLevels P;
    // when set, the Lagrange weights L1[4], L12[4] read from device memory instead of the
    // coefficient struct (steps replayed from a captured graph, whose arguments are frozen)
    const double *LW = nullptr;
    // herm_vel over L levels followed by lsum, unrolled over the (at most 4) levels so that no
    // array is indexed at run time
    __device__ __forceinline__ void w(const double *Lw, int L, double x, double y, double &kx, double &ky) const {
        Cell cc;
        find_cell(g, x, y, cc);
        double bx[4], by[4], dbx[4], dby[4];
        basis(cc.dx, bx);
        basis(cc.dy, by);
        basis_d(cc.dx, dbx);
        basis_d(cc.dy, dby);
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            if (l < L) {
                double v[16];
                load16(P.p[l], g.n, cc, v);
                const double uy = -hsum_v(v, dbx, by, g.hx, g.hy) * g.ihx;
                const double ux = hsum_v(v, bx, dby, g.hx, g.hy) * g.ihy;
                if (l == 0) {
                    kx = Lw[0] * ux;
                    ky = Lw[0] * uy;
                } else {
                    kx += Lw[l] * ux;
                    ky += Lw[l] * uy;
                }
            }
        }
    }
    __device__ __forceinline__ void w1(const cmm_advect_coeffs &c, double x, double y, double &kx, double &ky) const {
        if (LW) w(LW, c.l_order, x, y, kx, ky);
        else w(c.L1, c.l_order, x, y, kx, ky);
    }
    __device__ __forceinline__ void w12(const cmm_advect_coeffs &c, double x, double y, double &kx, double &ky) const {
        if (LW) w(LW + 4, c.l_order, x, y, kx, ky);
        else w(c.L12, c.l_order, x, y, kx, ky);
    }
    __device__ __forceinline__ void lev(int l, double x, double y, double &ux, double &uy) const {
        double u[2];
        herm_vel(g, PsiLevels{{P.p[l], P.p[l], P.p[l], P.p[l]}}, 1, x, y, u);
        ux = u[0];
        uy = u[1];
    }
};

// VelCombined reads the precombined fields C1 = sum_l L1[l] psi_l and C12 = sum_l L12[l] psi_l
// (Hermite interpolation is linear in the field, so this equals the weighted velocities up to
// rounding) and a copy C0 of level 0, all three node-interleaved: one 16-value gather per
// evaluation instead of one per level, issued as 8 16-byte loads instead of 16 8-byte loads.
// Only level 0 is read on its own by the combinable schemes.
template <bool P2>
struct VelCombined {
    DGrid g;
    const double *C1, *C12, *C0;
    __device__ __forceinline__ void w1(const cmm_advect_coeffs &, double x, double y, double &kx, double &ky) const {
        herm_vel_il<P2>(g, C1, x, y, kx, ky);
    }
    __device__ __forceinline__ void w12(const cmm_advect_coeffs &, double x, double y, double &kx, double &ky) const {
        herm_vel_il<P2>(g, C12, x, y, kx, ky);
    }
    __device__ __forceinline__ void lev(int, double x, double y, double &ux, double &uy) const {
        herm_vel_il<P2>(g, C0, x, y, ux, uy);
    }
};

// Two footpoints held by one lane: arithmetic component by component, in the order the scalar
// expressions give (so each component's result is that of the scalar code)
struct PointPair {
    double a, b;
};
__device__ __forceinline__ PointPair operator+(PointPair p, PointPair q) { return {p.a + q.a, p.b + q.b}; }
__device__ __forceinline__ PointPair operator-(PointPair p, PointPair q) { return {p.a - q.a, p.b - q.b}; }
__device__ __forceinline__ PointPair operator*(double s, PointPair p) { return {s * p.a, s * p.b}; }
__device__ __forceinline__ PointPair operator/(PointPair p, double s) { return {p.a / s, p.b / s}; }

// Backward footpoint integration for one start point; schemes of
// Cuda-CMM/src/numerical/cmm-timestep.cu (EulerExp :35-45, AB2 :115-137, Heun :68-85,
// RK3 :175-207, RK3Mod :357-386, RK4 :261-304, RK4Mod :441-470).
// T: double, or PointPair (two footpoints of one lane, component by component the same arithmetic)
template <int TIN, class V, class T = double>
__device__ __forceinline__ void step_backward(const V &vel, const cmm_advect_coeffs &c, T x0, T y0, T &xo, T &yo) {
    const double dt = c.dt;
    if constexpr (TIN == 10) {
        T kx, ky;
        vel.w1(c, x0, y0, kx, ky);
        xo = x0 - dt * kx;
        yo = y0 - dt * ky;
    } else if constexpr (TIN == 20) {
        T u0x, u0y, u1x, u1y, xp = x0, yp = y0;
        xo = x0;
        yo = y0;
        for (int it = 0; it < 10; ++it) {
            vel.lev(1, xp, yp, u1x, u1y);
            xp = xo - dt * u1x;
            yp = yo - dt * u1y;
            vel.lev(1, xp, yp, u1x, u1y);
            vel.lev(0, xo, yo, u0x, u0y);
            xo = x0 - dt * (1.5 * u0x - 0.5 * u1x);
            yo = y0 - dt * (1.5 * u0y - 0.5 * u1y);
        }
    } else if constexpr (TIN == 21) {
        T k1x, k1y, ux, uy;
        vel.w1(c, x0, y0, k1x, k1y);
        vel.lev(0, x0 - dt * k1x, y0 - dt * k1y, ux, uy);
        xo = x0 - dt * (k1x + ux) / 2;
        yo = y0 - dt * (k1y + uy) / 2;
    } else if constexpr (TIN == 30) {
        T k[6];
        vel.w1(c, x0, y0, k[0], k[1]);
        vel.w12(c, x0 - dt * k[0] / 2.0, y0 - dt * k[1] / 2.0, k[2], k[3]);
        vel.lev(0, x0 + dt * k[0] - 2 * dt * k[2], y0 + dt * k[1] - 2 * dt * k[3], k[4], k[5]);
        xo = x0 - dt * (k[0] + 4 * k[2] + k[4]) / 6.0;
        yo = y0 - dt * (k[1] + 4 * k[3] + k[5]) / 6.0;
    } else if constexpr (TIN == 31) {
        T k[6];
        vel.w1(c, x0, y0, k[0], k[1]);
        vel.lev(0, x0 - dt * k[0], y0 - dt * k[1], k[2], k[3]);
        vel.lev(1, x0 + 2 * dt * (-2 * k[0] + k[2]), y0 + 2 * dt * (-2 * k[1] + k[3]), k[4], k[5]);
        xo = x0 - dt * (5 * k[0] + 8 * k[2] - k[4]) / 12.0;
        yo = y0 - dt * (5 * k[1] + 8 * k[3] - k[5]) / 12.0;
    } else if constexpr (TIN == 40) {
        T k[8];
        vel.w1(c, x0, y0, k[0], k[1]);
        vel.w12(c, x0 - dt * k[0] / 2.0, y0 - dt * k[1] / 2.0, k[2], k[3]);
        vel.w12(c, x0 - dt * k[2] / 2.0, y0 - dt * k[3] / 2.0, k[4], k[5]);
        vel.lev(0, x0 - dt * k[4], y0 - dt * k[5], k[6], k[7]);
        xo = x0 - dt * (k[0] + 2 * k[2] + 2 * k[4] + k[6]) / 6.0;
        yo = y0 - dt * (k[1] + 2 * k[3] + 2 * k[5] + k[7]) / 6.0;
    } else {  // 41
        T k[8];
        vel.w1(c, x0, y0, k[0], k[1]);
        vel.lev(0, x0 - dt * k[0], y0 - dt * k[1], k[2], k[3]);
        vel.w12(c, x0 - dt * (3 * k[0] + k[2]) / 8.0, y0 - dt * (3 * k[1] + k[3]) / 8.0, k[4], k[5]);
        vel.lev(0, x0 + dt * (k[2] - 4 * k[4]) / 3.0, y0 + dt * (k[3] - 4 * k[5]) / 3.0, k[6], k[7]);
        xo = x0 - dt * (2 * k[0] - k[2] + 8 * k[4] + 3 * k[6]) / 12.0;
        yo = y0 - dt * (2 * k[1] - k[3] + 8 * k[5] + 3 * k[7]) / 12.0;
    }
}

// Forward footpoint integration of the forward map, the reference's forward schemes
// (Cuda-CMM/src/numerical/cmm-timestep.cu: EulerExp :28-33, Heun :48-65, AB2 :86-109, RK3 :143-174,
// RK4 :218-260, RK3Mod :314-337, RK4Mod :384-430).  As in the reference, RK4 weights its second
// stage with L1 and its fourth with L12, and RK4Mod is the backward formula.
template <int TIN, class V>
__device__ __forceinline__ void step_forward(const V &vel, const cmm_advect_coeffs &c, double x0, double y0,
                                             double &xo, double &yo) {
    const double dt = c.dt;
    if constexpr (TIN == 10) {
        double ux, uy;
        vel.lev(0, x0, y0, ux, uy);
        xo = x0 + dt * ux;
        yo = y0 + dt * uy;
    } else if constexpr (TIN == 20) {
        double u0x, u0y, u1x, u1y, xp = x0, yp = y0;
        xo = x0;
        yo = y0;
        for (int it = 0; it < 10; ++it) {
            vel.lev(1, xp, yp, u1x, u1y);
            xp = xo + dt * u1x;
            yp = yo + dt * u1y;
            vel.lev(1, xp, yp, u1x, u1y);
            vel.lev(0, xo, yo, u0x, u0y);
            xo = x0 + dt * (1.5 * u0x - 0.5 * u1x);
            yo = y0 + dt * (1.5 * u0y - 0.5 * u1y);
        }
    } else if constexpr (TIN == 21) {
        double k[4];
        vel.lev(0, x0, y0, k[0], k[1]);
        vel.w1(c, x0 + dt * k[0], y0 + dt * k[1], k[2], k[3]);
        xo = x0 + dt * (k[0] + k[2]) / 2;
        yo = y0 + dt * (k[1] + k[3]) / 2;
    } else if constexpr (TIN == 30) {
        double k[6];
        vel.lev(0, x0, y0, k[0], k[1]);
        vel.w12(c, x0 + dt * k[0] / 2.0, y0 + dt * k[1] / 2.0, k[2], k[3]);
        vel.w1(c, x0 - dt * k[0] + 2 * dt * k[2], y0 - dt * k[1] + 2 * dt * k[3], k[4], k[5]);
        xo = x0 + dt * (k[0] + 4 * k[2] + k[4]) / 6.0;
        yo = y0 + dt * (k[1] + 4 * k[3] + k[5]) / 6.0;
    } else if constexpr (TIN == 31) {
        double k[6];
        vel.lev(0, x0, y0, k[0], k[1]);
        vel.lev(1, x0 - dt * k[0], y0 - dt * k[1], k[2], k[3]);
        vel.lev(2, x0 - dt * (8 * k[0] + 2 * k[2]) / 5.0, y0 - dt * (8 * k[1] + 2 * k[3]) / 5.0, k[4], k[5]);
        xo = x0 + dt * (23 * k[0] - 16 * k[2] + 5 * k[4]) / 12.0;
        yo = y0 + dt * (23 * k[1] - 16 * k[3] + 5 * k[5]) / 12.0;
    } else if constexpr (TIN == 40) {
        double k[8];
        vel.lev(0, x0, y0, k[0], k[1]);
        vel.w1(c, x0 + dt * k[0] / 2.0, y0 + dt * k[1] / 2.0, k[2], k[3]);
        vel.w12(c, x0 + dt * k[2] / 2.0, y0 + dt * k[3] / 2.0, k[4], k[5]);
        vel.w12(c, x0 + dt * k[4], y0 + dt * k[5], k[6], k[7]);
        xo = x0 + dt * (k[0] + 2 * k[2] + 2 * k[4] + k[6]) / 6.0;
        yo = y0 + dt * (k[1] + 2 * k[3] + 2 * k[5] + k[7]) / 6.0;
    } else {  // 41
        step_backward<41>(vel, c, x0, y0, xo, yo);
    }
}

// Node-interleaved fields for the advection: C1[4j + p] = sum_l L1[l] psi_l[p*n + j], likewise C12
// with L12 (when set), and C0[4j + p] = psi_0[p*n + j] (when set); sums in level order like the
// weighted velocities.  One lane per node: plane reads coalesce across lanes, and each lane writes
// its 32-byte records with 16-byte stores.
// D1 / D12 (optional, plane-major Hermite fields): corrections added to the extrapolations C1 / C12
// LW (optional, device): the weights L1[4], L12[4] in place of the coefficient struct's
template <int L>
__global__ void __launch_bounds__(256) k_psi_combine(long n, PsiLevels P, cmm_advect_coeffs c, double *__restrict__ C1,
                                                     double *__restrict__ C12, double *__restrict__ C0,
                                                     const double *__restrict__ D1 = nullptr,
                                                     const double *__restrict__ D12 = nullptr,
                                                     const double *__restrict__ LW = nullptr) {
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    double w1[L], w12[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
        w1[l] = LW ? LW[l] : c.L1[l];
        w12[l] = LW ? LW[4 + l] : c.L12[l];
    }
    double v[L][4];
#pragma unroll
    for (int l = 0; l < L; ++l)
#pragma unroll
        for (int p = 0; p < 4; ++p) v[l][p] = P.p[l][p * n + j];
    auto comb = [&](const double *w, double *dst, const double *D) {
        double o[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            double a = w[0] * v[0][p];
#pragma unroll
            for (int l = 1; l < L; ++l) a += w[l] * v[l][p];
            if (D) a += D[p * n + j];
            o[p] = a;
        }
        double2 *d = reinterpret_cast<double2 *>(dst + 4 * j);
        d[0] = double2{o[0], o[1]};
        d[1] = double2{o[2], o[3]};
    };
    comb(w1, C1, D1);
    if (C12) comb(w12, C12, D12);
    if (C0) {
        double2 *d = reinterpret_cast<double2 *>(C0 + 4 * j);
        d[0] = double2{v[0][0], v[0][1]};
        d[1] = double2{v[0][2], v[0][3]};
    }
}

// Map advection with the footpoint stencil, Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:207-298.
// One lane per footpoint: G consecutive lanes (G >= number of footpoints) hold the footpoints of
// one coarse node, so a wavefront's gathers of psi and of the old map concentrate on the few
// cells around 64/G neighbouring nodes.  After the footpoints are traced back, lane j of the group
// forms output plane j by reading the group's positions with cross-lane shuffles and summing them
// in footpoint order with the stencil weights, the order of the reference's accumulation.
// the map node-interleaved for the advection's map evaluations: M8[8j + p] = ChiX plane p,
// M8[8j + 4 + p] = ChiY plane p
__global__ void __launch_bounds__(256) k_interleave_map(long n, const double *__restrict__ cx,
                                                        const double *__restrict__ cy, double *__restrict__ M8) {
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    double2 *d = reinterpret_cast<double2 *>(M8 + 8 * j);
    d[0] = double2{cx[j], cx[n + j]};
    d[1] = double2{cx[2 * n + j], cx[3 * n + j]};
    d[2] = double2{cy[j], cy[n + j]};
    d[3] = double2{cy[2 * n + j], cy[3 * n + j]};
}

// most footpoints of a stencil (6th order: 12)
constexpr int kMaxFoot = 12;

// P2 (with COMB): both grids have power-of-two point counts
// Output stage of the advection: lane k of a node's group forms output plane k (and k + G) by
// reading the group's traced positions with cross-lane shuffles and summing them in footpoint
// order with the stencil weights, the order of the reference's accumulation.
template <int G>
__device__ __forceinline__ void advect_store(const DGrid &gm, const cmm_advect_coeffs &c, int nfoot, int node,
                                             bool active, int k, double px, double py, double *__restrict__ ncx,
                                             double *__restrict__ ncy, double *__restrict__ M8out) {
    const int base = (threadIdx.x & 63) - k;  // first lane of the group within the wavefront
    // every lane takes part in every shuffle (the same trip count for all lanes); the footpoint
    // loop is unrolled to its maximum so that the stencil weights are read at constant indices
    // (a weight array indexed at run time would be copied to private memory).  The lane's plane
    // picks its weight by 0/1 factors, not by a select, which the compiler can turn into an index
    // into the four weight arrays at run time (0 * w + w' is w' exactly for finite weights).
#pragma unroll
    for (int r = 0; r < (8 + G - 1) / G; ++r) {
        const int o = k + r * G;
        const int plane = o & 3;
        const double f0 = plane == 0, f1 = plane == 1, f2 = plane == 2, f3 = plane == 3;
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m < kMaxFoot; ++m) {
            if (m < nfoot) {
                const double vx = __shfl(px, base + m, 64);
                const double vy = __shfl(py, base + m, 64);
                const double wm = ((f0 * c.c1[m] + f1 * c.cx[m]) + f2 * c.cy[m]) + f3 * c.cxy[m];
                acc += (o < 4 ? vx : vy) * wm;
            }
        }
        if (active && o < 8) {
            (o < 4 ? ncx : ncy)[(long)plane * gm.n + node] = acc;
            // the new map node-interleaved as well, for the next step's map evaluations
            if (M8out) M8out[8L * node + o] = acc;
        }
    }
}

template <int TIN, int G, bool COMB, bool P2 = false>
__global__ void __launch_bounds__(256) k_map_advect(DGrid gm, DGrid gp, const double *__restrict__ cx,
                                                    const double *__restrict__ cy, PsiLevels P, const double *C1,
                                                    const double *C12, const double *C0, const double *M8,
                                                    cmm_advect_coeffs c, double *__restrict__ ncx,
                                                    double *__restrict__ ncy, double *__restrict__ M8out,
                                                    int nblocks, const double *__restrict__ LW) {
    constexpr int NPB = 256 / G;  // nodes per block
    const int blk = xcd_block(blockIdx.x, nblocks);
    const int node = blk * NPB + threadIdx.x / G;
    const int k = threadIdx.x % G;
    const int nfoot = 4 + 4 * c.map_update_order_num;
    const bool active = node < gm.n;
    double px = 0.0, py = 0.0;
    if (active && k < nfoot) {
        const int ix = node % gm.nx, iy = node / gm.nx;
        const double xe = gm.x0 + ix * gm.hx + ((k >> 2) + 1) * c.eps * (1 - 2 * ((k >> 1) & 1));
        const double ye = gm.y0 + iy * gm.hy + ((k >> 2) + 1) * c.eps * (1 - 2 * (((k + 1) >> 1) & 1));
        double xf, yf;
        if constexpr (COMB) {
            step_backward<TIN>(VelCombined<P2>{gp, C1, C12, C0}, c, xe, ye, xf, yf);
            diffeo_il<P2>(gm, M8, xf, yf, px, py);
        } else {
            step_backward<TIN>(VelLevels{gp, P, LW}, c, xe, ye, xf, yf);
            diffeo(gm, cx, cy, xf, yf, px, py);
        }
    }
    advect_store<G>(gm, c, nfoot, node, active, k, px, py, ncx, ncy, M8out);
}

// ---- two footpoints per lane ----------------------------------------------------------------
// With the 8-footpoint stencil, lane k of a node's 4 lanes traces footpoints k and k + 4 (the
// same diagonal at eps and 2 eps from the node), which fall in the same psi cell, and then in the
// same map cell, most of the time: the second footpoint's corner records are loaded only when its
// cell differs, which spares the texture path about a third of its addresses.  The second load is
// held back behind the first footpoint's result (a zero index offset the optimiser cannot see
// through), so that it does not take registers of its own and cost occupancy.  Per footpoint the
// values and the operations are those of k_map_advect (the compiler's contraction of the pair's
// expressions can differ from the single footpoint's in the last bit).

// a zero that depends on v (issued after v is computed)
__device__ __forceinline__ int zero_after(double v) {
    int z = 0;
    asm volatile("" : "+v"(z) : "v"(__double2loint(v)));
    return z;
}

__device__ __forceinline__ Cell cell_after(Cell c, double v) {
    const int z = zero_after(v);
    c.i00 += z;
    c.i10 += z;
    c.i01 += z;
    c.i11 += z;
    return c;
}

template <bool P2>
struct VelPair {
    DGrid g;
    const double *C1, *C12, *C0;
    __device__ __forceinline__ void ev(const double *__restrict__ H4, PointPair x, PointPair y, PointPair &ux,
                                       PointPair &uy) const {
        Cell ca, cb;
        find_cell<P2>(g, x.a, y.a, ca);
        find_cell<P2>(g, x.b, y.b, cb);
        double v[16];
        load16_il(H4, ca, v);
        herm_vel_vals(g, ca, v, ux.a, uy.a);
        if (!same_cell(ca, cb)) load16_il(H4, cell_after(cb, ux.a), v);
        herm_vel_vals(g, cb, v, ux.b, uy.b);
    }
    __device__ __forceinline__ void w1(const cmm_advect_coeffs &, PointPair x, PointPair y, PointPair &kx,
                                       PointPair &ky) const {
        ev(C1, x, y, kx, ky);
    }
    __device__ __forceinline__ void w12(const cmm_advect_coeffs &, PointPair x, PointPair y, PointPair &kx,
                                        PointPair &ky) const {
        ev(C12, x, y, kx, ky);
    }
    __device__ __forceinline__ void lev(int, PointPair x, PointPair y, PointPair &ux, PointPair &uy) const {
        ev(C0, x, y, ux, uy);
    }
};

template <int TIN, bool P2>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_map_advect_pair(DGrid gm, DGrid gp, const double *C1, const double *C12,
                                                         const double *C0, const double *M8, cmm_advect_coeffs c,
                                                         double *__restrict__ ncx, double *__restrict__ ncy,
                                                         double *__restrict__ M8out, int nblocks) {
    constexpr int G = 4, NPB = 256 / G;
    const int blk = xcd_block(blockIdx.x, nblocks);
    const int node = blk * NPB + threadIdx.x / G;
    const int k = threadIdx.x % G;
    const bool active = node < gm.n;
    PointPair px{0.0, 0.0}, py{0.0, 0.0};
    if (active) {
        const int ix = node % gm.nx, iy = node / gm.nx;
        const int fa = k, fb = k + 4;
        const PointPair xe{gm.x0 + ix * gm.hx + ((fa >> 2) + 1) * c.eps * (1 - 2 * ((fa >> 1) & 1)),
                           gm.x0 + ix * gm.hx + ((fb >> 2) + 1) * c.eps * (1 - 2 * ((fb >> 1) & 1))};
        const PointPair ye{gm.y0 + iy * gm.hy + ((fa >> 2) + 1) * c.eps * (1 - 2 * (((fa + 1) >> 1) & 1)),
                           gm.y0 + iy * gm.hy + ((fb >> 2) + 1) * c.eps * (1 - 2 * (((fb + 1) >> 1) & 1))};
        PointPair xf, yf;
        step_backward<TIN>(VelPair<P2>{gp, C1, C12, C0}, c, xe, ye, xf, yf);
        Cell ca, cb;
        find_cell<P2>(gm, xf.a, yf.a, ca);
        find_cell<P2>(gm, xf.b, yf.b, cb);
        double vx[16], vy[16];
        load_map_il(M8, ca, vx, vy);
        diffeo_vals(gm, ca, vx, vy, px.a, py.a);
        if (!same_cell(ca, cb)) load_map_il(M8, cell_after(cb, px.a), vx, vy);
        diffeo_vals(gm, cb, vx, vy, px.b, py.b);
    }
    // output planes: footpoint m sits in lane m & 3 of the group, component m >> 2; lane k forms
    // plane k of X and of Y, summing in footpoint order like advect_store
    const int base = (threadIdx.x & 63) - k;
    double out[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int o = k + r * G;
        const int plane = o & 3;
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const double vx = __shfl(m < 4 ? px.a : px.b, base + (m & 3), 64);
            const double vy = __shfl(m < 4 ? py.a : py.b, base + (m & 3), 64);
            const double wm = plane == 0 ? c.c1[m] : plane == 1 ? c.cx[m] : plane == 2 ? c.cy[m] : c.cxy[m];
            acc += (o < 4 ? vx : vy) * wm;
        }
        if (active) (o < 4 ? ncx : ncy)[(long)plane * gm.n + node] = acc;
        out[r] = acc;
    }
    // the record of the node (X planes, then Y planes) with one 16-byte store per lane: lane k
    // stores entries 2k and 2k + 1, which lanes (2k) & 3 and (2k + 1) & 3 hold in component k >> 1
    if (M8out) {
        const double e0 = __shfl(out[0], base + ((2 * k) & 3), 64), e1 = __shfl(out[1], base + ((2 * k) & 3), 64);
        const double f0 = __shfl(out[0], base + ((2 * k + 1) & 3), 64), f1 = __shfl(out[1], base + ((2 * k + 1) & 3), 64);
        if (active)
            reinterpret_cast<double2 *>(M8out + 8L * node)[k] = double2{k < 2 ? e0 : e1, k < 2 ? f0 : f1};
    }
}

// ---- Vlasov-Poisson: the separable stream function --------------------------------------------
// Every psi level of the Vlasov-Poisson flow is psi_l(x, v) = phi_l(x) - H(v) on the psi grid
// (k_vp_psi_planes: the planes phi - H, phi', -H', 0).  The bicubic Hermite interpolant of such
// planes is the sum of the cubic Hermite interpolants of phi_l in x and of -H in v (the basis
// weights of each direction sum to one and their derivatives to zero), so the velocity of a
// Lagrange combination sum_l w_l psi_l is
//   u_x = d psi / dv = -(sum_l w_l) H'(v),   u_v = -d psi / dx = -(sum_l w_l phi_l)'(x),
// two 1D interpolations in tables of the psi grid's width.  Row 0 of a level's value plane holds
// phi_l(x) - H(v_0) (the constant drops out of the x derivative) and row 0 of its d/dx plane
// phi_l'(x).  Against the 2D evaluation of the planes the results differ by rounding only.

// the 1D tables of the combined fields: R (values along x) and D (d/dx along x) of C1 (w = L1),
// C12 (w = L12) and C0 (level 0), each np doubles at T + {0, 1, 2, 3, 4, 5} np, and the weight
// sums S[0] = sum L1, S[1] = sum L12 (LW: the weights from device memory, as k_psi_combine)
template <int L>
__device__ __forceinline__ void vp_combine1d_at(int x, int np, long N, const PsiLevels &P, const cmm_advect_coeffs &c,
                                                const double *__restrict__ LW, double *__restrict__ T,
                                                double *__restrict__ S) {
    double w1[L], w12[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
        w1[l] = LW ? LW[l] : c.L1[l];
        w12[l] = LW ? LW[4 + l] : c.L12[l];
    }
    if (x == 0) {
        double s1 = w1[0], s12 = w12[0];
#pragma unroll
        for (int l = 1; l < L; ++l) {
            s1 += w1[l];
            s12 += w12[l];
        }
        S[0] = s1;
        S[1] = s12;
    }
    if (x >= np) return;
    double r[L], d[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
        r[l] = P.p[l][x];
        d[l] = P.p[l][N + x];
    }
    double r1 = w1[0] * r[0], d1 = w1[0] * d[0], r12 = w12[0] * r[0], d12 = w12[0] * d[0];
#pragma unroll
    for (int l = 1; l < L; ++l) {
        r1 += w1[l] * r[l];
        d1 += w1[l] * d[l];
        r12 += w12[l] * r[l];
        d12 += w12[l] * d[l];
    }
    T[x] = r1;
    T[np + x] = d1;
    T[2 * np + x] = r12;
    T[3 * np + x] = d12;
    T[4 * np + x] = r[0];
    T[5 * np + x] = d[0];
}

template <int L>
__global__ void __launch_bounds__(256) k_vp_combine1d(int np, long N, PsiLevels P, cmm_advect_coeffs c,
                                                      const double *__restrict__ LW, double *__restrict__ T,
                                                      double *__restrict__ S) {
    vp_combine1d_at<L>(blockIdx.x * 256 + threadIdx.x, np, N, P, c, LW, T, S);
}

// one coordinate's cell, as find_cell takes it: lower node, upper node (wrapped) and offset
__device__ __forceinline__ void cell_1d(double x, double x0, double ih, int n, int lg, int &i0, int &i1, double &d) {
    const double f = (x - x0) * ih;
    double fl = floor(f);
    d = f - fl;
    fl = fmin(fmax(fl, -1.0e9), 1.0e9);
    const int i = (int)fl;
    i0 = i - floordiv_g(i, n, lg) * n;
    i1 = i0 + 1 == n ? 0 : i0 + 1;
}

struct VelSep {
    DGrid g;
    const double *T;  // k_vp_combine1d's tables
    const double *S;  // weight sums
    const double *H;  // H at [0, np), H' at [2 np, 3 np)
    __device__ __forceinline__ void ev1(const double *R, const double *D, double s, double x, double v, double &ux,
                                        double &uy) const {
        int i0, i1, j0, j1;
        double dx, dv;
        cell_1d(x, g.x0, g.ihx, g.nx, g.lgx, i0, i1, dx);
        cell_1d(v, g.y0, g.ihy, g.ny, g.lgy, j0, j1, dv);
        double dbx[4], dbv[4];
        basis_d(dx, dbx);
        basis_d(dv, dbv);
        const double *Hd = H + 2 * g.nx;
        uy = -(dbx[0] * R[i0] + dbx[1] * R[i1] + g.hx * (dbx[2] * D[i0] + dbx[3] * D[i1])) * g.ihx;
        ux = -s * (dbv[0] * H[j0] + dbv[1] * H[j1] + g.hy * (dbv[2] * Hd[j0] + dbv[3] * Hd[j1])) * g.ihy;
    }
    __device__ __forceinline__ void ev(int f, PointPair x, PointPair y, PointPair &ux, PointPair &uy) const {
        const int np = g.nx;
        const double s = f == 0 ? S[0] : f == 1 ? S[1] : 1.0;
        ev1(T + 2 * f * np, T + (2 * f + 1) * np, s, x.a, y.a, ux.a, uy.a);
        ev1(T + 2 * f * np, T + (2 * f + 1) * np, s, x.b, y.b, ux.b, uy.b);
    }
    __device__ __forceinline__ void w1(const cmm_advect_coeffs &, PointPair x, PointPair y, PointPair &kx,
                                       PointPair &ky) const {
        ev(0, x, y, kx, ky);
    }
    __device__ __forceinline__ void w12(const cmm_advect_coeffs &, PointPair x, PointPair y, PointPair &kx,
                                        PointPair &ky) const {
        ev(1, x, y, kx, ky);
    }
    __device__ __forceinline__ void lev(int, PointPair x, PointPair y, PointPair &ux, PointPair &uy) const {
        ev(2, x, y, ux, uy);
    }
};

// the 8-footpoint advection of a Vlasov-Poisson map with the separable velocity (two footpoints
// per lane and the map records as k_map_advect_pair)
template <int TIN>
__device__ __forceinline__ void advect_vp_block(int blk, const DGrid &gm, const DGrid &gp, const double *T,
                                                const double *S, const double *H, const double *M8,
                                                const cmm_advect_coeffs &c, double *__restrict__ ncx,
                                                double *__restrict__ ncy, double *__restrict__ M8out) {
    constexpr int G = 4, NPB = 256 / G;
    const int node = blk * NPB + threadIdx.x / G;
    const int k = threadIdx.x % G;
    const bool active = node < gm.n;
    PointPair px{0.0, 0.0}, py{0.0, 0.0};
    if (active) {
        const int ix = node % gm.nx, iy = node / gm.nx;
        const int fa = k, fb = k + 4;
        const PointPair xe{gm.x0 + ix * gm.hx + ((fa >> 2) + 1) * c.eps * (1 - 2 * ((fa >> 1) & 1)),
                           gm.x0 + ix * gm.hx + ((fb >> 2) + 1) * c.eps * (1 - 2 * ((fb >> 1) & 1))};
        const PointPair ye{gm.y0 + iy * gm.hy + ((fa >> 2) + 1) * c.eps * (1 - 2 * (((fa + 1) >> 1) & 1)),
                           gm.y0 + iy * gm.hy + ((fb >> 2) + 1) * c.eps * (1 - 2 * (((fb + 1) >> 1) & 1))};
        PointPair xf, yf;
        step_backward<TIN>(VelSep{gp, T, S, H}, c, xe, ye, xf, yf);
        Cell ca, cb;
        find_cell<true>(gm, xf.a, yf.a, ca);
        find_cell<true>(gm, xf.b, yf.b, cb);
        double vx[16], vy[16];
        load_map_il(M8, ca, vx, vy);
        diffeo_vals(gm, ca, vx, vy, px.a, py.a);
        if (!same_cell(ca, cb)) load_map_il(M8, cell_after(cb, px.a), vx, vy);
        diffeo_vals(gm, cb, vx, vy, px.b, py.b);
    }
    const int base = (threadIdx.x & 63) - k;
    double out[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int o = k + r * G;
        const int plane = o & 3;
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const double vx = __shfl(m < 4 ? px.a : px.b, base + (m & 3), 64);
            const double vy = __shfl(m < 4 ? py.a : py.b, base + (m & 3), 64);
            const double wm = plane == 0 ? c.c1[m] : plane == 1 ? c.cx[m] : plane == 2 ? c.cy[m] : c.cxy[m];
            acc += (o < 4 ? vx : vy) * wm;
        }
        if (active) (o < 4 ? ncx : ncy)[(long)plane * gm.n + node] = acc;
        out[r] = acc;
    }
    // the record of the node (X planes, then Y planes) with one 16-byte store per lane: lane k
    // stores entries 2k and 2k + 1, which lanes (2k) & 3 and (2k + 1) & 3 hold in component k >> 1
    if (M8out) {
        const double e0 = __shfl(out[0], base + ((2 * k) & 3), 64), e1 = __shfl(out[1], base + ((2 * k) & 3), 64);
        const double f0 = __shfl(out[0], base + ((2 * k + 1) & 3), 64), f1 = __shfl(out[1], base + ((2 * k + 1) & 3), 64);
        if (active)
            reinterpret_cast<double2 *>(M8out + 8L * node)[k] = double2{k < 2 ? e0 : e1, k < 2 ? f0 : f1};
    }
}

template <int TIN>
__global__ void __launch_bounds__(256) k_map_advect_vp(DGrid gm, DGrid gp, const double *T, const double *S,
                                                       const double *H, const double *M8, cmm_advect_coeffs c,
                                                       double *__restrict__ ncx, double *__restrict__ ncy,
                                                       double *__restrict__ M8out, int nblocks) {
    advect_vp_block<TIN>(xcd_block(blockIdx.x, nblocks), gm, gp, T, S, H, M8, c, ncx, ncy, M8out);
}

// The forward map (Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:254-274): the old forward map
// is evaluated at the footpoint first, then the point moves forward in time; psi per level.
template <int TIN, int G>
__global__ void __launch_bounds__(256) k_map_advect_fwd(DGrid gm, DGrid gp, const double *__restrict__ cx,
                                                        const double *__restrict__ cy, PsiLevels P, cmm_advect_coeffs c,
                                                        double *__restrict__ ncx, double *__restrict__ ncy, int nblocks) {
    constexpr int NPB = 256 / G;
    const int blk = xcd_block(blockIdx.x, nblocks);
    const int node = blk * NPB + threadIdx.x / G;
    const int k = threadIdx.x % G;
    const int nfoot = 4 + 4 * c.map_update_order_num;
    const bool active = node < gm.n;
    double px = 0.0, py = 0.0;
    if (active && k < nfoot) {
        const int ix = node % gm.nx, iy = node / gm.nx;
        const double xe = gm.x0 + ix * gm.hx + ((k >> 2) + 1) * c.eps * (1 - 2 * ((k >> 1) & 1));
        const double ye = gm.y0 + iy * gm.hy + ((k >> 2) + 1) * c.eps * (1 - 2 * (((k + 1) >> 1) & 1));
        double xf, yf;
        diffeo(gm, cx, cy, xe, ye, xf, yf);
        step_forward<TIN>(VelLevels{gp, P}, c, xf, yf, px, py);
    }
    advect_store<G>(gm, c, nfoot, node, active, k, px, py, ncx, ncy, nullptr);
}

// max |1 - det grad X| over cell centres of gc; each block reduces a strided set of 64x4 tiles,
// then one 64-bit atomic max on the bit pattern (order-preserving for non-negative doubles) per
// block.  Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:159-173 and the Thrust max-reduction of
// Cuda-CMM/src/simulation/cmm-simulation-host.cu:44-50.
__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

__global__ void __launch_bounds__(256) k_fill(double *p, long n, double v) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(256) k_incomp(DGrid gm, DGrid gc, const double *__restrict__ cx,
                                                const double *__restrict__ cy, unsigned long long *__restrict__ out,
                                                int tiles_x, int tiles_total) {
    double v = 0.0;
    for (int tt = blockIdx.x; tt < tiles_total; tt += gridDim.x) {
        const int t = xcd_block(tt, tiles_total);
        const int ix = (t % tiles_x) * 64 + threadIdx.x;
        const int iy = (t / tiles_x) * 4 + threadIdx.y;
        if (ix < gc.nx && iy < gc.ny) {
            const double x = gm.x0 + ix * gc.hx + 0.5 * gc.hx;
            const double y = gm.y0 + iy * gc.hy + 0.5 * gc.hy;
            v = fmax(v, fabs(1 - diffeo_det(gm, cx, cy, x, y)));
        }
    }
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __shared__ double red[4];
    if (threadIdx.x == 0) red[threadIdx.y] = v;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
    }
}

// max over cell centres of gc of |x - X_b(X_f(x))|, k_invertibility_check and invertibility_check
// (Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:177-198,
// Cuda-CMM/src/simulation/cmm-simulation-host.cu:52-60); the same reduction as k_incomp
__global__ void __launch_bounds__(256) k_invert(DGrid gb, DGrid gf, DGrid gc, const double *__restrict__ bx,
                                                const double *__restrict__ by, const double *__restrict__ fx,
                                                const double *__restrict__ fy, unsigned long long *__restrict__ out,
                                                int tiles_x, int tiles_total) {
    double v = 0.0;
    for (int tt = blockIdx.x; tt < tiles_total; tt += gridDim.x) {
        const int t = xcd_block(tt, tiles_total);
        const int ix = (t % tiles_x) * 64 + threadIdx.x;
        const int iy = (t / tiles_x) * 4 + threadIdx.y;
        if (ix < gc.nx && iy < gc.ny) {
            const double x = gc.x0 + ix * gc.hx + 0.5 * gc.hx;
            const double y = gc.y0 + iy * gc.hy + 0.5 * gc.hy;
            double xi, yi;
            diffeo(gf, fx, fy, x, y, xi, yi);
            diffeo(gb, bx, by, xi, yi, xi, yi);
            v = fmax(v, hypot(x - xi, y - yi));
        }
    }
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __shared__ double red[4];
    if (threadIdx.x == 0) red[threadIdx.y] = v;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
    }
}

// Same check when the check grid refines the map grid by an integer ratio r over the same domain,
// with the Hermite sums in separable form: a lane takes a TX x TY tile of check points inside one
// map cell (TX and TY divide r), loads the cell's corner values once, contracts them along x for
// the tile's TX columns and then along y per point.  The terms are those of hsum_v (derivative planes scaled by h, value planes with their
// wrap jumps) grouped by the x basis first, so the values agree with diffeo_det to rounding.  A
// tile whose first and last points fall in different cells is evaluated point by point.
template <int TX, int TY>
__device__ __forceinline__ double incomp_tensor_tiles(const DGrid &gm, const DGrid &gc, const double *__restrict__ cx,
                                                      const double *__restrict__ cy, int ntx, long t_begin,
                                                      long t_end) {
    double v = 0.0;
    for (long t0 = t_begin; t0 < t_end; t0 += 256) {
        const long t = t0 + threadIdx.x;
        if (t >= t_end) continue;
        const int ix0 = (int)(t % ntx) * TX, iy0 = (int)(t / ntx) * TY;
        double xs[TX], ys[TY];
#pragma unroll
        for (int a = 0; a < TX; ++a) xs[a] = gm.x0 + (ix0 + a) * gc.hx + 0.5 * gc.hx;
#pragma unroll
        for (int b = 0; b < TY; ++b) ys[b] = gm.y0 + (iy0 + b) * gc.hy + 0.5 * gc.hy;
        Cell c0, c1;
        find_cell(gm, xs[0], ys[0], c0);
        find_cell(gm, xs[TX - 1], ys[TY - 1], c1);
        if (!same_cell(c0, c1)) {
#pragma unroll 1
            for (int i = 0; i < TX * TY; ++i) {
                const int a = i % TX, b = i / TX;
                const double x = gm.x0 + (ix0 + a) * gc.hx + 0.5 * gc.hx, y = gm.y0 + (iy0 + b) * gc.hy + 0.5 * gc.hy;
                v = fmax(v, fabs(1 - diffeo_det(gm, cx, cy, x, y)));
            }
            continue;
        }
        double vx[16], vy[16];
        load16(cx, gm.n, c0, vx);
        load16(cy, gm.n, c0, vy);
        const double wjx[4] = {c0.wx0 * gm.jx, c0.wx1 * gm.jx, c0.wx0 * gm.jx, c0.wx1 * gm.jx};
        const double wjy[4] = {c0.wy0 * gm.jy, c0.wy0 * gm.jy, c0.wy1 * gm.jy, c0.wy1 * gm.jy};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            vx[q] += wjx[q];
            vy[q] += wjy[q];
        }
        // y basis of the tile's rows
        double by[TY][4], dby[TY][4];
#pragma unroll
        for (int b = 0; b < TY; ++b) {
            const double fy = (ys[b] - gm.y0) * gm.ihy;
            const double d = fy - floor(fy);
            basis(d, by[b]);
            basis_d(d, dby[b]);
        }
#pragma unroll
        for (int a = 0; a < TX; ++a) {
            const double fx = (xs[a] - gm.x0) * gm.ihx;
            const double d = fx - floor(fx);
            double bx[4], dbx[4];
            basis(d, bx);
            basis_d(d, dbx);
            // P[k][n]: x-contraction of field k (0: X with dbx, 1: X with bx, 2: Y with dbx, 3: Y
            // with bx) at y node n (0, 1) for the value planes (n) and the y-derivative planes (2 + n)
            double P[4][4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double *w = (k & 1) ? bx : dbx;
                const double *f = (k < 2) ? vx : vy;
#pragma unroll
                for (int n = 0; n < 2; ++n) {
                    P[k][n] = w[0] * f[2 * n] + w[1] * f[2 * n + 1] + gm.hx * (w[2] * f[4 + 2 * n] + w[3] * f[5 + 2 * n]);
                    P[k][2 + n] =
                        w[0] * f[8 + 2 * n] + w[1] * f[9 + 2 * n] + gm.hx * (w[2] * f[12 + 2 * n] + w[3] * f[13 + 2 * n]);
                }
            }
#pragma unroll
            for (int b = 0; b < TY; ++b) {
                auto ysum = [&](const double *p, const double *c) {
                    return c[0] * p[0] + c[1] * p[1] + gm.hy * (c[2] * p[2] + c[3] * p[3]);
                };
                const double Xx = ysum(P[0], by[b]) * gm.ihx;
                const double Xy = ysum(P[1], dby[b]) * gm.ihx;
                const double Yx = ysum(P[2], by[b]) * gm.ihy;
                const double Yy = ysum(P[3], dby[b]) * gm.ihy;
                v = fmax(v, fabs(1 - (Xx * Yy - Xy * Yx)));
            }
        }
    }
    return v;
}

// the block's maximum of v into *out (one 64-bit atomic max on the bit pattern; for non-negative
// doubles it preserves order); red: 4 doubles of LDS
__device__ __forceinline__ void block_max_to(double v, double *red, unsigned long long *out) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
    }
}

template <int TX, int TY>
__global__ void __launch_bounds__(256) k_incomp_tensor(DGrid gm, DGrid gc, const double *__restrict__ cx,
                                                       const double *__restrict__ cy,
                                                       unsigned long long *__restrict__ out, int r, int ntx,
                                                       long ntiles) {
    long t_begin, t_end;
    tile_range(ntiles, t_begin, t_end);
    const double v = incomp_tensor_tiles<TX, TY>(gm, gc, cx, cy, ntx, t_begin, t_end);
    __shared__ double red[4];
    block_max_to(v, red, out);
}

// ------------------------------------------------------------------------------------------------
// Initial conditions on the device, Cuda-CMM/src/simulation/cmm-init.cu:30-241 (+ new conditions)
struct ICParams {
    int ic;
    double prm[10];
};

__device__ double init_vorticity(double x, double y, const ICParams &p) {
    switch (p.ic) {
    // distribution functions f(x, v), d_init_distirbution_function (Cuda-CMM/src/simulation/cmm-init.cu:457-493)
    case CMM_IC_LANDAU_DAMPING:
        return (1 + 0.05 * cos(0.5 * x)) * exp(-y * y * 0.5) / sqrt(2 * kPi);
    case CMM_IC_TWO_STREAM:
        return 1 / (2 * sqrt(2 * kPi)) * (1 + 0.05 * cos(0.2 * x)) *
               (exp(-(y - 3) * (y - 3) * 0.5) + exp(-(y + 3) * (y + 3) * 0.5));
    case CMM_IC_NONLIN_LANDAU_DAMPING:
        return (1 + 0.5 * cos(0.5 * x)) * exp(-y * y * 0.5) / sqrt(2 * kPi);
    // ensemble conditions with a perturbation amplitude (no reference counterpart)
    case CMM_IC_LANDAU_AMP:
        return (1 + p.prm[0] * cos(0.5 * x)) * exp(-y * y * 0.5) / sqrt(2 * kPi);
    case CMM_IC_TWO_STREAM_V2:
        return (1 + p.prm[0] * cos(0.5 * x)) * y * y * exp(-y * y * 0.5) / sqrt(2 * kPi);
    // passive scalars, d_init_scalar (Cuda-CMM/src/simulation/cmm-init.cu:249-301)
    case CMM_IC_SCALAR_RECTANGLE: {
        x -= floor(x / kTwoPi) * kTwoPi;
        y -= floor(y / kTwoPi) * kTwoPi;
        const double cx0 = kPi, cy0 = kPi / 2.0, wx = 2 * kPi, wy = kPi;
        return (x > cx0 - wx / 2.0 && x < cx0 + wx / 2.0 && y > cy0 - wy / 2.0 && y < cy0 + wy / 2.0) ? 1.0 : 0.0;
    }
    case CMM_IC_SCALAR_GAUSSIAN: {
        const double mu = kPi, var = kPi / 4.0;
        return 1.0 / (kTwoPi * var * var) * exp(-((x - mu) * (x - mu) / (2 * var * var) + (y - mu) * (y - mu) / (2 * var * var)));
    }
    case CMM_IC_SCALAR_CIRCULAR_RING: {
        const double r = kPi / 2.0, rw = r / 8.0, s1 = 10.0, s2 = 10.0;
        const double d2 = (x - kPi) * (x - kPi) + (y - kPi) * (y - kPi);
        return (tanh(s1 * (d2 - (r + rw) * (r + rw))) - tanh(s2 * (d2 - (r - rw) * (r - rw)))) / 2.0;
    }
    case CMM_IC_4_NODES: {
        x -= floor(x / kTwoPi) * kTwoPi;
        y -= floor(y / kTwoPi) * kTwoPi;
        return cos(x) + cos(y) + 0.6 * cos(2 * x) + 0.2 * cos(3 * x);
    }
    case CMM_IC_QUADROPOLE: {
        double ret = 0;
        for (int jy = -2; jy <= 2; jy++)
            for (int jx = -2; jx <= 2; jx++) {
                double dx = x - kPi / 2 + jx * 2 * kPi, dy = y - kPi / 2 + jy * 2 * kPi;
                double A = 0.6258473, s = 0.5;
                double B = A / (s * s * s * s) * (dx * dy) * (dx * dx + dy * dy - 6 * s * s);
                double D = (dx * dx + dy * dy) / (2 * s * s);
                ret += B * exp(-D);
            }
        return ret;
    }
    case CMM_IC_ONE_VORTEX: {
        double sigma = kTwoPi * 0.125;
        return 1.0 / (kTwoPi * sigma * sigma) * exp(-((x - kPi) * (x - kPi) + (y - kPi) * (y - kPi)) / (2 * sigma * sigma));
    }
    case CMM_IC_SINGLE_SHEAR_LAYER:
        return (1 + 0.01 * cos(2 * x)) * exp(-50.0 * (y - kPi) * (y - kPi));
    case CMM_IC_TANH_SHEAR_LAYER: {
        double d = 14 * 2.0 / kTwoPi;
        double e = exp(-d * (y - kPi)) + exp(d * (y - kPi));
        return 5.0 * (1 + 0.01 * cos(2.0 * x)) * 4 / e / e;
    }
    case CMM_IC_SHIELDED_VORTEX: {
        double nu = 2e-1, nu_fac = 1 / (2 * nu * nu * nu), nu_c = 4 * nu, nu_s = 4 * nu;
        double xr = kPi - x, yr = kPi - y;
        return nu_fac * (nu_c - xr * xr - yr * yr) * exp(-(xr * xr + yr * yr) / nu_s);
    }
    case CMM_IC_TWO_COSINE: {
        x -= floor(x / kTwoPi) * kTwoPi;
        y -= floor(y / kTwoPi) * kTwoPi;
        return cos(x) * cos(y);
    }
    case CMM_IC_VORTEX_SHEETS: {
        double d = 1 / sqrt(p.prm[0]);
        double s = y - kPi - sin(x) / 2.0;
        return exp(-s * s / (2 * d * d)) / sqrt(2 * kPi) / d;
    }
    case CMM_IC_DOUBLE_SHEAR_LAYER: {
        const double delta = kPi / 15.0;
        double ret = 0.05 * cos(x);
        for (int im = -2; im <= 2; ++im) {
            double a = (y + im * kTwoPi - kPi / 2) / delta;
            double b = (y + im * kTwoPi - 3 * kPi / 2) / delta;
            double sa = 1.0 / cosh(a), sb = 1.0 / cosh(b);
            ret += (sa * sa - sb * sb) / delta;
        }
        return ret;
    }
    default:
        return 0.0;
    }
}

// ------------------------------------------------------------------------------------------------
// Row samplers feeding the R2C row transform (one value per grid point of the row grid).

// plain real array
struct SampleReal {
    const double *__restrict__ f;
    int nx;
    __device__ double operator()(int ix, int iy) const { return f[(long)iy * nx + ix]; }
};

// plain real array in FP32 (the split form's sampled vorticity)
struct SampleRealF {
    const float *__restrict__ f;
    int nx;
    __device__ double operator()(int ix, int iy) const { return (double)f[(long)iy * nx + ix]; }
};

// value of a Hermite field at the points of another grid (its own origin and spacing),
// k_h_sample (Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:122-135)
struct SampleHermite {
    DGrid gh, gt;
    const double *__restrict__ H;
    __device__ double operator()(int ix, int iy) const {
        return herm_eval(gh, H, gt.x0 + ix * gt.hx, gt.y0 + iy * gt.hy);
    }
};

// omega(X(x)) through the current map and the fine Hermite vorticity (mollifier 0/4/8),
// Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:402-461; positions use the map grid's origin.
// IL: the fine vorticity is given node-interleaved (W[4*i + plane]) instead of plane-major.
// The incompressibility check at R check points of one map column: the cell centres
// (xc, yc[r]) of the check grid, all inside map cell c (whose corner values vx, vy of the two map
// components are loaded as load16 gives them, without the wrap jumps).  The arithmetic of
// k_incomp_tensor: the value planes with their wrap jumps, contracted along x at xc, then along y
// per point; returns the largest |1 - det grad X|.
template <int R>
__device__ __forceinline__ double incomp_column(const DGrid &gm, const Cell &c, const double vx[16], const double vy[16],
                                                double xc, const double yc[R]) {
    const double wjx[4] = {c.wx0 * gm.jx, c.wx1 * gm.jx, c.wx0 * gm.jx, c.wx1 * gm.jx};
    const double wjy[4] = {c.wy0 * gm.jy, c.wy0 * gm.jy, c.wy1 * gm.jy, c.wy1 * gm.jy};
    const double fx = (xc - gm.x0) * gm.ihx;
    const double dx = fx - floor(fx);
    double bx[4], dbx[4];
    basis(dx, bx);
    basis_d(dx, dbx);
    double P[4][4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double *w = (k & 1) ? bx : dbx;
        const double *f = (k < 2) ? vx : vy;
        const double *wj = (k < 2) ? wjx : wjy;
#pragma unroll
        for (int n = 0; n < 2; ++n) {
            const double f0 = f[2 * n] + wj[2 * n], f1 = f[2 * n + 1] + wj[2 * n + 1];
            P[k][n] = w[0] * f0 + w[1] * f1 + gm.hx * (w[2] * f[4 + 2 * n] + w[3] * f[5 + 2 * n]);
            P[k][2 + n] = w[0] * f[8 + 2 * n] + w[1] * f[9 + 2 * n] + gm.hx * (w[2] * f[12 + 2 * n] + w[3] * f[13 + 2 * n]);
        }
    }
    double v = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const double fy = (yc[r] - gm.y0) * gm.ihy;
        const double d = fy - floor(fy);
        double by[4], dby[4];
        basis(d, by);
        basis_d(d, dby);
        auto ysum = [&](const double *q, const double *cw) {
            return cw[0] * q[0] + cw[1] * q[1] + gm.hy * (cw[2] * q[2] + cw[3] * q[3]);
        };
        const double Xx = ysum(P[0], by) * gm.ihx;
        const double Xy = ysum(P[1], dby) * gm.ihx;
        const double Yx = ysum(P[2], by) * gm.ihy;
        const double Yy = ysum(P[3], dby) * gm.ihy;
        v = fmax(v, fabs(1 - (Xx * Yy - Xy * Yx)));
    }
    return v;
}

template <bool IL>
struct SampleVortT {
    DGrid gm, gv, gf;
    const double *__restrict__ cx;
    const double *__restrict__ cy;
    const double *__restrict__ W;
    int molly;
    __device__ __forceinline__ double wval(double x, double y) const {
        if constexpr (IL) return herm_eval_il(gf, W, x, y);
        else return herm_eval(gf, W, x, y);
    }
    __device__ double operator()(int ix, int iy) const {
        const double x = gm.x0 + ix * gv.hx, y = gm.y0 + iy * gv.hy;
        double x2, y2, v;
        if (molly == 4) {
            diffeo(gm, cx, cy, x, y, x2, y2);
            v = wval(x2, y2) / 3.0;
            for (int m = 0; m < 4; ++m) {
                // the reference's (i_molly>>1+1)&1 parses as (i_molly>>2)&1
                x2 = x + gv.hx / 2.0 * ((m >> 2) & 1) * (-1 + 2 * (m & 1));
                y2 = y + gv.hy / 2.0 * ((m >> 1) & 1) * (-1 + 2 * (m & 1));
                diffeo(gm, cx, cy, x2, y2, x2, y2);
                v += wval(x2, y2) / 6.0;
            }
        } else if (molly == 8) {
            v = 0;
            for (int m = 0; m < 9; ++m) {
                x2 = x + gv.hx * (-1 + m % 3) / 2.0;
                y2 = y + gv.hy * (-1 + m / 3) / 2.0;
                diffeo(gm, cx, cy, x2, y2, x2, y2);
                // the reference's (1 + (i%3)&1) parses as ((1 + i%3) & 1)
                v += ((1 + (m % 3)) & 1) * ((1 + (m / 3)) & 1) * wval(x2, y2) / 16.0;
            }
        } else {
            diffeo(gm, cx, cy, x, y, x2, y2);
            v = wval(x2, y2);
        }
        return v;
    }
    // the check at the R cell centres (x + 1/2, y0 + r + 1/2) of the column, point by point
    template <int R>
    __device__ __forceinline__ double check_points(int x, int y0) const {
        double v = 0.0;
        const double xc = gm.x0 + x * gv.hx + 0.5 * gv.hx;
#pragma unroll 1
        for (int r = 0; r < R; ++r)
            v = fmax(v, fabs(1 - diffeo_det(gm, cx, cy, xc, gm.y0 + (y0 + r) * gv.hy + 0.5 * gv.hy)));
        return v;
    }
    // R points (x, y0 .. y0 + R - 1) of one column: the map cell corners are loaded once when all
    // R share the cell (the vorticity grid refines the map grid R times or more along y), and the
    // R vorticity gathers are independent.  The values equal operator().  CHK: also the
    // incompressibility check at the column's R cell centres of the same grid (the check grid
    // equals the vorticity grid), whose largest |1 - det grad X| goes to chk; the centres lie in
    // the nodes' map cell.  Returns false where the points were taken one by one (a sampling
    // stencil, or a column across map cells); the caller then checks the centres with
    // check_points, after the column's values are stored.
    template <int R, bool CHK = false>
    __device__ __forceinline__ bool column(int x, int y0, double *out, double &chk) const {
        if (molly != 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) out[r] = (*this)(x, y0 + r);
            return false;
        }
        const double xp = gm.x0 + x * gv.hx;
        Cell cf, cl;
        find_cell(gm, xp, gm.y0 + y0 * gv.hy, cf);
        find_cell(gm, xp, gm.y0 + (y0 + R - 1) * gv.hy, cl);
        if (!same_cell(cf, cl)) {
#pragma unroll
            for (int r = 0; r < R; ++r) out[r] = (*this)(x, y0 + r);
            return false;
        }
        double vx[16], vy[16];
        load16(cx, gm.n, cf, vx);
        load16(cy, gm.n, cf, vy);
        if constexpr (CHK) {
            double yc[R];
#pragma unroll
            for (int r = 0; r < R; ++r) yc[r] = gm.y0 + (y0 + r) * gv.hy + 0.5 * gv.hy;
            chk = fmax(chk, incomp_column<R>(gm, cf, vx, vy, xp + 0.5 * gv.hx, yc));
        }
        Cell wc[R];
        double wv[R][16];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Cell c;
            find_cell(gm, xp, gm.y0 + (y0 + r) * gv.hy, c);
            double px, py;
            diffeo_vals(gm, c, vx, vy, px, py);
            find_cell(gf, px, py, wc[r]);
            if constexpr (IL) load16_il(W, wc[r], wv[r]);
            else load16(W, gf.n, wc[r], wv[r]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) out[r] = herm_eval_vals(gf, wc[r], wv[r]);
        return true;
    }
    template <int R>
    __device__ __forceinline__ void column(int x, int y0, double *out) const {
        double chk = 0.0;
        column<R, false>(x, y0, out, chk);
    }
};
using SampleVort = SampleVortT<false>;
using SampleVortIL = SampleVortT<true>;


// Stacked maps as node records (op_pack_maps): one record per coarse node, so that a map
// evaluation gathers 4 records, 2 to 4 cache lines, instead of 32 values from 8 planes.
//   kStackRec64: 64 bytes, the ChiX planes then the ChiY planes (the advection's interleaved map);
//   kStackRec48: 48 bytes, the X and Y values in FP64, then dX/dx, dX/dy, d2X/dxdy, dY/dx, dY/dy,
//                d2Y/dxdy in FP32 and 8 unused bytes (split-precision engines).
template <int FMT>
__device__ __forceinline__ void load_map_rec(const char *__restrict__ M, const Cell &c, double vx[16], double vy[16]) {
    if constexpr (FMT == kStackRec64) {
        load_map_il(reinterpret_cast<const double *>(M), c, vx, vy);
    } else {
        const int idx[4] = {c.i00, c.i10, c.i01, c.i11};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const char *r = M + 48 * (long)idx[q];
            const double2 v = *reinterpret_cast<const double2 *>(r);
            const float4 d = *reinterpret_cast<const float4 *>(r + 16);
            const float2 e = *reinterpret_cast<const float2 *>(r + 32);
            vx[q] = v.x;
            vy[q] = v.y;
            vx[4 + q] = d.x;
            vx[8 + q] = d.y;
            vx[12 + q] = d.z;
            vy[4 + q] = d.w;
            vy[8 + q] = e.x;
            vy[12 + q] = e.y;
        }
    }
}

// bytes per node of a stack format
__host__ __device__ constexpr long stack_rec_bytes(int fmt) { return fmt == kStackRec48 ? 48 : 64; }

// maps m < n_maps read plane-major from (cx + m stride, cy + m stride) into records of format fmt at
// out + m N rec_bytes
__global__ void __launch_bounds__(256) k_pack_maps(const double *__restrict__ cx, const double *__restrict__ cy,
                                                   long stride, char *__restrict__ out, long N, int n_maps, int fmt) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= N * n_maps) return;
    const long m = t / N, i = t % N;
    const double *X = cx + stride * m, *Y = cy + stride * m;
    char *r = out + (m * N + i) * stack_rec_bytes(fmt);
    if (fmt == kStackRec48) {
        *reinterpret_cast<double2 *>(r) = double2{X[i], Y[i]};
        *reinterpret_cast<float4 *>(r + 16) = float4{(float)X[N + i], (float)X[2 * N + i], (float)X[3 * N + i], (float)Y[N + i]};
        *reinterpret_cast<float4 *>(r + 32) = float4{(float)Y[2 * N + i], (float)Y[3 * N + i], 0.0f, 0.0f};
    } else {
        double2 *d = reinterpret_cast<double2 *>(r);
        d[0] = double2{X[i], X[N + i]};
        d[1] = double2{X[2 * N + i], X[3 * N + i]};
        d[2] = double2{Y[i], Y[N + i]};
        d[3] = double2{Y[2 * N + i], Y[3 * N + i]};
    }
}

// Initial condition at X_1(...X_M(X_cur(x))): the current map then the stack from newest to
// oldest, wrapping after each layer (Cuda-CMM/src/simulation/cmm-simulation-host.cu:112-152,
// Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:302-373).  All layers are applied in
// registers; the stack is device resident, map m at stack + 8*Nc*m (X planes, then Y planes) or,
// as node records, at byte m*Nc*rec_bytes.
struct SampleCompose {
    DGrid gm, gt, gd;
    const double *__restrict__ cx;
    const double *__restrict__ cy;
    const double *__restrict__ stack;
    int nstack;
    int fmt;  // stack format: kStackFp64, kStackRec64 or kStackRec48 (the backward position() only)
    ICParams ic;
    int discrete;
    const double *__restrict__ disc;
    // One layer: the map's value at (x, y), wrapped into [x0, x0 + L).  On power-of-two map grids the
    // cell lookup takes its wrap counts by shifts, and the wrap after the layer multiplies by 1/L:
    // floor((p - x0) / L) and floor((p - x0) (1/L)) differ only for a point within rounding of a
    // period boundary, and then by one period, the same point of the torus.
    template <bool P2>
    __device__ __forceinline__ void layer(const double *planes, const char *rec, double ilx, double ily, double &x,
                                          double &y) const {
        Cell c;
        find_cell<P2>(gm, x, y, c);
        double vx[16], vy[16], px, py;
        if (rec) {
            if (fmt == kStackRec48) load_map_rec<kStackRec48>(rec, c, vx, vy);
            else load_map_rec<kStackRec64>(rec, c, vx, vy);
        } else {
            load16(planes, gm.n, c, vx);
            load16(planes + 4l * gm.n, gm.n, c, vy);
        }
        diffeo_vals(gm, c, vx, vy, px, py);
        x = px - floor((px - gm.x0) * ilx) * gm.lx;
        y = py - floor((py - gm.y0) * ily) * gm.ly;
    }
    template <bool P2>
    __device__ __forceinline__ void position_t(int ix, int iy, double &x, double &y) const {
        const double ilx = 1.0 / gm.lx, ily = 1.0 / gm.ly;
        x = gt.x0 + ix * gt.hx;
        y = gt.y0 + iy * gt.hy;
        // the current map (two plane-major fields)
        {
            Cell c;
            find_cell<P2>(gm, x, y, c);
            double vx[16], vy[16], px, py;
            load16(cx, gm.n, c, vx);
            load16(cy, gm.n, c, vy);
            diffeo_vals(gm, c, vx, vy, px, py);
            x = px - floor((px - gm.x0) * ilx) * gm.lx;
            y = py - floor((py - gm.y0) * ily) * gm.ly;
        }
        const char *recs = reinterpret_cast<const char *>(stack);
        for (int m = nstack - 1; m >= 0; --m) {
            if (fmt == kStackFp64) layer<P2>(stack + 8l * gm.n * m, nullptr, ilx, ily, x, y);
            else layer<P2>(nullptr, recs + stack_rec_bytes(fmt) * gm.n * m, ilx, ily, x, y);
        }
    }
    __device__ void position(int ix, int iy, double &x, double &y) const {
        if (gm.lgx >= 0 && gm.lgy >= 0) position_t<true>(ix, iy, x, y);
        else position_t<false>(ix, iy, x, y);
    }
    __device__ __forceinline__ void load_layer(const double *px_planes, const double *py_planes, const char *rec,
                                               const Cell &c, double vx[16], double vy[16]) const {
        if (rec) {
            if (fmt == kStackRec48) load_map_rec<kStackRec48>(rec, c, vx, vy);
            else load_map_rec<kStackRec64>(rec, c, vx, vy);
        } else {
            load16(px_planes, gm.n, c, vx);
            load16(py_planes, gm.n, c, vy);
        }
    }
    // NP points of one lane through one layer, in order: the corner records of a point's cell are
    // loaded unless the previous point's cell is the same one (neighbouring target points share a
    // map cell at most layers of a smooth stack, so the texture path sees fewer addresses).  Same
    // values and arithmetic per point as layer().
    template <bool P2, int NP>
    __device__ __forceinline__ void layer_n(const double *px_planes, const double *py_planes, const char *rec,
                                            double ilx, double ily, double *x, double *y) const {
        Cell c[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) find_cell<P2>(gm, x[p], y[p], c[p]);
        double vx[16], vy[16], px, py;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            if (p > 0) {
                // an index offset that is zero but depends on the previous point's result keeps the
                // compiler from issuing this point's loads early into registers of their own (which
                // costs occupancy); the empty asm only hides the zero from the optimiser
                int z = 0;
                asm volatile("" : "+v"(z) : "v"(__double2loint(x[p - 1])));
                c[p].i00 += z; c[p].i10 += z; c[p].i01 += z; c[p].i11 += z;
            }
            if (p == 0 || !same_cell(c[p - 1 < 0 ? 0 : p - 1], c[p])) load_layer(px_planes, py_planes, rec, c[p], vx, vy);
            diffeo_vals(gm, c[p], vx, vy, px, py);
            x[p] = px - floor((px - gm.x0) * ilx) * gm.lx;
            y[p] = py - floor((py - gm.y0) * ily) * gm.ly;
        }
    }
    // the NP points (ix, iy + p) of one lane
    template <bool P2, int NP>
    __device__ __forceinline__ void position_n_t(int ix, int iy, double *x, double *y) const {
        const double ilx = 1.0 / gm.lx, ily = 1.0 / gm.ly;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            x[p] = gt.x0 + ix * gt.hx;
            y[p] = gt.y0 + (iy + p) * gt.hy;
        }
        layer_n<P2, NP>(cx, cy, nullptr, ilx, ily, x, y);
        const char *recs = reinterpret_cast<const char *>(stack);
        for (int m = nstack - 1; m >= 0; --m) {
            if (fmt == kStackFp64)
                layer_n<P2, NP>(stack + 8l * gm.n * m, stack + 8l * gm.n * m + 4l * gm.n, nullptr, ilx, ily, x, y);
            else
                layer_n<P2, NP>(nullptr, nullptr, recs + stack_rec_bytes(fmt) * gm.n * m, ilx, ily, x, y);
        }
    }
    // forward order, Cuda-CMM/src/simulation/cmm-simulation-host.cu:128-148: the oldest stacked
    // forward map first, the current forward map last
    __device__ void position_forward(int ix, int iy, double &x, double &y) const {
        x = gt.x0 + ix * gt.hx;
        y = gt.y0 + iy * gt.hy;
        const long mstride = 8l * gm.n;
        for (int m = 0; m <= nstack; ++m) {
            const double *sx = m < nstack ? stack + mstride * m : cx;
            const double *sy = m < nstack ? sx + 4l * gm.n : cy;
            double px, py;
            diffeo(gm, sx, sy, x, y, px, py);
            x = px - floor((px - gm.x0) / gm.lx) * gm.lx;
            y = py - floor((py - gm.y0) / gm.ly) * gm.ly;
        }
    }
    __device__ double operator()(int ix, int iy) const {
        double x, y;
        position(ix, iy, x, y);
        return discrete ? herm_eval(gd, disc, x, y) : init_vorticity(x, y, ic);
    }
};

// omega through the map on the whole vorticity grid into a real row-major field f, for the row
// transform to read back.  A lane takes kSampleRows points of one column, so that consecutive
// lanes gather from consecutive vorticity nodes (coalesced 16-byte loads) and share map cells;
// without the FFT's LDS and barriers the gathers also keep more loads in flight than inside the
// row kernel.
constexpr int kSampleRows = 4;

// R: rows per lane (kSampleRows, or 2 where the vorticity grid refines the map grid only twice).
// CHK: the incompressibility check on the same grid's cell centres in the same pass (the sampler's
// column<R, true>), its maximum into chk_out with one 64-bit atomic max per block as k_incomp
// (chk_out cleared beforehand).
template <class S, class T = double, int R = kSampleRows, bool CHK = false>
__global__ void __launch_bounds__(256, 2) k_sample_field(S smp, int nx, int ny, T *__restrict__ f,
                                                         unsigned long long *__restrict__ chk_out) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    double chk = 0.0;
    if (t < (long)nx * (ny / R)) {
        const int x = (int)(t % nx), y0 = R * (int)(t / nx);
        double v[R];
        const bool fast = smp.template column<R, CHK>(x, y0, v, chk);
#pragma unroll
        for (int r = 0; r < R; ++r) f[(long)(y0 + r) * nx + x] = (T)v[r];
        if constexpr (CHK)
            if (!fast) chk = fmax(chk, smp.template check_points<R>(x, y0));
    }
    if constexpr (CHK) {
        for (int o = 32; o > 0; o >>= 1) chk = fmax(chk, __shfl_xor(chk, o, 64));
        __shared__ double red[4];
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = chk;
        __syncthreads();
        if (threadIdx.x == 0) {
            const double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
            atomicMax(chk_out, (unsigned long long)__double_as_longlong(m));
        }
    }
}

template <class S>
__device__ __forceinline__ void sample_quad(const S &smp, int x, int ya, int yb, cd *s) {
    for (int i = 0; i < 4; ++i) s[sw(x + i)] = cd{smp(x + i, ya), smp(x + i, yb)};
}

// ------------------------------------------------------------------------------------------------
// Workgroup FFT in LDS (length n <= kFftMaxN, power of two), kFftThreads lanes per transform.
constexpr int kFftMaxN = 4096;
constexpr int kFftThreads = 512;
constexpr int kFftPer = kFftMaxN / kFftThreads;  // elements per lane of a full-length transform

template <int R>
__device__ __forceinline__ void run_stage(cd *s, int n, int logn, int logNs, const cd *__restrict__ tw, bool inv) {
    constexpr int BPT = kFftPer / R > 0 ? kFftPer / R : 1;
    const int nb = n / R;
    cd v[BPT * R];
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        int j = threadIdx.x + b * kFftThreads;
        if (j < nb) stage_load<R>(s, v + b * R, j, nb);
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        int j = threadIdx.x + b * kFftThreads;
        if (j < nb) stage_store<R>(s, v + b * R, j, logNs, logn, tw, inv);
    }
    __syncthreads();
}

__device__ __forceinline__ void wg_fft(cd *s, int n, int logn, const cd *__restrict__ tw, bool inv) {
    int logNs = 0, rem = logn;
    while (rem > 0) {
        int rl = stage_radix_log(rem);
        if (rl == 3) run_stage<8>(s, n, logn, logNs, tw, inv);
        else if (rl == 2) run_stage<4>(s, n, logn, logNs, tw, inv);
        else run_stage<2>(s, n, logn, logNs, tw, inv);
        logNs += rl;
        rem -= rl;
    }
}

// Row R2C of two rows at once (z = row_a + i row_b), samples produced by the sampler.  Writes
// the half spectrum columns kx <= kx_max transposed: T[kx*ny + y].  Unnormalised.
template <class S>
__global__ void __launch_bounds__(kFftThreads, 4) k_row_r2c(S sampler, int nx, int ny, int logn, int kx_max,
                                                          const cd *__restrict__ tw, cd *__restrict__ T) {
    extern __shared__ cd s[];
    const int pair = xcd_block(blockIdx.x, gridDim.x);
    const int ya = 2 * pair, yb = ya + 1;
    // each lane samples 4 consecutive points of both rows (nx is a power of two >= 8)
    for (int x = 4 * threadIdx.x; x < nx; x += 4 * kFftThreads) sample_quad(sampler, x, ya, yb, s);
    __syncthreads();
    wg_fft(s, nx, logn, tw, false);
    for (int k = threadIdx.x; k <= kx_max; k += kFftThreads) {
        cd z = s[sw(k)];
        cd zc = s[sw((nx - k) & (nx - 1))];
        // X_a = (Z[k] + conj Z[n-k]) / 2,  X_b = (Z[k] - conj Z[n-k]) / (2i)
        cd xa = {(z.x + zc.x) * 0.5, (z.y - zc.y) * 0.5};
        cd xb = {(z.y + zc.y) * 0.5, -(z.x - zc.x) * 0.5};
        T[(long)k * ny + ya] = xa;
        T[(long)k * ny + yb] = xb;
    }
}

// Column stage.  For column kx <= kx_max: forward transform along y of T (length nin), scale
// by 1/norm, spectral operator, then two inverse transforms along y (length nout):
//   G0 = IFFT_y(P), G1 = IFFT_y(i ky P).
// mode 0 (Hermite, nin == nout): P = F.
// mode 1 (psi): P = -F/|k|^2 after the spectrum move between grids and the circular cut-off
//   (Cuda-CMM/src/numerical/cmm-fft.cu:294-323, :267-286, :55-78).
// mode 2 (spectrum given): T already holds the column spectrum; no forward transform, no scaling.
// mode 3 (Laplacian, nin == nout): P = -|k|^2 F.
struct ColParams {
    int nin, login, nout, logout, nx_in, nx_out;  // nx_* = grid sizes along x (for the move rule)
    double inv_norm;
    double fcut;
    double fkx, fky;  // 2 pi/(h*N) of the output grid along x and y
    int mode;
    int kx_max;
};

__global__ void __launch_bounds__(kFftThreads, 4) k_col(ColParams p, const cd *__restrict__ T,
                                                     const cd *__restrict__ tw_in, const cd *__restrict__ tw_out,
                                                     cd *__restrict__ G0, cd *__restrict__ G1) {
    extern __shared__ cd s[];
    const int kx = blockIdx.x;
    const cd *col = T + (long)kx * p.nin;
    for (int y = threadIdx.x; y < p.nin; y += kFftThreads) s[sw(y)] = col[y];
    __syncthreads();
    if (p.mode != 2) wg_fft(s, p.nin, p.login, tw_in, false);
    // build P for the output column in registers (each thread owns rows threadIdx.x + m*T)
    constexpr int MAXR = kFftPer;
    cd P[MAXR];
    const double kxv = p.fkx * kx;
#pragma unroll
    for (int m = 0; m < MAXR; ++m) {
        const int y = threadIdx.x + m * kFftThreads;
        cd v = {0.0, 0.0};
        if (y < p.nout) {
            if (p.mode == 1) {
                // spectrum move (only when the grids differ), Cuda-CMM/src/numerical/cmm-fft.cu:294-323
                bool zero = false;
                int ys = y;
                if (p.nout != p.nin || p.nx_out != p.nx_in) {
                    int mxo = kx + (kx > p.nx_out / 2.0) * (p.nx_out - 2 * kx);
                    int myo = y + (y > p.nout / 2.0) * (p.nout - 2 * y);
                    zero = (mxo > p.nx_in / 2.0) || (myo > p.nin / 2.0);
                    ys = y + (y > p.nout / 2.0) * (p.nin - p.nout);
                }
                if (!zero) {
                    cd f = s[sw(ys)];
                    v = {f.x * p.inv_norm, f.y * p.inv_norm};
                }
                // cut-off, Cuda-CMM/src/numerical/cmm-fft.cu:267-286 (DC removed)
                int mx = kx + (kx > p.nx_out / 2.0) * (p.nx_out - 2 * kx);
                int my = y + (y > p.nout / 2.0) * (p.nout - 2 * y);
                if ((mx * mx + my * my) > p.fcut * p.fcut || (kx == 0 && y == 0)) v = {0.0, 0.0};
                // inverse Laplacian, Cuda-CMM/src/numerical/cmm-fft.cu:55-78
                if (!(kx == 0 && y == 0)) {
                    double kyv = p.fky * (y - (y > p.nout / 2) * p.nout);
                    double k2 = kxv * kxv + kyv * kyv;
                    v = {-v.x / k2, -v.y / k2};
                }
            } else if (p.mode == 0) {
                cd f = s[sw(y)];
                v = {f.x * p.inv_norm, f.y * p.inv_norm};
            } else if (p.mode == 3) {
                // Laplacian, k_fft_lap_h (Cuda-CMM/src/numerical/cmm-fft.cu:28-51)
                cd f = s[sw(y)];
                const double kyv = p.fky * (y - (y > p.nout / 2) * p.nout);
                const double k2 = kxv * kxv + kyv * kyv;
                v = {-(f.x * p.inv_norm) * k2, -(f.y * p.inv_norm) * k2};
            } else {
                v = s[sw(y)];
            }
        }
        P[m] = v;
    }
    __syncthreads();
    // G0 = IFFT_y(P)
#pragma unroll
    for (int m = 0; m < MAXR; ++m) {
        const int y = threadIdx.x + m * kFftThreads;
        if (y < p.nout) s[sw(y)] = P[m];
    }
    __syncthreads();
    wg_fft(s, p.nout, p.logout, tw_out, true);
    cd *g0 = G0 + (long)kx * p.nout;
    for (int y = threadIdx.x; y < p.nout; y += kFftThreads) g0[y] = s[sw(y)];
    __syncthreads();
    // G1 = IFFT_y(i ky P), Cuda-CMM/src/numerical/cmm-fft.cu:169-187 (ky with the Nyquist index positive)
#pragma unroll
    for (int m = 0; m < MAXR; ++m) {
        const int y = threadIdx.x + m * kFftThreads;
        if (y < p.nout) {
            double kyv = p.fky * (y - (y > p.nout / 2) * p.nout);
            s[sw(y)] = cd{-P[m].y * kyv, P[m].x * kyv};
        }
    }
    __syncthreads();
    wg_fft(s, p.nout, p.logout, tw_out, true);
    cd *g1 = G1 + (long)kx * p.nout;
    for (int y = threadIdx.x; y < p.nout; y += kFftThreads) g1[y] = s[sw(y)];
}

// Row C2R producing the four Hermite planes of one row from the column results:
//   f = C2R(G0), fx = C2R(i kx G0), fy = C2R(G1), fxy = C2R(i kx G1)
// (Cuda-CMM/src/numerical/cmm-fft.cu:398-429).  Two real outputs come from one complex inverse
// transform of the Hermitian extensions A~ + i B~; the kx = 0 and kx = n/2 bins keep only their
// real parts (C2R semantics).  Columns above kx_max are zero.
__global__ void __launch_bounds__(kFftThreads, 4) k_row_c2r4(int nx, int ny, int logn, int kx_max, double fkx,
                                                          const cd *__restrict__ G0, const cd *__restrict__ G1,
                                                          const cd *__restrict__ tw, double *__restrict__ out) {
    extern __shared__ cd s[];
    const int y = xcd_block(blockIdx.x, gridDim.x);
    const long N = (long)nx * ny;
    const int h = nx / 2;
    for (int pass = 0; pass < 2; ++pass) {
        const cd *G = pass == 0 ? G0 : G1;
        for (int k = threadIdx.x; k <= h; k += kFftThreads) {
            cd A = {0.0, 0.0};
            if (k <= kx_max) A = G[(long)k * ny + y];
            const double kv = fkx * k;
            cd B = {-A.y * kv, A.x * kv};  // i kx A
            if (k == 0 || k == h) {
                s[sw(k)] = cd{A.x, B.x};
            } else {
                s[sw(k)] = cd{A.x - B.y, A.y + B.x};       // A + iB
                s[sw(nx - k)] = cd{A.x + B.y, -A.y + B.x};  // conj(A) + i conj(B)
            }
        }
        __syncthreads();
        wg_fft(s, nx, logn, tw, true);
        double *o0 = out + (2 * pass) * N + (long)y * nx;
        double *o1 = out + (2 * pass + 1) * N + (long)y * nx;
        for (int x = threadIdx.x; x < nx; x += kFftThreads) {
            cd v = s[sw(x)];
            o0[x] = v.x;
            o1[x] = v.y;
        }
        __syncthreads();
    }
}

// FP32 Hermite records whose value carries a low part (PsiRec32Pk): the value's leading FP32
// part sits in .x and an 8-bit signed code q of its remainder in the low byte of .w, whose upper
// 24 bits hold the cross derivative rounded to 15 explicit mantissa bits.  The remainder
// f - (float)f is at most half an FP32 ulp of the leading part; q counts it in units of 1/256 of
// that ulp (clamped to 127), so the value is kept to 2^-31 of itself or better.  The value plane
// of a stream function with |psi| of order one rounded to FP32 alone would move the velocity by
// 2^-24 |psi| 6/h, which the check's 1/h amplifies again; the cross derivative enters the
// velocity with weight h/7 or less, where its 2^-16 rounding stays at the level of the first
// derivatives' FP32 rounding.
__device__ __forceinline__ float4 pack_psi_rec(const double o[4]) {
    const float hi = (float)o[0];
    const unsigned e = (__float_as_uint(hi) >> 23) & 0xFFu;
    // 256 / ulp(hi) = 2^(158 - e), exact as a double for every biased exponent e
    const double inv = __longlong_as_double((long long)(1023 + 158 - (int)e) << 52);
    int q = (int)rint((o[0] - (double)hi) * inv);
    q = q < -128 ? -128 : q > 127 ? 127 : q;
    unsigned w = __float_as_uint((float)o[3]);
    w = ((w + 0x80u) & 0xFFFFFF00u) | ((unsigned)q & 0xFFu);
    return float4{hi, (float)o[1], (float)o[2], __uint_as_float(w)};
}

__device__ __forceinline__ void unpack_psi_rec(const float4 a, double &f, double &fxy) {
    const unsigned w = __float_as_uint(a.w);
    const unsigned e = (__float_as_uint(a.x) >> 23) & 0xFFu;
    const double s = __longlong_as_double((long long)(e + 865u) << 52);  // ulp(hi) / 256 = 2^(e - 158)
    f = (double)a.x + (double)(int)(signed char)(w & 0xFFu) * s;
    fxy = (double)__uint_as_float(w & 0xFFFFFF00u);
}

// k_row_c2r4 writing, besides the four planes of the new psi level, the next step's Lagrange-combined
// fields (k_psi_combine's C1, C12, C0, same weights and summation order) from the new level and
// the older ones: once a lane has formed the pass-1 planes of a node it reads back the pass-0 planes
// it wrote itself (its own earlier stores, in order) and combines.  FORM 0: FP64 node records (32
// bytes), 1: packed FP32 records (pack_psi_rec).  L: Lagrange order.
template <int FORM, int L>
__global__ void __launch_bounds__(kFftThreads, 4) k_row_c2r4_pre(int nx, int ny, int logn, int kx_max, double fkx,
                                                              const cd *__restrict__ G0, const cd *__restrict__ G1,
                                                              const cd *__restrict__ tw, double *__restrict__ out,
                                                              PsiPre pre) {
    extern __shared__ cd s[];
    constexpr int KM = kFftMaxN / kFftThreads;
    const int y = xcd_block(blockIdx.x, gridDim.x);
    const long N = (long)nx * ny;
    const int h = nx / 2;
    int z = 0;  // zero, made to depend on the previous node's result (see below)
    for (int pass = 0; pass < 2; ++pass) {
        const cd *G = pass == 0 ? G0 : G1;
        for (int k = threadIdx.x; k <= h; k += kFftThreads) {
            cd A = {0.0, 0.0};
            if (k <= kx_max) A = G[(long)k * ny + y];
            const double kv = fkx * k;
            cd B = {-A.y * kv, A.x * kv};  // i kx A
            if (k == 0 || k == h) {
                s[sw(k)] = cd{A.x, B.x};
            } else {
                s[sw(k)] = cd{A.x - B.y, A.y + B.x};       // A + iB
                s[sw(nx - k)] = cd{A.x + B.y, -A.y + B.x};  // conj(A) + i conj(B)
            }
        }
        __syncthreads();
        wg_fft(s, nx, logn, tw, true);
        double *o0 = out + (2 * pass) * N + (long)y * nx;
        double *o1 = out + (2 * pass + 1) * N + (long)y * nx;
#pragma unroll
        for (int q = 0; q < KM; ++q) {
            const int x = threadIdx.x + q * kFftThreads;
            if (x >= nx) continue;
            const cd v = s[sw(x)];
            o0[x] = v.x;
            o1[x] = v.y;
            if (pass == 0) continue;
            // the older levels' loads of a node are issued after the previous node's fields are
            // formed, so that they do not all take registers at once
            const long j = (long)y * nx + x + z;
            double lv[L][4];
            lv[0][0] = out[j];
            lv[0][1] = out[N + j];
            lv[0][2] = v.x;
            lv[0][3] = v.y;
#pragma unroll
            for (int l = 1; l < L; ++l)
#pragma unroll
                for (int p = 0; p < 4; ++p) lv[l][p] = pre.P.p[l][p * N + j];
            auto comb = [&](const double *w, void *dst) {
                double o[4];
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    double a = w[0] * lv[0][p];
#pragma unroll
                    for (int l = 1; l < L; ++l) a += w[l] * lv[l][p];
                    o[p] = a;
                }
                if constexpr (FORM == 0) {
                    double2 *d = reinterpret_cast<double2 *>(dst) + 2 * j;
                    d[0] = double2{o[0], o[1]};
                    d[1] = double2{o[2], o[3]};
                } else {
                    reinterpret_cast<float4 *>(dst)[j] = pack_psi_rec(o);
                }
            };
            comb(pre.w1, pre.C1);
            z = zero_after(lv[L - 1][3]);
            if (pre.C12) comb(pre.w12, pre.C12);
            if (pre.C0) {
                if constexpr (FORM == 0) {
                    double2 *d = reinterpret_cast<double2 *>(pre.C0) + 2 * j;
                    d[0] = double2{lv[0][0], lv[0][1]};
                    d[1] = double2{lv[0][2], lv[0][3]};
                } else {
                    reinterpret_cast<float4 *>(pre.C0)[j] = pack_psi_rec(lv[0]);
                }
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------------
// Point kernels used outside the fused transforms

template <class S>
__global__ void __launch_bounds__(256) k_sample_grid(S sampler, int nx, int ny, double *__restrict__ out) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= nx || iy >= ny) return;
    out[(long)iy * nx + ix] = sampler(ix, iy);
}

// FMT: the stack format, fixed at compile time (each format its own kernel)
template <int FMT>
__global__ void __launch_bounds__(256) k_compose_positions(SampleCompose sc, double *__restrict__ xy) {
    sc.fmt = FMT;
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= sc.gt.nx || iy >= sc.gt.ny) return;
    double x, y;
    sc.position(ix, iy, x, y);
    const long i = (long)iy * sc.gt.nx + ix;
    xy[2 * i] = x;
    xy[2 * i + 1] = y;
}

// NP points per lane, rows iy .. iy + NP - 1 of one column (the target grid's ny a multiple of NP)
template <int FMT, int NP>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_compose_multi(SampleCompose sc, double *__restrict__ xy) {
    sc.fmt = FMT;
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = NP * (blockIdx.y * blockDim.y + threadIdx.y);
    if (ix >= sc.gt.nx || iy >= sc.gt.ny) return;
    double x[NP], y[NP];
    if (sc.gm.lgx >= 0 && sc.gm.lgy >= 0) sc.position_n_t<true, NP>(ix, iy, x, y);
    else sc.position_n_t<false, NP>(ix, iy, x, y);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const long i = (long)(iy + p) * sc.gt.nx + ix;
        xy[2 * i] = x[p];
        xy[2 * i + 1] = y[p];
    }
}

// the initial condition (or the discrete field) at the composed positions, two rows per lane
// as k_compose_multi; out row-major on the target grid (ny even)
template <int FMT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_sample_init_multi(SampleCompose sc,
                                                                                                 double *__restrict__ out) {
    sc.fmt = FMT;
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = 2 * (blockIdx.y * blockDim.y + threadIdx.y);
    if (ix >= sc.gt.nx || iy >= sc.gt.ny) return;
    double x[2], y[2];
    if (sc.gm.lgx >= 0 && sc.gm.lgy >= 0) sc.position_n_t<true, 2>(ix, iy, x, y);
    else sc.position_n_t<false, 2>(ix, iy, x, y);
#pragma unroll
    for (int p = 0; p < 2; ++p)
        out[(long)(iy + p) * sc.gt.nx + ix] =
            sc.discrete ? herm_eval(sc.gd, sc.disc, x[p], y[p]) : init_vorticity(x[p], y[p], sc.ic);
}

__global__ void __launch_bounds__(256) k_forward_positions(SampleCompose sc, double *__restrict__ xy) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= sc.gt.nx || iy >= sc.gt.ny) return;
    double x, y;
    sc.position_forward(ix, iy, x, y);
    const long i = (long)iy * sc.gt.nx + ix;
    xy[2 * i] = x;
    xy[2 * i + 1] = y;
}

// Forwarded particles: k_h_sample_points_map (Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:97-118)
// once per map, stacked maps first..nstack-1 then (cur) the current map.  A point strictly inside
// the bounds b goes through the map and is wrapped by x - floor(x / Lx) Lx; others become NaN.
struct PointBounds {
    double x0, x1, y0, y1;
};

__global__ void __launch_bounds__(256) k_forward_points(DGrid gm, PointBounds b, const double *__restrict__ stack,
                                                        int first, int nstack, const double *__restrict__ cx,
                                                        const double *__restrict__ cy, int cur,
                                                        const double *__restrict__ in, double *__restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = in[2 * i], y = in[2 * i + 1];
    const long mstride = 8l * gm.n;
    for (int m = first; m < nstack + cur; ++m) {
        const double *sx = m < nstack ? stack + mstride * m : cx;
        const double *sy = m < nstack ? sx + 4l * gm.n : cy;
        if (x > b.x0 && x < b.x1 && y > b.y0 && y < b.y1) {
            double px, py;
            diffeo(gm, sx, sy, x, y, px, py);
            x = px - floor(px / gm.lx) * gm.lx;
            y = py - floor(py / gm.ly) * gm.ly;
        } else {
            x = y = __builtin_nan("");
        }
    }
    out[2 * i] = x;
    out[2 * i + 1] = y;
}

// Fluid particles, Cuda-CMM/src/numerical/cmm-particles.cu:119-276 (level 0 = t_{n+1}, 1 = t_n)
struct PartParams {
    double Lp[4];
    double dt;
    int L, scheme, n;
};

__global__ void __launch_bounds__(256) k_particles(DGrid g, PsiLevels P, PartParams q, double *__restrict__ pos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.n) return;
    double p0 = pos[2 * i], p1 = pos[2 * i + 1];
    const double dt = q.dt;
    const int L = q.L;
    PsiLevels P1 = {{P.p[1], P.p[1], P.p[1], P.p[1]}};
    PsiLevels P2 = {{P.p[2], P.p[2], P.p[2], P.p[2]}};
    double u[8], k[8];
    switch (q.scheme) {
    case 10:
        herm_vel(g, P1, 1, p0, p1, u);
        p0 += dt * u[0];
        p1 += dt * u[1];
        break;
    case 20:
        herm_vel(g, P1, 1, p0, p1, k);
        herm_vel(g, P, 1, p0 + dt * k[0], p1 + dt * k[1], k + 2);
        p0 += dt * (k[0] + k[2]) / 2;
        p1 += dt * (k[1] + k[3]) / 2;
        break;
    case 30:
        herm_vel(g, P1, 1, p0, p1, k);
        herm_vel(g, P, L, p0 + dt * k[0] / 2.0, p1 + dt * k[1] / 2.0, u);
        lsum(q.Lp, u, L, k[2], k[3]);
        herm_vel(g, P, 1, p0 - dt * k[0] + 2 * dt * k[2], p1 - dt * k[1] + 2 * dt * k[3], k + 4);
        p0 += dt * (k[0] + 4 * k[2] + k[4]) / 6.0;
        p1 += dt * (k[1] + 4 * k[3] + k[5]) / 6.0;
        break;
    case 40:
        herm_vel(g, P1, 1, p0, p1, k);
        herm_vel(g, P, L, p0 + dt * k[0] / 2.0, p1 + dt * k[1] / 2.0, u);
        lsum(q.Lp, u, L, k[2], k[3]);
        herm_vel(g, P, L, p0 + dt * k[2] / 2.0, p1 + dt * k[3] / 2.0, u);
        lsum(q.Lp, u, L, k[4], k[5]);
        herm_vel(g, P, 1, p0 + dt * k[4], p1 + dt * k[5], k + 6);
        p0 += dt * (k[0] + 2 * k[2] + 2 * k[4] + k[6]) / 6.0;
        p1 += dt * (k[1] + 2 * k[3] + 2 * k[5] + k[7]) / 6.0;
        break;
    case 31:
        herm_vel(g, P1, 1, p0, p1, k);
        herm_vel(g, P, 1, p0 + dt * k[0], p1 + dt * k[1], k + 2);
        herm_vel(g, P2, 1, p0 + dt * (k[0] - 2 * k[2]), p1 + dt * (k[1] - 2 * k[3]), k + 4);
        p0 += dt * (8 * k[0] + 5 * k[2] - k[4]) / 12.0;
        p1 += dt * (8 * k[1] + 5 * k[3] - k[5]) / 12.0;
        break;
    case 41:
        herm_vel(g, P1, 1, p0, p1, k);
        herm_vel(g, P, 1, p0 + dt * k[0], p1 + dt * k[1], k + 2);
        herm_vel(g, P, L, p0 + dt / 8.0 * (3 * k[0] + k[2]), p1 + dt / 8.0 * (3 * k[1] + k[3]), u);
        lsum(q.Lp, u, L, k[4], k[5]);
        herm_vel(g, P, 1, p0 + dt / 3.0 * (-k[2] + 4 * k[4]), p1 + dt / 3.0 * (-k[3] + 4 * k[5]), k + 6);
        p0 += dt * (2 * k[0] - k[2] + 8 * k[4] + 3 * k[6]) / 12.0;
        p1 += dt * (2 * k[1] - k[3] + 8 * k[5] + 3 * k[7]) / 12.0;
        break;
    default:
        p0 = p1 = 0;
    }
    pos[2 * i] = p0 - floor((p0 - g.x0) / g.lx) * g.lx;
    pos[2 * i + 1] = p1 - floor((p1 - g.y0) / g.ly) * g.ly;
}

// Inertial particles, Particle_advect_inertia (Cuda-CMM/src/numerical/cmm-particles.cu:280-551):
// position and velocity advance through the same stages, x' = v, v' = -(v - u(x, t)) / tau, with u
// from level 1 (t_n), level 0 (t_{n+1}), level 2 (t_n - dt, RK3Mod) or the Lagrange combination of
// the levels to t_n + dt/2 (Lp), the same sources as the fluid particles' stages.
__global__ void __launch_bounds__(256) k_particles_inertial(DGrid g, PsiLevels P, PartParams q, double tau,
                                                            double *__restrict__ pos, double *__restrict__ vel) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.n) return;
    const double x = pos[2 * i], y = pos[2 * i + 1];
    const double vx = vel[2 * i], vy = vel[2 * i + 1];
    const double dt = q.dt;
    const int L = q.L;
    // fluid velocity of one level, or of the combination to t_n + dt/2
    auto ulev = [&](int l, double px, double py, double &ux, double &uy) {
        double u[2];
        herm_vel(g, PsiLevels{{P.p[l], P.p[l], P.p[l], P.p[l]}}, 1, px, py, u);
        ux = u[0];
        uy = u[1];
    };
    auto uhalf = [&](double px, double py, double &ux, double &uy) {
        double u[8];
        herm_vel(g, P, L, px, py, u);
        lsum(q.Lp, u, L, ux, uy);
    };
    double xn, yn, wx, wy;
    double ux, uy;
    switch (q.scheme) {
    case 10: {
        ulev(1, x, y, ux, uy);
        xn = x + dt * vx;
        yn = y + dt * vy;
        wx = vx - dt * (vx - ux) / tau;
        wy = vy - dt * (vy - uy) / tau;
        break;
    }
    case 20: {
        ulev(1, x, y, ux, uy);
        const double a1x = -(vx - ux) / tau, a1y = -(vy - uy) / tau;
        const double v2x = vx + dt * a1x, v2y = vy + dt * a1y;
        ulev(0, x + dt * vx, y + dt * vy, ux, uy);
        const double a2x = -(v2x - ux) / tau, a2y = -(v2y - uy) / tau;
        xn = x + dt / 2.0 * (vx + v2x);
        yn = y + dt / 2.0 * (vy + v2y);
        wx = vx + dt / 2.0 * (a1x + a2x);
        wy = vy + dt / 2.0 * (a1y + a2y);
        break;
    }
    case 30: {
        ulev(1, x, y, ux, uy);
        const double a1x = -(vx - ux) / tau, a1y = -(vy - uy) / tau;
        const double v2x = vx + dt / 2.0 * a1x, v2y = vy + dt / 2.0 * a1y;
        uhalf(x + dt / 2.0 * vx, y + dt / 2.0 * vy, ux, uy);
        const double a2x = -(v2x - ux) / tau, a2y = -(v2y - uy) / tau;
        const double v3x = vx + dt * (-a1x + 2 * a2x), v3y = vy + dt * (-a1y + 2 * a2y);
        ulev(0, x + dt * (-vx + 2 * v2x), y + dt * (-vy + 2 * v2y), ux, uy);
        const double a3x = -(v3x - ux) / tau, a3y = -(v3y - uy) / tau;
        xn = x + dt / 6.0 * (vx + 4 * v2x + v3x);
        yn = y + dt / 6.0 * (vy + 4 * v2y + v3y);
        wx = vx + dt / 6.0 * (a1x + 4 * a2x + a3x);
        wy = vy + dt / 6.0 * (a1y + 4 * a2y + a3y);
        break;
    }
    case 40: {
        ulev(1, x, y, ux, uy);
        const double a1x = -(vx - ux) / tau, a1y = -(vy - uy) / tau;
        const double v2x = vx + dt / 2.0 * a1x, v2y = vy + dt / 2.0 * a1y;
        uhalf(x + dt / 2.0 * vx, y + dt / 2.0 * vy, ux, uy);
        const double a2x = -(v2x - ux) / tau, a2y = -(v2y - uy) / tau;
        const double v3x = vx + dt / 2.0 * a2x, v3y = vy + dt / 2.0 * a2y;
        uhalf(x + dt / 2.0 * v2x, y + dt / 2.0 * v2y, ux, uy);
        const double a3x = -(v3x - ux) / tau, a3y = -(v3y - uy) / tau;
        const double v4x = vx + dt * a3x, v4y = vy + dt * a3y;
        ulev(0, x + dt * v3x, y + dt * v3y, ux, uy);
        const double a4x = -(v4x - ux) / tau, a4y = -(v4y - uy) / tau;
        xn = x + dt / 6.0 * (vx + 2 * v2x + 2 * v3x + v4x);
        yn = y + dt / 6.0 * (vy + 2 * v2y + 2 * v3y + v4y);
        wx = vx + dt / 6.0 * (a1x + 2 * a2x + 2 * a3x + a4x);
        wy = vy + dt / 6.0 * (a1y + 2 * a2y + 2 * a3y + a4y);
        break;
    }
    case 31: {
        ulev(1, x, y, ux, uy);
        const double a1x = -(vx - ux) / tau, a1y = -(vy - uy) / tau;
        const double v2x = vx + dt * a1x, v2y = vy + dt * a1y;
        ulev(0, x + dt * vx, y + dt * vy, ux, uy);
        const double a2x = -(v2x - ux) / tau, a2y = -(v2y - uy) / tau;
        const double v3x = vx + dt * (a1x - 2 * a2x), v3y = vy + dt * (a1y - 2 * a2y);
        ulev(2, x + dt * (vx - 2 * v2x), y + dt * (vy - 2 * v2y), ux, uy);
        const double a3x = -(v3x - ux) / tau, a3y = -(v3y - uy) / tau;
        xn = x + dt / 12.0 * (8 * vx + 5 * v2x - v3x);
        yn = y + dt / 12.0 * (8 * vy + 5 * v2y - v3y);
        wx = vx + dt / 12.0 * (8 * a1x + 5 * a2x - a3x);
        wy = vy + dt / 12.0 * (8 * a1y + 5 * a2y - a3y);
        break;
    }
    case 41: {
        ulev(1, x, y, ux, uy);
        const double a1x = -(vx - ux) / tau, a1y = -(vy - uy) / tau;
        const double v2x = vx + dt * a1x, v2y = vy + dt * a1y;
        ulev(0, x + dt * vx, y + dt * vy, ux, uy);
        const double a2x = -(v2x - ux) / tau, a2y = -(v2y - uy) / tau;
        const double v3x = vx + dt / 8.0 * (3 * a1x + a2x), v3y = vy + dt / 8.0 * (3 * a1y + a2y);
        uhalf(x + dt / 8.0 * (3 * vx + v2x), y + dt / 8.0 * (3 * vy + v2y), ux, uy);
        const double a3x = -(v3x - ux) / tau, a3y = -(v3y - uy) / tau;
        const double v4x = vx + dt / 3.0 * (-a2x + 4 * a3x), v4y = vy + dt / 3.0 * (-a2y + 4 * a3y);
        ulev(0, x + dt / 3.0 * (-v2x + 4 * v3x), y + dt / 3.0 * (-v2y + 4 * v3y), ux, uy);
        const double a4x = -(v4x - ux) / tau, a4y = -(v4y - uy) / tau;
        xn = x + dt / 12.0 * (2 * vx - v2x + 8 * v3x + 3 * v4x);
        yn = y + dt / 12.0 * (2 * vy - v2y + 8 * v3y + 3 * v4y);
        wx = vx + dt / 12.0 * (2 * a1x - a2x + 8 * a3x + 3 * a4x);
        wy = vy + dt / 12.0 * (2 * a1y - a2y + 8 * a3y + 3 * a4y);
        break;
    }
    default:
        xn = yn = wx = wy = 0.0;
    }
    pos[2 * i] = xn - floor((xn - g.x0) / g.lx) * g.lx;
    pos[2 * i + 1] = yn - floor((yn - g.y0) / g.ly) * g.ly;
    vel[2 * i] = wx;
    vel[2 * i + 1] = wy;
}

// Start velocity of inertial particles: the fluid velocity of psi level 0 (fluid = 1) or zero,
// Particle_advect_inertia_init (Cuda-CMM/src/numerical/cmm-particles.cu:49-59) and the memset of
// Cuda-CMM/src/simulation/cmm-euler2d.cu:630
__global__ void __launch_bounds__(256) k_particles_vel_init(DGrid g, const double *__restrict__ psi0, int n, int fluid,
                                                            const double *__restrict__ pos, double *__restrict__ vel) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[2] = {0.0, 0.0};
    if (fluid) herm_vel(g, PsiLevels{{psi0, psi0, psi0, psi0}}, 1, pos[2 * i], pos[2 * i + 1], u);
    vel[2 * i] = u[0];
    vel[2 * i + 1] = u[1];
}

// (d psi/dx, d psi/dy) of psi level 0 at points strictly inside the grid's bounds, NaN outside:
// the velocity output of fluid particles, k_h_sample_points_dxdy
// (Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:138-155) with cmm-hermite.cu:343-371
__global__ void __launch_bounds__(256) k_points_dxdy(DGrid g, double x1, double y1, const double *__restrict__ psi0,
                                                     int n, const double *__restrict__ pos, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = pos[2 * i], y = pos[2 * i + 1];
    if (x > g.x0 && x < x1 && y > g.y0 && y < y1) {
        Cell c;
        find_cell(g, x, y, c);
        double bx[4], by[4], dbx[4], dby[4], v[16];
        basis(c.dx, bx);
        basis(c.dy, by);
        basis_d(c.dx, dbx);
        basis_d(c.dy, dby);
        load16(psi0, g.n, c, v);
        out[2 * i] = hsum_v(v, dbx, by, g.hx, g.hy) / g.hx;
        out[2 * i + 1] = hsum_v(v, bx, dby, g.hx, g.hy) / g.hy;
    } else {
        out[2 * i] = out[2 * i + 1] = __builtin_nan("");
    }
}

// ------------------------------------------------------------------------------------------------
// Vlasov-Poisson psi (x, v), evaluate_potential_from_density_hermite
// (Cuda-CMM/src/simulation/cmm-simulation-host.cu:385-471, :478-494, :534-571).
//   f(X(x)) on the vort grid -> I(x) = hv (f(x, v0) + sum_v f) - 1 (integral_v, whose output row
//   aliases the first input row, so v0 counts twice) -> phi_hat = F/(NX k^2), DC 0, zero padded to
//   the psi grid, C2R -> phi(x);  bump sigma(v), centred and scaled to max 1 -> H_hat = S/(NY k^2),
//   C2R -> H(v);  psi = phi(x) - H(v) -> Hermite planes.

// partial column sums over chunks of `rows` rows: part[c*nx + x] = sum_{y in chunk c} f[y*nx + x]
__global__ void __launch_bounds__(256) k_vp_colsum(const double *__restrict__ f, int nx, int ny, int rows,
                                                   double *__restrict__ part) {
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= nx) return;
    const int y0 = blockIdx.y * rows, y1 = min(ny, y0 + rows);
    double a = 0.0;
    for (int y = y0; y < y1; ++y) a += f[(long)y * nx + x];
    part[(long)blockIdx.y * nx + x] = a;
}

// block-wide sum / max over kFftThreads lanes (fixed tree order)
__device__ double block_reduce(double v, double *red, bool is_max) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kFftThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = is_max ? fmax(red[threadIdx.x], red[threadIdx.x + s])
                                                            : red[threadIdx.x] + red[threadIdx.x + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// s holds the forward transform of a real signal of length n_in (bins k <= n_in/2 valid).  Writes
// the Hermitian spectrum of length n_out with X[k] = s[k] * scale / (fk k)^2 for 1 <= k <= n_in/2
// (zero elsewhere and at DC; only the real part at n_out/2), then inverse-transforms in place:
// the C2R of the reference's zero-padded 1-D spectrum.
__device__ void inv_lap_c2r(cd *s, int n_in, int n_out, int logn_out, double scale, double fk,
                            const cd *__restrict__ tw_out) {
    constexpr int PER = kFftMaxN / 2 / kFftThreads + 1;
    cd keep[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int k = threadIdx.x + m * kFftThreads;
        keep[m] = (k <= n_in / 2 && k > 0) ? s[sw(k)] : cd{0.0, 0.0};
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n_out; i += kFftThreads) s[sw(i)] = cd{0.0, 0.0};
    __syncthreads();
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int k = threadIdx.x + m * kFftThreads;
        if (k > 0 && k <= n_in / 2 && k <= n_out / 2) {
            const double kk = fk * k, k2 = kk * kk;
            cd v{keep[m].x * scale / k2, keep[m].y * scale / k2};
            if (k == n_out / 2) {
                s[sw(k)] = cd{v.x, 0.0};
            } else {
                s[sw(k)] = v;
                s[sw(n_out - k)] = cd{v.x, -v.y};
            }
        }
    }
    __syncthreads();
    wg_fft(s, n_out, logn_out, tw_out, true);
}

// Spectral derivative of the n real values held in s[].x (any imaginary parts are ignored):
// forward transform, times i fk k / n with the Nyquist bin dropped, inverse transform; the
// derivative is left in s[].x.  The 2D Hermite transform of a field that depends on one
// coordinate only reduces to this along that coordinate: its column stage takes the Nyquist
// index positive and its C2R keeps the real part, which drops that bin's purely imaginary term.
__device__ void spec_deriv_1d(cd *s, int n, int logn, double fk, const cd *__restrict__ tw) {
    for (int i = threadIdx.x; i < n; i += kFftThreads) s[sw(i)].y = 0.0;
    __syncthreads();
    wg_fft(s, n, logn, tw, false);
    for (int k = threadIdx.x; k < n; k += kFftThreads) {
        const int kk = k < n / 2 ? k : (k == n / 2 ? 0 : k - n);
        const double fac = fk * kk / (double)n;
        const cd v = s[sw(k)];
        s[sw(k)] = cd{-v.y * fac, v.x * fac};
    }
    __syncthreads();
    wg_fft(s, n, logn, tw, true);
}

// two workgroups: phi and phi' (block 0), H and H' (block 1, independent of f), np values each,
// into phiH = [phi, H, phi', H']
// phi(x) and phi'(x) of one instance: I(x) = hv (f(x, v0) + sum_v f) - 1 (the row v0 once more,
// then the chunk sums in order), the 1D Poisson solve and zero padding to the psi grid, then the
// spectral derivative; into job.phi[0, np) and job.phi[2 np, 3 np)
__device__ void vp_phi_block(const VPJob &job, cd *s) {
    const VP1D &q = job.q;
    for (int x = threadIdx.x; x < q.nv; x += kFftThreads) {
        double a = job.f[x];
        for (int c = 0; c < q.nchunks; ++c) a += job.part[(long)c * q.nv + x];
        a *= q.hv;
        a += -1.0;
        s[sw(x)] = cd{a, 0.0};
    }
    __syncthreads();
    wg_fft(s, q.nv, q.lognv, job.tw_v, false);
    inv_lap_c2r(s, q.nv, q.np, q.lognp, 1.0 / (double)q.nv, q.fkx_v, job.tw_p);
    for (int x = threadIdx.x; x < q.np; x += kFftThreads) job.phi[x] = s[sw(x)].x;
    spec_deriv_1d(s, q.np, q.lognp, q.fkx_p, job.tw_p);
    for (int x = threadIdx.x; x < q.np; x += kFftThreads) job.phi[2 * q.np + x] = s[sw(x)].x;
}

__global__ void __launch_bounds__(kFftThreads) k_vp_phi(VPJob job) {
    extern __shared__ cd s[];
    vp_phi_block(job, s);
}

// the bump function over v on the psi grid, centred and scaled to maximum 1, its 1D Poisson solve
// H and the derivative H'; into H[0, np) and H[2 np, 3 np) (the layout of phi's buffer).  It
// depends on the grid only.
__global__ void __launch_bounds__(kFftThreads) k_vp_bump(VP1D q, const cd *__restrict__ tw_p, double *__restrict__ H) {
    extern __shared__ cd s[];
    __shared__ double red[kFftThreads];
    double sum = 0.0;
    for (int i = threadIdx.x; i < q.np; i += kFftThreads) {
        const double v = q.v0 + i * q.hv_p;
        const double b = 0.5 - 0.5 * tanh(2 * kPi * (fabs(v) - 0.4 * q.Lv) / (0.1 * q.Lv));
        s[sw(i)] = cd{b, 0.0};
        sum += b;
    }
    const double mean = block_reduce(sum, red, false) / (double)q.np;
    double mx = -1e300;
    for (int i = threadIdx.x; i < q.np; i += kFftThreads) {
        const double b = s[sw(i)].x - mean;
        s[sw(i)].x = b;
        mx = fmax(mx, b);
    }
    mx = block_reduce(mx, red, true);
    for (int i = threadIdx.x; i < q.np; i += kFftThreads) s[sw(i)].x /= mx;
    __syncthreads();
    wg_fft(s, q.np, q.lognp, tw_p, false);
    inv_lap_c2r(s, q.np, q.np, q.lognp, 1.0 / (double)q.np, q.fky_p, tw_p);
    for (int i = threadIdx.x; i < q.np; i += kFftThreads) H[i] = s[sw(i)].x;
    spec_deriv_1d(s, q.np, q.lognp, q.fky_p, tw_p);
    for (int i = threadIdx.x; i < q.np; i += kFftThreads) H[2 * q.np + i] = s[sw(i)].x;
}

// The Hermite planes of psi(x, v) = phi(x) - H(v) on the psi grid (k_assemble_psi_periodified and
// the Hermite transform of its result): psi, d/dx = phi', d/dv = -H', d2/dxdv = 0.  psi is a sum
// of one-variable functions, so its 2D spectrum lies on the two axes and these are the planes
// the 2D transform gives, without it.
// phi: phi at [0, np), phi' at [2 np, 3 np); H: H at [0, np), H' at [2 np, 3 np)
__global__ void __launch_bounds__(256) k_vp_psi_planes(const double *__restrict__ phi, const double *__restrict__ H,
                                                       int np, int lognp, double *__restrict__ out) {
    const long N = (long)np * np;
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= N) return;
    const int x = (int)(t & (np - 1)), y = (int)(t >> lognp);
    out[t] = phi[x] - H[y];
    out[N + t] = phi[2 * np + x];
    out[2 * N + t] = -H[2 * np + y];
    out[3 * N + t] = 0.0;
}


// ------------------------------------------------------------------------------------------------
// Host-side helpers: twiddle tables and workspaces, cached per device.

struct DeviceCache {
    std::map<int, cd *> tw;                 // n -> exp(-2 pi i m/n), m < n
    void *work = nullptr;
    size_t work_bytes = 0;
    unsigned long long *scalar = nullptr;
};

static std::mutex g_mutex;
static std::map<int, DeviceCache> g_cache;

static int current_device(int &dev) { return hip_check(hipGetDevice(&dev), "hipGetDevice"); }

int twiddles(int n, const cd **out) {
    int dev;
    if (int rc = current_device(dev)) return rc;
    std::lock_guard<std::mutex> lk(g_mutex);
    DeviceCache &c = g_cache[dev];
    auto it = c.tw.find(n);
    if (it != c.tw.end()) {
        *out = it->second;
        return CMM_OK;
    }
    std::vector<cd> h(n);
    for (int m = 0; m < n; ++m) {
        long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
        h[m] = cd{(double)cosl(a), (double)sinl(a)};
    }
    cd *d = nullptr;
    if (hipMalloc(&d, sizeof(cd) * n) != hipSuccess) {
        set_error("hipMalloc twiddles");
        return CMM_ERR_ALLOC;
    }
    if (int rc = hip_check(hipMemcpy(d, h.data(), sizeof(cd) * n, hipMemcpyHostToDevice), "hipMemcpy twiddles")) return rc;
    c.tw[n] = d;
    *out = d;
    return CMM_OK;
}

int workspace(size_t bytes, void **out) {
    int dev;
    if (int rc = current_device(dev)) return rc;
    std::lock_guard<std::mutex> lk(g_mutex);
    DeviceCache &c = g_cache[dev];
    if (c.work_bytes < bytes) {
        if (c.work) (void)hipFree(c.work);
        c.work = nullptr;
        c.work_bytes = 0;
        if (hipMalloc(&c.work, bytes) != hipSuccess) {
            set_error("hipMalloc workspace");
            return CMM_ERR_ALLOC;
        }
        c.work_bytes = bytes;
    }
    *out = c.work;
    return CMM_OK;
}

int scalar_slot(unsigned long long **out) {
    int dev;
    if (int rc = current_device(dev)) return rc;
    std::lock_guard<std::mutex> lk(g_mutex);
    DeviceCache &c = g_cache[dev];
    if (!c.scalar) {
        if (hipMalloc(&c.scalar, 256) != hipSuccess) {
            set_error("hipMalloc scalar");
            return CMM_ERR_ALLOC;
        }
    }
    *out = c.scalar;
    return CMM_OK;
}

static bool is_pow2(int n) { return n >= 2 && (n & (n - 1)) == 0; }
static int ilog2(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}
static bool grid_ok(const cmm_grid &g) { return g.nx > 0 && g.ny > 0 && g.x1 > g.x0 && g.y1 > g.y0; }
static int fft_grid_ok(const cmm_grid &g, const char *what) {
    if (!grid_ok(g) || !is_pow2(g.nx) || !is_pow2(g.ny) || g.nx > kFftMaxN || g.ny > kFftMaxN || g.nx < 8 ||
        g.ny < 8) {
        set_error(std::string(what) + ": FFT grids must be powers of two between 8 and 4096");
        return CMM_ERR_ARG;
    }
    return CMM_OK;
}

static int launch_check(const char *what) { return hip_check(hipGetLastError(), what); }

static int set_lds(const void *fn, size_t bytes) {
    if (bytes > 65536) return hip_check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes), "hipFuncSetAttribute");
    return CMM_OK;
}

// Spectral pipeline: row R2C (with a sampler) -> column stage -> row C2R of four planes.
struct SpecPlan {
    cmm_grid gin, gout;
    int mode;
    double fcut;
    int kx_max;
    size_t t_bytes, g_bytes;
};

static SpecPlan plan_spec(const cmm_grid &gin, const cmm_grid &gout, int mode, double fcut) {
    SpecPlan s;
    s.gin = gin;
    s.gout = gout;
    s.mode = mode;
    s.fcut = fcut;
    int k = gout.nx / 2;
    if (mode == 1) {
        k = std::min(k, gin.nx / 2);
        if (fcut < 0) k = -1;
        else k = std::min<double>(k, std::floor(fcut));
    }
    s.kx_max = k;
    const int ncol = k + 1 > 0 ? k + 1 : 1;
    s.t_bytes = sizeof(cd) * (size_t)ncol * gin.ny;
    s.g_bytes = sizeof(cd) * (size_t)ncol * gout.ny;
    return s;
}

size_t spec_workspace_bytes(const cmm_grid &gin, const cmm_grid &gout, int mode, double fcut) {
    SpecPlan s = plan_spec(gin, gout, mode, fcut);
    return s.t_bytes + 2 * s.g_bytes + 256;
}

// ev (optional): four events recorded before the row transform and after each of the three kernels
// tw_table (optional): twiddle tables of all power-of-two sizes 8..4096, table n at entry n - 8,
// in place of the cached ones (which are allocated on first use)
template <class S>
static int run_spec(const SpecPlan &sp, const S &sampler, double *out, void *work, hipStream_t st,
                    hipEvent_t *ev = nullptr, bool ev0 = true, const cd *tw_table = nullptr,
                    const PsiPre *pre = nullptr) {
    const cd *tw_in, *tw_out, *tw_x_in, *tw_x_out;
    if (tw_table) {
        tw_x_in = tw_table + (sp.gin.nx - 8);
        tw_in = tw_table + (sp.gin.ny - 8);
        tw_out = tw_table + (sp.gout.ny - 8);
        tw_x_out = tw_table + (sp.gout.nx - 8);
    } else {
        if (int rc = twiddles(sp.gin.nx, &tw_x_in)) return rc;
        if (int rc = twiddles(sp.gin.ny, &tw_in)) return rc;
        if (int rc = twiddles(sp.gout.ny, &tw_out)) return rc;
        if (int rc = twiddles(sp.gout.nx, &tw_x_out)) return rc;
    }
    cd *T = (cd *)work;
    cd *G0 = (cd *)((char *)work + sp.t_bytes);
    cd *G1 = (cd *)((char *)work + sp.t_bytes + sp.g_bytes);
    const int kxm = sp.kx_max;
    const DGrid dout = make_dgrid(sp.gout);
    if (ev && ev0) (void)hipEventRecord(ev[0], st);
    if (kxm >= 0) {
        if (sp.mode != 2) {
            size_t lds = sizeof(cd) * sp.gin.nx;
            if (int rc = set_lds((const void *)k_row_r2c<S>, lds)) return rc;
            hipLaunchKernelGGL(k_row_r2c<S>, dim3(sp.gin.ny / 2), dim3(kFftThreads), lds, st, sampler, sp.gin.nx, sp.gin.ny,
                               ilog2(sp.gin.nx), kxm, tw_x_in, T);
            if (int rc = launch_check("k_row_r2c")) return rc;
        }
        if (ev) (void)hipEventRecord(ev[1], st);
        ColParams p;
        p.nin = sp.gin.ny;
        p.login = ilog2(sp.gin.ny);
        p.nout = sp.gout.ny;
        p.logout = ilog2(sp.gout.ny);
        p.nx_in = sp.gin.nx;
        p.nx_out = sp.gout.nx;
        p.inv_norm = 1.0 / ((double)sp.gin.nx * (double)sp.gin.ny);
        p.fcut = sp.fcut;
        p.fkx = kTwoPi / (dout.hx * sp.gout.nx);
        p.fky = kTwoPi / (dout.hy * sp.gout.ny);
        p.mode = sp.mode;
        p.kx_max = kxm;
        size_t lds = sizeof(cd) * std::max(sp.gin.ny, sp.gout.ny);
        if (int rc = set_lds((const void *)k_col, lds)) return rc;
        hipLaunchKernelGGL(k_col, dim3(kxm + 1), dim3(kFftThreads), lds, st, p, T, tw_in, tw_out, G0, G1);
        if (int rc = launch_check("k_col")) return rc;
    } else if (ev) {
        (void)hipEventRecord(ev[1], st);
    }
    if (ev) (void)hipEventRecord(ev[2], st);
    if (pre) {
        // the new level and the next step's combined fields (pre->form, pre->L)
        size_t lds = sizeof(cd) * sp.gout.nx;
        const void *kern = nullptr;
#define CMM_PRE_K(F, LL) (const void *)k_row_c2r4_pre<F, LL>
        if (pre->form == 0) kern = pre->L == 2 ? CMM_PRE_K(0, 2) : pre->L == 3 ? CMM_PRE_K(0, 3) : CMM_PRE_K(0, 4);
        else kern = pre->L == 2 ? CMM_PRE_K(1, 2) : pre->L == 3 ? CMM_PRE_K(1, 3) : CMM_PRE_K(1, 4);
#undef CMM_PRE_K
        if (int rc = set_lds(kern, lds)) return rc;
#define CMM_PRE_L(F, LL)                                                                                               \
    hipLaunchKernelGGL((k_row_c2r4_pre<F, LL>), dim3(sp.gout.ny), dim3(kFftThreads), lds, st, sp.gout.nx, sp.gout.ny,  \
                       ilog2(sp.gout.nx), kxm, kTwoPi / (dout.hx * sp.gout.nx), G0, G1, tw_x_out, out, *pre)
        if (pre->form == 0) {
            if (pre->L == 2) CMM_PRE_L(0, 2);
            else if (pre->L == 3) CMM_PRE_L(0, 3);
            else CMM_PRE_L(0, 4);
        } else {
            if (pre->L == 2) CMM_PRE_L(1, 2);
            else if (pre->L == 3) CMM_PRE_L(1, 3);
            else CMM_PRE_L(1, 4);
        }
#undef CMM_PRE_L
        if (int rc = launch_check("k_row_c2r4_pre")) return rc;
    } else {
        size_t lds = sizeof(cd) * sp.gout.nx;
        if (int rc = set_lds((const void *)k_row_c2r4, lds)) return rc;
        hipLaunchKernelGGL(k_row_c2r4, dim3(sp.gout.ny), dim3(kFftThreads), lds, st, sp.gout.nx, sp.gout.ny,
                           ilog2(sp.gout.nx), kxm, kTwoPi / (dout.hx * sp.gout.nx), G0, G1, tw_x_out, out);
        if (int rc = launch_check("k_row_c2r4")) return rc;
    }
    if (ev) (void)hipEventRecord(ev[3], st);
    return CMM_OK;
}

static int get_work(void *work, size_t work_bytes, size_t need, void **w) {
    if (work) {
        if (work_bytes < need) {
            set_error("workspace too small");
            return CMM_ERR_ARG;
        }
        *w = work;
        return CMM_OK;
    }
    return workspace(need, w);
}

static ICParams make_ic(int ic, const double *prm) {
    ICParams p;
    p.ic = ic;
    for (int i = 0; i < 10; ++i) p.prm[i] = prm ? prm[i] : 0.0;
    return p;
}

static dim3 grid2d(int nx, int ny, dim3 b) { return dim3((nx + b.x - 1) / b.x, (ny + b.y - 1) / b.y); }

// ------------------------------------------------------------------------------------------------
// Operations shared by the C-ABI and the engine

int op_map_identity(const cmm_grid &g, double *cx, double *cy, hipStream_t st) {
    DGrid d = make_dgrid(g);
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_map_identity, grid2d(g.nx, g.ny, b), b, 0, st, d, cx, cy);
    return launch_check("k_map_identity");
}

// Schemes whose multi-level velocity evaluations use only the L1 / L12 weights can run on
// precombined fields; AB2 (20) and RK3Mod (31) read level 1 on its own.
static bool advect_combinable(const cmm_advect_coeffs &c) {
    const int t = c.time_integration_num;
    return c.l_order >= 2 && (t == 10 || t == 21 || t == 30 || t == 40 || t == 41);
}

static bool advect_needs_c12(const cmm_advect_coeffs &c) {
    const int t = c.time_integration_num;
    return t == 30 || t == 40 || t == 41;
}

static bool advect_needs_c0(const cmm_advect_coeffs &c) { return c.time_integration_num != 10; }

int op_advect_fields(const cmm_advect_coeffs &c) {
    if (!advect_combinable(c)) return 0;
    return 1 | (advect_needs_c12(c) ? 2 : 0) | (advect_needs_c0(c) ? 4 : 0);
}

size_t op_advect_workspace(const cmm_grid &gm, const cmm_grid &gp, const cmm_advect_coeffs &c) {
    if (!advect_combinable(c)) return 0;
    const int fields = 1 + advect_needs_c12(c) + advect_needs_c0(c);
    return fields * sizeof(double) * 4 * (size_t)gp.nx * gp.ny + sizeof(double) * 8 * (size_t)gm.nx * gm.ny;
}

int op_map_advect(const cmm_grid &gm, const cmm_grid &gp, const double *cx, const double *cy, const PsiLevels &P,
                  const cmm_advect_coeffs &c, double *ncx, double *ncy, void *work, size_t work_bytes,
                  hipStream_t st, hipEvent_t mid, const double *m8_in, double *m8_out, const double *corr1,
                  const double *corr12, const double *lw, const PsiPre *pre_in) {
    DGrid dm = make_dgrid(gm), dp = make_dgrid(gp);
    const int nfoot = 4 + 4 * c.map_update_order_num;
    const int G = nfoot <= 4 ? 4 : nfoot <= 8 ? 8 : 16;
    const int nb = (int)(((long)dm.n * G + 255) / 256);
    if (c.direction == 1) {
        if (mid) (void)hipEventRecord(mid, st);
#define CMM_FWD(T, GG) hipLaunchKernelGGL((k_map_advect_fwd<T, GG>), dim3(nb), dim3(256), 0, st, dm, dp, cx, cy, P, c, ncx, ncy, nb)
#define CMM_FWD_G(T)                                                                                                   \
    do {                                                                                                               \
        if (G == 4) CMM_FWD(T, 4);                                                                                     \
        else if (G == 8) CMM_FWD(T, 8);                                                                                \
        else CMM_FWD(T, 16);                                                                                           \
    } while (0)
        switch (c.time_integration_num) {
        case 10: CMM_FWD_G(10); break;
        case 20: CMM_FWD_G(20); break;
        case 21: CMM_FWD_G(21); break;
        case 30: CMM_FWD_G(30); break;
        case 31: CMM_FWD_G(31); break;
        case 40: CMM_FWD_G(40); break;
        case 41: CMM_FWD_G(41); break;
        default: set_error("unknown time integration"); return CMM_ERR_ARG;
        }
#undef CMM_FWD_G
#undef CMM_FWD
        return launch_check("k_map_advect_fwd");
    }
    const size_t need = op_advect_workspace(gm, gp, c);
    const bool comb = work && need > 0 && work_bytes >= need;
    if ((corr1 || corr12) && !comb) {
        set_error("map advection: a psi correction needs a combinable scheme and the combination workspace");
        return CMM_ERR_ARG;
    }
    if (pre_in && (!comb || corr1 || corr12 || lw || pre_in->form != 0)) {
        set_error("map advection: precombined fields need the combination path without corrections");
        return CMM_ERR_ARG;
    }
    const double *M8 = nullptr;
    double *C1 = nullptr, *C12 = nullptr, *C0 = nullptr;
    if (comb) {
        const long n = dp.n, n4 = 4L * n;
        double *next = static_cast<double *>(work);
        double *m8_work = next;
        next += 8L * dm.n;
        C1 = next;
        next += n4;
        if (advect_needs_c12(c)) { C12 = next; next += n4; }
        if (advect_needs_c0(c)) C0 = next;
        if (pre_in) {  // the fields the previous psi solve wrote for this step's weights
            C1 = static_cast<double *>(pre_in->C1);
            C12 = advect_needs_c12(c) ? static_cast<double *>(pre_in->C12) : nullptr;
            C0 = advect_needs_c0(c) ? static_cast<double *>(pre_in->C0) : nullptr;
        }
        if (m8_in) {
            M8 = m8_in;
        } else {
            hipLaunchKernelGGL(k_interleave_map, dim3((unsigned)((dm.n + 255) / 256)), dim3(256), 0, st, (long)dm.n, cx,
                               cy, m8_work);
            if (int rc = launch_check("k_interleave_map")) return rc;
            M8 = m8_work;
        }
        const dim3 g((unsigned)((n + 255) / 256));
        if (!pre_in) {
            switch (c.l_order) {
            case 2: hipLaunchKernelGGL(k_psi_combine<2>, g, dim3(256), 0, st, n, P, c, C1, C12, C0, corr1, corr12, lw); break;
            case 3: hipLaunchKernelGGL(k_psi_combine<3>, g, dim3(256), 0, st, n, P, c, C1, C12, C0, corr1, corr12, lw); break;
            default: hipLaunchKernelGGL(k_psi_combine<4>, g, dim3(256), 0, st, n, P, c, C1, C12, C0, corr1, corr12, lw); break;
            }
            if (int rc = launch_check("k_psi_combine")) return rc;
        }
    }
    if (mid) (void)hipEventRecord(mid, st);
    const bool p2 = dm.lgx >= 0 && dm.lgy >= 0 && dp.lgx >= 0 && dp.lgy >= 0;
    const int nb2 = (int)(((long)dm.n * 4 + 255) / 256);  // two footpoints per lane
#define CMM_ADV(T, GG, CB)                                                                                             \
    hipLaunchKernelGGL((k_map_advect<T, GG, CB>), dim3(nb), dim3(256), 0, st, dm, dp, cx, cy, P, C1, C12, C0, M8, c, ncx, \
                       ncy, m8_out, nb, lw)
#define CMM_ADV_G(T, CB)                                                                                               \
    do {                                                                                                               \
        if (G == 4) CMM_ADV(T, 4, CB);                                                                                 \
        else if (G == 8) CMM_ADV(T, 8, CB);                                                                            \
        else CMM_ADV(T, 16, CB);                                                                                       \
    } while (0)
#define CMM_ADV_T(T)                                                                                                   \
    do {                                                                                                               \
        if (comb && p2 && G == 8)                                                                                      \
            hipLaunchKernelGGL((k_map_advect_pair<T, true>), dim3(nb2), dim3(256), 0, st, dm, dp, C1, C12, C0, M8, c,  \
                               ncx, ncy, m8_out, nb2);                                                                \
        else if (comb) CMM_ADV_G(T, true);                                                                             \
        else CMM_ADV_G(T, false);                                                                                      \
    } while (0)
    switch (c.time_integration_num) {
    case 10: CMM_ADV_T(10); break;
    case 20: CMM_ADV_G(20, false); break;
    case 21: CMM_ADV_T(21); break;
    case 30: CMM_ADV_T(30); break;
    case 31: CMM_ADV_G(31, false); break;
    case 40: CMM_ADV_T(40); break;
    case 41: CMM_ADV_T(41); break;
    default: set_error("unknown time integration"); return CMM_ERR_ARG;
    }
#undef CMM_ADV_T
#undef CMM_ADV_G
#undef CMM_ADV
    return launch_check("k_map_advect");
}

// The Vlasov-Poisson advection with the separable velocity: the 8-footpoint stencil on power-of-two
// grids, a combinable scheme, levels built by k_vp_psi_planes; H: the bump function's H and H'
// (op_vp_bump).  Workspace: the 1D tables and, without m8_in, the interleaved map.
bool op_vp_advect_supported(const cmm_grid &gm, const cmm_grid &gp, const cmm_advect_coeffs &c) {
    const DGrid dm = make_dgrid(gm), dp = make_dgrid(gp);
    return c.direction != 1 && advect_combinable(c) && 4 + 4 * c.map_update_order_num == 8 && dm.lgx >= 0 &&
           dm.lgy >= 0 && dp.lgx >= 0 && dp.lgy >= 0 && c.l_order >= 2 && c.l_order <= 4;
}

int op_map_advect_vp(const cmm_grid &gm, const cmm_grid &gp, const PsiLevels &P, const cmm_advect_coeffs &c,
                     const double *cx, const double *cy, const double *m8_in, double *m8_out, const double *H,
                     double *ncx, double *ncy, void *work, size_t work_bytes, hipStream_t st, hipEvent_t mid,
                     const double *lw) {
    if (!op_vp_advect_supported(gm, gp, c)) {
        set_error("separable Vlasov-Poisson advection: needs the 8-footpoint stencil, a combinable scheme and "
                  "power-of-two grids");
        return CMM_ERR_ARG;
    }
    const DGrid dm = make_dgrid(gm), dp = make_dgrid(gp);
    const size_t need = sizeof(double) * (6 * (size_t)gp.nx + 2 + (m8_in ? 0 : 8 * (size_t)dm.n));
    void *w;
    if (int rc = get_work(work, work_bytes, need, &w)) return rc;
    double *T = static_cast<double *>(w), *S = T + 6 * gp.nx;
    const double *M8 = m8_in;
    if (!M8) {
        double *m8w = S + 2;
        hipLaunchKernelGGL(k_interleave_map, dim3((unsigned)((dm.n + 255) / 256)), dim3(256), 0, st, (long)dm.n, cx, cy,
                           m8w);
        if (int rc = launch_check("k_interleave_map")) return rc;
        M8 = m8w;
    }
    const dim3 g1((unsigned)((gp.nx + 255) / 256));
    switch (c.l_order) {
    case 2: hipLaunchKernelGGL(k_vp_combine1d<2>, g1, dim3(256), 0, st, gp.nx, (long)dp.n, P, c, lw, T, S); break;
    case 3: hipLaunchKernelGGL(k_vp_combine1d<3>, g1, dim3(256), 0, st, gp.nx, (long)dp.n, P, c, lw, T, S); break;
    default: hipLaunchKernelGGL(k_vp_combine1d<4>, g1, dim3(256), 0, st, gp.nx, (long)dp.n, P, c, lw, T, S); break;
    }
    if (int rc = launch_check("k_vp_combine1d")) return rc;
    if (mid) (void)hipEventRecord(mid, st);
    const int nb = (int)(((long)dm.n * 4 + 255) / 256);
    switch (c.time_integration_num) {
    case 10: hipLaunchKernelGGL(k_map_advect_vp<10>, dim3(nb), dim3(256), 0, st, dm, dp, T, S, H, M8, c, ncx, ncy, m8_out, nb); break;
    case 21: hipLaunchKernelGGL(k_map_advect_vp<21>, dim3(nb), dim3(256), 0, st, dm, dp, T, S, H, M8, c, ncx, ncy, m8_out, nb); break;
    case 30: hipLaunchKernelGGL(k_map_advect_vp<30>, dim3(nb), dim3(256), 0, st, dm, dp, T, S, H, M8, c, ncx, ncy, m8_out, nb); break;
    case 40: hipLaunchKernelGGL(k_map_advect_vp<40>, dim3(nb), dim3(256), 0, st, dm, dp, T, S, H, M8, c, ncx, ncy, m8_out, nb); break;
    case 41: hipLaunchKernelGGL(k_map_advect_vp<41>, dim3(nb), dim3(256), 0, st, dm, dp, T, S, H, M8, c, ncx, ncy, m8_out, nb); break;
    default: set_error("unknown time integration"); return CMM_ERR_ARG;
    }
    return launch_check("k_map_advect_vp");
}

// tile edge of the tiled check: the tuned one when it divides r, else the largest of 4, 3, 2, 1
// that divides r
static int incomp_tile(int r, const IncompTune &tune) {
    if (tune.tile > 0 && r % tune.tile == 0 && (tune.tile == 1 || tune.tile == 2 || tune.tile == 3 || tune.tile == 4))
        return tune.tile;
    return r % 4 == 0 ? 4 : r % 3 == 0 ? 3 : r % 2 == 0 ? 2 : 1;
}

int op_interleave_map(const cmm_grid &g, const double *cx, const double *cy, double *M8, hipStream_t st) {
    const long n = (long)g.nx * g.ny;
    hipLaunchKernelGGL(k_interleave_map, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, cx, cy, M8);
    return launch_check("k_interleave_map");
}

int op_incomp_launch(const cmm_grid &gm, const cmm_grid &gc, const double *cx, const double *cy,
                     unsigned long long *slot, hipStream_t st, IncompTune tune) {
    DGrid dm = make_dgrid(gm), dc = make_dgrid(gc);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, slot);
    dim3 b(64, 4);
    const int r = gc.nx / gm.nx;
    if (r >= 1 && gc.nx == r * gm.nx && gc.ny == r * gm.ny && gc.x0 == gm.x0 && gc.x1 == gm.x1 &&
        gc.y0 == gm.y0 && gc.y1 == gm.y1) {
        const int T = incomp_tile(r, tune);
        const int ntx = gc.nx / T;
        const long ntiles = (long)ntx * (gc.ny / T);
        const unsigned nb = (unsigned)std::max<long>(1, std::min<long>((ntiles + 255) / 256, tune.max_blocks));
        if (T == 4) hipLaunchKernelGGL((k_incomp_tensor<4, 4>), dim3(nb), dim3(256), 0, st, dm, dc, cx, cy, slot, r, ntx, ntiles);
        else if (T == 3) hipLaunchKernelGGL((k_incomp_tensor<3, 3>), dim3(nb), dim3(256), 0, st, dm, dc, cx, cy, slot, r, ntx, ntiles);
        else if (T == 2) hipLaunchKernelGGL((k_incomp_tensor<2, 2>), dim3(nb), dim3(256), 0, st, dm, dc, cx, cy, slot, r, ntx, ntiles);
        else hipLaunchKernelGGL((k_incomp_tensor<1, 1>), dim3(nb), dim3(256), 0, st, dm, dc, cx, cy, slot, r, ntx, ntiles);
        return launch_check("k_incomp_tensor");
    }
    const int tx = (gc.nx + 63) / 64, ty = (gc.ny + 3) / 4, tt = tx * ty;
    const int nb = std::min(tt, 2048);
    hipLaunchKernelGGL(k_incomp, dim3(nb), b, 0, st, dm, dc, cx, cy, slot, tx, tt);
    return launch_check("k_incomp");
}

int op_invert_launch(const cmm_grid &gb, const cmm_grid &gf, const cmm_grid &gc, const double *bx, const double *by,
                     const double *fx, const double *fy, unsigned long long *slot, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, slot);
    const int tx = (gc.nx + 63) / 64, ty = (gc.ny + 3) / 4, tt = tx * ty;
    hipLaunchKernelGGL(k_invert, dim3(std::min(tt, 2048)), dim3(64, 4), 0, st, make_dgrid(gb), make_dgrid(gf),
                       make_dgrid(gc), bx, by, fx, fy, slot, tx, tt);
    return launch_check("k_invert");
}

int op_incomp(const cmm_grid &gm, const cmm_grid &gc, const double *cx, const double *cy, double *out,
              hipStream_t st) {
    unsigned long long *slot;
    if (int rc = scalar_slot(&slot)) return rc;
    if (int rc = op_incomp_launch(gm, gc, cx, cy, slot, st)) return rc;
    unsigned long long h = 0;
    if (int rc = hip_check(hipMemcpyAsync(&h, slot, sizeof(h), hipMemcpyDeviceToHost, st), "hipMemcpyAsync")) return rc;
    if (int rc = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize")) return rc;
    double v;
    std::memcpy(&v, &h, sizeof(v));
    *out = v;
    return CMM_OK;
}

int op_invert(const cmm_grid &gb, const cmm_grid &gf, const cmm_grid &gc, const double *bx, const double *by,
              const double *fx, const double *fy, double *out, hipStream_t st) {
    unsigned long long *slot;
    if (int rc = scalar_slot(&slot)) return rc;
    if (int rc = op_invert_launch(gb, gf, gc, bx, by, fx, fy, slot, st)) return rc;
    unsigned long long h = 0;
    if (int rc = hip_check(hipMemcpyAsync(&h, slot, sizeof(h), hipMemcpyDeviceToHost, st), "hipMemcpyAsync")) return rc;
    if (int rc = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize")) return rc;
    std::memcpy(out, &h, sizeof(double));
    return CMM_OK;
}

// spectral buffers and the sampled vorticity field
size_t op_psi_workspace(const cmm_grid &gv, const cmm_grid &gp, double fcut) {
    return spec_workspace_bytes(gv, gp, 1, fcut) + sizeof(double) * (size_t)gv.nx * gv.ny;
}

// whether the incompressibility check on grid gc can run inside the psi sampler on grid gv: the same
// grid, refining the map grid gm by a multiple of the sampler's rows per lane over the same domain
bool op_check_fusable(const cmm_grid &gm, const cmm_grid &gv, const cmm_grid &gc) {
    const int r = gv.ny / gm.ny;
    return gc.nx == gv.nx && gc.ny == gv.ny && gc.x0 == gv.x0 && gc.x1 == gv.x1 && gc.y0 == gv.y0 &&
           gc.y1 == gv.y1 && gv.x0 == gm.x0 && gv.x1 == gm.x1 && gv.y0 == gm.y0 && gv.y1 == gm.y1 &&
           gv.nx % gm.nx == 0 && gv.ny == r * gm.ny && r % kSampleRows == 0;
}

// The psi solve in two parts: op_psi_sample writes omega(X(x)) on the vorticity grid into the
// workspace (with chk_slot: the incompressibility check of the same map on the same grid in the
// same pass, op_check_fusable, its maximum as k_incomp's bit pattern in *chk_slot);
// op_psi_spectral transforms that field into psi.  Work on the same stream in between must leave
// the field alone.
int op_psi_sample(const cmm_grid &gm, const cmm_grid &gv, const cmm_grid &gf, const cmm_grid &gp, const double *cx,
                  const double *cy, const double *W, int molly, double fcut, void *work, size_t work_bytes,
                  hipStream_t st, hipEvent_t *ev, const double *W4, unsigned long long *chk_slot) {
    if (int rc = fft_grid_ok(gv, "vort grid")) return rc;
    if (int rc = fft_grid_ok(gp, "psi grid")) return rc;
    SpecPlan sp = plan_spec(gv, gp, 1, fcut);
    const size_t base = sp.t_bytes + 2 * sp.g_bytes + 256;
    void *w;
    if (int rc = get_work(work, work_bytes, base + sizeof(double) * (size_t)gv.nx * gv.ny, &w)) return rc;
    if (chk_slot && !op_check_fusable(gm, gv, gv)) {
        set_error("psi sampler: the check cannot run in the sampler's pass on these grids");
        return CMM_ERR_ARG;
    }
    // omega(X(x)) on the vorticity grid into a real field after the spectral buffers, for the
    // plain row transform; the sampling counts to the row transform's profile slot
    double *f = reinterpret_cast<double *>(static_cast<char *>(w) + base);
    const long lanes = (long)gv.nx * (gv.ny / kSampleRows);
    const dim3 grid((unsigned)((lanes + 255) / 256));
    auto launch = [&](auto smp, const double *field) {
        smp.gm = make_dgrid(gm);
        smp.gv = make_dgrid(gv);
        smp.gf = make_dgrid(gf);
        smp.cx = cx;
        smp.cy = cy;
        smp.W = field;
        smp.molly = molly;
        if (chk_slot) {
            hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, chk_slot);
            hipLaunchKernelGGL((k_sample_field<decltype(smp), double, kSampleRows, true>), grid, dim3(256), 0, st, smp,
                               gv.nx, gv.ny, f, chk_slot);
        } else {
            hipLaunchKernelGGL(k_sample_field<decltype(smp)>, grid, dim3(256), 0, st, smp, gv.nx, gv.ny, f, nullptr);
        }
    };
    if (ev) (void)hipEventRecord(ev[0], st);
    if (W4) launch(SampleVortIL{}, W4);
    else launch(SampleVort{}, W);
    return launch_check("k_sample_field");
}

int op_psi_spectral(const cmm_grid &gv, const cmm_grid &gp, double fcut, double *psi, void *work, size_t work_bytes,
                    hipStream_t st, hipEvent_t *ev, const PsiPre *pre) {
    SpecPlan sp = plan_spec(gv, gp, 1, fcut);
    const size_t base = sp.t_bytes + 2 * sp.g_bytes + 256;
    void *w;
    if (int rc = get_work(work, work_bytes, base + sizeof(double) * (size_t)gv.nx * gv.ny, &w)) return rc;
    double *f = reinterpret_cast<double *>(static_cast<char *>(w) + base);
    return run_spec(sp, SampleReal{f, gv.nx}, psi, w, st, ev, false, nullptr, pre);
}

int op_psi_from_vort(const cmm_grid &gm, const cmm_grid &gv, const cmm_grid &gf, const cmm_grid &gp, const double *cx,
                     const double *cy, const double *W, int molly, double fcut, double *psi, void *work,
                     size_t work_bytes, hipStream_t st, hipEvent_t *ev, const double *W4, const PsiPre *pre) {
    if (int rc = op_psi_sample(gm, gv, gf, gp, cx, cy, W, molly, fcut, work, work_bytes, st, ev, W4, nullptr))
        return rc;
    return op_psi_spectral(gv, gp, fcut, psi, work, work_bytes, st, ev, pre);
}

// ------------------------------------------------------------------------------------------------
// Allocation-free psi solve (op_psi_from_vort_noalloc).  `scratch` is a file-scope device array of
// kScratchBytes that every GPU the library runs on holds until the process ends: the twiddle tables
// of every power-of-two size 8..4096 (table n at entry n - 8, exp(-2 pi i m/n), m < n, from long
// double as twiddles() makes them), then the psi solve's workspace.  Its device address on a GPU
// comes from hipGetSymbolAddress, taken once per device with that device current, when the first
// call on it sets the tables up; the array's name is never passed to a kernel.  The workspace needs
// no clearing: every kernel writes what the next one reads.  Calls that overlap on one GPU race.
constexpr size_t kScratchTwBytes = 131072;  // 8184 complex doubles, padded to 256-byte blocks
constexpr size_t kScratchBytes = (size_t)192 << 20;
static_assert(kScratchBytes - kScratchTwBytes == CMM_PSI_NOALLOC_MAX_WORKSPACE, "scratch limit of include/cmm.h");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kScratchDevices = 64;
static char *g_scratch_ptr[kScratchDevices];  // device address of scratch per device (under g_mutex)

static int scratch_setup(char **out) {
    int dev;
    if (int rc = current_device(dev)) return rc;
    if (dev < 0 || dev >= kScratchDevices) {
        set_error("allocation-free psi solve: device ordinal beyond 63");
        return CMM_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        if (int rc = hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress scratch")) return rc;
        static cd h_tw[8184];  // host copy of the tables (static: no allocation on this path)
        for (int n = 8; n <= kFftMaxN; n *= 2)
            for (int m = 0; m < n; ++m) {
                long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
                h_tw[n - 8 + m] = cd{(double)cosl(a), (double)sinl(a)};
            }
        if (int rc = hip_check(hipMemcpy(ptr, h_tw, sizeof(h_tw), hipMemcpyHostToDevice), "hipMemcpy scratch twiddles"))
            return rc;
        g_scratch_ptr[dev] = static_cast<char *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return CMM_OK;
}

size_t op_psi_noalloc_max_bytes() { return kScratchBytes - kScratchTwBytes; }

int op_psi_from_vort_noalloc(const cmm_grid &gm, const cmm_grid &gv, const cmm_grid &gf, const cmm_grid &gp,
                             const double *cx, const double *cy, const double *W, int molly, double fcut, double *psi,
                             hipStream_t st) {
    if (int rc = fft_grid_ok(gv, "vort grid")) return rc;
    if (int rc = fft_grid_ok(gp, "psi grid")) return rc;
    SpecPlan sp = plan_spec(gv, gp, 1, fcut);
    const size_t base = sp.t_bytes + 2 * sp.g_bytes + 256;
    const size_t need = base + sizeof(double) * (size_t)gv.nx * gv.ny;
    if (need > op_psi_noalloc_max_bytes()) {
        set_error("allocation-free psi solve: the workspace of these grids and cut-off (" + std::to_string(need) +
                  " bytes) exceeds the static scratch array (" + std::to_string(op_psi_noalloc_max_bytes()) + " bytes)");
        return CMM_ERR_ARG;
    }
    char *base_ptr = nullptr;
    if (int rc = scratch_setup(&base_ptr)) return rc;
    const cd *tw = reinterpret_cast<const cd *>(base_ptr);
    void *w = base_ptr + kScratchTwBytes;
    double *f = reinterpret_cast<double *>(static_cast<char *>(w) + base);
    const long lanes = (long)gv.nx * (gv.ny / kSampleRows);
    SampleVort smp;
    smp.gm = make_dgrid(gm);
    smp.gv = make_dgrid(gv);
    smp.gf = make_dgrid(gf);
    smp.cx = cx;
    smp.cy = cy;
    smp.W = W;
    smp.molly = molly;
    hipLaunchKernelGGL(k_sample_field<SampleVort>, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, st, smp, gv.nx,
                       gv.ny, f, nullptr);
    if (int rc = launch_check("k_sample_field")) return rc;
    return run_spec(sp, SampleReal{f, gv.nx}, psi, w, st, nullptr, false, tw);
}

// node-interleaved copy of a Hermite field: H4[4j + p] = H[p*n + j]
__global__ void __launch_bounds__(256) k_interleave4(long n, const double *__restrict__ H, double *__restrict__ H4) {
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    double2 *d = reinterpret_cast<double2 *>(H4 + 4 * j);
    d[0] = double2{H[j], H[n + j]};
    d[1] = double2{H[2 * n + j], H[3 * n + j]};
}

int op_interleave4(const cmm_grid &g, const double *H, double *H4, hipStream_t st) {
    const long n = (long)g.nx * g.ny;
    hipLaunchKernelGGL(k_interleave4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, H, H4);
    return launch_check("k_interleave4");
}

size_t op_hermite_workspace(const cmm_grid &g) { return spec_workspace_bytes(g, g, 0, 0.0); }

// Hermite planes of F (op 0), of the inverse Laplacian -F/|k|^2 with DC 0 (op 1: i_laplacian,
// Cuda-CMM/src/numerical/cmm-fft.cu:441-447) or of the Laplacian -|k|^2 F (op 2: laplacian,
// :432-438) of a real field; planes 1 and 2 are grad_x and grad_y of the result (:450-474).
int op_spectral_from_real(const cmm_grid &g, const double *real, int op, double *herm, void *work, size_t work_bytes,
                          hipStream_t st) {
    if (int rc = fft_grid_ok(g, "grid")) return rc;
    if (op < 0 || op > 2) {
        set_error("spectral operator must be 0 (Hermite), 1 (inverse Laplacian) or 2 (Laplacian)");
        return CMM_ERR_ARG;
    }
    const int mode = op == 0 ? 0 : op == 1 ? 1 : 3;
    SpecPlan sp = plan_spec(g, g, mode, op == 1 ? 1e300 : 0.0);
    void *w;
    if (int rc = get_work(work, work_bytes, sp.t_bytes + 2 * sp.g_bytes + 256, &w)) return rc;
    SampleReal s{real, g.nx};
    return run_spec(sp, s, herm, w, st);
}

int op_hermite_from_real(const cmm_grid &g, const double *real, double *herm, void *work, size_t work_bytes,
                         hipStream_t st) {
    if (int rc = fft_grid_ok(g, "grid")) return rc;
    SpecPlan sp = plan_spec(g, g, 0, 0.0);
    void *w;
    if (int rc = get_work(work, work_bytes, sp.t_bytes + 2 * sp.g_bytes + 256, &w)) return rc;
    SampleReal s{real, g.nx};
    return run_spec(sp, s, herm, w, st);
}

int op_hermite_from_spectrum(const cmm_grid &g, const double *spec_cols, double *herm, void *work,
                             size_t work_bytes, hipStream_t st) {
    if (int rc = fft_grid_ok(g, "grid")) return rc;
    SpecPlan sp = plan_spec(g, g, 2, 0.0);
    void *w;
    if (int rc = get_work(work, work_bytes, sp.t_bytes + 2 * sp.g_bytes + 256, &w)) return rc;
    if (int rc = hip_check(hipMemcpyAsync(w, spec_cols, sp.t_bytes, hipMemcpyDeviceToDevice, st), "hipMemcpyAsync")) return rc;
    SampleReal s{nullptr, g.nx};
    return run_spec(sp, s, herm, w, st);
}

// the sampler launch of op_translate_ic and op_sample_init
static int launch_sample_init(const SampleCompose &s, const cmm_grid &gt, double *out, hipStream_t st) {
    dim3 b(64, 4);
    if (gt.ny % 2 == 0) {
        const dim3 g((gt.nx + b.x - 1) / b.x, (gt.ny / 2 + b.y - 1) / b.y);
        if (s.fmt == kStackRec48) hipLaunchKernelGGL(k_sample_init_multi<kStackRec48>, g, b, 0, st, s, out);
        else if (s.fmt == kStackRec64) hipLaunchKernelGGL(k_sample_init_multi<kStackRec64>, g, b, 0, st, s, out);
        else hipLaunchKernelGGL(k_sample_init_multi<kStackFp64>, g, b, 0, st, s, out);
        return launch_check("k_sample_init_multi");
    }
    hipLaunchKernelGGL(k_sample_grid<SampleCompose>, dim3((gt.nx + b.x - 1) / b.x, (gt.ny + b.y - 1) / b.y), b, 0, st,
                       s, gt.nx, gt.ny, out);
    return launch_check("k_sample_grid");
}

int op_translate_ic(const cmm_grid &gm, const cmm_grid &gf, const double *cx, const double *cy, const double *stack,
                    int nstack, int ic, const double *prm, int discrete, const cmm_grid &gd, const double *disc,
                    double *herm, void *work, size_t work_bytes, hipStream_t st, int stack_fmt) {
    if (int rc = fft_grid_ok(gf, "fine grid")) return rc;
    SpecPlan sp = plan_spec(gf, gf, 0, 0.0);
    const size_t base = sp.t_bytes + 2 * sp.g_bytes + 256;
    void *w;
    if (int rc = get_work(work, work_bytes, op_translate_workspace(gf), &w)) return rc;
    SampleCompose s;
    s.gm = make_dgrid(gm);
    s.gt = make_dgrid(gf);
    s.gd = make_dgrid(gd);
    s.cx = cx;
    s.cy = cy;
    s.stack = stack;
    s.nstack = nstack;
    s.fmt = stack_fmt;
    s.ic = make_ic(ic, prm);
    s.discrete = discrete;
    s.disc = disc;
    // the initial condition at the composed positions into a real field after the spectral
    // buffers, one point per lane at full occupancy, then the plain row transform of that field:
    // the composition's dependent gathers need the waves that the row FFT's LDS leaves too few of
    double *f = reinterpret_cast<double *>(static_cast<char *>(w) + base);
    if (int rc = launch_sample_init(s, gf, f, st)) return rc;
    return run_spec(sp, SampleReal{f, gf.nx}, herm, w, st);
}

size_t op_translate_workspace(const cmm_grid &g) {
    return spec_workspace_bytes(g, g, 0, 0.0) + sizeof(double) * (size_t)g.nx * g.ny;
}

int op_pack_maps(const cmm_grid &gm, const double *cx, const double *cy, long stride, int n_maps, int fmt, void *out,
                 hipStream_t st) {
    if (n_maps <= 0) return CMM_OK;
    if (fmt != kStackRec64 && fmt != kStackRec48) {
        set_error("stack records: unknown format");
        return CMM_ERR_ARG;
    }
    const long N = (long)gm.nx * gm.ny;
    hipLaunchKernelGGL(k_pack_maps, dim3((unsigned)((N * n_maps + 255) / 256)), dim3(256), 0, st, cx, cy, stride,
                       static_cast<char *>(out), N, n_maps, fmt);
    return launch_check("k_pack_maps");
}

int op_compose_positions(const cmm_grid &gm, const cmm_grid &gt, const double *cx, const double *cy,
                         const double *stack, int nstack, double *xy, hipStream_t st, int stack_fmt) {
    SampleCompose s;
    s.gm = make_dgrid(gm);
    s.gt = make_dgrid(gt);
    s.gd = s.gt;
    s.cx = cx;
    s.cy = cy;
    s.stack = stack;
    s.nstack = nstack;
    s.fmt = stack_fmt;
    s.ic = make_ic(0, nullptr);
    s.discrete = 0;
    s.disc = nullptr;
    dim3 b(64, 4);
    // two points per lane (33 stacked 1024^2 maps composed onto 4096^2: FP64 records 4.4 -> 3.9 ms,
    // 48-byte records 3.7 -> 3.4 ms)
    if (gt.ny % 2 == 0) {
        const dim3 g = grid2d(gt.nx, gt.ny / 2, b);
        if (stack_fmt == kStackRec48) hipLaunchKernelGGL((k_compose_multi<kStackRec48, 2>), g, b, 0, st, s, xy);
        else if (stack_fmt == kStackRec64) hipLaunchKernelGGL((k_compose_multi<kStackRec64, 2>), g, b, 0, st, s, xy);
        else hipLaunchKernelGGL((k_compose_multi<kStackFp64, 2>), g, b, 0, st, s, xy);
        return launch_check("k_compose_multi");
    }
    if (stack_fmt == kStackRec48) hipLaunchKernelGGL(k_compose_positions<kStackRec48>, grid2d(gt.nx, gt.ny, b), b, 0, st, s, xy);
    else if (stack_fmt == kStackRec64) hipLaunchKernelGGL(k_compose_positions<kStackRec64>, grid2d(gt.nx, gt.ny, b), b, 0, st, s, xy);
    else hipLaunchKernelGGL(k_compose_positions<kStackFp64>, grid2d(gt.nx, gt.ny, b), b, 0, st, s, xy);
    return launch_check("k_compose_positions");
}

int op_forward_positions(const cmm_grid &gm, const cmm_grid &gt, const double *cx, const double *cy,
                         const double *stack, int nstack, double *xy, hipStream_t st) {
    SampleCompose s;
    s.gm = make_dgrid(gm);
    s.gt = make_dgrid(gt);
    s.gd = s.gt;
    s.cx = cx;
    s.cy = cy;
    s.stack = stack;
    s.nstack = nstack;
    s.fmt = kStackFp64;
    s.ic = make_ic(0, nullptr);
    s.discrete = 0;
    s.disc = nullptr;
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_forward_positions, grid2d(gt.nx, gt.ny, b), b, 0, st, s, xy);
    return launch_check("k_forward_positions");
}

int op_forward_points(const cmm_grid &gm, const cmm_grid &gb, const double *stack, int first, int nstack,
                      const double *cx, const double *cy, int cur, const double *in, double *out, int n, hipStream_t st) {
    if (n <= 0) return CMM_OK;
    PointBounds b{gb.x0, gb.x1, gb.y0, gb.y1};
    const int f = std::min(std::max(first, 0), nstack);  // maps f..nstack-1, then the current one
    hipLaunchKernelGGL(k_forward_points, dim3((n + 255) / 256), dim3(256), 0, st, make_dgrid(gm), b, stack, f, nstack,
                       cx, cy, cur ? 1 : 0, in, out, n);
    return launch_check("k_forward_points");
}

int op_sample_init(const cmm_grid &gm, const cmm_grid &gt, const double *cx, const double *cy, const double *stack,
                   int nstack, int ic, const double *prm, int discrete, const cmm_grid &gd, const double *disc,
                   double *out, hipStream_t st, int stack_fmt) {
    SampleCompose s;
    s.gm = make_dgrid(gm);
    s.gt = make_dgrid(gt);
    s.gd = make_dgrid(gd);
    s.cx = cx;
    s.cy = cy;
    s.stack = stack;
    s.nstack = nstack;
    s.fmt = stack_fmt;
    s.ic = make_ic(ic, prm);
    s.discrete = discrete;
    s.disc = disc;
    return launch_sample_init(s, gt, out, st);
}

static constexpr int kVpRows = 64;

static int vp_check(const cmm_grid &gv, const cmm_grid &gp) {
    if (int rc = fft_grid_ok(gv, "vort grid")) return rc;
    if (int rc = fft_grid_ok(gp, "psi grid")) return rc;
    if (gv.nx > gp.nx || gv.ny > gp.ny) {
        // the reference stops with error 1013 (Cuda-CMM/src/simulation/cmm-simulation-host.cu:462-465)
        set_error("vorticity grid should be smaller than phi_out grid (vort grid finer than psi grid)");
        return CMM_ERR_ARG;
    }
    if (gp.nx != gp.ny) {
        // reference error 20230904 (Cuda-CMM/src/simulation/cmm-simulation-host.cu:555)
        set_error("velocity periodification assumes NX==NY on the psi grid");
        return CMM_ERR_ARG;
    }
    return CMM_OK;
}

size_t op_vp_psi_workspace(const cmm_grid &gv, const cmm_grid &gp) {
    const size_t nchunks = (gv.ny + kVpRows - 1) / kVpRows;
    return sizeof(double) * ((size_t)gv.nx * gv.ny + nchunks * gv.nx + 4 * (size_t)gp.nx);
}

static VP1D vp_params(const cmm_grid &gv, const cmm_grid &gp) {
    const DGrid dv = make_dgrid(gv), dp = make_dgrid(gp);
    VP1D q;
    q.nv = gv.nx;
    q.lognv = ilog2(gv.nx);
    q.np = gp.nx;
    q.lognp = ilog2(gp.nx);
    q.nchunks = (gv.ny + kVpRows - 1) / kVpRows;
    q.hv = dv.hy;
    q.fkx_v = kTwoPi / (dv.hx * gv.nx);
    q.fky_p = kTwoPi / (dp.hy * gp.ny);
    q.fkx_p = kTwoPi / (dp.hx * gp.nx);
    q.v0 = gp.y0;
    q.hv_p = dp.hy;
    q.Lv = gp.y1 - gp.y0;
    return q;
}

int op_vp_bump(const cmm_grid &gv, const cmm_grid &gp, double *H, hipStream_t st) {
    if (int rc = vp_check(gv, gp)) return rc;
    const cd *tw_p;
    if (int rc = twiddles(gp.nx, &tw_p)) return rc;
    const VP1D q = vp_params(gv, gp);
    const size_t lds = sizeof(cd) * gp.nx;
    if (int rc = set_lds((const void *)k_vp_bump, lds)) return rc;
    hipLaunchKernelGGL(k_vp_bump, dim3(1), dim3(kFftThreads), lds, st, q, tw_p, H);
    return launch_check("k_vp_bump");
}

int op_vp_psi_from_f(const cmm_grid &gm, const cmm_grid &gv, const cmm_grid &gf, const cmm_grid &gp, const double *cx,
                     const double *cy, const double *W, int molly, double *psi, void *work, size_t work_bytes,
                     hipStream_t st, const double *W4, const double *H) {
    if (int rc = vp_check(gv, gp)) return rc;
    void *w;
    if (int rc = get_work(work, work_bytes, op_vp_psi_workspace(gv, gp), &w)) return rc;
    const VP1D q = vp_params(gv, gp);
    double *f = static_cast<double *>(w);
    double *part = f + (size_t)gv.nx * gv.ny;
    double *phiH = part + (size_t)q.nchunks * gv.nx;  // phi, H, phi', H' (H, H' when not given)
    const int r = gv.ny / gm.ny;
    if (W4 && molly == 0 && gv.ny == r * gm.ny && r >= 2 && gv.x0 == gm.x0 && gv.y0 == gm.y0) {
        // omega(X(x)) by the column sampler on the node-interleaved field: 4 rows per lane, or 2
        // where the vorticity grid refines the map grid twice, so a lane's rows share a map cell
        SampleVortIL smp;
        smp.gm = make_dgrid(gm);
        smp.gv = make_dgrid(gv);
        smp.gf = make_dgrid(gf);
        smp.cx = cx;
        smp.cy = cy;
        smp.W = W4;
        smp.molly = 0;
        if (r >= kSampleRows) {
            const long lanes = (long)gv.nx * (gv.ny / kSampleRows);
            hipLaunchKernelGGL((k_sample_field<SampleVortIL, double, kSampleRows>), dim3((unsigned)((lanes + 255) / 256)),
                               dim3(256), 0, st, smp, gv.nx, gv.ny, f, nullptr);
        } else {
            const long lanes = (long)gv.nx * (gv.ny / 2);
            hipLaunchKernelGGL((k_sample_field<SampleVortIL, double, 2>), dim3((unsigned)((lanes + 255) / 256)),
                               dim3(256), 0, st, smp, gv.nx, gv.ny, f, nullptr);
        }
        if (int rc = launch_check("k_sample_field")) return rc;
    } else {
        if (int rc = op_sample_vort(gm, gv, gf, cx, cy, W, molly, f, st)) return rc;
    }
    hipLaunchKernelGGL(k_vp_colsum, dim3((gv.nx + 255) / 256, q.nchunks), dim3(256), 0, st, f, gv.nx, gv.ny, kVpRows,
                       part);
    if (int rc = launch_check("k_vp_colsum")) return rc;
    const cd *tw_v, *tw_p;
    if (int rc = twiddles(gv.nx, &tw_v)) return rc;
    if (int rc = twiddles(gp.nx, &tw_p)) return rc;
    if (!H) {  // the bump function into the workspace: H at phiH + np, H' at phiH + 3 np
        const size_t lds = sizeof(cd) * gp.nx;
        if (int rc = set_lds((const void *)k_vp_bump, lds)) return rc;
        hipLaunchKernelGGL(k_vp_bump, dim3(1), dim3(kFftThreads), lds, st, q, tw_p, phiH + gp.nx);
        if (int rc = launch_check("k_vp_bump")) return rc;
    }
    VPJob job{q, f, part, tw_v, tw_p, phiH};
    const size_t lds = sizeof(cd) * std::max(gv.nx, gp.nx);
    if (int rc = set_lds((const void *)k_vp_phi, lds)) return rc;
    hipLaunchKernelGGL(k_vp_phi, dim3(1), dim3(kFftThreads), lds, st, job);
    if (int rc = launch_check("k_vp_phi")) return rc;
    return op_vp_psi_planes(gp, phiH, H ? H : phiH + gp.nx, psi, st);
}

size_t op_vp_efield_workspace(const cmm_grid &g) {
    const size_t nchunks = (g.ny + kVpRows - 1) / kVpRows;
    return sizeof(double) * (nchunks * g.nx + 5 * (size_t)g.nx);
}

// phi(x) and phi'(x) of a distribution f given on the grid g itself (the sample outputs): the
// column integral I(x) = hv sum_v f - 1 from a zeroed accumulator, the 1D Poisson solve and the
// spectral derivative on g (vp_phi_block with an all-zero extra row); phi into phi3[0, nx),
// phi' into phi3[2 nx, 3 nx)
int op_vp_efield(const cmm_grid &g, const double *f, double *phi3, void *work, size_t work_bytes, hipStream_t st) {
    if (int rc = vp_check(g, g)) return rc;
    void *w;
    if (int rc = get_work(work, work_bytes, op_vp_efield_workspace(g), &w)) return rc;
    const VP1D q = vp_params(g, g);
    double *part = static_cast<double *>(w);
    double *zero = part + (size_t)q.nchunks * g.nx;
    hipLaunchKernelGGL(k_fill, dim3((g.nx + 255) / 256), dim3(256), 0, st, zero, (long)g.nx, 0.0);
    if (int rc = launch_check("k_fill")) return rc;
    hipLaunchKernelGGL(k_vp_colsum, dim3((g.nx + 255) / 256, q.nchunks), dim3(256), 0, st, f, g.nx, g.ny, kVpRows,
                       part);
    if (int rc = launch_check("k_vp_colsum")) return rc;
    const cd *tw;
    if (int rc = twiddles(g.nx, &tw)) return rc;
    VPJob job{q, zero, part, tw, tw, phi3};
    const size_t lds = sizeof(cd) * g.nx;
    if (int rc = set_lds((const void *)k_vp_phi, lds)) return rc;
    hipLaunchKernelGGL(k_vp_phi, dim3(1), dim3(kFftThreads), lds, st, job);
    return launch_check("k_vp_phi");
}

int op_vp_psi_planes(const cmm_grid &gp, const double *phi, const double *H, double *psi, hipStream_t st) {
    const long N = (long)gp.nx * gp.ny;
    hipLaunchKernelGGL(k_vp_psi_planes, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, phi, H, gp.nx,
                       ilog2(gp.nx), psi);
    return launch_check("k_vp_psi_planes");
}

int op_sample_vort(const cmm_grid &gm, const cmm_grid &gt, const cmm_grid &gf, const double *cx, const double *cy,
                   const double *W, int molly, double *out, hipStream_t st) {
    SampleVort s;
    s.gm = make_dgrid(gm);
    s.gv = make_dgrid(gt);
    s.gf = make_dgrid(gf);
    s.cx = cx;
    s.cy = cy;
    s.W = W;
    s.molly = molly;
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_sample_grid<SampleVort>, grid2d(gt.nx, gt.ny, b), b, 0, st, s, gt.nx, gt.ny, out);
    return launch_check("k_sample_grid");
}

int op_sample_hermite(const cmm_grid &gh, const double *H, const cmm_grid &gt, double *out, hipStream_t st) {
    SampleHermite s{make_dgrid(gh), make_dgrid(gt), H};
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_sample_grid<SampleHermite>, grid2d(gt.nx, gt.ny, b), b, 0, st, s, gt.nx, gt.ny, out);
    return launch_check("k_sample_grid");
}

int op_particles(const cmm_grid &gp, const PsiLevels &P, int L, const double *Lp, double dt, int scheme, int n,
                 double *pos, hipStream_t st) {
    if (n <= 0) return CMM_OK;
    PartParams q;
    for (int i = 0; i < 4; ++i) q.Lp[i] = (Lp && i < L) ? Lp[i] : 0.0;
    q.dt = dt;
    q.L = L;
    q.scheme = scheme;
    q.n = n;
    hipLaunchKernelGGL(k_particles, dim3((n + 255) / 256), dim3(256), 0, st, make_dgrid(gp), P, q, pos);
    return launch_check("k_particles");
}

int op_particles_inertial(const cmm_grid &gp, const PsiLevels &P, int L, const double *Lp, double dt, int scheme,
                          double tau, int n, double *pos, double *vel, hipStream_t st) {
    if (n <= 0) return CMM_OK;
    PartParams q;
    for (int i = 0; i < 4; ++i) q.Lp[i] = (Lp && i < L) ? Lp[i] : 0.0;
    q.dt = dt;
    q.L = L;
    q.scheme = scheme;
    q.n = n;
    hipLaunchKernelGGL(k_particles_inertial, dim3((n + 255) / 256), dim3(256), 0, st, make_dgrid(gp), P, q, tau, pos, vel);
    return launch_check("k_particles_inertial");
}

int op_particles_vel_init(const cmm_grid &gp, const double *psi0, int n, int fluid, const double *pos, double *vel,
                          hipStream_t st) {
    if (n <= 0) return CMM_OK;
    hipLaunchKernelGGL(k_particles_vel_init, dim3((n + 255) / 256), dim3(256), 0, st, make_dgrid(gp), psi0, n, fluid,
                       pos, vel);
    return launch_check("k_particles_vel_init");
}

int op_points_dxdy(const cmm_grid &gp, const double *psi0, int n, const double *pos, double *out, hipStream_t st) {
    if (n <= 0) return CMM_OK;
    hipLaunchKernelGGL(k_points_dxdy, dim3((n + 255) / 256), dim3(256), 0, st, make_dgrid(gp), gp.x1, gp.y1, psi0, n,
                       pos, out);
    return launch_check("k_points_dxdy");
}

// ================================================================================================
// Split-precision form.
//
// The map is stored as the identity plus a displacement D = X - identity, one 48-byte record of
// three float4 per node:
//   M[3j]     = (Dx hi, Dx lo, Dy hi, Dy lo)        value planes as FP32 pairs (hi + lo, ~48 bits)
//   M[3j + 1] = (dX/dx - 1, dX/dy, dY/dx, dY/dy - 1)
//   M[3j + 2] = (d2X/dxdy, d2Y/dxdy, 0, 0)
// with physical derivatives and no wrap jumps (the displacement is periodic).  The value planes
// keep a low part because they reach the Jacobian through the basis derivative (1/h times their
// rounding) and the advection carries them from step to step; the derivative planes enter both
// with weight h and keep one FP32 rounding.  The Hermite fields the gathers read (the
// Lagrange-combined psi fields, the fine vorticity) are FP32 records (f, fx, fy, fxy), one
// 16-byte load per corner.  The arithmetic stays FP64: values are widened on load, the identity
// is added back exactly (X(p) = p + D(p), since bicubic Hermite reproduces the identity), the
// footpoint integration and the stencil sums accumulate in FP64, and the Jacobian check forms
// 1 - det from the displacement derivatives, -(a + d + a d - b c), without the cancellation of
// 1 - det.  The psi levels, the FFTs and the Poisson divide stay FP64.

__device__ __forceinline__ void load16_f4(const float4 *__restrict__ H, const Cell &c, double v[16]) {
    const int idx[4] = {c.i00, c.i10, c.i01, c.i11};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 a = H[idx[q]];
        v[q] = a.x;
        v[4 + q] = a.y;
        v[8 + q] = a.z;
        v[12 + q] = a.w;
    }
}

// corner values of the split map; LO: add the low parts of the value planes
template <bool LO = true>
__device__ __forceinline__ void load_map32(const float4 *__restrict__ M, const Cell &c, double vx[16], double vy[16]) {
    const int idx[4] = {c.i00, c.i10, c.i01, c.i11};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 *r = M + 3 * (long)idx[q];
        const float4 a = r[0], b = r[1], d = r[2];
        vx[q] = LO ? (double)a.x + (double)a.y : (double)a.x;
        vy[q] = LO ? (double)a.z + (double)a.w : (double)a.z;
        vx[4 + q] = b.x;
        vx[8 + q] = b.y;
        vy[4 + q] = b.z;
        vy[8 + q] = b.w;
        vx[12 + q] = d.x;
        vy[12 + q] = d.y;
    }
}

// X(p) = p + D(p) with the corner displacements of p's cell given
__device__ __forceinline__ void diffeo32_vals(const DGrid &g, const Cell &c, const double vx[16], const double vy[16],
                                              double x, double y, double &xo, double &yo) {
    double bx[4], by[4];
    basis(c.dx, bx);
    basis(c.dy, by);
    xo = x + hsum_v(vx, bx, by, g.hx, g.hy);
    yo = y + hsum_v(vy, bx, by, g.hx, g.hy);
}

template <bool P2 = false, bool LO = true>
__device__ __forceinline__ void diffeo32(const DGrid &g, const float4 *__restrict__ M, double x, double y, double &xo,
                                         double &yo) {
    Cell c;
    find_cell<P2>(g, x, y, c);
    double vx[16], vy[16];
    load_map32<LO>(M, c, vx, vy);
    diffeo32_vals(g, c, vx, vy, x, y, xo, yo);
}

// |1 - det grad X| at (x, y) from the displacement derivatives (the derivative scalings of diffeo_det)
__device__ __forceinline__ double incomp32_point(const DGrid &g, const float4 *__restrict__ M, double x, double y) {
    Cell c;
    find_cell(g, x, y, c);
    double bx[4], by[4], dbx[4], dby[4], vx[16], vy[16];
    basis(c.dx, bx);
    basis(c.dy, by);
    basis_d(c.dx, dbx);
    basis_d(c.dy, dby);
    load_map32(M, c, vx, vy);
    const double a = hsum_v(vx, dbx, by, g.hx, g.hy) * g.ihx;
    const double cc = hsum_v(vy, dbx, by, g.hx, g.hy) * g.ihy;
    const double b = hsum_v(vx, bx, dby, g.hx, g.hy) * g.ihx;
    const double d = hsum_v(vy, bx, dby, g.hx, g.hy) * g.ihy;
    return fabs(a + d + (a * d - b * cc));
}

// the four corner records of a cell
__device__ __forceinline__ void load4_rec(const float4 *__restrict__ H, const Cell &c, float4 r[4]) {
    r[0] = H[c.i00];
    r[1] = H[c.i10];
    r[2] = H[c.i01];
    r[3] = H[c.i11];
}

// herm_vel_vals on the four corner records of a cell, each converted where it is used: the same
// terms summed in the same order (corners 00, 10, 01, 11 within each plane's partial sum), with
// four float4 records live instead of 16 doubles; PK: packed records (pack_psi_rec)
template <bool PK>
__device__ __forceinline__ void herm_vel_rec(const DGrid &g, const Cell &c, const float4 r[4], double &ux, double &uy) {
    double bx[4], by[4], dbx[4], dby[4];
    basis(c.dx, bx);
    basis(c.dy, by);
    basis_d(c.dx, dbx);
    basis_d(c.dy, dby);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int xa = q & 1, yb = q >> 1;
        double f, fxy;
        if constexpr (PK) {
            unpack_psi_rec(r[q], f, fxy);
        } else {
            f = r[q].x;
            fxy = r[q].w;
        }
        const double fx = r[q].y, fy = r[q].z;
        a0 += by[yb] * dbx[xa] * f;
        a1 += by[yb] * dbx[2 + xa] * fx;
        a2 += by[2 + yb] * dbx[xa] * fy;
        a3 += by[2 + yb] * dbx[2 + xa] * fxy;
        b0 += dby[yb] * bx[xa] * f;
        b1 += dby[yb] * bx[2 + xa] * fx;
        b2 += dby[2 + yb] * bx[xa] * fy;
        b3 += dby[2 + yb] * bx[2 + xa] * fxy;
    }
    double oa = a0, ob = b0;
    oa += a1 * g.hx;
    ob += b1 * g.hx;
    oa += a2 * g.hy;
    ob += b2 * g.hy;
    a3 *= g.hx;
    b3 *= g.hx;
    a3 *= g.hy;
    b3 *= g.hy;
    oa += a3;
    ob += b3;
    uy = -oa * g.ihx;
    ux = ob * g.ihy;
}

// velocity sources of the footpoint integration on FP32 combined fields (see VelCombined); PK:
// packed records carrying the values' low parts
template <bool P2, bool PK>
struct VelCombined32 {
    DGrid g;
    const float4 *C1, *C12, *C0;
    __device__ __forceinline__ void ev(const float4 *__restrict__ H, double x, double y, double &ux, double &uy) const {
        Cell c;
        find_cell<P2>(g, x, y, c);
        float4 r[4];
        load4_rec(H, c, r);
        herm_vel_rec<PK>(g, c, r, ux, uy);
    }
    __device__ __forceinline__ void w1(const cmm_advect_coeffs &, double x, double y, double &kx, double &ky) const {
        ev(C1, x, y, kx, ky);
    }
    __device__ __forceinline__ void w12(const cmm_advect_coeffs &, double x, double y, double &kx, double &ky) const {
        ev(C12, x, y, kx, ky);
    }
    __device__ __forceinline__ void lev(int, double x, double y, double &ux, double &uy) const {
        ev(C0, x, y, ux, uy);
    }
};

// k_psi_combine with FP32 records: the sums in FP64, rounded once; PK: packed records that keep
// the values' low parts (pack_psi_rec)
template <int L, bool PK>
__global__ void __launch_bounds__(256) k_psi_combine32(long n, PsiLevels P, cmm_advect_coeffs c, float4 *__restrict__ C1,
                                                       float4 *__restrict__ C12, float4 *__restrict__ C0,
                                                       const double *__restrict__ D1, const double *__restrict__ D12,
                                                       const double *__restrict__ LW) {
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    double w1[L], w12[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
        w1[l] = LW ? LW[l] : c.L1[l];
        w12[l] = LW ? LW[4 + l] : c.L12[l];
    }
    double v[L][4];
#pragma unroll
    for (int l = 0; l < L; ++l)
#pragma unroll
        for (int p = 0; p < 4; ++p) v[l][p] = P.p[l][p * n + j];
    auto put = [&](const double *o, float4 *dst) {
        if constexpr (PK) dst[j] = pack_psi_rec(o);
        else dst[j] = float4{(float)o[0], (float)o[1], (float)o[2], (float)o[3]};
    };
    auto comb = [&](const double *w, float4 *dst, const double *D) {
        double o[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            double a = w[0] * v[0][p];
#pragma unroll
            for (int l = 1; l < L; ++l) a += w[l] * v[l][p];
            if (D) a += D[p * n + j];
            o[p] = a;
        }
        put(o, dst);
    };
    comb(w1, C1, D1);
    if (C12) comb(w12, C12, D12);
    if (C0) put(v[0], C0);
}

// Output stage of the split advection: lane k of a node's group forms displacement plane k (and
// k + G) from the group's footpoint displacements q_m = X(footpoint m) - node (FP64), in footpoint
// order with the stencil weights.  The stencil sums of the derivative planes annihilate the node
// coordinate (their weights sum to zero) and the value weights sum to one, so these are the
// planes of X - identity; the identity's unit derivatives are removed before rounding to FP32.
// record position of plane o (0-3: X planes, 4-7: Y planes) in the 12 floats of a node
__device__ __forceinline__ int split_slot(int o) {
    const int p = o & 3, y = o >> 2;
    return p == 0 ? 2 * y : p == 3 ? 8 + y : 4 + 2 * y + (p - 1);
}

template <int G>
__device__ __forceinline__ void advect_store32(const cmm_advect_coeffs &c, int nfoot, int node, bool active, int k,
                                               double qx, double qy, float *__restrict__ Mout) {
    const int base = (threadIdx.x & 63) - k;
#pragma unroll
    for (int r = 0; r < (8 + G - 1) / G; ++r) {
        const int o = k + r * G;
        const int plane = o & 3;
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m < kMaxFoot; ++m) {
            if (m < nfoot) {
                const double vx = __shfl(qx, base + m, 64);
                const double vy = __shfl(qy, base + m, 64);
                const double wm = plane == 0 ? c.c1[m] : plane == 1 ? c.cx[m] : plane == 2 ? c.cy[m] : c.cxy[m];
                acc += (o < 4 ? vx : vy) * wm;
            }
        }
        if (o == 1 || o == 6) acc -= 1.0;
        if (active && o < 8) {
            const float hi = (float)acc;
            float *rec = Mout + 12L * node;
            rec[split_slot(o)] = hi;
            if (plane == 0) rec[split_slot(o) + 1] = (float)(acc - (double)hi);
            else if (plane == 3) rec[10 + (o >> 2)] = 0.0f;
        }
    }
}

// PSI: the form of the combined psi fields (a PsiRecForm)
template <int TIN, int G, bool P2, int PSI>
__global__ void __launch_bounds__(256) k_map_advect32(DGrid gm, DGrid gp, const void *C1, const void *C12,
                                                      const void *C0, const float4 *__restrict__ M,
                                                      cmm_advect_coeffs c, float *__restrict__ Mout, int nblocks) {
    constexpr int NPB = 256 / G;
    const int blk = xcd_block(blockIdx.x, nblocks);
    const int node = blk * NPB + threadIdx.x / G;
    const int k = threadIdx.x % G;
    const int nfoot = 4 + 4 * c.map_update_order_num;
    const bool active = node < gm.n;
    double qx = 0.0, qy = 0.0;
    if (active && k < nfoot) {
        const int ix = node % gm.nx, iy = node / gm.nx;
        const double xn = gm.x0 + ix * gm.hx, yn = gm.y0 + iy * gm.hy;
        const double xe = xn + ((k >> 2) + 1) * c.eps * (1 - 2 * ((k >> 1) & 1));
        const double ye = yn + ((k >> 2) + 1) * c.eps * (1 - 2 * (((k + 1) >> 1) & 1));
        double xf, yf, px, py;
        if constexpr (PSI == PsiRec64)
            step_backward<TIN>(VelCombined<P2>{gp, (const double *)C1, (const double *)C12, (const double *)C0}, c, xe,
                               ye, xf, yf);
        else
            step_backward<TIN>(VelCombined32<P2, PSI == PsiRec32Pk>{gp, (const float4 *)C1, (const float4 *)C12,
                                                                    (const float4 *)C0},
                               c, xe, ye, xf, yf);
    diffeo32<P2>(gm, M, xf, yf, px, py);
        qx = px - xn;
        qy = py - yn;
    }
    advect_store32<G>(c, nfoot, node, active, k, qx, qy, Mout);
}

// k_map_advect32 with two footpoints per lane (footpoints k and k + 4 of the 8-footpoint stencil),
// as k_map_advect_pair: the second footpoint's psi and map records are loaded only when its cell
// differs from the first one's
template <bool P2, bool PK>
struct VelPair32 {
    DGrid g;
    const float4 *C1, *C12, *C0;
    __device__ __forceinline__ void ev(const float4 *__restrict__ H, PointPair x, PointPair y, PointPair &ux,
                                       PointPair &uy) const {
        Cell ca, cb;
        find_cell<P2>(g, x.a, y.a, ca);
        find_cell<P2>(g, x.b, y.b, cb);
        float4 r[4];
        load4_rec(H, ca, r);
        herm_vel_rec<PK>(g, ca, r, ux.a, uy.a);
        if (!same_cell(ca, cb)) load4_rec(H, cell_after(cb, ux.a), r);
        herm_vel_rec<PK>(g, cb, r, ux.b, uy.b);
    }
    __device__ __forceinline__ void w1(const cmm_advect_coeffs &, PointPair x, PointPair y, PointPair &kx,
                                       PointPair &ky) const {
        ev(C1, x, y, kx, ky);
    }
    __device__ __forceinline__ void w12(const cmm_advect_coeffs &, PointPair x, PointPair y, PointPair &kx,
                                        PointPair &ky) const {
        ev(C12, x, y, kx, ky);
    }
    __device__ __forceinline__ void lev(int, PointPair x, PointPair y, PointPair &ux, PointPair &uy) const {
        ev(C0, x, y, ux, uy);
    }
};

// k_map_advect32_pair's block body with the velocity source V (k_map_advect32_pair keeps its own
// copy: as a shared body the compiler schedules it a few per cent slower)
template <int TIN, class V>
__device__ __forceinline__ void advect32_pair_block(int blk, const DGrid &gm, const V &vel,
                                                    const float4 *__restrict__ M, const cmm_advect_coeffs &c,
                                                    float *__restrict__ Mout) {
    constexpr int G = 4, NPB = 256 / G;
    const int node = blk * NPB + threadIdx.x / G;
    const int k = threadIdx.x % G;
    const bool active = node < gm.n;
    PointPair qx{0.0, 0.0}, qy{0.0, 0.0};
    if (active) {
        const int ix = node % gm.nx, iy = node / gm.nx;
        const double xn = gm.x0 + ix * gm.hx, yn = gm.y0 + iy * gm.hy;
        const int fa = k, fb = k + 4;
        const PointPair xe{xn + ((fa >> 2) + 1) * c.eps * (1 - 2 * ((fa >> 1) & 1)),
                           xn + ((fb >> 2) + 1) * c.eps * (1 - 2 * ((fb >> 1) & 1))};
        const PointPair ye{yn + ((fa >> 2) + 1) * c.eps * (1 - 2 * (((fa + 1) >> 1) & 1)),
                           yn + ((fb >> 2) + 1) * c.eps * (1 - 2 * (((fb + 1) >> 1) & 1))};
        PointPair xf, yf;
        step_backward<TIN>(vel, c, xe, ye, xf, yf);
        Cell ca, cb;
        find_cell<true>(gm, xf.a, yf.a, ca);
        find_cell<true>(gm, xf.b, yf.b, cb);
        double vx[16], vy[16], px, py;
        load_map32(M, ca, vx, vy);
        diffeo32_vals(gm, ca, vx, vy, xf.a, yf.a, px, py);
        qx.a = px - xn;
        qy.a = py - yn;
        if (!same_cell(ca, cb)) load_map32(M, cell_after(cb, qx.a), vx, vy);
        diffeo32_vals(gm, cb, vx, vy, xf.b, yf.b, px, py);
        qx.b = px - xn;
        qy.b = py - yn;
    }
    // output planes as advect_store32: footpoint m in lane m & 3, component m >> 2
    const int base = (threadIdx.x & 63) - k;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int o = k + r * G;
        const int plane = o & 3;
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const double vx = __shfl(m < 4 ? qx.a : qx.b, base + (m & 3), 64);
            const double vy = __shfl(m < 4 ? qy.a : qy.b, base + (m & 3), 64);
            const double wm = plane == 0 ? c.c1[m] : plane == 1 ? c.cx[m] : plane == 2 ? c.cy[m] : c.cxy[m];
            acc += (o < 4 ? vx : vy) * wm;
        }
        if (o == 1 || o == 6) acc -= 1.0;
        if (active) {
            const float hi = (float)acc;
            float *rec = Mout + 12L * node;
            rec[split_slot(o)] = hi;
            if (plane == 0) rec[split_slot(o) + 1] = (float)(acc - (double)hi);
            else if (plane == 3) rec[10 + (o >> 2)] = 0.0f;
        }
    }
}

template <int TIN, bool PK>
__global__ void __launch_bounds__(256)
k_map_advect32_pair(DGrid gm, DGrid gp, const float4 *C1, const float4 *C12, const float4 *C0,
                    const float4 *__restrict__ M, cmm_advect_coeffs c,
                    float *__restrict__ Mout, int nblocks) {
    constexpr int G = 4, NPB = 256 / G;
    const int blk = xcd_block(blockIdx.x, nblocks);
    const int node = blk * NPB + threadIdx.x / G;
    const int k = threadIdx.x % G;
    const bool active = node < gm.n;
    PointPair qx{0.0, 0.0}, qy{0.0, 0.0};
    if (active) {
        const int ix = node % gm.nx, iy = node / gm.nx;
        const double xn = gm.x0 + ix * gm.hx, yn = gm.y0 + iy * gm.hy;
        const int fa = k, fb = k + 4;
        const PointPair xe{xn + ((fa >> 2) + 1) * c.eps * (1 - 2 * ((fa >> 1) & 1)),
                           xn + ((fb >> 2) + 1) * c.eps * (1 - 2 * ((fb >> 1) & 1))};
        const PointPair ye{yn + ((fa >> 2) + 1) * c.eps * (1 - 2 * (((fa + 1) >> 1) & 1)),
                           yn + ((fb >> 2) + 1) * c.eps * (1 - 2 * (((fb + 1) >> 1) & 1))};
        PointPair xf, yf;
        step_backward<TIN>(VelPair32<true, PK>{gp, C1, C12, C0}, c, xe, ye, xf, yf);
        Cell ca, cb;
        find_cell<true>(gm, xf.a, yf.a, ca);
        find_cell<true>(gm, xf.b, yf.b, cb);
        double vx[16], vy[16], px, py;
        load_map32(M, ca, vx, vy);
        diffeo32_vals(gm, ca, vx, vy, xf.a, yf.a, px, py);
        qx.a = px - xn;
        qy.a = py - yn;
        if (!same_cell(ca, cb)) load_map32(M, cell_after(cb, qx.a), vx, vy);
        diffeo32_vals(gm, cb, vx, vy, xf.b, yf.b, px, py);
        qx.b = px - xn;
        qy.b = py - yn;
    }
    // output planes as advect_store32: footpoint m in lane m & 3, component m >> 2
    const int base = (threadIdx.x & 63) - k;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int o = k + r * G;
        const int plane = o & 3;
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const double vx = __shfl(m < 4 ? qx.a : qx.b, base + (m & 3), 64);
            const double vy = __shfl(m < 4 ? qy.a : qy.b, base + (m & 3), 64);
            const double wm = plane == 0 ? c.c1[m] : plane == 1 ? c.cx[m] : plane == 2 ? c.cy[m] : c.cxy[m];
            acc += (o < 4 ? vx : vy) * wm;
        }
        if (o == 1 || o == 6) acc -= 1.0;
        if (active) {
            const float hi = (float)acc;
            float *rec = Mout + 12L * node;
            rec[split_slot(o)] = hi;
            if (plane == 0) rec[split_slot(o) + 1] = (float)(acc - (double)hi);
            else if (plane == 3) rec[10 + (o >> 2)] = 0.0f;
        }
    }
}

// the split map of a Vlasov-Poisson flow advected with the separable velocity (k_map_advect_vp's
// velocity from k_vp_combine1d's tables, the split records as k_map_advect32_pair)
template <int TIN>
__global__ void __launch_bounds__(256) k_map_advect32_vp(DGrid gm, DGrid gp, const double *T, const double *S,
                                                         const double *H, const float4 *__restrict__ M,
                                                         cmm_advect_coeffs c, float *__restrict__ Mout, int nblocks) {
    advect32_pair_block<TIN>(xcd_block(blockIdx.x, nblocks), gm, VelSep{gp, T, S, H}, M, c, Mout);
}

// k_incomp_tensor on the split map: the same separable sums over the displacement records
template <int TX, int TY>
__global__ void __launch_bounds__(256) k_incomp_tensor32(DGrid gm, DGrid gc, const float4 *__restrict__ M,
                                                         unsigned long long *__restrict__ out, int ntx, long ntiles) {
    double v = 0.0;
    long t_begin, t_end;
    tile_range(ntiles, t_begin, t_end);
    for (long t0 = t_begin; t0 < t_end; t0 += 256) {
        const long t = t0 + threadIdx.x;
        if (t >= t_end) continue;
        const int ix0 = (int)(t % ntx) * TX, iy0 = (int)(t / ntx) * TY;
        double xs[TX], ys[TY];
#pragma unroll
        for (int a = 0; a < TX; ++a) xs[a] = gm.x0 + (ix0 + a) * gc.hx + 0.5 * gc.hx;
#pragma unroll
        for (int b = 0; b < TY; ++b) ys[b] = gm.y0 + (iy0 + b) * gc.hy + 0.5 * gc.hy;
        Cell c0, c1;
        find_cell(gm, xs[0], ys[0], c0);
        find_cell(gm, xs[TX - 1], ys[TY - 1], c1);
        if (!same_cell(c0, c1)) {
#pragma unroll 1
            for (int i = 0; i < TX * TY; ++i) v = fmax(v, incomp32_point(gm, M, xs[i % TX], ys[i / TX]));
            continue;
        }
        double vx[16], vy[16];
        load_map32(M, c0, vx, vy);
        double by[TY][4], dby[TY][4];
#pragma unroll
        for (int b = 0; b < TY; ++b) {
            const double fy = (ys[b] - gm.y0) * gm.ihy;
            const double d = fy - floor(fy);
            basis(d, by[b]);
            basis_d(d, dby[b]);
        }
#pragma unroll
        for (int a = 0; a < TX; ++a) {
            const double fx = (xs[a] - gm.x0) * gm.ihx;
            const double d = fx - floor(fx);
            double bx[4], dbx[4];
            basis(d, bx);
            basis_d(d, dbx);
            double P[4][4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double *w = (k & 1) ? bx : dbx;
                const double *f = (k < 2) ? vx : vy;
#pragma unroll
                for (int n = 0; n < 2; ++n) {
                    P[k][n] = w[0] * f[2 * n] + w[1] * f[2 * n + 1] + gm.hx * (w[2] * f[4 + 2 * n] + w[3] * f[5 + 2 * n]);
                    P[k][2 + n] =
                        w[0] * f[8 + 2 * n] + w[1] * f[9 + 2 * n] + gm.hx * (w[2] * f[12 + 2 * n] + w[3] * f[13 + 2 * n]);
                }
            }
#pragma unroll
            for (int b = 0; b < TY; ++b) {
                auto ysum = [&](const double *p, const double *cw) {
                    return cw[0] * p[0] + cw[1] * p[1] + gm.hy * (cw[2] * p[2] + cw[3] * p[3]);
                };
                const double ea = ysum(P[0], by[b]) * gm.ihx;
                const double eb = ysum(P[1], dby[b]) * gm.ihx;
                const double ec = ysum(P[2], by[b]) * gm.ihy;
                const double ed = ysum(P[3], dby[b]) * gm.ihy;
                v = fmax(v, fabs(ea + ed + (ea * ed - eb * ec)));
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
    }
}

// the check point by point (check grids that do not refine the map grid by an integer ratio)
__global__ void __launch_bounds__(256) k_incomp32(DGrid gm, DGrid gc, const float4 *__restrict__ M,
                                                  unsigned long long *__restrict__ out, int tiles_x, int tiles_total) {
    double v = 0.0;
    for (int tt = blockIdx.x; tt < tiles_total; tt += gridDim.x) {
        const int t = xcd_block(tt, tiles_total);
        const int ix = (t % tiles_x) * 64 + threadIdx.x;
        const int iy = (t / tiles_x) * 4 + threadIdx.y;
        if (ix < gc.nx && iy < gc.ny)
            v = fmax(v, incomp32_point(gm, M, gm.x0 + ix * gc.hx + 0.5 * gc.hx, gm.y0 + iy * gc.hy + 0.5 * gc.hy));
    }
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __shared__ double red[4];
    if (threadIdx.x == 0) red[threadIdx.y] = v;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
    }
}

// omega(X(x)) through the split map and the FP32 fine vorticity records; SampleVortT's mollifiers
// and column sampling with the split evaluations.  The positions use the displacement without its
// low parts: their error, 2^-24 |X - x|, moves the sampled vorticity by that times |grad omega|.
// incomp_column on the split map: the corner values as load_map32 gives them (displacement, value
// planes with their low parts), the arithmetic of k_incomp_tensor32
template <int R>
__device__ __forceinline__ double incomp32_column(const DGrid &gm, const double vx[16], const double vy[16], double xc,
                                                  const double yc[R]) {
    const double fx = (xc - gm.x0) * gm.ihx;
    const double dx = fx - floor(fx);
    double bx[4], dbx[4];
    basis(dx, bx);
    basis_d(dx, dbx);
    double P[4][4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double *w = (k & 1) ? bx : dbx;
        const double *f = (k < 2) ? vx : vy;
#pragma unroll
        for (int n = 0; n < 2; ++n) {
            P[k][n] = w[0] * f[2 * n] + w[1] * f[2 * n + 1] + gm.hx * (w[2] * f[4 + 2 * n] + w[3] * f[5 + 2 * n]);
            P[k][2 + n] = w[0] * f[8 + 2 * n] + w[1] * f[9 + 2 * n] + gm.hx * (w[2] * f[12 + 2 * n] + w[3] * f[13 + 2 * n]);
        }
    }
    double v = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const double fy = (yc[r] - gm.y0) * gm.ihy;
        const double d = fy - floor(fy);
        double by[4], dby[4];
        basis(d, by);
        basis_d(d, dby);
        auto ysum = [&](const double *q, const double *cw) {
            return cw[0] * q[0] + cw[1] * q[1] + gm.hy * (cw[2] * q[2] + cw[3] * q[3]);
        };
        const double ea = ysum(P[0], by) * gm.ihx;
        const double eb = ysum(P[1], dby) * gm.ihx;
        const double ec = ysum(P[2], by) * gm.ihy;
        const double ed = ysum(P[3], dby) * gm.ihy;
        v = fmax(v, fabs(ea + ed + (ea * ed - eb * ec)));
    }
    return v;
}

struct SampleVort32 {
    DGrid gm, gv, gf;
    const float4 *__restrict__ M;
    const float4 *__restrict__ W;
    int molly;
    __device__ __forceinline__ double wval(double x, double y) const {
        Cell c;
        find_cell(gf, x, y, c);
        double v[16];
        load16_f4(W, c, v);
        return herm_eval_vals(gf, c, v);
    }
    __device__ double operator()(int ix, int iy) const {
        const double x = gm.x0 + ix * gv.hx, y = gm.y0 + iy * gv.hy;
        double x2, y2, v;
        if (molly == 4) {
            diffeo32<false, true>(gm, M, x, y, x2, y2);
            v = wval(x2, y2) / 3.0;
            for (int m = 0; m < 4; ++m) {
                x2 = x + gv.hx / 2.0 * ((m >> 2) & 1) * (-1 + 2 * (m & 1));
                y2 = y + gv.hy / 2.0 * ((m >> 1) & 1) * (-1 + 2 * (m & 1));
                diffeo32<false, true>(gm, M, x2, y2, x2, y2);
                v += wval(x2, y2) / 6.0;
            }
        } else if (molly == 8) {
            v = 0;
            for (int m = 0; m < 9; ++m) {
                x2 = x + gv.hx * (-1 + m % 3) / 2.0;
                y2 = y + gv.hy * (-1 + m / 3) / 2.0;
                diffeo32<false, true>(gm, M, x2, y2, x2, y2);
                v += ((1 + (m % 3)) & 1) * ((1 + (m / 3)) & 1) * wval(x2, y2) / 16.0;
            }
        } else {
            diffeo32<false, true>(gm, M, x, y, x2, y2);
            v = wval(x2, y2);
        }
        return v;
    }
    template <int R>
    __device__ __forceinline__ double check_points(int x, int y0) const {
        double v = 0.0;
        const double xc = gm.x0 + x * gv.hx + 0.5 * gv.hx;
#pragma unroll 1
        for (int r = 0; r < R; ++r) v = fmax(v, incomp32_point(gm, M, xc, gm.y0 + (y0 + r) * gv.hy + 0.5 * gv.hy));
        return v;
    }
    // as SampleVortT::column (also its return value); CHK: the check at the column's cell
    // centres.  The map's value planes are taken with their low parts (the check needs them: they
    // reach the Jacobian through the basis derivative).
    template <int R, bool CHK = false>
    __device__ __forceinline__ bool column(int x, int y0, double *out, double &chk) const {
        if (molly != 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) out[r] = (*this)(x, y0 + r);
            return false;
        }
        const double xp = gm.x0 + x * gv.hx;
        Cell cf, cl;
        find_cell(gm, xp, gm.y0 + y0 * gv.hy, cf);
        find_cell(gm, xp, gm.y0 + (y0 + R - 1) * gv.hy, cl);
        if (!same_cell(cf, cl)) {
#pragma unroll
            for (int r = 0; r < R; ++r) out[r] = (*this)(x, y0 + r);
            return false;
        }
        double vx[16], vy[16];
        load_map32(M, cf, vx, vy);
        if constexpr (CHK) {
            double yc[R];
#pragma unroll
            for (int r = 0; r < R; ++r) yc[r] = gm.y0 + (y0 + r) * gv.hy + 0.5 * gv.hy;
            chk = fmax(chk, incomp32_column<R>(gm, vx, vy, xp + 0.5 * gv.hx, yc));
        }
        Cell wc[R];
        double wv[R][16];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Cell c;
            const double yp = gm.y0 + (y0 + r) * gv.hy;
            find_cell(gm, xp, yp, c);
            double px, py;
            diffeo32_vals(gm, c, vx, vy, xp, yp, px, py);
            find_cell(gf, px, py, wc[r]);
            load16_f4(W, wc[r], wv[r]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) out[r] = herm_eval_vals(gf, wc[r], wv[r]);
        return true;
    }
    template <int R>
    __device__ __forceinline__ void column(int x, int y0, double *out) const {
        double chk = 0.0;
        column<R, false>(x, y0, out, chk);
    }
};

// plane-major FP64 map <-> split records
__global__ void __launch_bounds__(256) k_map_to32(DGrid g, const double *__restrict__ cx, const double *__restrict__ cy,
                                                  float4 *__restrict__ M) {
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= g.n) return;
    const long n = g.n;
    const double xn = g.x0 + (int)(j % g.nx) * g.hx, yn = g.y0 + (int)(j / g.nx) * g.hy;
    const double dx = cx[j] - xn, dy = cy[j] - yn;
    const float hx = (float)dx, hy = (float)dy;
    M[3 * j] = float4{hx, (float)(dx - (double)hx), hy, (float)(dy - (double)hy)};
    M[3 * j + 1] = float4{(float)(cx[n + j] - 1.0), (float)cx[2 * n + j], (float)cy[n + j], (float)(cy[2 * n + j] - 1.0)};
    M[3 * j + 2] = float4{(float)cx[3 * n + j], (float)cy[3 * n + j], 0.0f, 0.0f};
}

__global__ void __launch_bounds__(256) k_map_from32(DGrid g, const float4 *__restrict__ M, double *__restrict__ cx,
                                                    double *__restrict__ cy) {
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= g.n) return;
    const long n = g.n;
    const double xn = g.x0 + (int)(j % g.nx) * g.hx, yn = g.y0 + (int)(j / g.nx) * g.hy;
    const float4 a = M[3 * j], b = M[3 * j + 1], d = M[3 * j + 2];
    cx[j] = xn + ((double)a.x + (double)a.y);
    cx[n + j] = 1.0 + (double)b.x;
    cx[2 * n + j] = b.y;
    cx[3 * n + j] = d.x;
    cy[j] = yn + ((double)a.z + (double)a.w);
    cy[n + j] = b.z;
    cy[2 * n + j] = 1.0 + (double)b.w;
    cy[3 * n + j] = d.y;
}

// FP32 node records of a plane-major FP64 Hermite field
__global__ void __launch_bounds__(256) k_interleave4_32(long n, const double *__restrict__ H, float4 *__restrict__ H4) {
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    H4[j] = float4{(float)H[j], (float)H[n + j], (float)H[2 * n + j], (float)H[3 * n + j]};
}

int op_map_to32(const cmm_grid &g, const double *cx, const double *cy, float *M, hipStream_t st) {
    const DGrid d = make_dgrid(g);
    hipLaunchKernelGGL(k_map_to32, dim3((unsigned)((d.n + 255) / 256)), dim3(256), 0, st, d, cx, cy, (float4 *)M);
    return launch_check("k_map_to32");
}

int op_map_from32(const cmm_grid &g, const float *M, double *cx, double *cy, hipStream_t st) {
    const DGrid d = make_dgrid(g);
    hipLaunchKernelGGL(k_map_from32, dim3((unsigned)((d.n + 255) / 256)), dim3(256), 0, st, d, (const float4 *)M, cx, cy);
    return launch_check("k_map_from32");
}

int op_interleave4_32(const cmm_grid &g, const double *H, float *H4, hipStream_t st) {
    const long n = (long)g.nx * g.ny;
    hipLaunchKernelGGL(k_interleave4_32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, H, (float4 *)H4);
    return launch_check("k_interleave4_32");
}

bool op_advect32_supported(const cmm_advect_coeffs &c) { return c.direction != 1 && advect_combinable(c); }

// bytes of one combined psi field in form f on n nodes, rounded up to 256
size_t psi_rec_bytes(int f, long n) {
    const size_t per = f == PsiRec64 ? 4 * sizeof(double) : sizeof(float4);
    return (per * (size_t)n + 255) / 256 * 256;
}

size_t op_advect32_workspace(const cmm_grid &gp, const cmm_advect_coeffs &c, int psi_form) {
    if (!op_advect32_supported(c)) return 0;
    const int fields = 1 + advect_needs_c12(c) + advect_needs_c0(c);
    return fields * psi_rec_bytes(psi_form, (long)gp.nx * gp.ny);
}

int op_map_advect32(const cmm_grid &gm, const cmm_grid &gp, const PsiLevels &P, const cmm_advect_coeffs &c,
                    const float *M_in, float *M_out, void *work, size_t work_bytes, hipStream_t st, hipEvent_t mid,
                    int psi_form, const double *corr1, const double *corr12, const double *lw, const PsiPre *pre_in) {
    if (!op_advect32_supported(c)) {
        set_error("split-precision advection: the time scheme reads psi levels on their own (AB2, RK3Mod), "
                  "needs two or more psi levels, and is backward only");
        return CMM_ERR_ARG;
    }
    if (psi_form != PsiRec32 && psi_form != PsiRec32Pk && psi_form != PsiRec64) {
        set_error("split-precision advection: unknown psi record form");
        return CMM_ERR_ARG;
    }
    const size_t need = op_advect32_workspace(gp, c, psi_form);
    void *w;
    if (int rc = get_work(work, work_bytes, need, &w)) return rc;
    DGrid dm = make_dgrid(gm), dp = make_dgrid(gp);
    const long n = dp.n;
    const size_t rec = psi_rec_bytes(psi_form, n);
    char *next = static_cast<char *>(w);
    void *C1 = next, *C12 = nullptr, *C0 = nullptr;
    next += rec;
    if (advect_needs_c12(c)) { C12 = next; next += rec; }
    if (advect_needs_c0(c)) C0 = next;
    if (pre_in) {  // the fields the previous psi solve wrote for this step's weights
        if (corr1 || corr12 || lw || pre_in->form != (psi_form == PsiRec64 ? 0 : 1) || psi_form == PsiRec32) {
            set_error("split advection: precombined fields need the packed or FP64 form without corrections");
            return CMM_ERR_ARG;
        }
        C1 = pre_in->C1;
        C12 = advect_needs_c12(c) ? pre_in->C12 : nullptr;
        C0 = advect_needs_c0(c) ? pre_in->C0 : nullptr;
    }
    const dim3 g((unsigned)((n + 255) / 256));
    if (!pre_in && psi_form == PsiRec64) {
        double *d1 = (double *)C1, *d12 = (double *)C12, *d0 = (double *)C0;
        switch (c.l_order) {
        case 2: hipLaunchKernelGGL(k_psi_combine<2>, g, dim3(256), 0, st, n, P, c, d1, d12, d0, corr1, corr12, lw); break;
        case 3: hipLaunchKernelGGL(k_psi_combine<3>, g, dim3(256), 0, st, n, P, c, d1, d12, d0, corr1, corr12, lw); break;
        default: hipLaunchKernelGGL(k_psi_combine<4>, g, dim3(256), 0, st, n, P, c, d1, d12, d0, corr1, corr12, lw); break;
        }
    } else if (!pre_in) {
        float4 *f1 = (float4 *)C1, *f12 = (float4 *)C12, *f0 = (float4 *)C0;
#define CMM_C32(LL, PK)                                                                                                \
    hipLaunchKernelGGL((k_psi_combine32<LL, PK>), g, dim3(256), 0, st, n, P, c, f1, f12, f0, corr1, corr12, lw)
        const bool pk = psi_form == PsiRec32Pk;
        switch (c.l_order) {
        case 2: if (pk) CMM_C32(2, true); else CMM_C32(2, false); break;
        case 3: if (pk) CMM_C32(3, true); else CMM_C32(3, false); break;
        default: if (pk) CMM_C32(4, true); else CMM_C32(4, false); break;
        }
#undef CMM_C32
    }
    if (int rc = launch_check("k_psi_combine")) return rc;
    if (mid) (void)hipEventRecord(mid, st);
    const int nfoot = 4 + 4 * c.map_update_order_num;
    const int G = nfoot <= 4 ? 4 : nfoot <= 8 ? 8 : 16;
    const int nb = (int)(((long)dm.n * G + 255) / 256);
    const bool p2 = dm.lgx >= 0 && dm.lgy >= 0 && dp.lgx >= 0 && dp.lgy >= 0;
    const float4 *Mi = (const float4 *)M_in;
#define CMM_A32(T, GG, PP, S)                                                                                          \
    hipLaunchKernelGGL((k_map_advect32<T, GG, PP, S>), dim3(nb), dim3(256), 0, st, dm, dp, C1, C12, C0, Mi, c, M_out, \
                       nb)
    const int nb2 = (int)(((long)dm.n * 4 + 255) / 256);  // two footpoints per lane
#define CMM_A32P(T, LOV)                                                                                               \
    hipLaunchKernelGGL((k_map_advect32_pair<T, LOV>), dim3(nb2), dim3(256), 0, st, dm, dp, (const float4 *)C1,          \
                       (const float4 *)C12, (const float4 *)C0, Mi, c, M_out, nb2)
#define CMM_A32_S(T, S)                                                                                                \
    do {                                                                                                               \
        if (G == 8 && p2 && S == PsiRec32Pk) CMM_A32P(T, true);                                                       \
        else if (G == 8 && p2 && S == PsiRec32) CMM_A32P(T, false);                                                   \
        else if (G == 8 && p2) CMM_A32(T, 8, true, S);                                                                 \
        else if (G == 4) CMM_A32(T, 4, false, S);                                                                      \
        else if (G == 8) CMM_A32(T, 8, false, S);                                                                      \
        else CMM_A32(T, 16, false, S);                                                                                 \
    } while (0)
#define CMM_A32_T(T)                                                                                                   \
    do {                                                                                                               \
        if (psi_form == PsiRec64) CMM_A32_S(T, PsiRec64);                                                              \
        else if (psi_form == PsiRec32Pk) CMM_A32_S(T, PsiRec32Pk);                                                     \
        else CMM_A32_S(T, PsiRec32);                                                                                   \
    } while (0)
    switch (c.time_integration_num) {
    case 10: CMM_A32_T(10); break;
    case 21: CMM_A32_T(21); break;
    case 30: CMM_A32_T(30); break;
    case 40: CMM_A32_T(40); break;
    case 41: CMM_A32_T(41); break;
    default: set_error("unknown time integration"); return CMM_ERR_ARG;
    }
#undef CMM_A32_T
#undef CMM_A32_S
#undef CMM_A32P
#undef CMM_A32
    return launch_check("k_map_advect32");
}

int op_map_advect32_vp(const cmm_grid &gm, const cmm_grid &gp, const PsiLevels &P, const cmm_advect_coeffs &c,
                       const float *M_in, float *M_out, const double *H, void *work, size_t work_bytes,
                       hipStream_t st, hipEvent_t mid, const double *lw) {
    if (!op_vp_advect_supported(gm, gp, c)) {
        set_error("separable Vlasov-Poisson advection: needs the 8-footpoint stencil, a combinable scheme and "
                  "power-of-two grids");
        return CMM_ERR_ARG;
    }
    const DGrid dm = make_dgrid(gm), dp = make_dgrid(gp);
    void *w;
    if (int rc = get_work(work, work_bytes, sizeof(double) * (6 * (size_t)gp.nx + 2), &w)) return rc;
    double *T = static_cast<double *>(w), *S = T + 6 * gp.nx;
    const dim3 g1((unsigned)((gp.nx + 255) / 256));
    switch (c.l_order) {
    case 2: hipLaunchKernelGGL(k_vp_combine1d<2>, g1, dim3(256), 0, st, gp.nx, (long)dp.n, P, c, lw, T, S); break;
    case 3: hipLaunchKernelGGL(k_vp_combine1d<3>, g1, dim3(256), 0, st, gp.nx, (long)dp.n, P, c, lw, T, S); break;
    default: hipLaunchKernelGGL(k_vp_combine1d<4>, g1, dim3(256), 0, st, gp.nx, (long)dp.n, P, c, lw, T, S); break;
    }
    if (int rc = launch_check("k_vp_combine1d")) return rc;
    if (mid) (void)hipEventRecord(mid, st);
    const int nb = (int)(((long)dm.n * 4 + 255) / 256);
    const float4 *Mi = reinterpret_cast<const float4 *>(M_in);
#define CMM_A32VP(T_)                                                                                                  \
    hipLaunchKernelGGL(k_map_advect32_vp<T_>, dim3(nb), dim3(256), 0, st, dm, dp, T, S, H, Mi, c, M_out, nb)
    switch (c.time_integration_num) {
    case 10: CMM_A32VP(10); break;
    case 21: CMM_A32VP(21); break;
    case 30: CMM_A32VP(30); break;
    case 40: CMM_A32VP(40); break;
    case 41: CMM_A32VP(41); break;
    default: set_error("unknown time integration"); return CMM_ERR_ARG;
    }
#undef CMM_A32VP
    return launch_check("k_map_advect32_vp");
}

int op_incomp32_launch(const cmm_grid &gm, const cmm_grid &gc, const float *Mf, unsigned long long *slot,
                       hipStream_t st, IncompTune tune) {
    DGrid dm = make_dgrid(gm), dc = make_dgrid(gc);
    const float4 *M = (const float4 *)Mf;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, slot);
    const int r = gc.nx / gm.nx;
    if (r >= 1 && gc.nx == r * gm.nx && gc.ny == r * gm.ny && gc.x0 == gm.x0 && gc.x1 == gm.x1 &&
        gc.y0 == gm.y0 && gc.y1 == gm.y1) {
        const int T = incomp_tile(r, tune);
        const int ntx = gc.nx / T;
        const long ntiles = (long)ntx * (gc.ny / T);
        const unsigned nb = (unsigned)std::max<long>(1, std::min<long>((ntiles + 255) / 256, tune.max_blocks));
        if (T == 4) hipLaunchKernelGGL((k_incomp_tensor32<4, 4>), dim3(nb), dim3(256), 0, st, dm, dc, M, slot, ntx, ntiles);
        else if (T == 3) hipLaunchKernelGGL((k_incomp_tensor32<3, 3>), dim3(nb), dim3(256), 0, st, dm, dc, M, slot, ntx, ntiles);
        else if (T == 2) hipLaunchKernelGGL((k_incomp_tensor32<2, 2>), dim3(nb), dim3(256), 0, st, dm, dc, M, slot, ntx, ntiles);
        else hipLaunchKernelGGL((k_incomp_tensor32<1, 1>), dim3(nb), dim3(256), 0, st, dm, dc, M, slot, ntx, ntiles);
        return launch_check("k_incomp_tensor32");
    }
    const int tx = (gc.nx + 63) / 64, ty = (gc.ny + 3) / 4, tt = tx * ty;
    hipLaunchKernelGGL(k_incomp32, dim3(std::min(tt, 2048)), dim3(64, 4), 0, st, dm, dc, M, slot, tx, tt);
    return launch_check("k_incomp32");
}

int op_incomp32(const cmm_grid &gm, const cmm_grid &gc, const float *M, double *out, hipStream_t st) {
    unsigned long long *slot;
    if (int rc = scalar_slot(&slot)) return rc;
    if (int rc = op_incomp32_launch(gm, gc, M, slot, st)) return rc;
    unsigned long long h = 0;
    if (int rc = hip_check(hipMemcpyAsync(&h, slot, sizeof(h), hipMemcpyDeviceToHost, st), "hipMemcpyAsync")) return rc;
    if (int rc = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize")) return rc;
    std::memcpy(out, &h, sizeof(double));
    return CMM_OK;
}

// the split psi solve in the two parts of op_psi_sample / op_psi_spectral
int op_psi_sample32(const cmm_grid &gm, const cmm_grid &gv, const cmm_grid &gf, const cmm_grid &gp, const float *M,
                    const float *W4, int molly, double fcut, void *work, size_t work_bytes, hipStream_t st,
                    hipEvent_t *ev, unsigned long long *chk_slot) {
    if (int rc = fft_grid_ok(gv, "vort grid")) return rc;
    if (int rc = fft_grid_ok(gp, "psi grid")) return rc;
    SpecPlan sp = plan_spec(gv, gp, 1, fcut);
    const size_t base = sp.t_bytes + 2 * sp.g_bytes + 256;
    void *w;
    if (int rc = get_work(work, work_bytes, base + sizeof(double) * (size_t)gv.nx * gv.ny, &w)) return rc;
    if (chk_slot && !op_check_fusable(gm, gv, gv)) {
        set_error("psi sampler: the check cannot run in the sampler's pass on these grids");
        return CMM_ERR_ARG;
    }
    // the sampled vorticity in FP32: its rounding (2^-24 relative) reaches psi through 1/|k|^2
    float *f = reinterpret_cast<float *>(static_cast<char *>(w) + base);
    const long lanes = (long)gv.nx * (gv.ny / kSampleRows);
    const dim3 grid((unsigned)((lanes + 255) / 256));
    SampleVort32 smp{make_dgrid(gm), make_dgrid(gv), make_dgrid(gf), (const float4 *)M, (const float4 *)W4, molly};
    if (ev) (void)hipEventRecord(ev[0], st);
    if (chk_slot) {
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, chk_slot);
        hipLaunchKernelGGL((k_sample_field<SampleVort32, float, kSampleRows, true>), grid, dim3(256), 0, st, smp, gv.nx,
                           gv.ny, f, chk_slot);
    } else {
        hipLaunchKernelGGL((k_sample_field<SampleVort32, float>), grid, dim3(256), 0, st, smp, gv.nx, gv.ny, f, nullptr);
    }
    return launch_check("k_sample_field");
}

int op_psi_spectral32(const cmm_grid &gv, const cmm_grid &gp, double fcut, double *psi, void *work, size_t work_bytes,
                      hipStream_t st, hipEvent_t *ev, const PsiPre *pre) {
    SpecPlan sp = plan_spec(gv, gp, 1, fcut);
    const size_t base = sp.t_bytes + 2 * sp.g_bytes + 256;
    void *w;
    if (int rc = get_work(work, work_bytes, base + sizeof(double) * (size_t)gv.nx * gv.ny, &w)) return rc;
    float *f = reinterpret_cast<float *>(static_cast<char *>(w) + base);
    return run_spec(sp, SampleRealF{f, gv.nx}, psi, w, st, ev, false, nullptr, pre);
}

int op_psi_from_vort32(const cmm_grid &gm, const cmm_grid &gv, const cmm_grid &gf, const cmm_grid &gp, const float *M,
                       const float *W4, int molly, double fcut, double *psi, void *work, size_t work_bytes,
                       hipStream_t st, hipEvent_t *ev, const PsiPre *pre) {
    if (int rc = op_psi_sample32(gm, gv, gf, gp, M, W4, molly, fcut, work, work_bytes, st, ev, nullptr)) return rc;
    return op_psi_spectral32(gv, gp, fcut, psi, work, work_bytes, st, ev, pre);
}

int op_sample_vort32(const cmm_grid &gm, const cmm_grid &gt, const cmm_grid &gf, const float *M, const float *W4,
                     int molly, double *out, hipStream_t st) {
    SampleVort32 s{make_dgrid(gm), make_dgrid(gt), make_dgrid(gf), (const float4 *)M, (const float4 *)W4, molly};
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_sample_grid<SampleVort32>, grid2d(gt.nx, gt.ny, b), b, 0, st, s, gt.nx, gt.ny, out);
    return launch_check("k_sample_grid");
}

// ------------------------------------------------------------------------------------------------
// Ensembles: grouped launches.  One launch per kind of work covers every instance of a
// Vlasov-Poisson ensemble: its workgroups walk the concatenated blocks of all instances (the blocks
// of instance j are items start[j] .. start[j+1] - 1), each workgroup a contiguous range of items,
// the ranges dealt to the XCDs in contiguous runs (xcd_block).  With a capped grid the launch is
// persistent (each workgroup takes several items); with one workgroup per item it is a plain
// grouped launch.  Each item runs the block body of the single-instance kernel with that
// instance's arguments, so every instance computes what its own launches compute, bit for bit,
// whatever the grid.

struct GroupMap {
    int n;
    int start[kGroupMax + 1];
};


// the instance of an item; the same for the whole workgroup, which readfirstlane states, so that
// the instance's arguments are read with scalar loads
__device__ __forceinline__ int group_find(const GroupMap &g, int item) {
    int j = 0;
    while (j + 1 < g.n && g.start[j + 1] <= item) ++j;
    return __builtin_amdgcn_readfirstlane(j);
}

// the items [lo, hi) of this workgroup
__device__ __forceinline__ void group_items(int total, int &lo, int &hi) {
    const int per = (total + (int)gridDim.x - 1) / (int)gridDim.x;
    lo = xcd_block(blockIdx.x, gridDim.x) * per;
    hi = min(total, lo + per);
}

struct GAdvVP {
    DGrid gm, gp;
    const double *T, *S, *H, *M8;
    cmm_advect_coeffs c;
    double *ncx, *ncy, *M8out;
};
struct GComb {
    int np;
    long N;
    PsiLevels P;
    cmm_advect_coeffs c;
    double *T, *S;
};
struct GChk {
    DGrid gm, gc;
    const double *cx, *cy;
    unsigned long long *slot;
    int ntx, nblocks;
    long ntiles;
};
struct GSample {
    SampleVortIL smp;
    int nx, ny;
    double *f;
};
struct GColsum {
    const double *f;
    int nx, ny, nbx;
    double *part;
};
struct GPlanes {
    const double *phi, *H;
    int np, lognp;
    double *out;
};

template <int L>
__global__ void __launch_bounds__(256) k_grp_vp_combine1d(const GComb *__restrict__ jobs, GroupMap g) {
    int lo, hi;
    group_items(g.start[g.n], lo, hi);
    for (int it = lo; it < hi; ++it) {
        const int j = group_find(g, it);
        const GComb &a = jobs[j];
        vp_combine1d_at<L>((it - g.start[j]) * 256 + threadIdx.x, a.np, a.N, a.P, a.c, nullptr, a.T, a.S);
    }
}

template <int TIN>
__global__ void __launch_bounds__(256) k_grp_advect_vp(const GAdvVP *__restrict__ jobs, GroupMap g) {
    int lo, hi;
    group_items(g.start[g.n], lo, hi);
    for (int it = lo; it < hi; ++it) {
        const int j = group_find(g, it);
        const GAdvVP &a = jobs[j];
        advect_vp_block<TIN>(it - g.start[j], a.gm, a.gp, a.T, a.S, a.H, a.M8, a.c, a.ncx, a.ncy, a.M8out);
    }
}

__global__ void k_grp_zero(unsigned long long *slots, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) slots[i] = 0ull;
}

template <int TX, int TY>
__global__ void __launch_bounds__(256) k_grp_incomp(const GChk *__restrict__ jobs, GroupMap g) {
    __shared__ double red[4];
    int lo, hi;
    group_items(g.start[g.n], lo, hi);
    for (int it = lo; it < hi; ++it) {
        const int j = group_find(g, it);
        const GChk &a = jobs[j];
        long t_begin, t_end;
        tile_range_of(a.ntiles, it - g.start[j], a.nblocks, t_begin, t_end);
        const double v = incomp_tensor_tiles<TX, TY>(a.gm, a.gc, a.cx, a.cy, a.ntx, t_begin, t_end);
        block_max_to(v, red, a.slot);
        __syncthreads();  // red is written again by the next item
    }
}

template <int R>
__global__ void __launch_bounds__(256, 2) k_grp_sample(const GSample *__restrict__ jobs, GroupMap g) {
    int lo, hi;
    group_items(g.start[g.n], lo, hi);
    for (int it = lo; it < hi; ++it) {
        const int j = group_find(g, it);
        const SampleVortIL smp = jobs[j].smp;
        const int nx = jobs[j].nx, ny = jobs[j].ny;
        double *f = jobs[j].f;
        const long t = (long)(it - g.start[j]) * 256 + threadIdx.x;
        if (t < (long)nx * (ny / R)) {
            const int x = (int)(t % nx), y0 = R * (int)(t / nx);
            double v[R];
            smp.template column<R>(x, y0, v);
#pragma unroll
            for (int r = 0; r < R; ++r) f[(long)(y0 + r) * nx + x] = v[r];
        }
    }
}

__global__ void __launch_bounds__(256) k_grp_colsum(const GColsum *__restrict__ jobs, GroupMap g) {
    int lo, hi;
    group_items(g.start[g.n], lo, hi);
    for (int it = lo; it < hi; ++it) {
        const int j = group_find(g, it);
        const GColsum &a = jobs[j];
        const int lb = it - g.start[j];
        const int x = (lb % a.nbx) * 256 + threadIdx.x, c = lb / a.nbx;
        if (x >= a.nx) continue;
        const int y0 = c * kVpRows, y1 = min(a.ny, y0 + kVpRows);
        double s = 0.0;
        for (int y = y0; y < y1; ++y) s += a.f[(long)y * a.nx + x];
        a.part[(long)c * a.nx + x] = s;
    }
}

// one workgroup per instance
__global__ void __launch_bounds__(kFftThreads) k_grp_vp_phi(const VPJob *__restrict__ jobs) {
    extern __shared__ cd s[];
    vp_phi_block(jobs[blockIdx.x], s);
}

__global__ void __launch_bounds__(256) k_grp_psi_planes(const GPlanes *__restrict__ jobs, GroupMap g) {
    int lo, hi;
    group_items(g.start[g.n], lo, hi);
    for (int it = lo; it < hi; ++it) {
        const int j = group_find(g, it);
        const GPlanes &a = jobs[j];
        const long N = (long)a.np * a.np;
        const long t = (long)(it - g.start[j]) * 256 + threadIdx.x;
        if (t >= N) continue;
        const int x = (int)(t & (a.np - 1)), y = (int)(t >> a.lognp);
        a.out[t] = a.phi[x] - a.H[y];
        a.out[N + t] = a.phi[2 * a.np + x];
        a.out[2 * N + t] = -a.H[2 * a.np + y];
        a.out[3 * N + t] = 0.0;
    }
}

static size_t group_align(size_t b) { return (b + 255) / 256 * 256; }

size_t op_group_vp_job_bytes(int n) {
    return group_align(sizeof(GAdvVP) * n) + group_align(sizeof(GComb) * n) + group_align(sizeof(GChk) * n) +
           group_align(sizeof(GSample) * n) + group_align(sizeof(GColsum) * n) + group_align(sizeof(VPJob) * n) +
           group_align(sizeof(GPlanes) * n);
}

// the sampler's rows per lane and the check's tile edge of one instance (0 where its grids do not
// take the column sampler or the tensor check)
static int group_rows(const GroupInstance &a) {
    const int r = a.gv.ny / a.gm.ny;
    if (!a.W4 || a.gv.ny != r * a.gm.ny || r < 2 || a.gv.x0 != a.gm.x0 || a.gv.y0 != a.gm.y0) return 0;
    return r >= kSampleRows ? kSampleRows : 2;
}

static int group_tile(const GroupInstance &a) {
    const int r = a.gc.nx / a.gm.nx;
    if (!(r >= 1 && a.gc.nx == r * a.gm.nx && a.gc.ny == r * a.gm.ny && a.gc.x0 == a.gm.x0 && a.gc.x1 == a.gm.x1 &&
          a.gc.y0 == a.gm.y0 && a.gc.y1 == a.gm.y1))
        return 0;
    return incomp_tile(r, a.tune);
}

// The grouped check runs every instance with one tile edge, and the tile kernels round the
// determinant near 1 differently (about 1e-14), so the instances' own tiles must agree for each to
// get its own launch's value.
bool op_group_vp_alike(const GroupInstance *inst, int n) {
    if (n < 1 || n > kGroupMax) return false;
    const int R = group_rows(inst[0]), TT = group_tile(inst[0]);
    if (R == 0 || TT == 0) return false;
    for (int i = 0; i < n; ++i) {
        const GroupInstance &a = inst[i];
        if (group_rows(a) != R || group_tile(a) != TT) return false;
        if (a.c.time_integration_num != inst[0].c.time_integration_num || a.c.l_order != inst[0].c.l_order)
            return false;
    }
    return true;
}

bool op_group_vp_eligible(const GroupInstance *inst, int n) {
    if (!op_group_vp_alike(inst, n)) return false;
    for (int i = 0; i < n; ++i) {
        const GroupInstance &a = inst[i];
        if (a.parts != 7 || !a.m8_in || !a.H || !op_vp_advect_supported(a.gm, a.gp, a.c) || vp_check(a.gv, a.gp))
            return false;
        // the workspaces of op_map_advect_vp (with its interleaved input map) and op_vp_psi_from_f
        if (a.adv_work_bytes < sizeof(double) * (6 * (size_t)a.gp.nx + 2) ||
            a.psi_work_bytes < op_vp_psi_workspace(a.gv, a.gp))
            return false;
    }
    set_error("");
    return true;
}

int op_group_vp_step(const GroupInstance *inst, int n, void *dev_jobs, void *host_jobs,
                     unsigned long long *dev_slots, unsigned long long *host_slots, hipStream_t st, int max_blocks) {
    if (!op_group_vp_eligible(inst, n)) {
        set_error("grouped ensemble step: the instances differ in scheme, order or grid ratios");
        return CMM_ERR_ARG;
    }
    char *h = static_cast<char *>(host_jobs);
    GAdvVP *adv = reinterpret_cast<GAdvVP *>(h);
    GComb *comb = reinterpret_cast<GComb *>(h + group_align(sizeof(GAdvVP) * n));
    GChk *chk = reinterpret_cast<GChk *>(reinterpret_cast<char *>(comb) + group_align(sizeof(GComb) * n));
    GSample *smp = reinterpret_cast<GSample *>(reinterpret_cast<char *>(chk) + group_align(sizeof(GChk) * n));
    GColsum *col = reinterpret_cast<GColsum *>(reinterpret_cast<char *>(smp) + group_align(sizeof(GSample) * n));
    VPJob *phi = reinterpret_cast<VPJob *>(reinterpret_cast<char *>(col) + group_align(sizeof(GColsum) * n));
    GPlanes *pl = reinterpret_cast<GPlanes *>(reinterpret_cast<char *>(phi) + group_align(sizeof(VPJob) * n));
    GroupMap g_comb{n, {}}, g_adv{n, {}}, g_chk{n, {}}, g_smp{n, {}}, g_col{n, {}}, g_pl{n, {}};
    const int R = group_rows(inst[0]), TT = group_tile(inst[0]);
    size_t lds_phi = 0;
    for (int i = 0; i < n; ++i) {
        const GroupInstance &a = inst[i];
        const DGrid dm = make_dgrid(a.gm), dp = make_dgrid(a.gp), dc = make_dgrid(a.gc);
        // op_map_advect_vp's tables in its workspace
        double *T = static_cast<double *>(a.adv_work), *S = T + 6 * a.gp.nx;
        comb[i] = GComb{a.gp.nx, (long)dp.n, a.P, a.c, T, S};
        g_comb.start[i + 1] = g_comb.start[i] + (a.gp.nx + 255) / 256;
        adv[i] = GAdvVP{dm, dp, T, S, a.H, a.m8_in, a.c, a.ncx, a.ncy, a.m8_out};
        g_adv.start[i + 1] = g_adv.start[i] + (int)(((long)dm.n * 4 + 255) / 256);
        // op_incomp_launch's tile split
        const int ntx = a.gc.nx / TT;
        const long ntiles = (long)ntx * (a.gc.ny / TT);
        const int nb = (int)std::max<long>(1, std::min<long>((ntiles + 255) / 256, a.tune.max_blocks));
        chk[i] = GChk{dm, dc, a.ccx, a.ccy, dev_slots + i, ntx, nb, ntiles};
        g_chk.start[i + 1] = g_chk.start[i] + nb;
        // op_vp_psi_from_f's workspace: f, the chunk sums, phi
        const VP1D q = vp_params(a.gv, a.gp);
        double *f = static_cast<double *>(a.psi_work);
        double *part = f + (size_t)a.gv.nx * a.gv.ny;
        double *phiH = part + (size_t)q.nchunks * a.gv.nx;
        SampleVortIL sv;
        sv.gm = dm;
        sv.gv = make_dgrid(a.gv);
        sv.gf = make_dgrid(a.gf);
        sv.cx = a.pcx;
        sv.cy = a.pcy;
        sv.W = a.W4;
        sv.molly = 0;
        smp[i] = GSample{sv, a.gv.nx, a.gv.ny, f};
        g_smp.start[i + 1] = g_smp.start[i] + (int)(((long)a.gv.nx * (a.gv.ny / R) + 255) / 256);
        const int nbx = (a.gv.nx + 255) / 256;
        col[i] = GColsum{f, a.gv.nx, a.gv.ny, nbx, part};
        g_col.start[i + 1] = g_col.start[i] + nbx * q.nchunks;
        const cd *tw_v, *tw_p;
        if (int rc = twiddles(a.gv.nx, &tw_v)) return rc;
        if (int rc = twiddles(a.gp.nx, &tw_p)) return rc;
        phi[i] = VPJob{q, f, part, tw_v, tw_p, phiH};
        lds_phi = std::max(lds_phi, sizeof(cd) * std::max(a.gv.nx, a.gp.nx));
        pl[i] = GPlanes{phiH, a.H, a.gp.nx, ilog2(a.gp.nx), a.psi};
        g_pl.start[i + 1] = g_pl.start[i] + (int)(((long)a.gp.nx * a.gp.ny + 255) / 256);
    }
    const size_t bytes = op_group_vp_job_bytes(n);
    if (int rc = hip_check(hipMemcpyAsync(dev_jobs, host_jobs, bytes, hipMemcpyHostToDevice, st), "hipMemcpyAsync (jobs)"))
        return rc;
    char *d = static_cast<char *>(dev_jobs);
    auto dev = [&](const void *hp) { return d + (static_cast<const char *>(hp) - h); };
    // max_blocks > 0: persistent launches of at most that many workgroups; else one per item
    auto grid = [&](const GroupMap &g) {
        return dim3((unsigned)std::max(1, max_blocks > 0 ? std::min(g.start[g.n], max_blocks) : g.start[g.n]));
    };
    switch (inst[0].c.l_order) {
    case 2: hipLaunchKernelGGL(k_grp_vp_combine1d<2>, grid(g_comb), dim3(256), 0, st, (const GComb *)dev(comb), g_comb); break;
    case 3: hipLaunchKernelGGL(k_grp_vp_combine1d<3>, grid(g_comb), dim3(256), 0, st, (const GComb *)dev(comb), g_comb); break;
    default: hipLaunchKernelGGL(k_grp_vp_combine1d<4>, grid(g_comb), dim3(256), 0, st, (const GComb *)dev(comb), g_comb); break;
    }
    if (int rc = launch_check("k_grp_vp_combine1d")) return rc;
    const GAdvVP *da = (const GAdvVP *)dev(adv);
    switch (inst[0].c.time_integration_num) {
    case 10: hipLaunchKernelGGL(k_grp_advect_vp<10>, grid(g_adv), dim3(256), 0, st, da, g_adv); break;
    case 21: hipLaunchKernelGGL(k_grp_advect_vp<21>, grid(g_adv), dim3(256), 0, st, da, g_adv); break;
    case 30: hipLaunchKernelGGL(k_grp_advect_vp<30>, grid(g_adv), dim3(256), 0, st, da, g_adv); break;
    case 40: hipLaunchKernelGGL(k_grp_advect_vp<40>, grid(g_adv), dim3(256), 0, st, da, g_adv); break;
    case 41: hipLaunchKernelGGL(k_grp_advect_vp<41>, grid(g_adv), dim3(256), 0, st, da, g_adv); break;
    default: set_error("unknown time integration"); return CMM_ERR_ARG;
    }
    if (int rc = launch_check("k_grp_advect_vp")) return rc;
    hipLaunchKernelGGL(k_grp_zero, dim3((n + 63) / 64), dim3(64), 0, st, dev_slots, n);
    const GChk *dc = (const GChk *)dev(chk);
    if (TT == 4) hipLaunchKernelGGL((k_grp_incomp<4, 4>), grid(g_chk), dim3(256), 0, st, dc, g_chk);
    else if (TT == 3) hipLaunchKernelGGL((k_grp_incomp<3, 3>), grid(g_chk), dim3(256), 0, st, dc, g_chk);
    else if (TT == 2) hipLaunchKernelGGL((k_grp_incomp<2, 2>), grid(g_chk), dim3(256), 0, st, dc, g_chk);
    else hipLaunchKernelGGL((k_grp_incomp<1, 1>), grid(g_chk), dim3(256), 0, st, dc, g_chk);
    if (int rc = launch_check("k_grp_incomp")) return rc;
    if (int rc = hip_check(hipMemcpyAsync(host_slots, dev_slots, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost, st),
                           "hipMemcpyAsync (E_inc)"))
        return rc;
    const GSample *ds = (const GSample *)dev(smp);
    if (R == kSampleRows) hipLaunchKernelGGL((k_grp_sample<kSampleRows>), grid(g_smp), dim3(256), 0, st, ds, g_smp);
    else hipLaunchKernelGGL((k_grp_sample<2>), grid(g_smp), dim3(256), 0, st, ds, g_smp);
    if (int rc = launch_check("k_grp_sample")) return rc;
    hipLaunchKernelGGL(k_grp_colsum, grid(g_col), dim3(256), 0, st, (const GColsum *)dev(col), g_col);
    if (int rc = launch_check("k_grp_colsum")) return rc;
    if (int rc = set_lds((const void *)k_grp_vp_phi, lds_phi)) return rc;
    hipLaunchKernelGGL(k_grp_vp_phi, dim3(n), dim3(kFftThreads), lds_phi, st, (const VPJob *)dev(phi));
    if (int rc = launch_check("k_grp_vp_phi")) return rc;
    hipLaunchKernelGGL(k_grp_psi_planes, grid(g_pl), dim3(256), 0, st, (const GPlanes *)dev(pl), g_pl);
    return launch_check("k_grp_psi_planes");
}

}  // namespace cmm
