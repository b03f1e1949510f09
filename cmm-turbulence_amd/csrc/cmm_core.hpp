// Device-side core of the CMM engine: grid descriptor, periodic cell lookup, cubic Hermite
// evaluation and the workgroup FFT building blocks.  Everything here is __host__ __device__
// so the index arithmetic can also be exercised by host-side unit tests.
#pragma once

#include <stdint.h>
#include <math.h>
#include "../../include/cmm_types.h"

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define CMM_HD __host__ __device__ __forceinline__
#else
#define CMM_HD inline
#endif

namespace cmm {

constexpr double kPi = 3.141592653589793238462643383279502884197169399375105820974944592307816406286;
constexpr double kTwoPi = 6.283185307179586476925286766559005768394338798750211641949889184615632812572;

// Grid as the kernels see it.  h = (x1-x0)/(float)nx as in Cuda-CMM/src/grid/cmm-grid2d.cu:27-28;
// jx/jy are the jumps nx*h added to a map's value plane on a periodic wrap
// (Cuda-CMM/src/numerical/cmm-hermite.cu:450), lx/ly the domain lengths used to wrap positions
// (Cuda-CMM/src/simulation/cmm-simulation-kernel.cu:316-320).  ihx/ihy = 1/h; lgx/lgy = log2 of
// the point counts when they are powers of two (else -1).
struct DGrid {
    int nx, ny, n, lgx, lgy;
    double hx, hy, ihx, ihy, x0, y0, jx, jy, lx, ly;
};

inline int log2_or_neg(int n) {
    if (n <= 0 || (n & (n - 1))) return -1;
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

inline DGrid make_dgrid(const cmm_grid &g) {
    DGrid d;
    d.nx = g.nx; d.ny = g.ny; d.n = g.nx * g.ny;
    d.lgx = log2_or_neg(g.nx); d.lgy = log2_or_neg(g.ny);
    d.hx = (g.x1 - g.x0) / (float)g.nx;
    d.hy = (g.y1 - g.y0) / (float)g.ny;
    d.ihx = 1.0 / d.hx; d.ihy = 1.0 / d.hy;
    d.x0 = g.x0; d.y0 = g.y0;
    d.jx = g.nx * d.hx; d.jy = g.ny * d.hy;
    d.lx = g.x1 - g.x0; d.ly = g.y1 - g.y0;
    return d;
}

// Four corner indices of the cell holding (x,y), fractional offsets, and the periodic wrap counts
// of the corners.  Cuda-CMM/src/numerical/cmm-hermite.cu:152-164 / :191-208.  The wrap is a true
// floor division (identical to the reference except at exact negative multiples of the grid
// size, where the reference indexes outside the array); the floor is clamped to +-1e9 cells so
// that non-finite positions still address memory inside the grid.  The cell coordinate is
// (x - x0) * (1/h): it can differ from the reference's (x - x0) / h in the last bit, which moves a
// point lying on a cell edge to the neighbouring cell, where the C1 Hermite interpolant takes the
// same value.
struct Cell {
    int i00, i10, i01, i11;
    int wx0, wx1, wy0, wy1;
    double dx, dy;
};

CMM_HD int floordiv(int a, int n) { return a >= 0 ? a / n : -((-a - 1) / n) - 1; }
CMM_HD int floordiv_g(int a, int n, int lg) { return lg >= 0 ? (a >> lg) : floordiv(a, n); }

// P2: the grid's point counts are known to be powers of two (lgx, lgy >= 0), so the wrap counts
// are shifts, with no branch on the grid
template <bool P2 = false>
CMM_HD void find_cell(const DGrid &g, double x, double y, Cell &c) {
    double fx = (x - g.x0) * g.ihx, fy = (y - g.y0) * g.ihy;
    double flx = floor(fx), fly = floor(fy);
    c.dx = fx - flx;
    c.dy = fy - fly;
    flx = fmin(fmax(flx, -1.0e9), 1.0e9);
    fly = fmin(fmax(fly, -1.0e9), 1.0e9);
    int ix = (int)flx, iy = (int)fly;
    if constexpr (P2) {
        c.wx0 = ix >> g.lgx;
        c.wy0 = iy >> g.lgy;
    } else {
        c.wx0 = floordiv_g(ix, g.nx, g.lgx);
        c.wy0 = floordiv_g(iy, g.ny, g.lgy);
    }
    int rx0 = ix - c.wx0 * g.nx, ry0 = iy - c.wy0 * g.ny;
    int rx1 = rx0 + 1, ry1 = ry0 + 1;
    c.wx1 = c.wx0; c.wy1 = c.wy0;
    if (rx1 == g.nx) { rx1 = 0; c.wx1++; }
    if (ry1 == g.ny) { ry1 = 0; c.wy1++; }
    c.i00 = ry0 * g.nx + rx0;
    c.i10 = ry0 * g.nx + rx1;
    c.i01 = ry1 * g.nx + rx0;
    c.i11 = ry1 * g.nx + rx1;
}

// Cubic Hermite basis in the f3/f4 form of Cuda-CMM/src/numerical/cmm-hermite.h:27-30:
// b[0], b[1] weight the corner values, b[2], b[3] the corner derivatives (unit cell).
CMM_HD void basis(double d, double b[4]) {
    double e = 1 - d;
    b[0] = (3 - 2 * e) * e * e;
    b[1] = (3 - 2 * d) * d * d;
    b[2] = -((e - 1) * e * e);
    b[3] = (d - 1) * d * d;
}
CMM_HD void basis_d(double d, double b[4]) {
    double e = 1 - d;
    b[0] = -((6 - 6 * e) * e);
    b[1] = (6 - 6 * d) * d;
    b[2] = (3 * e - 2) * e;
    b[3] = (3 * d - 2) * d;
}

// The 16 values a cell evaluation needs: planes [f, fx, fy, fxy] at corners [00, 10, 01, 11].
CMM_HD void load16(const double *__restrict__ H, int N, const Cell &c, double v[16]) {
    const long n = N;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const double *Hp = H + p * n;
        v[4 * p + 0] = Hp[c.i00];
        v[4 * p + 1] = Hp[c.i10];
        v[4 * p + 2] = Hp[c.i01];
        v[4 * p + 3] = Hp[c.i11];
    }
}

// Tensor-product sum over the planes of one Hermite field (Cuda-CMM/src/numerical/cmm-hermite.cu:31-45);
// derivative planes are physical, so they are scaled by h.  wjump adds the wrap jump of a map's
// value plane per corner (Cuda-CMM/src/numerical/cmm-hermite.cu:67-90).
template <bool WARP>
CMM_HD double hsum_t(const double v[16], const double bx[4], const double by[4], double hx, double hy,
                     const double *wj) {
    double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
    if constexpr (WARP) {
        t0 += by[0] * bx[0] * (v[0] + wj[0]); t0 += by[0] * bx[1] * (v[1] + wj[1]);
        t0 += by[1] * bx[0] * (v[2] + wj[2]); t0 += by[1] * bx[1] * (v[3] + wj[3]);
    } else {
        t0 += by[0] * bx[0] * v[0]; t0 += by[0] * bx[1] * v[1];
        t0 += by[1] * bx[0] * v[2]; t0 += by[1] * bx[1] * v[3];
    }
    t1 += by[0] * bx[2] * v[4]; t1 += by[0] * bx[3] * v[5];
    t1 += by[1] * bx[2] * v[6]; t1 += by[1] * bx[3] * v[7];
    t2 += by[2] * bx[0] * v[8]; t2 += by[2] * bx[1] * v[9];
    t2 += by[3] * bx[0] * v[10]; t2 += by[3] * bx[1] * v[11];
    t3 += by[2] * bx[2] * v[12]; t3 += by[2] * bx[3] * v[13];
    t3 += by[3] * bx[2] * v[14]; t3 += by[3] * bx[3] * v[15];
    double out = t0;
    out += t1 * hx;
    out += t2 * hy;
    t3 *= hx;
    t3 *= hy;
    out += t3;
    return out;
}

CMM_HD double hsum_v(const double v[16], const double bx[4], const double by[4], double hx, double hy) {
    return hsum_t<false>(v, bx, by, hx, hy, nullptr);
}
CMM_HD double hsum_v(const double v[16], const double bx[4], const double by[4], double hx, double hy,
                     const double wj[4]) {
    return hsum_t<true>(v, bx, by, hx, hy, wj);
}

// Value of a Hermite field, Cuda-CMM/src/numerical/cmm-hermite.cu:323-331
CMM_HD double herm_eval(const DGrid &g, const double *__restrict__ H, double x, double y) {
    Cell c;
    find_cell(g, x, y, c);
    double bx[4], by[4], v[16];
    basis(c.dx, bx);
    basis(c.dy, by);
    load16(H, g.n, c, v);
    return hsum_v(v, bx, by, g.hx, g.hy);
}

// load16 of a field stored node-interleaved (H4[4*i + plane]): two 16-byte loads per corner
CMM_HD void load16_il(const double *__restrict__ H4, const Cell &c, double v[16]) {
    const int idx[4] = {c.i00, c.i10, c.i01, c.i11};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double2 *r = reinterpret_cast<const double2 *>(H4 + 4 * (long)idx[q]);
        const double2 a = r[0], b = r[1];
        v[q] = a.x;
        v[4 + q] = a.y;
        v[8 + q] = b.x;
        v[12 + q] = b.y;
    }
}

// herm_eval at a point whose cell c has been found, with the corner values given
CMM_HD double herm_eval_vals(const DGrid &g, const Cell &c, const double v[16]) {
    double bx[4], by[4];
    basis(c.dx, bx);
    basis(c.dy, by);
    return hsum_v(v, bx, by, g.hx, g.hy);
}

// herm_eval of a field stored node-interleaved (H4[4*i + plane]); same arithmetic as herm_eval
CMM_HD double herm_eval_il(const DGrid &g, const double *__restrict__ H4, double x, double y) {
    Cell c;
    find_cell(g, x, y, c);
    double v[16];
    load16_il(H4, c, v);
    return herm_eval_vals(g, c, v);
}

// Velocity (d psi/dy, -d psi/dx) of up to four psi levels at one point,
// Cuda-CMM/src/numerical/cmm-hermite.cu:396-420.  u[2l] = dpsi_l/dy, u[2l+1] = -dpsi_l/dx.
struct PsiLevels {
    const double *p[4];
};

// The next step's Lagrange-combined psi fields written by the psi solve's last transform
// (k_row_c2r4_pre): P.p[l], l = 1 .. L - 1, the older levels (level 0 is the transform's output),
// the weights of C1 (w1) and C12 (w12), and the fields (C12, C0 may be null); form 0: FP64
// records, 1: packed FP32 records.
struct PsiPre {
    PsiLevels P;
    int L;
    int form;
    double w1[4], w12[4];
    void *C1, *C12, *C0;
};

CMM_HD void herm_vel(const DGrid &g, const PsiLevels &P, int nl, double x, double y, double *u) {
    Cell c;
    find_cell(g, x, y, c);
    double bx[4], by[4], dbx[4], dby[4];
    basis(c.dx, bx);
    basis(c.dy, by);
    basis_d(c.dx, dbx);
    basis_d(c.dy, dby);
    // unrolled to the 4 levels a PsiLevels holds, so that u is written at constant indices (an
    // array indexed at run time would go to private memory)
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        if (l >= nl) break;
        double v[16];
        load16(P.p[l], g.n, c, v);
        u[2 * l + 1] = -hsum_v(v, dbx, by, g.hx, g.hy) * g.ihx;
        u[2 * l] = hsum_v(v, bx, dby, g.hx, g.hy) * g.ihy;
    }
}

// herm_vel of one field at a point whose cell c has been found, with the corner values given
CMM_HD void herm_vel_vals(const DGrid &g, const Cell &c, const double v[16], double &ux, double &uy) {
    double bx[4], by[4], dbx[4], dby[4];
    basis(c.dx, bx);
    basis(c.dy, by);
    basis_d(c.dx, dbx);
    basis_d(c.dy, dby);
    uy = -hsum_v(v, dbx, by, g.hx, g.hy) * g.ihx;
    ux = hsum_v(v, bx, dby, g.hx, g.hy) * g.ihy;
}

// herm_vel of one field stored node-interleaved (H4[4*i + plane]): the four planes of a corner
// are one 32-byte record read with two 16-byte loads; same arithmetic as herm_vel.
template <bool P2 = false>
CMM_HD void herm_vel_il(const DGrid &g, const double *__restrict__ H4, double x, double y, double &ux, double &uy) {
    Cell c;
    find_cell<P2>(g, x, y, c);
    double v[16];
    load16_il(H4, c, v);
    herm_vel_vals(g, c, v, ux, uy);
}

// Map evaluation with periodic jumps, Cuda-CMM/src/numerical/cmm-hermite.cu:446-460
CMM_HD void diffeo(const DGrid &g, const double *__restrict__ Hx, const double *__restrict__ Hy, double x, double y,
                   double &xo, double &yo) {
    Cell c;
    find_cell(g, x, y, c);
    double bx[4], by[4], v[16];
    basis(c.dx, bx);
    basis(c.dy, by);
    const double wjx[4] = {c.wx0 * g.jx, c.wx1 * g.jx, c.wx0 * g.jx, c.wx1 * g.jx};
    const double wjy[4] = {c.wy0 * g.jy, c.wy0 * g.jy, c.wy1 * g.jy, c.wy1 * g.jy};
    load16(Hx, g.n, c, v);
    xo = hsum_v(v, bx, by, g.hx, g.hy, wjx);
    load16(Hy, g.n, c, v);
    yo = hsum_v(v, bx, by, g.hx, g.hy, wjy);
}

// diffeo() at a point whose cell c has been found, with the corner values of both fields given
CMM_HD void diffeo_vals(const DGrid &g, const Cell &c, const double vx[16], const double vy[16], double &xo,
                        double &yo) {
    double bx[4], by[4];
    basis(c.dx, bx);
    basis(c.dy, by);
    const double wjx[4] = {c.wx0 * g.jx, c.wx1 * g.jx, c.wx0 * g.jx, c.wx1 * g.jx};
    const double wjy[4] = {c.wy0 * g.jy, c.wy0 * g.jy, c.wy1 * g.jy, c.wy1 * g.jy};
    xo = hsum_v(vx, bx, by, g.hx, g.hy, wjx);
    yo = hsum_v(vy, bx, by, g.hx, g.hy, wjy);
}

CMM_HD bool same_cell(const Cell &a, const Cell &b) { return a.i00 == b.i00 && a.wx0 == b.wx0 && a.wy0 == b.wy0; }

// the corner values of a map stored node-interleaved (M8[8*i + p]: planes of X, then of Y), each
// corner one 64-byte record read with four 16-byte loads
CMM_HD void load_map_il(const double *__restrict__ M8, const Cell &c, double vx[16], double vy[16]) {
    const int idx[4] = {c.i00, c.i10, c.i01, c.i11};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double2 *r = reinterpret_cast<const double2 *>(M8 + 8 * (long)idx[q]);
        const double2 a = r[0], b = r[1], d = r[2], e = r[3];
        vx[q] = a.x;
        vx[4 + q] = a.y;
        vx[8 + q] = b.x;
        vx[12 + q] = b.y;
        vy[q] = d.x;
        vy[4 + q] = d.y;
        vy[8 + q] = e.x;
        vy[12 + q] = e.y;
    }
}

// diffeo() on a map stored node-interleaved; same arithmetic as diffeo
template <bool P2 = false>
CMM_HD void diffeo_il(const DGrid &g, const double *__restrict__ M8, double x, double y, double &xo, double &yo) {
    Cell c;
    find_cell<P2>(g, x, y, c);
    double vx[16], vy[16];
    load_map_il(M8, c, vx, vy);
    diffeo_vals(g, c, vx, vy, xo, yo);
}

// det(grad X), Cuda-CMM/src/numerical/cmm-hermite.cu:479-503
CMM_HD double diffeo_det(const DGrid &g, const double *__restrict__ Hx, const double *__restrict__ Hy, double x,
                         double y) {
    Cell c;
    find_cell(g, x, y, c);
    double bx[4], by[4], dbx[4], dby[4], vx[16], vy[16];
    basis(c.dx, bx);
    basis(c.dy, by);
    basis_d(c.dx, dbx);
    basis_d(c.dy, dby);
    const double wjx[4] = {c.wx0 * g.jx, c.wx1 * g.jx, c.wx0 * g.jx, c.wx1 * g.jx};
    const double wjy[4] = {c.wy0 * g.jy, c.wy0 * g.jy, c.wy1 * g.jy, c.wy1 * g.jy};
    load16(Hx, g.n, c, vx);
    load16(Hy, g.n, c, vy);
    const double Xx = hsum_v(vx, dbx, by, g.hx, g.hy, wjx) * g.ihx;
    const double Yx = hsum_v(vy, dbx, by, g.hx, g.hy, wjy) * g.ihy;
    const double Xy = hsum_v(vx, bx, dby, g.hx, g.hy, wjx) * g.ihx;
    const double Yy = hsum_v(vy, bx, dby, g.hx, g.hy, wjy) * g.ihy;
    return Xx * Yy - Xy * Yx;
}

CMM_HD double wrap(double v, double o, double L) { return v - floor((v - o) / L) * L; }

// ------------------------------------------------------------------------------------------------
// Workgroup FFT building blocks (Stockham autosort, radix 8/4/2, complex double).
// A transform of length n lives in LDS as n complex values at swizzled slots sw(i); each stage
// reads R values at stride n/R, twiddles, applies an R-point DFT and writes them at stride Ns.
struct cd {
    double x, y;
};

CMM_HD cd cmul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
CMM_HD cd cadd(cd a, cd b) { return {a.x + b.x, a.y + b.y}; }
CMM_HD cd csub(cd a, cd b) { return {a.x - b.x, a.y - b.y}; }
// multiply by -i (forward) or +i (inverse)
CMM_HD cd mul_mi(cd a, bool inv) { return inv ? cd{-a.y, a.x} : cd{a.y, -a.x}; }

// XOR swizzle inside aligned groups of 64 slots: spreads stride-8 accesses over the LDS banks.
CMM_HD int sw(int i) { return i ^ ((i >> 3) & 7); }

template <int R>
CMM_HD void dft(cd *v, bool inv);

template <>
CMM_HD void dft<2>(cd *v, bool) {
    cd a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <>
CMM_HD void dft<4>(cd *v, bool inv) {
    cd t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    cd t2 = cadd(v[1], v[3]), t3 = mul_mi(csub(v[1], v[3]), inv);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}

template <>
CMM_HD void dft<8>(cd *v, bool inv) {
    const double r = 0.70710678118654752440084436210484903928;
    cd e[4] = {v[0], v[2], v[4], v[6]};
    cd o[4] = {v[1], v[3], v[5], v[7]};
    dft<4>(e, inv);
    dft<4>(o, inv);
    // o[k] *= exp(-+ 2 pi i k / 8)
    cd w1 = inv ? cd{r, r} : cd{r, -r};
    cd w3 = inv ? cd{-r, r} : cd{-r, -r};
    o[1] = cmul(o[1], w1);
    o[2] = mul_mi(o[2], inv);
    o[3] = cmul(o[3], w3);
    for (int k = 0; k < 4; ++k) {
        v[k] = cadd(e[k], o[k]);
        v[k + 4] = csub(e[k], o[k]);
    }
}

// Radix sequence for a length 2^logn: radix 8 while three or more bits remain, then 4 or 2.
CMM_HD int stage_radix_log(int remaining) { return remaining >= 3 ? 3 : remaining; }

// Load the R inputs of butterfly j (stride n/R).
template <int R>
CMM_HD void stage_load(const cd *s, cd *v, int j, int nb) {
    for (int r = 0; r < R; ++r) v[r] = s[sw(j + r * nb)];
}

// Twiddle, DFT and store butterfly j.  Ns = product of the earlier radices (power of two),
// tw[m] = exp(-2 pi i m / n); the twiddle for input r is W_{Ns R}^{(j mod Ns) r}.
template <int R>
CMM_HD void stage_store(cd *s, cd *v, int j, int logNs, int logn, const cd *tw, bool inv) {
    const int Ns = 1 << logNs;
    const int k = j & (Ns - 1);
    const int rlog = (R == 8) ? 3 : (R == 4 ? 2 : 1);
    const int shift = logn - logNs - rlog;
    if (logNs > 0) {  // the first stage has k = 0: all twiddles are 1
        // W^r for r = 1, 2, 4 from the table, the others as products of those (within 2 ulp):
        // three table reads per butterfly instead of R - 1
        cd w1 = tw[k << shift];
        if (inv) w1.y = -w1.y;
        v[1] = cmul(v[1], w1);
        if constexpr (R >= 4) {
            cd w2 = tw[(2 * k) << shift];
            if (inv) w2.y = -w2.y;
            v[2] = cmul(v[2], w2);
            const cd w3 = cmul(w1, w2);
            v[3] = cmul(v[3], w3);
            if constexpr (R == 8) {
                cd w4 = tw[(4 * k) << shift];
                if (inv) w4.y = -w4.y;
                v[4] = cmul(v[4], w4);
                v[5] = cmul(v[5], cmul(w1, w4));
                v[6] = cmul(v[6], cmul(w2, w4));
                v[7] = cmul(v[7], cmul(w3, w4));
            }
        }
    }
    dft<R>(v, inv);
    const int base = ((j >> logNs) << (logNs + rlog)) + k;
    for (int r = 0; r < R; ++r) s[sw(base + r * Ns)] = v[r];
}

}  // namespace cmm
