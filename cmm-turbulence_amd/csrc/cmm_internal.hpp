// Internal (C++) interface between the kernels, the C-ABI layer and the engine.
#pragma once

#include <hip/hip_runtime.h>
#include <string>

#include "cmm_core.hpp"

namespace cmm {

void set_error(const std::string &s);
const char *last_error();
int hip_check(hipError_t e, const char *what);

int twiddles(int n, const cd **out);
int workspace(size_t bytes, void **out);
int scalar_slot(unsigned long long **out);

int op_map_identity(const cmm_grid &g, double *cx, double *cy, hipStream_t st);
size_t op_advect_workspace(const cmm_grid &gm, const cmm_grid &gp, const cmm_advect_coeffs &c);
int op_map_advect(const cmm_grid &gm, const cmm_grid &gp, const double *cx, const double *cy, const PsiLevels &P,
                  const cmm_advect_coeffs &c, double *ncx, double *ncy, void *work, size_t work_bytes,
                  hipStream_t st, hipEvent_t mid = nullptr, const double *m8_in = nullptr, double *m8_out = nullptr,
                  const double *corr1 = nullptr, const double *corr12 = nullptr, const double *lw = nullptr,
                  const PsiPre *pre_in = nullptr);
// (pre_in: combined fields written by the previous psi solve for these weights, used in place of
// the combination kernel)
// the combined psi fields a scheme's advection reads: 0 (the scheme reads the levels on their own),
// else 1 (C1) | 2 (C12) | 4 (C0)
int op_advect_fields(const cmm_advect_coeffs &c);
// tile (0: the largest of 4, 3, 2, 1 dividing the refinement ratio) and cap on the number of
// blocks of the tiled check; set by the autotuner
struct IncompTune {
    int tile = 0, max_blocks = 8192;
};
int op_incomp_launch(const cmm_grid &gm, const cmm_grid &gc, const double *cx, const double *cy,
                     unsigned long long *slot, hipStream_t st, IncompTune tune = IncompTune());
// M8[8j + p] = ChiX plane p, M8[8j + 4 + p] = ChiY plane p (the advection's interleaved map)
int op_interleave_map(const cmm_grid &g, const double *cx, const double *cy, double *M8, hipStream_t st);
int op_incomp(const cmm_grid &gm, const cmm_grid &gc, const double *cx, const double *cy, double *out,
              hipStream_t st);
// max |x - X_b(X_f(x))| over the cell centres of gc (forward map on gf, backward on gb)
int op_invert_launch(const cmm_grid &gb, const cmm_grid &gf, const cmm_grid &gc, const double *bx, const double *by,
                     const double *fx, const double *fy, unsigned long long *slot, hipStream_t st);
int op_invert(const cmm_grid &gb, const cmm_grid &gf, const cmm_grid &gc, const double *bx, const double *by,
              const double *fx, const double *fy, double *out, hipStream_t st);
size_t op_psi_workspace(const cmm_grid &gv, const cmm_grid &gp, double fcut);
int op_psi_from_vort(const cmm_grid &gm, const cmm_grid &gv, const cmm_grid &gf, const cmm_grid &gp, const double *cx,
                     const double *cy, const double *W, int molly, double fcut, double *psi, void *work,
                     size_t work_bytes, hipStream_t st, hipEvent_t *ev = nullptr, const double *W4 = nullptr,
                     const PsiPre *pre = nullptr);
// op_psi_from_vort in two parts: the sampling of omega(X(x)) into the workspace, optionally with
// the incompressibility check on the same grid in the same pass (chk_slot, cleared by the call;
// possible when op_check_fusable(gm, gv, gc) holds for the check grid gc), and the spectral part
bool op_check_fusable(const cmm_grid &gm, const cmm_grid &gv, const cmm_grid &gc);
int op_psi_sample(const cmm_grid &gm, const cmm_grid &gv, const cmm_grid &gf, const cmm_grid &gp, const double *cx,
                  const double *cy, const double *W, int molly, double fcut, void *work, size_t work_bytes,
                  hipStream_t st, hipEvent_t *ev, const double *W4, unsigned long long *chk_slot);
// (pre: the transform also writes the next step's combined fields, k_row_c2r4_pre)
int op_psi_spectral(const cmm_grid &gv, const cmm_grid &gp, double fcut, double *psi, void *work, size_t work_bytes,
                    hipStream_t st, hipEvent_t *ev, const PsiPre *pre = nullptr);
// op_psi_from_vort with its workspace and twiddle tables in a static device array (no allocation);
// op_psi_noalloc_max_bytes() is the largest workspace it takes
size_t op_psi_noalloc_max_bytes();
int op_psi_from_vort_noalloc(const cmm_grid &gm, const cmm_grid &gv, const cmm_grid &gf, const cmm_grid &gp,
                             const double *cx, const double *cy, const double *W, int molly, double fcut, double *psi,
                             hipStream_t st);
// H4[4j + p] = H[p*n + j] (node-interleaved copy of a Hermite field)
int op_interleave4(const cmm_grid &g, const double *H, double *H4, hipStream_t st);
size_t op_hermite_workspace(const cmm_grid &g);
// op_translate_ic: the spectral buffers and the sampled initial condition (8 bytes per point)
size_t op_translate_workspace(const cmm_grid &g);
int op_hermite_from_real(const cmm_grid &g, const double *real, double *herm, void *work, size_t work_bytes,
                         hipStream_t st);
int op_hermite_from_spectrum(const cmm_grid &g, const double *spec_cols, double *herm, void *work,
                             size_t work_bytes, hipStream_t st);
// Stack formats: kStackFp64 = eight FP64 planes per map (ChiX planes, then ChiY planes), map m at
// stack + 8*Nc*m; kStackRec64 / kStackRec48 = node records of op_pack_maps (64 bytes FP64; 48 bytes
// with FP32 derivatives), map m at byte m*Nc*64 / m*Nc*48.  The values equal CMM_STACK_RECORDS_*.
constexpr int kStackFp64 = 0, kStackRec64 = 1, kStackRec48 = 2;
int op_translate_ic(const cmm_grid &gm, const cmm_grid &gf, const double *cx, const double *cy, const double *stack,
                    int nstack, int ic, const double *prm, int discrete, const cmm_grid &gd, const double *disc,
                    double *herm, void *work, size_t work_bytes, hipStream_t st, int stack_fmt = kStackFp64);
int op_compose_positions(const cmm_grid &gm, const cmm_grid &gt, const double *cx, const double *cy,
                         const double *stack, int nstack, double *xy, hipStream_t st, int stack_fmt = kStackFp64);
// n_maps maps read plane-major from (cx + m*stride, cy + m*stride) into node records of format fmt
// (kStackRec64 or kStackRec48) at out
int op_pack_maps(const cmm_grid &gm, const double *cx, const double *cy, long stride, int n_maps, int fmt, void *out,
                 hipStream_t st);
int op_forward_positions(const cmm_grid &gm, const cmm_grid &gt, const double *cx, const double *cy,
                         const double *stack, int nstack, double *xy, hipStream_t st);
int op_forward_points(const cmm_grid &gm, const cmm_grid &gb, const double *stack, int first, int nstack,
                      const double *cx, const double *cy, int cur, const double *in, double *out, int n, hipStream_t st);
int op_sample_init(const cmm_grid &gm, const cmm_grid &gt, const double *cx, const double *cy, const double *stack,
                   int nstack, int ic, const double *prm, int discrete, const cmm_grid &gd, const double *disc,
                   double *out, hipStream_t st, int stack_fmt = kStackFp64);
// Vlasov-Poisson psi: the 1D problems of one instance
struct VP1D {
    int nv, lognv;       // vort grid along x
    int np, lognp;       // psi grid along x (= along v)
    int nchunks;
    double hv;           // vort grid spacing along v
    double fkx_v;        // 2 pi / (hx NX) of the vort grid
    double fky_p;        // 2 pi / (hy NY) of the psi grid
    double fkx_p;        // 2 pi / (hx NX) of the psi grid
    double v0, hv_p, Lv; // psi grid v origin, spacing, length
};
// phi(x) of one instance from its column sums: f (the sampled field, row v0 read), part (chunk
// sums), twiddles of the vort and psi grids; phi written at [0, np), phi' at [2 np, 3 np)
struct VPJob {
    VP1D q;
    const double *f, *part;
    const cd *tw_v, *tw_p;
    double *phi;
};
size_t op_vp_psi_workspace(const cmm_grid &gv, const cmm_grid &gp);
// W4 (optional): the fine field node-interleaved, for the column sampler; H (optional): the bump
// function's H and H' (op_vp_bump), else computed into the workspace
int op_vp_psi_from_f(const cmm_grid &gm, const cmm_grid &gv, const cmm_grid &gf, const cmm_grid &gp, const double *cx,
                     const double *cy, const double *W, int molly, double *psi, void *work, size_t work_bytes,
                     hipStream_t st, const double *W4 = nullptr, const double *H = nullptr);
// ---- Ensembles: one Vlasov-Poisson step of many instances in grouped, persistent launches ----
// The arguments one instance's step gives its launches (those of op_map_advect_vp,
// op_incomp_launch and op_vp_psi_from_f), recorded by the engine instead of launched.
constexpr int kGroupMax = 64;
struct GroupInstance {
    // advection
    cmm_grid gm, gp;
    PsiLevels P;
    cmm_advect_coeffs c;
    const double *m8_in;
    double *m8_out, *ncx, *ncy;
    const double *H;
    void *adv_work;
    size_t adv_work_bytes;
    // check of the advected map on the fine grid
    cmm_grid gc;
    const double *ccx, *ccy;
    IncompTune tune;
    // psi of the advected map
    cmm_grid gv, gf;
    const double *pcx, *pcy, *W4;
    double *psi;
    void *psi_work;
    size_t psi_work_bytes;
    int parts;  // recorded parts: 1 advection, 2 check, 4 psi
};
// whether n instances can share grouped launches, from their grids, check tiles, time scheme and
// Lagrange order alone (gm, gp, gv, gc, W4, tune, c.time_integration_num, c.l_order): the same
// scheme, order, sampler rows and check tile edge
bool op_group_vp_alike(const GroupInstance *inst, int n);
// whether n recorded instances can run in one grouped step (op_group_vp_alike; the separable
// advection, the tensor check and the column sampler, their workspaces)
bool op_group_vp_eligible(const GroupInstance *inst, int n);
size_t op_group_vp_job_bytes(int n);
// The step of the n instances: every kind of launch of every instance in one launch, in the order
// of a single instance's step (1D combination, advection, check, sampling, column sums, phi, psi
// planes); max_blocks > 0 caps the grid (persistent workgroups), else one workgroup per block of
// work.  The job tables go up in one copy from host_jobs (pinned) to dev_jobs; E_inc of instance
// i lands in dev_slots[i] and is copied to host_slots[i] (pinned) after the check.
int op_group_vp_step(const GroupInstance *inst, int n, void *dev_jobs, void *host_jobs,
                     unsigned long long *dev_slots, unsigned long long *host_slots, hipStream_t st,
                     int max_blocks);
// H(v) at [0, np) and H'(v) at [2 np, 3 np) of H (3 np doubles): the grid-only part of psi
int op_vp_bump(const cmm_grid &gv, const cmm_grid &gp, double *H, hipStream_t st);
int op_vp_psi_planes(const cmm_grid &gp, const double *phi, const double *H, double *psi, hipStream_t st);
// phi and phi' of a distribution given on the grid g (sample outputs): into phi3[0, nx) and
// phi3[2 nx, 3 nx); workspace op_vp_efield_workspace(g)
size_t op_vp_efield_workspace(const cmm_grid &g);
int op_vp_efield(const cmm_grid &g, const double *f, double *phi3, void *work, size_t work_bytes, hipStream_t st);
// Map advection of a Vlasov-Poisson engine from the separable psi levels (1D tables of the combined
// phi and of the bump H instead of 2D gathers); H: op_vp_bump's array
bool op_vp_advect_supported(const cmm_grid &gm, const cmm_grid &gp, const cmm_advect_coeffs &c);
int op_map_advect_vp(const cmm_grid &gm, const cmm_grid &gp, const PsiLevels &P, const cmm_advect_coeffs &c,
                     const double *cx, const double *cy, const double *m8_in, double *m8_out, const double *H,
                     double *ncx, double *ncy, void *work, size_t work_bytes, hipStream_t st, hipEvent_t mid,
                     const double *lw);
// the same separable advection of a split map (identity + FP32 displacement records)
int op_map_advect32_vp(const cmm_grid &gm, const cmm_grid &gp, const PsiLevels &P, const cmm_advect_coeffs &c,
                       const float *M_in, float *M_out, const double *H, void *work, size_t work_bytes,
                       hipStream_t st, hipEvent_t mid, const double *lw);
int op_spectral_from_real(const cmm_grid &g, const double *real, int op, double *herm, void *work, size_t work_bytes,
                          hipStream_t st);
int op_sample_vort(const cmm_grid &gm, const cmm_grid &gt, const cmm_grid &gf, const double *cx, const double *cy,
                   const double *W, int molly, double *out, hipStream_t st);
// value plane of the Hermite field H (grid gh) at the points of gt
int op_sample_hermite(const cmm_grid &gh, const double *H, const cmm_grid &gt, double *out, hipStream_t st);
int op_particles(const cmm_grid &gp, const PsiLevels &P, int L, const double *Lp, double dt, int scheme, int n,
                 double *pos, hipStream_t st);
// inertial particles (positions and velocities), their start velocity (fluid = 1: psi level 0's
// velocity, else zero), and (d psi/dx, d psi/dy) at points (NaN outside the grid's bounds)
int op_particles_inertial(const cmm_grid &gp, const PsiLevels &P, int L, const double *Lp, double dt, int scheme,
                          double tau, int n, double *pos, double *vel, hipStream_t st);
int op_particles_vel_init(const cmm_grid &gp, const double *psi0, int n, int fluid, const double *pos, double *vel,
                          hipStream_t st);
int op_points_dxdy(const cmm_grid &gp, const double *psi0, int n, const double *pos, double *out, hipStream_t st);

// Split-precision form: the map as identity + FP32 displacement records (M[8j + p]: planes of
// X - x, then of Y - y, identity derivatives removed), FP32 Hermite records for the gathered
// fields, FP64 arithmetic, FP64 psi levels and FFTs.
int op_map_to32(const cmm_grid &g, const double *cx, const double *cy, float *M, hipStream_t st);
int op_map_from32(const cmm_grid &g, const float *M, double *cx, double *cy, hipStream_t st);
int op_interleave4_32(const cmm_grid &g, const double *H, float *H4, hipStream_t st);
bool op_advect32_supported(const cmm_advect_coeffs &c);
// The form of the Lagrange-combined psi fields that the split advection gathers: FP32 records
// (value, d/dx, d/dy, d2/dxdy), the same with an 8-bit code of the value's low part in the low byte
// of d2/dxdy (the value of a stream function of order one loses 2^-24 |psi| in FP32, which the
// velocity's 6/h amplifies), or
// FP64 records (the Vlasov-Poisson flow, whose -H(v) is large and smooth).
enum PsiRecForm { PsiRec32 = 0, PsiRec32Pk = 1, PsiRec64 = 2 };
size_t op_advect32_workspace(const cmm_grid &gp, const cmm_advect_coeffs &c, int psi_form = PsiRec32Pk);
int op_map_advect32(const cmm_grid &gm, const cmm_grid &gp, const PsiLevels &P, const cmm_advect_coeffs &c,
                    const float *M_in, float *M_out, void *work, size_t work_bytes, hipStream_t st,
                    hipEvent_t mid = nullptr, int psi_form = PsiRec32Pk, const double *corr1 = nullptr,
                    const double *corr12 = nullptr, const double *lw = nullptr, const PsiPre *pre_in = nullptr);
int op_incomp32_launch(const cmm_grid &gm, const cmm_grid &gc, const float *M, unsigned long long *slot,
                       hipStream_t st, IncompTune tune = IncompTune());
int op_incomp32(const cmm_grid &gm, const cmm_grid &gc, const float *M, double *out, hipStream_t st);
int op_psi_from_vort32(const cmm_grid &gm, const cmm_grid &gv, const cmm_grid &gf, const cmm_grid &gp, const float *M,
                       const float *W4, int molly, double fcut, double *psi, void *work, size_t work_bytes,
                       hipStream_t st, hipEvent_t *ev = nullptr, const PsiPre *pre = nullptr);
int op_psi_sample32(const cmm_grid &gm, const cmm_grid &gv, const cmm_grid &gf, const cmm_grid &gp, const float *M,
                    const float *W4, int molly, double fcut, void *work, size_t work_bytes, hipStream_t st,
                    hipEvent_t *ev, unsigned long long *chk_slot);
int op_psi_spectral32(const cmm_grid &gv, const cmm_grid &gp, double fcut, double *psi, void *work, size_t work_bytes,
                      hipStream_t st, hipEvent_t *ev, const PsiPre *pre = nullptr);
int op_sample_vort32(const cmm_grid &gm, const cmm_grid &gt, const cmm_grid &gf, const float *M, const float *W4,
                     int molly, double *out, hipStream_t st);

}  // namespace cmm
