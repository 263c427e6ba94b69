// mgp_kernels.h -- host-side launchers for the kernels in mgp_kernels.hip.
// Every launcher enqueues on the given stream and returns immediately.
#pragma once

#include <hip/hip_runtime.h>

#include "mgp2d.h"
#include "mgp_common.h"

namespace mgp {

void launch_fill_ghosts(hipStream_t s, double *phi, const Grid &g, const Geo &geo, const Bc &bc, int homog,
                        int do_neum);
void launch_laplacian(hipStream_t s, double *out, const double *phi, const double *rhs, const Grid &g,
                      const Geo &geo, const Bc &bc, int homog, int jscale);
void launch_invdiags(hipStream_t s, double *out, const Grid &g, const Geo &geo, const Bc &bc, int homog);
// out = rhs - L[phi] (or L[phi] if rhs is null) on a homogeneous level; Dirichlet
// ghosts are synthesised in the kernel, so no fill is needed first.
void launch_residual_synth(hipStream_t s, double *out, const double *phi, const double *rhs, const Grid &g,
                           const Geo &geo, const Bc &bc);
void launch_dc_invdiags(hipStream_t s, double *out, const Grid &g, const Geo &geo);
void launch_dc_laplacian(hipStream_t s, double *out, const double *phi, const double *rhs, const Grid &g,
                         const Geo &geo);
void launch_dc_modrhs(hipStream_t s, double *modrhs, const double *rhs, const double *phi, const Grid &g,
                      const Geo &geo);
void launch_gs_mc_interior(hipStream_t s, double *phi, const double *rhs, const double *invd, const Grid &g,
                           const Geo &geo, double omega, int ci, int cj);
// The four interior colours of one multi-colour sweep in one launch over LDS
// tiles (bit-identical to the four launch_gs_mc_interior calls in colour order).
void launch_gs_mc_interior_tiled(hipStream_t s, double *phi, const double *rhs, const double *invd, const Grid &g,
                                 const Geo &geo, double omega);
void launch_gs_mc_ring(hipStream_t s, double *phi, const double *rhs, const double *invd, const Grid &g,
                       const Geo &geo, const Bc &bc, int homog, double omega);
void launch_gs_lex(hipStream_t s, double *phi, const double *rhs, const double *invd, const Grid &g,
                   const Geo &geo, const Bc &bc, int homog, double omega, int sweeps);
void launch_dc_rb(hipStream_t s, double *phi, const double *modrhs, const double *invd, const Grid &g,
                  const Geo &geo, double omega, int colour);
void launch_dc_lex(hipStream_t s, double *phi, const double *modrhs, const double *invd, const Grid &g,
                   const Geo &geo, const Bc &bc, int homog, double omega, int sweeps);
// DC relax_gs, red-black (MGPoisson2D_DD.f90:227-285): modrhs, then `sweeps`
// times fill ghosts, colour 0, colour 1.  On LAH Neumann/Dirichlet sides the
// sweeps run in LDS tiles, several per load (mgp_rbtile.hip), ping-ponging
// through tmp; otherwise one launch per colour.  mod and tmp are level-sized
// scratch; the result is bit-identical either way.  may_end_in_tmp: when the
// result lies in tmp, leave it there and return true (the caller swaps the
// two arrays' roles) instead of copying it back to phi.
bool dc_rb_relax(hipStream_t s, double *phi, double *tmp, double *mod, const double *rhs, const double *invd,
                 const Grid &g, const Geo &geo, const Bc &bc, int homog, double omega, int sweeps,
                 bool may_end_in_tmp = false);
// -1: tile shape timed per level (default); 0: one launch per colour; k >= 1:
// tile shape k - 1 of dc_rb_tile_shapes().  Returns the previous setting.
int dc_rb_tile_choice(int choice);
int dc_rb_tile_shapes();
// One tile launch of S sweeps with the modified rhs formed in the kernel,
// phi -> out (measurement); returns the tile shape used.
int dc_rb_tile_once(hipStream_t s, const double *phi, double *out, const double *rhs, const double *invd,
                    const Grid &g, const Geo &geo, const Bc &bc, double omega, int S);
void launch_restrict_cc(hipStream_t s, double *crse, const double *fine, const Grid &cg, int fpitch, int rx,
                        int ry, double *e, const double *invd);
void launch_restrict_xf(hipStream_t s, double *crse, const double *fine, const Grid &cg, int fpitch, int rx,
                        int ry);
void launch_restrict_yf(hipStream_t s, double *crse, const double *fine, const Grid &cg, int fpitch, int rx,
                        int ry);
void launch_prolong2x2(hipStream_t s, double *fine, int fpitch, const double *crse, const double *crse_x3,
                       const Grid &cg, int order);
void launch_extrap3(hipStream_t s, double *dst, const double *src, const Grid &g);
void launch_einit(hipStream_t s, double *e, const double *r, const double *invd, const Grid &g);
void launch_axpy_ghosted(hipStream_t s, double *y, const double *x, double a, const Grid &g);
void launch_lincomb(hipStream_t s, double *out, double a, const double *y, double b, const double *x,
                    const Grid &g, int ghosted);
void launch_setval(hipStream_t s, double *y, double v, const Grid &g, int ghosted);
// out = a * b / out = a / b over valid cells
void launch_lincomb_mul(hipStream_t s, double *out, const double *a, const double *b, const Grid &g, int unused);
void launch_lincomb_div(hipStream_t s, double *out, const double *a, const double *b, const Grid &g, int unused);
// kind 0: sum w a b; 1: sum w a^2; 2: max |a|; 3: sum w a; 4: sum w |a|^p.
// w = vol (or vol*J).  Writes the scalar to out (device); partial needs
// reduce_blocks(g) doubles.
int reduce_blocks(const Grid &g);
void launch_reduce(hipStream_t s, double *out, double *partial, const double *a, const double *b,
                   const double *J, double vol, const Grid &g, int kind, int p = 1);
// extrapolate_ghosts of order 0..3, in place (one workgroup).
void launch_extrapolate(hipStream_t s, double *phi, const Grid &g, int order);

void launch_seq_mean(hipStream_t s, double *out, const double *fine, const double *J, const Grid &g, double dxprod);
void launch_sub_mean(hipStream_t s, double *fine, const double *sums, const Grid &g);

// BiCGStab device scalars
enum { SC_RHO = 0, SC_RHO_NEW = 1, SC_ALPHA = 2, SC_OMEGA = 3, SC_BETA = 4, SC_RV = 5, SC_TS = 6, SC_TT = 7,
       SC_RN = 8, SC_COUNT = 16 };
void launch_bicg_scalars(hipStream_t s, double *sc, int op);
void launch_bicg_update(hipStream_t s, double *out, const double *x, const double *y, const double *sc,
                        const Grid &g, int op);

struct BottomArgs {
    Grid g;
    Geo geo;
    Bc bc;
    double *phi, *rhs, *invd;
    double *e, *pt, *st, *r, *rt, *p, *t, *v, *mod;
    double *scal;
    int dc, order, max_iters, max_restarts, homog;
    double tol;
};
void launch_bicgstab2(hipStream_t s, const BottomArgs &a);

// Line relaxation along x (fast solver, mgp_xline.hip).
// Ring-cell stencil table of the homogeneous operator: ring_table_elems(g) doubles.
int ring_table_elems(const Grid &g);
void launch_ring_coef(hipStream_t s, double *K, const Grid &g, const Geo &geo, const Bc &bc);
// rc = 1x2 (y) average of the fine residual f - L e (table at ring cells); ec = 0.
void launch_residual_restrict_y(hipStream_t s, double *rc, double *ec, const double *e, const double *f,
                                const double *K, const Grid &gf, const Grid &gc, const Geo &geo, int off = 0);
// FP32 form (mixed-precision preconditioner): same layout and formulas.
void launch_residual_restrict_y(hipStream_t s, float *rc, float *ec, const float *e, const float *f, const float *K,
                                const Grid &gf, const Grid &gc, const GeoF &geo, int off = 0);
// Flexible GMRES vector kernels (mgp_gmres.hip); V holds vectors i < nv at
// V + i * stride (padded level layout, valid cells only).
// out[i] = V_i . w (plain sum over the valid cells); partial: mdot_blocks(g) * nv doubles.
int mdot_blocks(const Grid &g);
void launch_mdot(hipStream_t s, double *out, double *partial, const double *V, size_t stride, const double *w, int nv,
                 const Grid &g);
// y += sign * sum_i h[i] V_i (h on the device).
void launch_maxpy(hipStream_t s, double *y, const double *V, size_t stride, const double *h, int nv, double sign,
                  const Grid &g);
// v = src / sqrt(*nrm2) (nrm2 on the device).
void launch_scale_rsqrt(hipStream_t s, double *v, const double *src, const double *nrm2, const Grid &g);
// Rows the line relaxation supports: up to 4111 cells solved whole (8207 in
// FP32), longer ones cut into 2, 4, .. equal segments (xline_segments, the
// fewest that fit), each solved as a line, the even segments of a colour
// before the odd ones.  xline_supported: some form exists in FP64.
bool xline_supported(int nx);
// Rows a 256-thread workgroup solves whole (the batched kernels' only form).
bool xline_single_line(int nx);
// Rows solved whole for elements of `elem` bytes.
bool xline_whole(int nx, size_t elem);
// Segments per row for elements of `elem` bytes: 1 for a whole row, 0 if nx
// has no supported cut.
int xline_segments(int nx, size_t elem);
// One y-semicoarsening step from a level of ny rows (levels of at least
// min_rows rows): the coarse row count, 0 to stop.  An even count halves.  An
// odd count of at least 2 min_rows + 1 rows with a Neumann y side leaves that
// side's row out (the top one when both are Neumann) and halves the rest:
// coarse row J then covers fine rows off + 2J, off + 2J + 1 (*off = 1 when the
// bottom row is left out, else 0); the coarse grid's side there lies one row
// in and keeps the Neumann condition, and the left-out row takes the
// correction of the coarse row next to it.  Without this the hierarchy stops
// at the odd count (49 rows for ny = 784) and the coarsest level's sweeps
// bound the convergence of the whole solve.
inline int semicoarsen_step(int ny, int min_rows, bool neum_lo, bool neum_hi, int &off) {
    off = 0;
    if (ny % 2 == 0) return ny / 2 >= min_rows ? ny / 2 : 0;
    if (ny < 2 * min_rows + 1 || !(neum_lo || neum_hi)) return 0;
    off = neum_hi ? 0 : 1;
    return (ny - 1) / 2;
}
// Sweeps on the coarsest level of the semicoarsening hierarchy: at least
// `sb`.  An x-line zebra sweep reduces errors that are smooth in y like one
// Gauss-Seidel sweep in y, so a coarsest level of many rows needs more of
// them (ny = 16 k with k odd and Dirichlet y sides leaves k rows), and so
// does a strongly stretched grid with Neumann y sides, whose lowest mode is
// nearly constant in y (208 x 448 at beta = 6, unwarped: 30 BiCGStab
// iterations with 3 sweeps on its 7 coarsest rows, 19 with 8, 7 with 128).
//  * The single-problem solver: half the level's rows, up to 64.  Its
//    sweeps are latency-bound launches: 8 on the 4 coarsest rows of the
//    2048^2 bench problem saved 4 of 70 V-cycles but cost 0.23 ms of a
//    2.2 ms cycle.
//  * The batch (batched = true): 8 below a finer level.  Its coarsest
//    sweeps are one chain of launches for all problems; the batched mix ran
//    5 % faster with 8 than with half the rows up to 64, and 4 % slower
//    with at least 12.
// A hierarchy of one level (an odd row count from the start, at Dirichlet y
// sides) solves on its only level: half its rows, up to 64, in both.
inline int coarsest_sweeps(int sb, int ny, bool only_level, bool batched) {
    const int half = ny / 2 < 64 ? ny / 2 : 64;
    const int want = batched && !only_level ? 8 : half;
    return sb > want ? sb : want;
}
// Shape of the batched line solve of an nx-cell row: threads per row (64 up
// to 512 cells, 128 up to 1024, else 256) and chunk length C (2, 4, 8 or 16;
// 0: too long).  The default rule; the batch autotuner may pick another.
void xline_batch_shape(int nx, int &C, int &threads);
// Chunk length of the batched line solve of an nx-cell row with `threads`
// threads per row (64, 128 or 256), 0 if that shape has no kernel or its LDS
// exceeds 160 KB for elements of `elem` bytes.
int xline_batch_chunk(int nx, int threads, size_t elem);
// dynamic LDS of one x-line workgroup for elements of `elem` bytes
size_t xline_lds_bytes(int nx, size_t elem);
// the same for chunk length C and `threads` threads per row
size_t xline_lds_bytes_for(int nx, int C, size_t elem, int threads);
// Relax rows j = row0 + k*rowstep, k < nrows (one workgroup each; no row of a
// launch may read another: interior rows at least two apart, boundary rows
// three).  The first workgroup then also relaxes row extra0, the last one row
// extra1 (< 0: none), each after its own row.
void launch_xline(hipStream_t s, double *e, const double *f, const double *K, const Grid &g, const Geo &geo,
                  int row0, int rowstep, int nrows, int extra0, int extra1);
void launch_xline(hipStream_t s, float *e, const float *f, const float *K, const Grid &g, const GeoF &geo, int row0,
                  int rowstep, int nrows, int extra0, int extra1);
// ef += 1x2 (y) prolongation of ec (prolong_y_cell); off: the fine row of
// coarse row 0's first row (semicoarsen_step)
void launch_prolong_y(hipStream_t s, double *ef, const double *ec, const Grid &gf, const Grid &gc, double slo,
                      double shi, int off = 0);
void launch_prolong_y(hipStream_t s, float *ef, const float *ec, const Grid &gf, const Grid &gc, double slo,
                      double shi, int off = 0);
// Precision conversions over a level array.  ghosted: also the ghost ring.
void launch_setval(hipStream_t s, float *y, float v, const Grid &g, int ghosted);
void launch_convert(hipStream_t s, float *out, const double *in, const Grid &g, int ghosted);
void launch_convert(hipStream_t s, double *out, const float *in, const Grid &g, int ghosted);
// out[k] = (float)in[k], k < n (flat arrays: stencil tables, whole padded levels)
void launch_convert_flat(hipStream_t s, float *out, const double *in, size_t n);
// Start of an FP32 V-cycle: f32 = (float)f over valid cells, e32 = 0 over
// valid and ghost cells.
void launch_begin32(hipStream_t s, float *f32, float *e32, const double *f, const Grid &g);
void launch_axpbypcz(hipStream_t s, double *out, double a, const double *x, double b, const double *y, double c,
                     const double *z, const Grid &g);
void launch_relax_block(hipStream_t s, const BottomArgs &a, double omega, int sweeps, int homog);

// Fourth-order finite-volume operator (mgp_fourth.hip), homogeneous Neumann sides.
// out = rhs - L4 phi (or L4 phi if rhs is null) over the valid cells; nx, ny >= 5.
void launch_lap4(hipStream_t s, double *out, const double *phi, const double *rhs, const Grid &g, const Geo4 &m);
// Face-averaged metrics of an MGP_MAP_WARPED mapping (3-point Gauss-Legendre over each face).
struct Geo4W {
    double *A4, *Bx4, *dA, *dBx, *C4, *By4, *dC, *dBy;
};
void launch_metric4(hipStream_t s, const mgp_mapping &m, const Grid &g, const Geo4W &out);

// Refinement loop (mgp_refine.hip).
// out = rhs - L (hi + lo) in double-double, rounded once (homogeneous Dirichlet / Neumann sides).
void launch_residual_dd(hipStream_t s, double *out, const double *hi, const double *lo, const double *rhs,
                        const Grid &g, const Geo &geo, const Bc &bc);
// (hi, lo) += x in double-double, then x = 0, over the valid cells.
void launch_fold_dd(hipStream_t s, double *hi, double *lo, double *x, const Grid &g);
// rhs -= J sum(rhs) / sum(J) (compensated sums into sums[0..1]; partial: reduce_blocks(g) doubles).
void launch_project_jmean(hipStream_t s, double *rhs, const double *J, double *sums, double *partial,
                          const Grid &g);
// dst(j, i) = src(i, j) for 0 <= i < ni, 0 <= j < nj (level layouts of pitch spitch / dpitch).
void launch_transpose(hipStream_t s, double *dst, int dpitch, const double *src, int spitch, int ni, int nj);
void launch_transpose(hipStream_t s, float *dst, int dpitch, const double *src, int spitch, int ni, int nj);
void launch_transpose(hipStream_t s, double *dst, int dpitch, const float *src, int spitch, int ni, int nj);

// Flux tools (mgp_flux.hip).  Face fields use the level layout: x-face (i, j)
// and y-face (i, j) at ix(pitch, i, j).
// xflux, yflux = J grad phi with Neumann faces from the BCs; phi's ghosts filled first.
void launch_grad(hipStream_t s, double *xflux, double *yflux, const double *phi, const Grid &g, const Geo &geo,
                 const Bc &bc, int homog);
// xflux -= J grad phi, yflux -= J grad phi (same face values as launch_grad).
void launch_grad_subtract(hipStream_t s, double *xflux, double *yflux, const double *phi, const Grid &g,
                          const Geo &geo, const Bc &bc, int homog);
// One compute_pd field (nodedir 1: x-faces, 2: y-faces; dir: derivative direction).
void launch_pd(hipStream_t s, double *pd, const double *phi, const Grid &g, int dir, int nodedir);
void launch_div(hipStream_t s, double *div, const double *xflux, const double *yflux, const Grid &g);
void launch_boundary_fluxes(hipStream_t s, double *xflux, double *yflux, const Grid &g, const Bc &bc, int homog);
// integrate2d_bdry: *out (device) = sum dx (xflux(nx, j) - xflux(0, j)) + sum dy (yflux(i, ny) - yflux(i, 0)),
// summed in the reference's order.
void launch_integrate_bdry(hipStream_t s, double *out, const double *xflux, const double *yflux, const Grid &g);
// gpnorm before its root: p == 0 max |phi|, else sum dy phi^p over columns cx0, cx1 (rows 0 .. ny-1) then
// sum dx phi^p over rows ry0, ry1 (columns 0 .. nx-1), in the reference's order; indices -1 .. n.
void launch_gpnorm_sum(hipStream_t s, double *out, const double *phi, const Grid &g, int p, int cx0, int cx1, int ry0,
                       int ry1);

}  // namespace mgp
