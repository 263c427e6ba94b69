// mgp_batch.hip -- the batched fast solver (mgp_batch.h).
//
// Every problem runs the KRYLOV_MG iteration of mgp_krylov.hip: FP64
// right-preconditioned BiCGStab on the reference operator with homogeneous
// BCs, one y-semicoarsening x-line V-cycle per preconditioner application,
// restarts from the true residual.  What the batch changes is the launch
// structure: one launch per operation covers every problem that takes part
// in it, the workgroups finding their problem (and level) through tables
// built once.  Per-problem state (scalars, convergence, restarts) lives in
// device arrays and a per-problem active flag; problems that have converged
// keep their results while the others go on.
//
// Reductions are per problem and independent of the batch: problem p's
// partial sums come from nb(p) row blocks fixed by its size and are summed
// in block order, so a problem's iterates are the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <tuple>

#include "mgp_batch.h"
#include "mgp_geometry.h"
#include "mgp_kernels.h"
#include "mgp_reduce.h"
#include "mgp_stencil.h"

namespace mgp {

namespace {

constexpr int MIN_COARSE_NY = 4;

// The cells of a tile (64 x BT_ROWS), each thread taking four rows 4 apart,
// so that a thread's four cells' loads go out together.
template <typename F>
__device__ __forceinline__ void tile_cells(const BTile &t, const Grid &g, F &&f) {
    const int i = t.tx * 64 + (threadIdx.x & 63);
    if (i >= g.nx) return;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int j = t.ty * BT_ROWS + (threadIdx.x >> 6) + 4 * r;
        if (j < g.ny) f(i, j);
    }
}

// v[out] = (v[rhs] -) A v[in], A the reference operator with the problem's
// homogeneous BCs (ghosts synthesised from the valid cells).
__global__ void __launch_bounds__(256) k_b_apply(const BTile *tiles, const BProb *P, const int *mask, int out, int in,
                                                 int rhs) {
    const BTile t = tiles[blockIdx.x];
    if (!mask[t.item]) return;
    const BProb *p = P + t.item;
    const Grid g = p->g;
    const Geo geo = p->geo;
    const Bc bc = p->bc;
    const double *phi = p->v[in];
    double *o = p->v[out];
    const double *r = rhs >= 0 ? p->v[rhs] : nullptr;
    tile_cells(t, g, [&](int i, int j) {
        const size_t c = ix(g.pitch, i, j);
        const bool interior = i > 0 && i < g.nx - 1 && j > 0 && j < g.ny - 1;
        const double v = interior ? cell_lap_interior(CFld{phi, g.pitch}, g, geo, i, j)
                                  : cell_lap(synth(phi, g, bc), g, geo, bc, 1, i, j);
        o[c] = r ? r[c] - v : v;
    });
}

// Partial sums of vol * v[a] * v[b] over the rows b, b + nb, ... of a problem
// into partial[b]; with c >= 0 also those of vol * v[a] * v[c] into
// partial[nb + b], in the same order (so each equals its own pass's sum).
__global__ void __launch_bounds__(256) k_b_reduce_partial(const BTile *tiles, const BProb *P, const int *mask, int a,
                                                          int b, int c2) {
    __shared__ double ss[256], sc[256];
    const BTile t = tiles[blockIdx.x];
    if (!mask[t.item]) return;  // uniform over the workgroup
    const BProb *p = P + t.item;
    const Grid g = p->g;
    const double vol = g.dx * g.dy;
    const double *A = p->v[a], *B = p->v[b], *C = c2 >= 0 ? p->v[c2] : nullptr;
    const int nb = p->nb;
    Acc acc{0.0, 0.0}, acc2{0.0, 0.0};
    // The thread's cells, in the order of the sum: rows tx, tx + nb, .. and in
    // each row cells tid, tid + 256, .. (M of them).  Their loads go out Q at
    // a time ahead of the adds, so that Q pairs are in flight per thread; the
    // sum keeps that order.
    constexpr int Q = 8;
    const int tid = threadIdx.x;
    const int M = tid < g.nx ? (g.nx - 1 - tid) / 256 + 1 : 0;
    const int R = t.tx < g.ny ? (g.ny - 1 - t.tx) / nb + 1 : 0;
    const int total = M * R;
    int k = 0, m = 0;  // row and cell index of the next load
    for (int e0 = 0; e0 < total; e0 += Q) {
        double va[Q], vb[Q], vc[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            va[q] = 0.0;
            vb[q] = 0.0;
            vc[q] = 0.0;
            if (e0 + q < total) {
                const size_t c = ix(g.pitch, tid + 256 * m, t.tx + nb * k);
                va[q] = A[c];
                vb[q] = B[c];
                if (C) vc[q] = C[c];
                if (++m == M) {
                    m = 0;
                    ++k;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < Q; ++q)
            if (e0 + q < total) {
                acc.add(vol * (va[q] * vb[q]));
                if (C) acc2.add(vol * (va[q] * vc[q]));
            }
    }
    const double v = block_combine(ss, sc, acc, 0.0, 0);
    if (threadIdx.x == 0) p->partial[t.tx] = v;
    if (C) {
        const double v2 = block_combine(ss, sc, acc2, 0.0, 0);
        if (threadIdx.x == 0) p->partial[nb + t.tx] = v2;
    }
}

// One workgroup per problem: sc[slot] = sum of the nb partials in block order
// (and sc[slot2] of the second product's, slot2 >= 0).
__global__ void __launch_bounds__(256) k_b_reduce_final(const BProb *P, const int *mask, int slot, int slot2) {
    __shared__ double ss[256], sc[256];
    if (!mask[blockIdx.x]) return;
    const BProb *p = P + blockIdx.x;
    for (int h = 0; h < (slot2 >= 0 ? 2 : 1); ++h) {
        Acc acc{0.0, 0.0};
        for (int k = threadIdx.x; k < p->nb; k += blockDim.x) acc.add(p->partial[h * p->nb + k]);
        const double v = block_combine(ss, sc, acc, 0.0, 0);
        if (threadIdx.x == 0) p->sc[h ? slot2 : slot] = v;
    }
}

// BiCGStab scalars of every masked problem (the ops of k_bicg_scalars).
__global__ void k_b_scalars(const BProb *P, const int *mask, int n, int op) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || !mask[k]) return;
    double *sc = P[k].sc;
    if (op == 0) {
        sc[SC_BETA] = (sc[SC_RHO_NEW] / sc[SC_RHO]) * (sc[SC_ALPHA] / sc[SC_OMEGA]);
        sc[SC_RHO] = sc[SC_RHO_NEW];
    } else if (op == 1) {
        sc[SC_ALPHA] = sc[SC_RHO] / sc[SC_RV];
    } else if (op == 2) {
        sc[SC_OMEGA] = sc[SC_TT] > 0.0 ? sc[SC_TS] / sc[SC_TT] : 0.0;
    } else {
        sc[SC_RHO] = 1.0;
        sc[SC_ALPHA] = 1.0;
        sc[SC_OMEGA] = 1.0;
        sc[SC_BETA] = 0.0;
    }
}

// Vector updates (the ops of k_bicg_update):
//  0: p = r + beta (p - omega v)   1: s = r - alpha v   2: phi += alpha phat
//  3: phi += omega shat            4: r = s - omega t
// With L (FP32 cycles), ops 0 and 1 also start the cycle that preconditions
// the vector they produce: the finest level's f32 = (float) p or s, e32 = 0.
__global__ void __launch_bounds__(256) k_b_update(const BTile *tiles, const BProb *P, const int *mask, int op,
                                                  const BLev *L) {
    const BTile t = tiles[blockIdx.x];
    if (!mask[t.item]) return;
    const BProb *p = P + t.item;
    const Grid g = p->g;
    const double *sc = p->sc;
    double *const *v = p->v;
    const double alpha = sc[SC_ALPHA], beta = sc[SC_BETA], omega = sc[SC_OMEGA];
    float *f32 = L ? L[p->lev0].f32 : nullptr, *e32 = L ? L[p->lev0].e32 : nullptr;
    tile_cells(t, g, [&](int i, int j) {
        const size_t c = ix(g.pitch, i, j);
        double out;
        switch (op) {
        case 0: out = v[BV_P][c] = v[BV_R][c] + beta * (v[BV_P][c] - omega * v[BV_V][c]); break;
        case 1: out = v[BV_S][c] = v[BV_R][c] - alpha * v[BV_V][c]; break;
        case 2: out = v[BV_PHI][c] = v[BV_PHI][c] + alpha * v[BV_PHAT][c]; break;
        case 3: out = v[BV_PHI][c] = v[BV_PHI][c] + omega * v[BV_SHAT][c]; break;
        default: out = v[BV_R][c] = v[BV_S][c] - omega * v[BV_T][c]; break;
        }
        if (f32 && op <= 1) {
            f32[c] = (float)out;
            e32[c] = 0.0f;
        }
    });
}

// v[dst] = v[src] (src >= 0) or 0 over the valid cells.
__global__ void __launch_bounds__(256) k_b_copy(const BTile *tiles, const BProb *P, const int *mask, int dst, int src) {
    const BTile t = tiles[blockIdx.x];
    if (!mask[t.item]) return;
    const BProb *p = P + t.item;
    const Grid g = p->g;
    double *d = p->v[dst];
    const double *sv = src >= 0 ? p->v[src] : nullptr;
    tile_cells(t, g, [&](int i, int j) {
        const size_t c = ix(g.pitch, i, j);
        d[c] = sv ? sv[c] : 0.0;
    });
}

// End of an FP32 cycle: v[PHAT or SHAT] = (double) e32 of the finest level.
__global__ void __launch_bounds__(256) k_b_end32(const BTile *tiles, const BProb *P, const BLev *L, const int *mask,
                                                 int sel) {
    const BTile t = tiles[blockIdx.x];
    if (!mask[t.item]) return;
    const BProb *p = P + t.item;
    const Grid g = p->g;
    const float *e32 = L[p->lev0].e32;
    double *dst = p->v[sel ? BV_SHAT : BV_PHAT];
    tile_cells(t, g, [&](int i, int j) {
        const size_t c = ix(g.pitch, i, j);
        dst[c] = (double)e32[c];
    });
}

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

void tiles_of(std::vector<BTile> &out, int item, int nx, int ny) {
    for (int ty = 0; ty < (ny + BT_ROWS - 1) / BT_ROWS; ++ty)
        for (int tx = 0; tx < (nx + 63) / 64; ++tx) out.push_back(BTile{item, tx, ty, 0});
}

}  // namespace

// Appends a table to the set's pinned staging at a 256-byte boundary; until
// build_tables uploads the set, the table's p holds its offset there.
BatchKMG::Table BatchKMG::stage(TableSet &ts, const void *host, size_t bytes, int n) {
    Table t;
    t.n = n;
    if (n == 0) return t;
    const size_t off = round256(ts.used);
    if (off + bytes > ts.hcap) {
        const size_t cap = std::max(off + bytes, 2 * ts.hcap);
        unsigned char *h = nullptr;
        if (hipHostMalloc((void **)&h, cap, hipHostMallocDefault) != hipSuccess || !h) {
            (void)hipGetLastError();
            g_last_error = "batch: pinned staging of " + std::to_string(cap) + " bytes for the launch tables";
            throw Failure(MGP_ERR_NOMEM);
        }
        if (ts.used) memcpy(h, ts.host, ts.used);
        if (ts.host) (void)hipHostFree(ts.host);
        ts.host = h;
        ts.hcap = cap;
    }
    memcpy(ts.host + off, host, bytes);
    ts.used = off + bytes;
    t.p = reinterpret_cast<void *>(off);
    return t;
}

// Device bytes one problem adds to a batch (its part of the arena, in the
// constructor's layout, plus its scalars and descriptors).
size_t BatchKMG::problem_bytes(const Problem &q, const mgp_solver_opts &o) {
    const bool mx = o.method == MGP_METHOD_KRYLOV_MG32;
    const int maxlev = o.maxdepth < 0 ? 64 : o.maxdepth + 1;
    size_t b = 0;
    const size_t n0 = level_elems(q.nx, q.ny);
    b += BV_COUNT * round256(n0 * sizeof(double));
    const int nb = std::max(1, std::min(q.ny, (q.nx * q.ny + 8191) / 8192));
    b += round256((size_t)2 * nb * sizeof(double));
    int ny = q.ny, lev = 0;
    size_t levels = 0;
    while (true) {
        const size_t nl = level_elems(q.nx, ny);
        b += 5 * round256(nl * sizeof(double));
        if (lev) b += 2 * round256(nl * sizeof(double));
        const Grid gl = make_grid(q.nx, ny, q.dx, q.dy);
        b += round256((size_t)ring_table_elems(gl) * sizeof(double));
        if (mx) {
            b += 4 * round256(nl * sizeof(float));
            b += round256((size_t)ring_table_elems(gl) * sizeof(float));
            b += 2 * round256(nl * sizeof(float));
        }
        ++lev;
        ++levels;
        int off = 0;
        const int nyc = lev >= maxlev ? 0 : semicoarsen_step(ny, MIN_COARSE_NY, q.bc.type[YLO] == BC_NEUM,
                                                              q.bc.type[YHI] == BC_NEUM, off);
        if (nyc == 0) break;
        ny = nyc;
    }
    return b + SC_COUNT * sizeof(double) + sizeof(BProb) + 2 * sizeof(int) + levels * sizeof(BLev);
}

BatchKMG::BatchKMG(hipStream_t s_, const std::vector<Problem> &probs, const mgp_solver_opts &o) : s(s_) {
    const int n = (int)probs.size();
    if (n <= 0) {
        g_last_error = "batch: at least one problem";
        throw Failure(MGP_ERR_ARG);
    }
    mixed = o.method == MGP_METHOD_KRYLOV_MG32;
    sd = o.smooth_down > 0 ? o.smooth_down : 1;
    su = o.smooth_up > 0 ? o.smooth_up : 1;
    sb = o.smooth_bottom > 0 ? o.smooth_bottom : 1;
    const int maxlev = o.maxdepth < 0 ? 64 : o.maxdepth + 1;
    for (const Problem &q : probs) {
        if (q.nx < 4 || q.ny < 4 || !(q.dx > 0.0) || !(q.dy > 0.0)) {
            g_last_error = "batch: every problem needs nx, ny >= 4 and dx, dy > 0";
            throw Failure(MGP_ERR_ARG);
        }
        if (!xline_single_line(q.nx)) {
            g_last_error = "batch: row too long for one workgroup's line solve (nx <= 4111)";
            throw Failure(MGP_ERR_UNSUPPORTED);
        }
        for (int k = 0; k < 4; ++k)
            if (q.bc.type[k] != BC_DIRI && q.bc.type[k] != BC_NEUM) {
                g_last_error = "batch: Dirichlet and Neumann sides only";
                throw Failure(MGP_ERR_UNSUPPORTED);
            }
    }

    // level counts and the arena
    nlev.resize(n);
    size_t bytes = 0;
    auto take = [&](size_t elems, size_t elem = sizeof(double)) {
        const size_t off = bytes;
        bytes += round256(elems * elem);
        return off;
    };
    struct Off {
        size_t v[BV_COUNT], partial;
        std::vector<size_t> J, Jxx, Jxy, Jyx, Jyy, e, r, K;
        std::vector<size_t> f32m[4], K32, e32, r32;  // FP32 copies (mixed precision)
    };
    std::vector<Off> off(n);
    for (int p = 0; p < n; ++p) {
        const Problem &q = probs[p];
        int ny = q.ny, lev = 0;
        Off &O = off[p];
        const size_t n0 = level_elems(q.nx, q.ny);
        for (int k = 0; k < BV_COUNT; ++k) O.v[k] = take(n0);
        const int nb = std::max(1, std::min(q.ny, (q.nx * q.ny + 8191) / 8192));
        O.partial = take((size_t)2 * nb);  // two products' partials per reduction pass
        while (true) {
            const size_t nl = level_elems(q.nx, ny);
            O.J.push_back(take(nl));
            O.Jxx.push_back(take(nl));
            O.Jxy.push_back(take(nl));
            O.Jyx.push_back(take(nl));
            O.Jyy.push_back(take(nl));
            O.e.push_back(lev ? take(nl) : 0);
            O.r.push_back(lev ? take(nl) : 0);
            const Grid gl = make_grid(q.nx, ny, q.dx, q.dy);
            O.K.push_back(take((size_t)ring_table_elems(gl)));
            if (mixed) {
                for (int m = 0; m < 4; ++m) O.f32m[m].push_back(take(nl, sizeof(float)));
                O.K32.push_back(take((size_t)ring_table_elems(gl), sizeof(float)));
                O.e32.push_back(take(nl, sizeof(float)));
                O.r32.push_back(take(nl, sizeof(float)));
            }
            ++lev;
            int off = 0;
            const int nyc = lev >= maxlev ? 0 : semicoarsen_step(ny, MIN_COARSE_NY, q.bc.type[YLO] == BC_NEUM,
                                                                  q.bc.type[YHI] == BC_NEUM, off);
            if (nyc == 0) break;
            ny = nyc;
        }
        nlev[p] = lev;
        depth = std::max(depth, lev);
    }
    arena = dev_alloc_zero(bytes);
    scal = static_cast<double *>(dev_alloc_zero((size_t)n * SC_COUNT * sizeof(double)));
    char *base = static_cast<char *>(arena);
    auto at = [&](size_t o) { return reinterpret_cast<double *>(base + o); };
    auto atf = [&](size_t o) { return reinterpret_cast<float *>(base + o); };

    // problems, levels, metrics and stencil tables
    P.resize(n);
    for (int p = 0; p < n; ++p) {
        const Problem &q = probs[p];
        const Off &O = off[p];
        BProb &B = P[p];
        B.g = make_grid(q.nx, q.ny, q.dx, q.dy);
        for (int k = 0; k < 4; ++k) {
            B.bc.type[k] = q.bc.type[k];
            B.bc.mode[k] = MODE_UNIFORM;
            B.bc.val[k] = 0.0;
            B.bc.data[k] = nullptr;
        }
        for (int k = 0; k < BV_COUNT; ++k) B.v[k] = at(O.v[k]);
        B.sc = scal + (size_t)p * SC_COUNT;
        B.partial = at(O.partial);
        B.nb = std::max(1, std::min(q.ny, (q.nx * q.ny + 8191) / 8192));
        B.lev0 = (int)L.size();
        int ny = q.ny;
        double dy = q.dy;
        const Bc lbc = mg_bc(B.bc, MODE_UNIFORM);
        for (int d = 0; d < nlev[p]; ++d) {
            BLev l;
            l.g = make_grid(q.nx, ny, q.dx, dy);
            l.geo = Geo{at(O.J[d]), at(O.Jxx[d]), at(O.Jxy[d]), at(O.Jyx[d]), at(O.Jyy[d])};
            double *J = at(O.J[d]), *Jxx = at(O.Jxx[d]), *Jxy = at(O.Jxy[d]), *Jyx = at(O.Jyx[d]),
                   *Jyy = at(O.Jyy[d]);
            if (d == 0) {
                if (q.map) {
                    launch_metric(s, *q.map, l.g, J, Jxx, Jxy, Jyx, Jyy);
                } else {
                    const mgp_geo &h = *q.geo;
                    upload_block(s, J, l.g, h.J, 0, 0, q.nx, q.ny);
                    upload_block(s, Jxx, l.g, h.Jxx, 0, 0, q.nx + 1, q.ny);
                    upload_block(s, Jxy, l.g, h.Jxy, 0, 0, q.nx + 1, q.ny);
                    upload_block(s, Jyx, l.g, h.Jyx, 0, 0, q.nx, q.ny + 1);
                    upload_block(s, Jyy, l.g, h.Jyy, 0, 0, q.nx, q.ny + 1);
                }
                l.e[0] = B.v[BV_PHAT];
                l.e[1] = B.v[BV_SHAT];
                l.f[0] = B.v[BV_P];
                l.f[1] = B.v[BV_S];
                B.geo = l.geo;
            } else {
                // from the fine level's rows off.. (off = 1: its bottom row is left out)
                const BLev &f = L.back();
                const size_t sh = (size_t)f.off * f.g.pitch;
                launch_restrict_cc(s, J, f.geo.J + sh, l.g, f.g.pitch, 1, 2, nullptr, nullptr);
                launch_restrict_xf(s, Jxx, f.geo.Jxx + sh, l.g, f.g.pitch, 1, 2);
                launch_restrict_xf(s, Jxy, f.geo.Jxy + sh, l.g, f.g.pitch, 1, 2);
                launch_restrict_yf(s, Jyx, f.geo.Jyx + sh, l.g, f.g.pitch, 1, 2);
                launch_restrict_yf(s, Jyy, f.geo.Jyy + sh, l.g, f.g.pitch, 1, 2);
                l.e[0] = l.e[1] = at(O.e[d]);
                l.f[0] = l.f[1] = at(O.r[d]);
            }
            launch_ring_coef(s, at(O.K[d]), l.g, l.geo, lbc);
            l.K = at(O.K[d]);
            l.geo32 = GeoF{nullptr, nullptr, nullptr, nullptr, nullptr};
            l.K32 = nullptr;
            l.e32 = l.f32 = nullptr;
            if (mixed) {
                // FP32 copies of the whole padded arrays (formed in FP64, rounded once)
                const size_t nl = level_elems(q.nx, ny);
                const double *src[4] = {Jxx, Jxy, Jyx, Jyy};
                for (int m = 0; m < 4; ++m) launch_convert_flat(s, atf(O.f32m[m][d]), src[m], nl);
                launch_convert_flat(s, atf(O.K32[d]), l.K, (size_t)ring_table_elems(l.g));
                l.geo32 = GeoF{nullptr, atf(O.f32m[0][d]), atf(O.f32m[1][d]), atf(O.f32m[2][d]), atf(O.f32m[3][d])};
                l.K32 = atf(O.K32[d]);
                l.e32 = atf(O.e32[d]);
                l.f32 = atf(O.r32[d]);
            }
            l.slo = B.bc.type[YLO] == BC_DIRI ? -1.0 : 1.0;
            l.shi = B.bc.type[YHI] == BC_DIRI ? -1.0 : 1.0;
            l.prob = p;
            l.nt = 0;  // set with the chunk class below
            l.sweeps = d == nlev[p] - 1 ? coarsest_sweeps(sb, ny, nlev[p] == 1, true) : 0;
            l.off = 0;
            const int nyc = semicoarsen_step(ny, MIN_COARSE_NY, B.bc.type[YLO] == BC_NEUM, B.bc.type[YHI] == BC_NEUM,
                                             l.off);
            L.push_back(l);
            ny = nyc;
            dy *= 2.0;
        }
    }

    // each problem's line solves use its own chunk length, whatever the batch
    // (the partition of a row into chunks decides the rounding of its solve);
    // set_tuning / tune change it per (nx, BC class)
    Cp.resize(n);
    NTp.resize(n);
    for (int p = 0; p < n; ++p) {
        xline_batch_shape(P[p].g.nx, Cp[p], NTp[p]);
        for (int d = 0; d < nlev[p]; ++d) L[P[p].lev0 + d].nt = P[p].g.nx / Cp[p];
    }

    dP = static_cast<BProb *>(dev_alloc_zero(n * sizeof(BProb)));
    dL = static_cast<BLev *>(dev_alloc_zero(L.size() * sizeof(BLev)));
    hip_check(hipMemcpyAsync(dP, P.data(), n * sizeof(BProb), hipMemcpyHostToDevice, s), "problem table");
    hip_check(hipMemcpyAsync(dL, L.data(), L.size() * sizeof(BLev), hipMemcpyHostToDevice, s), "level table");
    active = static_cast<int *>(dev_alloc_zero(n * sizeof(int)));
    mask = static_cast<int *>(dev_alloc_zero(n * sizeof(int)));
    hip_check(hipHostMalloc((void **)&host_sc, (size_t)n * SC_COUNT * sizeof(double)), "hipHostMalloc");
    use_graph = o.use_graph != 0;
    {
        // streams for the chunk classes of a phase (default priority)
        std::map<int, int> classes;
        for (int p = 0; p < n; ++p) classes[Cp[p] * 1024 + NTp[p]] = 1;
        const int nc = (int)classes.size();
        hip_check(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming), "hipEventCreate");
        for (int c = 0; c < nc; ++c) {
            hipStream_t st = nullptr;
            hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
            aux.push_back(st);
            hipEvent_t ev = nullptr;
            hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
            join_ev.push_back(ev);
        }
    }
    act_.assign(n, 0);
    it_.assign(n, 0);
    oc_.assign(n, MGP_MAXITERS);
    maxit_.assign(n, 0);
    tol_.assign(n, 0.0);
    r0_.assign(n, 0.0);
    base_.assign(n, 1.0);
    footprint_ = bytes + (size_t)n * (SC_COUNT * sizeof(double) + sizeof(BProb) + 2 * sizeof(int)) +
                 L.size() * sizeof(BLev);
    std::vector<int> all(n);
    for (int p = 0; p < n; ++p) all[p] = p;
    build_tables(full, all);
}

// Launch tables of the problems `which` (and the captured iteration over them).
void BatchKMG::build_tables(TableSet &ts, const std::vector<int> &which) {
    ts.nprob = (int)which.size();
    ts.depth = 0;
    for (int p : which) ts.depth = std::max(ts.depth, nlev[p]);
    std::vector<BTile> t0h, tredh;
    for (int p : which) {
        tiles_of(t0h, p, P[p].g.nx, P[p].g.ny);
        for (int b = 0; b < P[p].nb; ++b) tredh.push_back(BTile{p, b, 0, 0});
    }
    ts.used = 0;
    ts.t0 = stage(ts, t0h.data(), t0h.size() * sizeof(BTile), (int)t0h.size());
    ts.tred = stage(ts, tredh.data(), tredh.size() * sizeof(BTile), (int)tredh.size());
    ts.tr.assign(ts.depth, Table());
    ts.tp.assign(ts.depth, Table());
    ts.bottom.assign(ts.depth, 0);
    for (int d = 0; d < ts.depth; ++d) {
        std::map<std::tuple<int, int, int, int>, std::vector<XlTask>> tabs;  // (set, colour, C, threads)
        std::map<std::tuple<int, int, int, int>, size_t> lds;
        std::vector<BTile> trh, tph;
        for (int p : which) {
            if (d >= nlev[p]) continue;
            const int lev = P[p].lev0 + d;
            const int ny = L[lev].g.ny, nx = L[lev].g.nx;
            const int set = d == nlev[p] - 1 ? SET_BOTTOM : SET_MID;
            const int C = Cp[p], NT = NTp[p];
            if (set == SET_BOTTOM) ts.bottom[d] = std::max(ts.bottom[d], L[lev].sweeps);
            const size_t need = xline_lds_bytes_for(nx, C, mixed ? sizeof(float) : sizeof(double), NT);
            auto table = [&](int colour) -> std::vector<XlTask> & {
                const auto key = std::make_tuple(set, colour, C, NT);
                size_t &b = lds[key];
                b = std::max(b, need);
                return tabs[key];
            };
            // the order of xline_zebra: even interior rows, odd interior rows
            // (the first and last relaxing rows 0 and ny-1 after their own),
            // else rows 0 and ny-1 in a launch of their own
            const int n_even = ny >= 4 ? (ny - 4) / 2 + 1 : 0;
            const int n_odd = ny >= 3 ? (ny - 3) / 2 + 1 : 0;
            if (n_even) {
                std::vector<XlTask> &v = table(0);
                for (int k = 0; k < n_even; ++k) v.push_back(XlTask{lev, 2 + 2 * k, -1, 0});
            }
            if (n_odd) {
                std::vector<XlTask> &v = table(1);
                for (int k = 0; k < n_odd; ++k) {
                    int extra = -1;
                    if (n_odd >= 2 && k == 0) extra = 0;
                    if (n_odd >= 2 && k == n_odd - 1) extra = ny - 1;
                    v.push_back(XlTask{lev, 1 + 2 * k, extra, 0});
                }
            }
            if (n_odd < 2) {
                std::vector<XlTask> &v = table(2);
                v.push_back(XlTask{lev, 0, -1, 0});
                v.push_back(XlTask{lev, ny - 1, -1, 0});
            }
            if (set == SET_MID) {
                tiles_of(trh, lev, nx, L[lev + 1].g.ny);
                tiles_of(tph, lev, nx, ny);
            }
        }
        for (auto &kv : tabs) {
            XlTable x;
            x.d = d;
            x.set = std::get<0>(kv.first);
            x.colour = std::get<1>(kv.first);
            x.t = stage(ts, kv.second.data(), kv.second.size() * sizeof(XlTask), (int)kv.second.size());
            x.t.C = std::get<2>(kv.first);
            x.t.threads = std::get<3>(kv.first);
            x.t.lds = lds[kv.first];
            ts.xl.push_back(x);
        }
        ts.tr[d] = stage(ts, trh.data(), trh.size() * sizeof(BTile), (int)trh.size());
        ts.tp[d] = stage(ts, tph.data(), tph.size() * sizeof(BTile), (int)tph.size());
    }
    // one upload of the staged set; the offsets become device addresses
    if (ts.used > ts.dcap) {
        if (ts.dev) dev_free(ts.dev);
        ts.dev = nullptr;
        ts.dcap = 0;
        ts.dev = dev_alloc_zero(ts.used);
        ts.dcap = ts.used;
    }
    auto place = [&](Table &t) {
        if (t.n) t.p = static_cast<unsigned char *>(ts.dev) + reinterpret_cast<size_t>(t.p);
    };
    place(ts.t0);
    place(ts.tred);
    for (XlTable &x : ts.xl) place(x.t);
    for (Table &t : ts.tr) place(t);
    for (Table &t : ts.tp) place(t);
    if (ts.used) hip_check(hipMemcpyAsync(ts.dev, ts.host, ts.used, hipMemcpyHostToDevice, s), "table upload");
    std::sort(ts.xl.begin(), ts.xl.end(), [](const XlTable &a, const XlTable &b) {
        return std::make_tuple(a.d, a.set, a.colour, a.t.C, a.t.threads) <
               std::make_tuple(b.d, b.set, b.colour, b.t.C, b.t.threads);
    });
    hip_check(hipStreamSynchronize(s), "batch tables");
    if (use_graph) {
        TableSet *const saved = T;
        T = &ts;
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "capture");
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            T = saved;
            throw;
        }
        T = saved;
        hip_check(hipStreamEndCapture(s, &g), "capture end");
        hip_check(hipGraphInstantiate(&ts.graph, g, nullptr, nullptr, 0), "graph instantiate");
        (void)hipGraphDestroy(g);
    }
}

void BatchKMG::release(TableSet &ts) {
    (void)hipStreamSynchronize(s);
    if (ts.graph) (void)hipGraphExecDestroy(ts.graph);
    ts.graph = nullptr;
    ts.t0 = ts.tred = Table();
    ts.xl.clear();
    ts.tr.clear();
    ts.tp.clear();
    ts.bottom.clear();
    ts.nprob = ts.depth = 0;
    ts.used = 0;
}

void BatchKMG::free_buffers(TableSet &ts) {
    release(ts);
    if (ts.dev) dev_free(ts.dev);
    if (ts.host) (void)hipHostFree(ts.host);
    ts.dev = nullptr;
    ts.host = nullptr;
    ts.dcap = ts.hcap = 0;
}

BatchKMG::~BatchKMG() {
    free_buffers(part);
    free_buffers(full);
    for (hipStream_t st : aux) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    for (hipEvent_t ev : join_ev) (void)hipEventDestroy(ev);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (dP) dev_free(dP);
    if (dL) dev_free(dL);
    if (active) dev_free(active);
    if (mask) dev_free(mask);
    if (arena) dev_free(arena);
    if (scal) dev_free(scal);
    if (host_sc) (void)hipHostFree(host_sc);
}

void BatchKMG::zebra(int d, int set, int sel, int k) {
    for (int colour = 0; colour < 3; ++colour) {
        std::vector<const XlTable *> tabs;
        for (const XlTable &x : T->xl)
            if (x.d == d && x.set == set && x.colour == colour) tabs.push_back(&x);
        if (tabs.empty()) continue;
        if (tabs.size() == 1 || tabs.size() > aux.size()) {
            for (const XlTable *x : tabs)
                launch_xline_batch(s, (const XlTask *)x->t.p, x->t.n, dL, active, sel, x->t.C, x->t.threads,
                                   x->t.lds, k, mixed);
            continue;
        }
        // one launch per chunk class, side by side
        hip_check(hipEventRecord(fork_ev, s), "fork");
        for (size_t i = 0; i < tabs.size(); ++i) {
            hip_check(hipStreamWaitEvent(aux[i], fork_ev, 0), "fork wait");
            launch_xline_batch(aux[i], (const XlTask *)tabs[i]->t.p, tabs[i]->t.n, dL, active, sel, tabs[i]->t.C,
                               tabs[i]->t.threads, tabs[i]->t.lds, k, mixed);
            hip_check(hipEventRecord(join_ev[i], aux[i]), "join");
        }
        for (size_t i = 0; i < tabs.size(); ++i) hip_check(hipStreamWaitEvent(s, join_ev[i], 0), "join wait");
    }
}

// e = M f for the pair `sel` of every active problem (phat from p: 0, shat
// from s: 1): the V-cycle of mgp_krylov.hip, level by level over the batch.
// Problems whose coarsest level is d make their bottom sweeps there while
// the others go on down.
void BatchKMG::precondition(int sel) {
    // FP32 cycles: the update that produced p or s has set the finest
    // level's f32 and zeroed its e32 (k_b_update with the level table)
    if (!mixed) k_b_copy<<<T->t0.n, 256, 0, s>>>((const BTile *)T->t0.p, dP, active, sel ? BV_SHAT : BV_PHAT, -1);
    const int depth = T->depth;
    for (int d = 0; d < depth; ++d) {
        for (int k = 0; k < sd; ++k) zebra(d, SET_MID, sel, -1);
        // each problem ending at level d makes its own number of sweeps there
        for (int k = 0; k < T->bottom[d]; ++k) zebra(d, SET_BOTTOM, sel, k);
        if (d + 1 < depth) launch_restrict_batch(s, (const BTile *)T->tr[d].p, T->tr[d].n, dL, active, sel, mixed);
    }
    for (int d = depth - 2; d >= 0; --d) {
        launch_prolong_batch(s, (const BTile *)T->tp[d].p, T->tp[d].n, dL, active, sel, mixed);
        for (int k = 0; k < su; ++k) zebra(d, SET_MID, sel, -1);
    }
    if (mixed) k_b_end32<<<T->t0.n, 256, 0, s>>>((const BTile *)T->t0.p, dP, dL, active, sel);
}

void BatchKMG::reduce(int a, int b, int slot, const int *m, int c, int slot2) {
    k_b_reduce_partial<<<T->tred.n, 256, 0, s>>>((const BTile *)T->tred.p, dP, m, a, b, c);
    k_b_reduce_final<<<count(), 256, 0, s>>>(dP, m, slot, c >= 0 ? slot2 : -1);
}

// One BiCGStab iteration of every active problem (the body of
// KrylovMG::solve); ends with ||r||^2 vol in each problem's SC_RN.
void BatchKMG::body() {
    const int n = count();
    const BTile *tl = (const BTile *)T->t0.p;
    const int nt = T->t0.n;
    const int sb_ = (n + 63) / 64;
    const BLev *l32 = mixed ? dL : nullptr;
    reduce(BV_RHAT, BV_R, SC_RHO_NEW, active);
    k_b_scalars<<<sb_, 64, 0, s>>>(dP, active, n, 0);
    k_b_update<<<nt, 256, 0, s>>>(tl, dP, active, 0, l32);
    precondition(0);
    k_b_apply<<<nt, 256, 0, s>>>(tl, dP, active, BV_V, BV_PHAT, -1);
    reduce(BV_RHAT, BV_V, SC_RV, active);
    k_b_scalars<<<sb_, 64, 0, s>>>(dP, active, n, 1);
    k_b_update<<<nt, 256, 0, s>>>(tl, dP, active, 1, l32);
    k_b_update<<<nt, 256, 0, s>>>(tl, dP, active, 2, nullptr);
    precondition(1);
    k_b_apply<<<nt, 256, 0, s>>>(tl, dP, active, BV_T, BV_SHAT, -1);
    reduce(BV_T, BV_S, SC_TS, active, BV_T, SC_TT);  // t.s and t.t in one pass
    k_b_scalars<<<sb_, 64, 0, s>>>(dP, active, n, 2);
    k_b_update<<<nt, 256, 0, s>>>(tl, dP, active, 3, nullptr);
    k_b_update<<<nt, 256, 0, s>>>(tl, dP, active, 4, nullptr);
    reduce(BV_R, BV_R, SC_RN, active);
}

// r = rhs - A phi and ||r||^2 vol into SC_RN, for the masked problems.
void BatchKMG::true_residual(const int *m) {
    k_b_apply<<<T->t0.n, 256, 0, s>>>((const BTile *)T->t0.p, dP, m, BV_R, BV_PHI, BV_RHS);
    reduce(BV_R, BV_R, SC_RN, m);
}

// New BiCGStab run from the current r for the masked problems.
void BatchKMG::restart(const int *m) {
    const int n = count();
    const BTile *tl = (const BTile *)T->t0.p;
    const int nt = T->t0.n;
    k_b_copy<<<nt, 256, 0, s>>>(tl, dP, m, BV_RHAT, BV_R);
    k_b_copy<<<nt, 256, 0, s>>>(tl, dP, m, BV_P, -1);
    k_b_copy<<<nt, 256, 0, s>>>(tl, dP, m, BV_V, -1);
    k_b_scalars<<<(n + 63) / 64, 64, 0, s>>>(dP, m, n, 3);
}

void BatchKMG::upload_flags(int *dst, const std::vector<int> &h) {
    hip_check(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice, s), "flags");
}

void BatchKMG::fetch_scalars() {
    hip_check(hipMemcpyAsync(host_sc, scal, (size_t)count() * SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, s),
              "scalars");
    hip_check(hipStreamSynchronize(s), "sync");
}

// Starts the problems `which` (their rhs already set): phi from zero, or from
// its current value where warm[k] != 0; tolerance tol[k] relative to ||rhs||
// and maxit[k] iterations.  Problems whose initial residual already meets the
// tolerance finish at once (converged, 0 iterations).
void BatchKMG::start(const std::vector<int> &which, const std::vector<int> &warm, const std::vector<double> &tol,
                     const std::vector<int> &maxit) {
    const int n = count();
    std::vector<int> msk(n, 0), cold(n, 0);
    for (size_t q = 0; q < which.size(); ++q) {
        const int p = which[q];
        msk[p] = 1;
        cold[p] = warm[q] ? 0 : 1;
        it_[p] = 0;
        oc_[p] = MGP_MAXITERS;
        tol_[p] = tol[q];
        maxit_[p] = maxit[q];
        base_[p] = 1.0;
    }
    const BTile *tl = (const BTile *)T->t0.p;
    upload_flags(mask, cold);
    k_b_copy<<<T->t0.n, 256, 0, s>>>(tl, dP, mask, BV_PHI, -1);
    upload_flags(mask, msk);
    reduce(BV_RHS, BV_RHS, SC_RN, mask);
    fetch_scalars();
    std::vector<int> zero(n, 0);
    bool any_zero = false;
    for (int p : which) {
        r0_[p] = std::sqrt(host_sc[p * SC_COUNT + SC_RN]);
        if (r0_[p] == 0.0) zero[p] = any_zero = true;
    }
    if (any_zero) {
        // rhs = 0: the solution is phi = 0 (a warm phi is not kept)
        upload_flags(mask, zero);
        k_b_copy<<<T->t0.n, 256, 0, s>>>(tl, dP, mask, BV_PHI, -1);
        upload_flags(mask, msk);
    }
    true_residual(mask);
    fetch_scalars();
    for (int p : which) {
        const double rel = r0_[p] > 0.0 ? std::sqrt(host_sc[p * SC_COUNT + SC_RN]) / r0_[p] : 0.0;
        if (rel <= tol_[p]) {
            oc_[p] = MGP_CONVERGED;
            msk[p] = 0;
        } else if (maxit_[p] <= 0) {
            oc_[p] = MGP_MAXITERS;
            msk[p] = 0;
        } else {
            act_[p] = 1;
            base_[p] = rel;
        }
    }
    upload_flags(mask, msk);
    restart(mask);
    upload_flags(active, act_);
}

// One BiCGStab iteration of every running problem, then their convergence
// checks (true residual where the recursive one says done, or has fallen by
// the refinement factor: restart from it).  Returns the problems that
// finished in this step.
std::vector<int> BatchKMG::step() {
    launch_step();
    hip_check(hipStreamSynchronize(s), "sync");
    return finish_step();
}

// The iteration and the copy of the scalars to the host, on the stream.
void BatchKMG::launch_step() {
    if (T->graph) hip_check(hipGraphLaunch(T->graph, s), "graph launch");
    else body();
    hip_check(hipMemcpyAsync(host_sc, scal, (size_t)count() * SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, s),
              "scalars");
}

bool BatchKMG::step_done() const { return hipStreamQuery(s) == hipSuccess; }

// After launch_step has completed: the convergence checks (and restarts).
std::vector<int> BatchKMG::finish_step() {
    const int n = count();
    // iterative refinement factors (as KrylovMG::solve)
    const double refine = mixed ? 1e-3 : 1e-6;
    std::vector<int> done, msk(n, 0);
    bool check = false;
    for (int p = 0; p < n; ++p) {
        if (!act_[p]) continue;
        ++it_[p];
        const double rn = std::sqrt(host_sc[p * SC_COUNT + SC_RN]) / r0_[p];
        const double omega = host_sc[p * SC_COUNT + SC_OMEGA];
        if (rn <= tol_[p] || omega == 0.0 || !std::isfinite(rn) || rn <= refine * base_[p]) {
            msk[p] = 1;
            check = true;
        } else if (it_[p] >= maxit_[p]) {
            act_[p] = 0;
            done.push_back(p);
        }
    }
    if (check) {
        upload_flags(mask, msk);
        true_residual(mask);
        fetch_scalars();
        for (int p = 0; p < n; ++p) {
            if (!msk[p]) continue;
            const double tr = std::sqrt(host_sc[p * SC_COUNT + SC_RN]) / r0_[p];
            if (tr <= tol_[p] || !std::isfinite(tr) || it_[p] >= maxit_[p]) {
                oc_[p] = tr <= tol_[p] ? MGP_CONVERGED : (std::isfinite(tr) ? MGP_MAXITERS : MGP_DIVERGED);
                act_[p] = 0;
                msk[p] = 0;
                done.push_back(p);
            } else {
                base_[p] = tr;  // restart from the true residual
            }
        }
        upload_flags(mask, msk);
        restart(mask);
    }
    if (!done.empty()) upload_flags(active, act_);
    return done;
}

double BatchKMG::time_fine_xline(int reps, double *bytes) {
    if (T != &full) {
        release(part);
        T = &full;
    }
    const int n = count();
    // the even interior rows of level 0 (colour 0 of the first zebra phase)
    std::vector<const XlTable *> tabs;
    for (const XlTable &x : T->xl)
        if (x.d == 0 && x.set == SET_MID && x.colour == 0) tabs.push_back(&x);
    const double es = mixed ? 4.0 : 8.0;
    double b = 0.0;
    for (int p = 0; p < n; ++p) {
        if (nlev[p] < 2) continue;  // a single-level problem's rows are in the bottom set
        const Grid &g = P[p].g;
        const double R = g.ny >= 4 ? (g.ny - 4) / 2 + 1 : 0;
        // e on rows 1 .. 2R+1 and written on R rows; f, Jxx, Jxy on R rows;
        // Jyx, Jyy on 2R face rows; 9 table entries for the 2R ring cells
        b += g.nx * es * ((2.0 * R + 1.0) + R + R + R + 2.0 * R + 2.0 * R + R) + R * 2.0 * 9.0 * es;
    }
    *bytes = b;
    std::vector<int> all(n, 1);
    upload_flags(active, all);
    auto launch_set = [&]() {
        if (tabs.size() == 1 || tabs.size() > aux.size()) {
            for (const XlTable *x : tabs)
                launch_xline_batch(s, (const XlTask *)x->t.p, x->t.n, dL, active, 0, x->t.C, x->t.threads, x->t.lds,
                                   -1, mixed);
            return;
        }
        hip_check(hipEventRecord(fork_ev, s), "fork");
        for (size_t i = 0; i < tabs.size(); ++i) {
            hip_check(hipStreamWaitEvent(aux[i], fork_ev, 0), "fork wait");
            launch_xline_batch(aux[i], (const XlTask *)tabs[i]->t.p, tabs[i]->t.n, dL, active, 0, tabs[i]->t.C,
                               tabs[i]->t.threads, tabs[i]->t.lds, -1, mixed);
            hip_check(hipEventRecord(join_ev[i], aux[i]), "join");
        }
        for (size_t i = 0; i < tabs.size(); ++i) hip_check(hipStreamWaitEvent(s, join_ev[i], 0), "join wait");
    };
    hipEvent_t a = nullptr, e = nullptr;
    hip_check(hipEventCreate(&a), "event");
    hip_check(hipEventCreate(&e), "event");
    launch_set();  // warm
    hip_check(hipEventRecord(a, s), "event");
    for (int k = 0; k < reps; ++k) launch_set();
    hip_check(hipEventRecord(e, s), "event");
    hip_check(hipEventSynchronize(e), "event");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, a, e), "event");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(e);
    upload_flags(active, act_);
    hip_check(hipStreamSynchronize(s), "sync");
    return ms / reps;
}

int BatchKMG::bc_class(const Bc &bc) {
    int c = 0;
    for (int k = 0; k < 4; ++k)
        if (bc.type[k] == BC_DIRI) c |= 1 << k;
    return c;
}

std::vector<BatchKMG::TuneEntry> BatchKMG::tuning() const {
    std::map<std::pair<int, int>, int> keys;
    for (int p = 0; p < count(); ++p) keys[{P[p].g.nx, bc_class(P[p].bc)}] = NTp[p];
    std::vector<TuneEntry> out;
    for (const auto &kv : keys) out.push_back(TuneEntry{kv.first.first, kv.first.second, mixed ? 1 : 0, kv.second});
    return out;
}

// After a change of shapes: the level table's chunk counts, enough streams
// for the chunk classes of a phase, the launch tables and the captured graph.
void BatchKMG::apply_shapes() {
    const int n = count();
    for (int p = 0; p < n; ++p)
        for (int d = 0; d < nlev[p]; ++d) L[P[p].lev0 + d].nt = P[p].g.nx / Cp[p];
    hip_check(hipMemcpyAsync(dL, L.data(), L.size() * sizeof(BLev), hipMemcpyHostToDevice, s), "level table");
    std::map<int, int> classes;
    for (int p = 0; p < n; ++p) classes[Cp[p] * 1024 + NTp[p]] = 1;
    while (aux.size() < classes.size()) {
        hipStream_t st = nullptr;
        hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        aux.push_back(st);
        hipEvent_t ev = nullptr;
        hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
        join_ev.push_back(ev);
    }
    release(part);
    release(full);
    T = &full;
    std::vector<int> all(n);
    for (int p = 0; p < n; ++p) all[p] = p;
    build_tables(full, all);
}

int BatchKMG::set_tuning(const std::vector<TuneEntry> &t) {
    const size_t elem = mixed ? sizeof(float) : sizeof(double);
    std::map<std::pair<int, int>, int> want;
    for (const TuneEntry &e : t)
        if (e.fp32 == (mixed ? 1 : 0)) want[{e.nx, e.bc}] = e.threads;
    int changed = 0;
    for (int p = 0; p < count(); ++p) {
        const auto it = want.find({P[p].g.nx, bc_class(P[p].bc)});
        if (it == want.end() || it->second == NTp[p]) continue;
        const int C = xline_batch_chunk(P[p].g.nx, it->second, elem);
        if (C == 0) {
            g_last_error = "batch set_tuning: no line kernel with " + std::to_string(it->second) +
                           " threads for rows of " + std::to_string(P[p].g.nx) + " cells";
            throw Failure(MGP_ERR_ARG);
        }
        Cp[p] = C;
        NTp[p] = it->second;
        ++changed;
    }
    if (changed) apply_shapes();
    return changed;
}

// Mean ms of `reps` launch sets of the finest-level even-row line relaxation
// of every problem with shapes (C[p], NT[p]): one launch per chunk class, side
// by side on the forked streams, as inside a solve.  Sets the level table's
// finest-level chunk counts to those shapes (apply_shapes restores them).
double BatchKMG::time_shapes(const std::vector<int> &C, const std::vector<int> &NT, int reps) {
    const int n = count();
    const size_t elem = mixed ? sizeof(float) : sizeof(double);
    std::map<std::pair<int, int>, std::vector<XlTask>> cls;
    std::map<std::pair<int, int>, size_t> lds;
    for (int p = 0; p < n; ++p) {
        const int nx = P[p].g.nx, ny = P[p].g.ny, lev = P[p].lev0;
        L[lev].nt = nx / C[p];
        if (ny < 4) continue;
        auto &t = cls[{C[p], NT[p]}];
        for (int k = 0; k < (ny - 4) / 2 + 1; ++k) t.push_back(XlTask{lev, 2 + 2 * k, -1, 0});
        size_t &l = lds[{C[p], NT[p]}];
        l = std::max(l, xline_lds_bytes_for(nx, C[p], elem, NT[p]));
    }
    hip_check(hipMemcpyAsync(dL, L.data(), L.size() * sizeof(BLev), hipMemcpyHostToDevice, s), "level table");
    while (aux.size() < cls.size()) {
        hipStream_t st = nullptr;
        hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        aux.push_back(st);
        hipEvent_t ev = nullptr;
        hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
        join_ev.push_back(ev);
    }
    struct Cls {
        void *t;
        int n, C, NT;
        size_t lds;
    };
    std::vector<Cls> v;
    for (auto &kv : cls) {
        void *d = dev_alloc_zero(kv.second.size() * sizeof(XlTask));
        v.push_back(Cls{d, (int)kv.second.size(), kv.first.first, kv.first.second, lds[kv.first]});
        hip_check(hipMemcpyAsync(d, kv.second.data(), kv.second.size() * sizeof(XlTask), hipMemcpyHostToDevice, s),
                  "tasks");
    }
    auto launch_set = [&]() {
        hip_check(hipEventRecord(fork_ev, s), "fork");
        for (size_t i = 0; i < v.size(); ++i) {
            hip_check(hipStreamWaitEvent(aux[i], fork_ev, 0), "fork wait");
            launch_xline_batch(aux[i], (const XlTask *)v[i].t, v[i].n, dL, active, 0, v[i].C, v[i].NT, v[i].lds, -1,
                               mixed);
            hip_check(hipEventRecord(join_ev[i], aux[i]), "join");
        }
        for (size_t i = 0; i < v.size(); ++i) hip_check(hipStreamWaitEvent(s, join_ev[i], 0), "join wait");
    };
    hipEvent_t a = nullptr, e = nullptr;
    float ms = 0.0f;
    try {
        hip_check(hipEventCreate(&a), "event");
        hip_check(hipEventCreate(&e), "event");
        launch_set();  // warm
        hip_check(hipEventRecord(a, s), "event");
        for (int r = 0; r < reps; ++r) launch_set();
        hip_check(hipEventRecord(e, s), "event");
        hip_check(hipEventSynchronize(e), "event");
        hip_check(hipEventElapsedTime(&ms, a, e), "event");
    } catch (...) {
        (void)hipStreamSynchronize(s);
        for (Cls &c : v) dev_free(c.t);
        if (a) (void)hipEventDestroy(a);
        if (e) (void)hipEventDestroy(e);
        throw;
    }
    hip_check(hipStreamSynchronize(s), "sync");
    for (Cls &c : v) dev_free(c.t);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(e);
    return ms / reps;
}

// Coordinate search over row-length bands (rows of up to 128, 256, 512, 1024
// and 4111 cells): for each band in turn, each width its rows all have a
// kernel for is timed with the whole batch's finest-level set (the other
// bands as chosen so far), and a width replaces the current one when it is at
// least 2 % faster.  Timing the whole set, not a band alone, weighs a width by
// what it does to the launches that run beside it.
std::vector<BatchKMG::TuneEntry> BatchKMG::tune(int reps) {
    if (reps < 1) {
        g_last_error = "batch tune: reps >= 1";
        throw Failure(MGP_ERR_ARG);
    }
    if (T != &full) {
        release(part);
        T = &full;
    }
    const int n = count();
    const size_t elem = mixed ? sizeof(float) : sizeof(double);
    std::vector<int> all(n, 1);
    upload_flags(active, all);
    std::vector<int> C = Cp, NT = NTp;
    const int edges[6] = {0, 128, 256, 512, 1024, 1 << 30};
    for (int b = 0; b < 5; ++b) {
        std::vector<int> mem;
        for (int p = 0; p < n; ++p)
            if (P[p].g.nx > edges[b] && P[p].g.nx <= edges[b + 1]) mem.push_back(p);
        if (mem.empty()) continue;
        double best = time_shapes(C, NT, reps);
        for (int w : {64, 128, 256}) {  // the batched line kernels' widths
            std::vector<int> c2 = C, t2 = NT;
            bool ok = true, same = true;
            for (int p : mem) {
                c2[p] = xline_batch_chunk(P[p].g.nx, w, elem);
                t2[p] = w;
                ok = ok && c2[p] != 0;
                same = same && NT[p] == w;
            }
            if (!ok || same) continue;
            const double ms = time_shapes(c2, t2, reps);
            if (ms < 0.98 * best) {
                best = ms;
                C = c2;
                NT = t2;
            }
        }
    }
    Cp = C;
    NTp = NT;
    upload_flags(active, act_);
    apply_shapes();
    return tuning();
}

int BatchKMG::running() const {
    int r = 0;
    for (int a : act_) r += a;
    return r;
}

int BatchKMG::solve(double tol, int maxiters, int zerophi, int *iters, int *outcome) {
    const int n = count();
    if (T != &full) {
        release(part);
        T = &full;
    }
    std::vector<int> all(n);
    for (int p = 0; p < n; ++p) all[p] = p;
    start(all, std::vector<int>(n, zerophi ? 0 : 1), std::vector<double>(n, tol), std::vector<int>(n, maxiters));
    // Finished problems' workgroups return at once, but each still costs its
    // dispatch: about 7 us per finished problem per iteration on MI355X (a
    // 1024-problem table set with one problem running takes 8.8 ms per
    // iteration, the same problem in a set of its own 1.2 ms).  Rebuilding the
    // tables over the running problems (and capturing their graph) costs
    // about 4 ms + 60 us per problem.  The tables are rebuilt once the
    // dispatches wasted since the last build have cost about one rebuild.
    constexpr double DEAD_US = 7.0, REBUILD_US = 4000.0, REBUILD_PER_PROBLEM_US = 60.0;
    double wasted = 0.0;
    while (running() > 0) {
        step();
        std::vector<int> live;
        for (int p = 0; p < n; ++p)
            if (act_[p]) live.push_back(p);
        wasted += DEAD_US * (T->nprob - (int)live.size());
        if (!live.empty() && wasted >= REBUILD_US + REBUILD_PER_PROBLEM_US * (double)live.size()) {
            T = &full;  // intact, whatever happens to part below
            release(part);
            build_tables(part, live);
            T = &part;
            wasted = 0.0;
        }
    }
    hip_check(hipStreamSynchronize(s), "sync");
    int most = 0;
    for (int p = 0; p < n; ++p) {
        if (iters) iters[p] = it_[p];
        if (outcome) outcome[p] = oc_[p];
        most = std::max(most, it_[p]);
    }
    last_cycles = 2 * most;
    return last_cycles;
}

// ---------------------------------------------------------------------------
// Serving loop
// ---------------------------------------------------------------------------
// Size classes of the serving loop: slots of up to 2^22 cells, and larger ones.
static int size_class(const BatchKMG::Problem &q) { return (long long)q.nx * q.ny > (1LL << 22) ? 1 : 0; }

Server::Server(const std::vector<BatchKMG::Problem> &slots, const mgp_solver_opts &o, double budget)
    : opts_(o), budget_(budget) {
    for (const BatchKMG::Problem &q : slots) {
        desc_.push_back(q);
        if (q.map) {
            maps_.push_back(std::make_unique<mgp_mapping>(*q.map));
            desc_.back().map = maps_.back().get();
        }
    }
    where.resize(slots.size());
    std::vector<std::vector<BatchKMG::Problem>> per(2);
    std::vector<std::vector<int>> ids(2);
    for (size_t k = 0; k < slots.size(); ++k) {
        const int c = size_class(slots[k]);
        ids[c].push_back((int)k);
        per[c].push_back(slots[k]);
    }
    try {
        for (int c = 0; c < 2; ++c) {
            if (per[c].empty()) continue;
            Class C;
            hip_check(hipStreamCreateWithFlags(&C.s, hipStreamNonBlocking), "hipStreamCreate");
            cls.push_back(C);
            cls.back().B = new BatchKMG(cls.back().s, per[c], o);
            cls.back().slot = ids[c];
            for (size_t q = 0; q < ids[c].size(); ++q) where[ids[c][q]] = {(int)cls.size() - 1, (int)q};
        }
    } catch (...) {
        for (Class &C : cls) {
            delete C.B;
            if (C.s) (void)hipStreamDestroy(C.s);
        }
        throw;
    }
    if (budget > 0.0 && (double)footprint() > budget) {
        g_last_error = "server: the resident hierarchies need " + std::to_string(footprint()) +
                       " bytes of HBM, over the budget of " + std::to_string((long long)budget);
        for (Class &C : cls) {
            delete C.B;
            if (C.s) (void)hipStreamDestroy(C.s);
        }
        throw Failure(MGP_ERR_NOMEM);
    }
    owner.assign(slots.size(), 0);
}

void Server::free_class(Class &C) {
    if (C.s) (void)hipStreamSynchronize(C.s);
    delete C.B;
    C.B = nullptr;
    if (C.s) (void)hipStreamDestroy(C.s);
    C.s = nullptr;
    C.slot.clear();
    C.in_flight = false;
}

Server::~Server() {
    for (Class &C : cls) free_class(C);
}

size_t Server::footprint() const {
    size_t b = 0;
    for (const Class &C : cls)
        if (C.B) b += C.B->footprint();
    return b;
}

int Server::add_slot(const BatchKMG::Problem &q, const mgp_mapping *map) {
    if (!map) {
        g_last_error = "server add_slot: the geometry needs a mapping";
        throw Failure(MGP_ERR_ARG);
    }
    BatchKMG::Problem p = q;
    auto m = std::make_unique<mgp_mapping>(*map);
    p.map = m.get();
    const size_t need = BatchKMG::problem_bytes(p, opts_);
    if (budget_ > 0.0 && (double)(footprint() + need) > budget_) {
        g_last_error = "server add_slot: the geometry's hierarchy needs " + std::to_string(need) + " bytes; with the " +
                       std::to_string(footprint()) + " resident that exceeds the budget of " +
                       std::to_string((long long)budget_);
        throw Failure(MGP_ERR_NOMEM);
    }
    Class C;
    hip_check(hipStreamCreateWithFlags(&C.s, hipStreamNonBlocking), "hipStreamCreate");
    try {
        C.B = new BatchKMG(C.s, std::vector<BatchKMG::Problem>{p}, opts_);
    } catch (...) {
        (void)hipStreamDestroy(C.s);
        throw;
    }
    const int slot = slots();
    C.slot = {slot};
    cls.push_back(C);
    maps_.push_back(std::move(m));
    desc_.push_back(p);
    where.push_back({(int)cls.size() - 1, 0});
    owner.push_back(0);
    return slot;
}

void Server::evict(int slot) {
    if (!resident(slot)) {
        g_last_error = "server evict: the slot is not resident";
        throw Failure(MGP_ERR_ARG);
    }
    bool queued_for = false;
    for (const Req &r : queue) queued_for = queued_for || r.slot == slot;
    Class &c = cls[where[slot].first];
    if (owner[slot] != 0 || queued_for || c.in_flight || c.B->running() > 0) {
        g_last_error = "server evict: the slot (or its batch) has requests queued or running";
        throw Failure(MGP_ERR_ARG);
    }
    if (c.slot.size() == 1) {
        free_class(c);
        where[slot] = {-1, -1};
        return;
    }
    // rebuild the batch over the other slots, keeping their warm starts
    std::vector<int> keep;
    std::vector<BatchKMG::Problem> probs;
    std::vector<std::vector<double>> phi;
    for (size_t q = 0; q < c.slot.size(); ++q) {
        if (c.slot[q] == slot) continue;
        keep.push_back(c.slot[q]);
        probs.push_back(desc_[c.slot[q]]);
        const Grid &g = c.B->grid((int)q);
        phi.emplace_back((size_t)g.nx * g.ny);
        download_block(c.s, phi.back().data(), c.B->vec((int)q, BV_PHI), g, 0, 0, g.nx, g.ny);
    }
    hip_check(hipStreamSynchronize(c.s), "server evict");
    BatchKMG *nb = new BatchKMG(c.s, probs, opts_);
    for (size_t q = 0; q < keep.size(); ++q) {
        const Grid &g = nb->grid((int)q);
        upload_block(c.s, nb->vec((int)q, BV_PHI), g, phi[q].data(), 0, 0, g.nx, g.ny);
    }
    hip_check(hipStreamSynchronize(c.s), "server evict");
    delete c.B;
    c.B = nb;
    c.slot = keep;
    for (size_t q = 0; q < keep.size(); ++q) where[keep[q]].second = (int)q;
    where[slot] = {-1, -1};
}

long long Server::submit(int slot, const double *rhs, double tol, int maxiters, int warm) {
    if (slot < 0 || slot >= slots() || !rhs || !(tol >= 0.0) || maxiters < 0) {
        g_last_error = "server submit: bad slot, rhs, tol or maxiters";
        throw Failure(MGP_ERR_ARG);
    }
    if (!resident(slot)) {
        g_last_error = "server submit: the slot was evicted";
        throw Failure(MGP_ERR_ARG);
    }
    const Grid &g = cls[where[slot].first].B->grid(where[slot].second);
    Req r;
    r.ticket = next++;
    r.slot = slot;
    r.tol = tol;
    r.maxiters = maxiters;
    r.warm = warm;
    r.rhs.assign(rhs, rhs + (size_t)g.nx * g.ny);
    queue.push_back(std::move(r));
    return queue.back().ticket;
}

// The oldest queued request of every idle slot of class c starts.
void Server::admit(Class &c, int ci) {
    std::vector<int> which, warm, maxit;
    std::vector<double> tol;
    std::vector<Req> taken;  // kept until the uploads from their rhs have completed
    for (size_t q = 0; q < queue.size();) {
        Req &r = queue[q];
        if (where[r.slot].first != ci || owner[r.slot] != 0) {
            ++q;
            continue;
        }
        owner[r.slot] = r.ticket;
        taken.push_back(std::move(r));
        queue.erase(queue.begin() + (long)q);
    }
    if (taken.empty()) return;
    for (const Req &r : taken) {
        const int local = where[r.slot].second;
        const Grid &g = c.B->grid(local);
        upload_block(c.s, c.B->vec(local, BV_RHS), g, r.rhs.data(), 0, 0, g.nx, g.ny);
        which.push_back(local);
        warm.push_back(r.warm);
        tol.push_back(r.tol);
        maxit.push_back(r.maxiters);
    }
    c.B->start(which, warm, tol, maxit);
    hip_check(hipStreamSynchronize(c.s), "server admit");
    std::vector<int> at_once;
    for (int p : which)
        if (!c.B->is_running(p)) at_once.push_back(p);  // met the tolerance at the start
    retire(c, at_once);
}

void Server::retire(Class &c, const std::vector<int> &finished) {
    for (int p : finished) {
        const Grid &g = c.B->grid(p);
        Res r;
        r.phi.resize((size_t)g.nx * g.ny);
        download_block(c.s, r.phi.data(), c.B->vec(p, BV_PHI), g, 0, 0, g.nx, g.ny);
        r.iters = c.B->iterations(p);
        r.outcome = c.B->outcome(p);
        const int slot = c.slot[p];
        done.emplace_back(owner[slot], std::move(r));
        owner[slot] = 0;
    }
    if (!finished.empty()) hip_check(hipStreamSynchronize(c.s), "server results");
}

int Server::step() {
    // Every class whose iteration has finished takes its results, admits
    // queued requests and launches its next iteration; a class whose
    // iteration still runs is left alone.  Returns once at least one class
    // has finished an iteration (or none has work).
    bool progressed = false;
    while (true) {
        bool any_in_flight = false;
        for (size_t ci = 0; ci < cls.size(); ++ci) {
            Class &c = cls[ci];
            if (!c.B) continue;  // every slot of it evicted
            if (c.in_flight) {
                if (!c.B->step_done()) {
                    any_in_flight = true;
                    continue;
                }
                c.in_flight = false;
                progressed = true;
                retire(c, c.B->finish_step());
            }
            admit(c, (int)ci);
            if (c.B->running() > 0) {
                c.B->launch_step();
                c.in_flight = true;
                any_in_flight = true;
            }
        }
        if (progressed || !any_in_flight) break;
        std::this_thread::yield();
    }
    int running = 0;
    for (const Class &c : cls)
        if (c.B) running += c.B->running();
    return running;
}

int Server::result(long long ticket, double *phi, int *iters, int *outcome) {
    for (size_t q = 0; q < done.size(); ++q)
        if (done[q].first == ticket) {
            const Res &r = done[q].second;
            if (phi) std::copy(r.phi.begin(), r.phi.end(), phi);
            if (iters) *iters = r.iters;
            if (outcome) *outcome = r.outcome;
            done.erase(done.begin() + (long)q);
            return 1;
        }
    return 0;
}

}  // namespace mgp
