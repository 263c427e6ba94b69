// mgp_xline.hip -- x-line relaxation of the resident fast solver on gfx950.
//
// A sweep of one row colour solves, for every row j of the colour,
//     T_j delta = f - L e     (row j),   e(:, j) += delta
// where L is the reference 9-point operator with homogeneous BCs (ghosts
// eliminated) and T_j its tridiagonal part along the row: the coefficients
// of e(i-1,j), e(i,j), e(i+1,j) in (L e)(i,j).
//
// Interior cells evaluate L and T from the face metrics already loaded for
// the residual (cell_lap_interior's stencil, differentiated by hand).  Ring
// cells (i = 0, nx-1 or j = 0, ny-1), whose stencils carry the boundary
// extrapolations and the Dirichlet ghost rule, use an explicit coefficient
// table probed once per level with unit impulses (k_ring_coef).
//
// One workgroup (256 threads) per row:
//  1. coalesced pass: each thread evaluates d = f - L e and the row
//     coefficients for cells i = tid + 256 u (all loads issued up front);
//  2. the values go through LDS into contiguous chunks of C or C+1 cells
//     per thread (padded stride C+1: no bank conflicts);
//  3. partition method in registers: each chunk is eliminated towards its
//     two ends, leaving a tridiagonal system of 2 nt unknowns;
//  4. parallel cyclic reduction of that system in LDS (log2(2 nt) steps);
//  5. back substitution in registers, transpose back, coalesced e update.
// Rows are mapped to workgroups so that the rows an XCD processes are
// contiguous (workgroups are dealt to the 8 XCDs round robin): neighbouring
// rows share their halo rows through that XCD's L2.
#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>

#include "mgp_batch.h"
#include "mgp_host.h"
#include "mgp_kernels.h"
#include "mgp_stencil.h"

namespace mgp {

namespace {

// Threads per row of the line solve: 256; in the batched launches (many
// rows) 64 for rows of at most 512 cells and 128 up to 1024 cells, so that
// a CU holds more rows at once (a short row then occupies one or two waves).
constexpr int XL_THREADS = 256;
constexpr int XL_SHORT = 64;
constexpr int XL_SHORT_MAX_NX = 512;
constexpr int XL_MID = 128;
constexpr int XL_MID_MAX_NX = 1024;
// FP32 rows of 4112 .. 8207 cells: 512 threads, chunks of 16 (139 KB of LDS),
// so the mixed-precision cycle solves rows of 8192 cells whole.
constexpr int XL_WIDE = 512;
constexpr int RING_PLANES = 9;

// Impulse whose ghosts follow the homogeneous fill of the MG levels
// (Dirichlet order 3 / LAH negate; Neumann sides contribute no ghost).
struct ImpulseFilled {
    int ci, cj;
    Grid g;
    int t[4], m[4];
    __device__ double v(int i, int j) const { return (i == ci && j == cj) ? 1.0 : 0.0; }
    __device__ static double dg(int type, int mode, double p0, double p1, double p2) {
        if (type != BC_DIRI) return 0.0;
        if (mode == MODE_LAH) return -p0;
        return -3.0 * p0 + p1 - (1.0 / 5.0) * p2;
    }
    __device__ double operator()(int i, int j) const {
        const int nx = g.nx, ny = g.ny;
        const bool xin = i >= 0 && i < nx, yin = j >= 0 && j < ny;
        if (xin && yin) return v(i, j);
        if (!xin && !yin) return 0.0;
        if (i < 0) return dg(t[XLO], m[XLO], v(0, j), v(1, j), v(2, j));
        if (i >= nx) return dg(t[XHI], m[XHI], v(nx - 1, j), v(nx - 2, j), v(nx - 3, j));
        if (j < 0) return dg(t[YLO], m[YLO], v(i, 0), v(i, 1), v(i, 2));
        return dg(t[YHI], m[YHI], v(i, ny - 1), v(i, ny - 2), v(i, ny - 3));
    }
};

// Device-memory pointers typed as global: the compiler then knows they
// cannot alias LDS and may issue loads ahead of LDS stores (with generic
// pointers every LDS store would wait for the loads before it).
template <typename T>
using gconst = const __attribute__((address_space(1))) T;
template <typename T>
__device__ __forceinline__ gconst<T> *gptr(const T *p) {
    return (gconst<T> *)p;
}

__host__ __device__ inline int ring_count(int nx, int ny) { return 2 * nx + 2 * (ny - 2); }

// Every ring-cell stencil lies in a 3x3 window of valid cells: the window
// reaches inward (two cells) across each boundary side the cell touches and
// one cell either way along it.  Window origin (w0, v0) of cell (i, j):
//   w0 = 0 on column 0, -2 on column nx-1, else -1   (likewise v0 for rows).
// Ring cell index r: rows j = 0 (r = i) and ny-1 (r = nx + i), then columns
// i = 0 (r = 2 nx + j - 1) and nx-1 (r = 2 nx + ny - 2 + j - 1), 1 <= j <= ny-2.
// Table plane q = 3 a + b holds the coefficient of e(i + w0 + b, j + v0 + a).
__device__ __forceinline__ int window_origin(int i, int n) { return i == 0 ? 0 : (i == n - 1 ? -2 : -1); }

__device__ inline void ring_cell(int nx, int ny, int r, int &i, int &j) {
    if (r < 2 * nx) {
        const bool top = r >= nx;
        i = top ? r - nx : r;
        j = top ? ny - 1 : 0;
    } else {
        const int rr = r - 2 * nx;
        const bool right = rr >= ny - 2;
        i = right ? nx - 1 : 0;
        j = (right ? rr - (ny - 2) : rr) + 1;
    }
}

__device__ __forceinline__ int ring_index(int nx, int ny, int i, int j) {
    if (j == 0) return i;
    if (j == ny - 1) return nx + i;
    return 2 * nx + (i == 0 ? 0 : ny - 2) + (j - 1);
}

// (L e)(i, j) at a ring cell from the table.  All nine table entries and
// field values are loaded before any is used (one memory round trip).
template <typename T>
__device__ __forceinline__ T ring_lap(gconst<T> *e, gconst<T> *K, int nring, const Grid &g, int i, int j) {
    const int w0 = window_origin(i, g.nx), v0 = window_origin(j, g.ny);
    const int r = ring_index(g.nx, g.ny, i, j);
    T kv[9], ev[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        kv[q] = K[(size_t)q * nring + r];
        ev[q] = e[ix(g.pitch, i + w0 + q % 3, j + v0 + q / 3)];
    }
    T lap = 0;
#pragma unroll
    for (int q = 0; q < 9; ++q) lap += kv[q] * ev[q];
    return lap;
}

// Coefficients of e(i-1, j), e(i, j), e(i+1, j) at a ring cell (0 where the
// neighbour is outside the window, i.e. not a valid cell).
template <typename T>
__device__ __forceinline__ void ring_row_coef(gconst<T> *K, int nring, const Grid &g, int i, int j, T &a, T &b,
                                              T &c) {
    const int w0 = window_origin(i, g.nx), v0 = window_origin(j, g.ny);
    const int r = ring_index(g.nx, g.ny, i, j);
    const int base = 3 * (-v0);  // window row of j itself
    a = w0 <= -1 ? K[(size_t)(base + (-1 - w0)) * nring + r] : T(0);
    b = K[(size_t)(base + (0 - w0)) * nring + r];
    c = w0 + 2 >= 1 ? K[(size_t)(base + (1 - w0)) * nring + r] : T(0);
}

// K[q * nring + r]: coefficient of e(i + w0 + q%3, j + v0 + q/3) in the
// homogeneous L at ring cell r (impulse response with the ghost rule).
__global__ void __launch_bounds__(256) k_ring_coef(double *K, Grid g, Geo geo, Bc bc) {
    const int nring = ring_count(g.nx, g.ny);
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= RING_PLANES * nring) return;
    const int q = id / nring, r = id % nring;
    int i, j;
    ring_cell(g.nx, g.ny, r, i, j);
    ImpulseFilled imp;
    imp.g = g;
    for (int k = 0; k < 4; ++k) {
        imp.t[k] = bc.type[k];
        imp.m[k] = bc.mode[k];
    }
    imp.ci = i + window_origin(i, g.nx) + q % 3;
    imp.cj = j + window_origin(j, g.ny) + q / 3;
    K[id] = cell_lap(imp, g, geo, bc, 1, i, j);
}

// Position of cell i in the chunk layout: the first rem chunks hold C+1
// cells, the others C; chunk t starts at LDS slot t*(C+1).
template <int C>
__device__ __forceinline__ int chunk_pos(int i, int rem) {
    const int big = rem * (C + 1);
    if (i < big) return i;
    const int k = i - big;
    return big + (k / C) * (C + 1) + k % C;
}

// Row j of the workgroup's launch.  XCD-aware: XCD x (= blockIdx % 8) takes a
// contiguous block of q8 or q8+1 slots.
__device__ __forceinline__ int launch_slot() {
    const int nb = gridDim.x, b = blockIdx.x, x = b & 7;
    const int q8 = nb >> 3, r8 = nb & 7;
    return x * q8 + (x < r8 ? x : r8) + (b >> 3);
}

template <typename T>
struct RowCtx {
    int nx, ny, P, nt, rem, nring;
    int i0, nxg;  // the cells solved: i0 .. i0+nx-1 of a row of nxg cells (a whole row: 0, nx)
    T idx, idy, idx2, idy2, ixy4;
};

// The inverse spacings are formed in FP64 and rounded once to T.
template <typename T>
__device__ __forceinline__ RowCtx<T> row_ctx(const Grid &g, int nt, int C) {
    RowCtx<T> r;
    r.nx = g.nx;
    r.i0 = 0;
    r.nxg = g.nx;
    r.ny = g.ny;
    r.P = g.pitch;
    r.nt = nt;
    r.rem = g.nx - nt * C;
    r.nring = ring_count(g.nx, g.ny);
    const double idx = 1.0 / g.dx, idy = 1.0 / g.dy;
    r.idx = (T)idx;
    r.idy = (T)idy;
    r.idx2 = (T)(idx * idx);
    r.idy2 = (T)(idy * idy);
    r.ixy4 = (T)(0.25 * idx * idy);
    return r;
}

// Coalesced pass over row j: for cells i = tid + 256 u, the residual
// d = f - L e and the row coefficients a, b, c, written into the chunk layout
// (Ld, La, Lb, Lc).  Interior cells use the metric stencil, ring cells the
// table.  Lanes past the row end evaluate the last cell and store into the
// chunk layout's last slot, which no chunk uses (the last chunk has C cells):
// the loop stays branch-free.
//
// RG cells' loads are issued together before their stencils are evaluated,
// so RG memory round trips overlap instead of following one another (one
// cell at a time, a coarse-level row pass took 5.4 us of a 10.6 us row
// solve).  The batched kernels take the RG of batch_rg: no more than fits
// their VGPR budget, since they rely on many workgroups per CU.
// Buffer-addressed reads of a field of a level: one 32-bit byte offset per
// cell in a vector register, the row offsets (multiples of the pitch) in
// scalar registers and the column neighbours in the instruction's immediate
// offset.  Flat loads need a 64-bit address per array and neighbour: in the
// batched mid-row kernel 350 of its 2700 vector instructions were 64-bit
// address arithmetic.  A buffer of 2^31 - 1 bytes from the start of the
// row above the one solved: the reads span three rows, so none is clipped.
template <typename T>
struct Buf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ explicit Buf(const T *p)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(p), 0, 0x7fffffff, 0x00020000)) {}
    __device__ __forceinline__ T operator()(int vo, int so) const {
        if constexpr (sizeof(T) == 4)
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
        else
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
    }
};

template <int C, int NT, bool SEG = false, int RG = 1, typename T, typename GM>
__device__ __forceinline__ void row_pass(const RowCtx<T> &X, const T *e_, const T *f_, const T *K_, const Grid &g,
                                         const GM &geo_, int j, T *Ld, T *La, T *Lb, T *Lc) {
    constexpr int U = C + 1;
    gconst<T> *e = gptr(e_), *f = gptr(f_), *K = gptr(K_);
    // SEG: the cells solved may be a segment of the row (xline_rows); whole
    // rows compile without the segment terms
    const int tid = threadIdx.x, nx = X.nx, ny = X.ny, P = X.P, i0 = SEG ? X.i0 : 0, nxg = SEG ? X.nxg : X.nx;
    const T fourth = T(0.25);
    const int dump = X.nt * (C + 1) - 1;
    // A segment of a longer row is solved as a line of its own: its couplings
    // to the cells beyond its ends (in the neighbouring segments) stay in the
    // residual and leave its tridiagonal system (a = 0 at its first cell,
    // c = 0 at its last), as in block Gauss-Seidel over the segments.
    const bool cut_w = SEG && i0 > 0, cut_e = SEG && i0 + nx < nxg;
    if (j != 0 && j != ny - 1) {
        // the buffers start at cell (-1, j-1) of each array (j is the same
        // for the whole workgroup), so the offsets are small and never
        // negative whatever the array's size
        const size_t r0 = ix(P, 0, j - 1) - 1;
        const Buf<T> bE(e_ + r0), bJxx(geo_.Jxx + r0), bJxy(geo_.Jxy + r0), bJyy(geo_.Jyy + r0),
            bJyx(geo_.Jyx + r0), bF(f_ + r0);
        struct CellIn {
            T p, pe, pw, pn, ps, ene, enw, ese, esw, Jxx_e, Jxx_w, Jxy_e, Jxy_w, Jyy_n, Jyy_s, Jyx_n, Jyx_s, fv;
        };
#pragma unroll
        for (int u0 = 0; u0 < U; u0 += RG) {
            CellIn in[RG];
#pragma unroll
            for (int r = 0; r < RG; ++r) {
                const int u = u0 + r;
                if (u >= U) break;
                const int iraw = tid + NT * u;
                const int il = iraw < nx ? iraw : nx - 1;
                const size_t c = ix(P, i0 + il, j);
                // interior stencil; for i = 0, nx-1 the reads stay inside the
                // padded level and the values are replaced below
                // bytes from cell (i-1, j-1) (i elements past the buffer's
                // start); rows j and j+1 one and two pitches on
                constexpr int ES = (int)sizeof(T);
                const int vo = (int)((c - P - 1 - r0) * ES), s1 = P * ES, s2 = 2 * P * ES;
                in[r].esw = bE(vo, 0);
                in[r].ps = bE(vo + ES, 0);
                in[r].ese = bE(vo + 2 * ES, 0);
                in[r].pw = bE(vo, s1);
                in[r].p = bE(vo + ES, s1);
                in[r].pe = bE(vo + 2 * ES, s1);
                in[r].enw = bE(vo, s2);
                in[r].pn = bE(vo + ES, s2);
                in[r].ene = bE(vo + 2 * ES, s2);
                in[r].Jxx_w = bJxx(vo + ES, s1);
                in[r].Jxx_e = bJxx(vo + 2 * ES, s1);
                in[r].Jxy_w = bJxy(vo + ES, s1);
                in[r].Jxy_e = bJxy(vo + 2 * ES, s1);
                in[r].Jyy_s = bJyy(vo + ES, s1);
                in[r].Jyy_n = bJyy(vo + ES, s2);
                in[r].Jyx_s = bJyx(vo + ES, s1);
                in[r].Jyx_n = bJyx(vo + ES, s2);
                in[r].fv = bF(vo + ES, s1);
            }
#pragma unroll
            for (int r = 0; r < RG; ++r) {
                const int u = u0 + r;
                if (u >= U) break;
                const int iraw = tid + NT * u;
                const int il = iraw < nx ? iraw : nx - 1;
                const CellIn &v = in[r];
                const T xfe = v.Jxx_e * (v.pe - v.p) * X.idx + v.Jxy_e * fourth * (v.ene - v.ese + v.pn - v.ps) * X.idy;
                const T xfw = v.Jxx_w * (v.p - v.pw) * X.idx + v.Jxy_w * fourth * (v.pn - v.ps + v.enw - v.esw) * X.idy;
                const T yfn = v.Jyy_n * (v.pn - v.p) * X.idy + v.Jyx_n * fourth * (v.ene - v.enw + v.pe - v.pw) * X.idx;
                const T yfs = v.Jyy_s * (v.p - v.ps) * X.idy + v.Jyx_s * fourth * (v.pe - v.pw + v.ese - v.esw) * X.idx;
                const int o = iraw < nx ? chunk_pos<C>(il, X.rem) : dump;
                Ld[o] = v.fv - ((xfe - xfw) * X.idx + (yfn - yfs) * X.idy);
                La[o] = cut_w && il == 0 ? T(0) : v.Jxx_w * X.idx2 - (v.Jyx_n - v.Jyx_s) * X.ixy4;
                Lb[o] = -(v.Jxx_e + v.Jxx_w) * X.idx2 - (v.Jyy_n + v.Jyy_s) * X.idy2;
                Lc[o] = cut_e && il == nx - 1 ? T(0) : v.Jxx_e * X.idx2 + (v.Jyx_n - v.Jyx_s) * X.ixy4;
            }
        }
        // the two ring cells of the row (i = 0, nxg-1), where they lie in
        // the cells solved: explicit stencil table
        for (int side = 0; side < 2; ++side) {
            const int iend = side ? nxg - 1 : 0, il = iend - i0;
            if (il < 0 || il >= nx || il % NT != tid) continue;
            const int o = chunk_pos<C>(il, X.rem);
            Ld[o] = f[ix(P, iend, j)] - ring_lap(e, K, X.nring, g, iend, j);
            ring_row_coef(K, X.nring, g, iend, j, La[o], Lb[o], Lc[o]);
        }
    } else {
        // boundary row: every cell from the table
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int iraw = tid + NT * u;
            const int il = iraw < nx ? iraw : nx - 1;
            const int i = i0 + il;
            const T fv = f[ix(P, i, j)];
            const T lap = ring_lap(e, K, X.nring, g, i, j);
            T a, b, c;
            ring_row_coef(K, X.nring, g, i, j, a, b, c);
            const int o = iraw < nx ? chunk_pos<C>(il, X.rem) : dump;
            Ld[o] = fv - lap;
            La[o] = cut_w && il == 0 ? T(0) : a;
            Lb[o] = b;
            Lc[o] = cut_e && il == nx - 1 ? T(0) : c;
        }
    }
}

// Coalesced update e(:, j) += x, x in the chunk layout (e re-read).
template <int C, int NT, bool SEG = false, typename T>
__device__ __forceinline__ void row_update_add(const RowCtx<T> &X, T *e, int j, const T *Lx) {
#pragma unroll
    for (int u = 0; u < C + 1; ++u) {
        const int i = threadIdx.x + NT * u;
        if (i < X.nx) e[ix(X.P, (SEG ? X.i0 : 0) + i, j)] += Lx[chunk_pos<C>(i, X.rem)];
    }
}

// ---------------------------------------------------------------------------
// Interface solve of the row system, one wave.  The 2 NT interface rows (2
// per thread of the workgroup; NT = 256: 512, L = 8 per lane; NT = 64: 128,
// L = 2) sit in LDS lane-major: row q at iface_pos(q) = (q % L) * 64 + q / L,
// so lane l reads its rows L l .. L l + L-1 without bank conflicts.  Lane l
//  1. eliminates its L rows to 2 (the partition step of the chunks): its
//     first row in terms of x(Ll-1), x(Ll), x(Ll+L-1), its last in terms of
//     x(Ll), x(Ll+L-1), x(Ll+L);
//  2. takes part in parallel cyclic reduction of those 128 rows (7 steps,
//     neighbours exchanged with cross-lane shuffles, no barrier);
//  3. evaluates its L-2 inner unknowns from the first and last ones (their
//     linear forms were set up in step 1) and writes x(Ll .. Ll+L-1) to X in
//     natural order.
// ---------------------------------------------------------------------------
template <int NT>
constexpr int iface_rows() {
    return 2 * NT;  // interface rows, padded
}

// Quotients of the line solves: v_rcp_f64 refined by one Newton step (error
// of order eps^2 relative, three dependent instructions instead of the
// correctly rounded division sequence).  The line solves are part of the
// preconditioner, so they need accuracy, not bit-exact rounding.
__device__ __forceinline__ double rcp(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}
// FP32: v_rcp_f32 alone (about 1 ulp).
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
template <typename T>
__device__ __forceinline__ T quot(T a, T b) {
    return a * rcp(b);
}

template <int NT>
__device__ __forceinline__ int iface_pos(int q) {
    constexpr int L = NT / 32;
    return (q % L) * 64 + q / L;
}

template <typename T>
struct Row4 {
    T a, b, c, d;
};

template <typename T>
__device__ __forceinline__ Row4<T> shfl_up4(const Row4<T> &r, int s) {
    return Row4<T>{__shfl_up(r.a, s), __shfl_up(r.b, s), __shfl_up(r.c, s), __shfl_up(r.d, s)};
}
template <typename T>
__device__ __forceinline__ Row4<T> shfl_down4(const Row4<T> &r, int s) {
    return Row4<T>{__shfl_down(r.a, s), __shfl_down(r.b, s), __shfl_down(r.c, s), __shfl_down(r.d, s)};
}

// One PCR update of row r from its neighbours m (q - st) and p (q + st);
// a missing neighbour contributes nothing.
template <typename T>
__device__ __forceinline__ Row4<T> pcr_update(const Row4<T> &r, const Row4<T> &m, bool has_m, const Row4<T> &p,
                                              bool has_p) {
    const T k1 = has_m ? quot(r.a, m.b) : T(0), k2 = has_p ? quot(r.c, p.b) : T(0);
    Row4<T> o;
    o.a = has_m ? -m.a * k1 : T(0);
    o.c = has_p ? -p.c * k2 : T(0);
    o.b = r.b - (has_m ? m.c * k1 : T(0)) - (has_p ? p.a * k2 : T(0));
    o.d = r.d - (has_m ? m.d * k1 : T(0)) - (has_p ? p.d * k2 : T(0));
    return o;
}

template <int NT, typename T>
__device__ __forceinline__ void interface_solve(T *IA, T *IB, const T *IC, T *ID, T *IX, int l) {
    constexpr int L = NT / 32;
    T a[L], b[L], c[L], d[L];
#pragma unroll
    for (int k = 0; k < L; ++k) {
        const int q = k * 64 + l;
        a[k] = IA[q];
        b[k] = IB[q];
        c[k] = IC[q];
        d[k] = ID[q];
    }
    // 1. partition: first row (backward sweep), last row (forward sweep)
    T bu = b[L - 2], cu = c[L - 2], du = d[L - 2];
#pragma unroll
    for (int k = L - 3; k >= 0; --k) {
        const T r = quot(c[k], bu);
        bu = b[k] - r * a[k + 1];
        cu = -r * cu;
        du = d[k] - r * du;
    }
    Row4<T> f{a[0], bu, cu, du};
#pragma unroll
    for (int k = 2; k < L; ++k) {
        const T r = quot(a[k], b[k - 1]);
        a[k] = -r * a[k - 1];
        b[k] = b[k] - r * c[k - 1];
        d[k] = d[k] - r * d[k - 1];
    }
    Row4<T> g{a[L - 1], b[L - 1], c[L - 1], d[L - 1]};
    // inner unknowns as x_k = G_k + H_k x_first + M_k x_last, kept in this
    // lane's own LDS slots (IA, IB, ID) for step 3, so that the PCR runs with
    // only the two lane rows in registers and step 3 has no dependent chain
    {
        T G = 0, H = 0, M = 1;
#pragma unroll
        for (int k = L - 2; k >= 1; --k) {
            const T ib = rcp(b[k]);
            G = (d[k] - c[k] * G) * ib;
            H = (-a[k] - c[k] * H) * ib;
            M = (-c[k] * M) * ib;
            const int q = k * 64 + l;
            IA[q] = G;
            IB[q] = H;
            ID[q] = M;
        }
    }
    asm volatile("" ::: "memory");
    // 2. PCR over the 128 lane rows, Q = 2l (first) and 2l+1 (last)
#pragma unroll
    for (int st = 1; st < 128; st *= 2) {
        Row4<T> fm, fp, gm, gp;
        if (st == 1) {
            fm = shfl_up4(g, 1);   // Q-1 of the first row: the previous lane's last
            fp = g;                // Q+1: this lane's last
            gm = f;                // Q-1 of the last row: this lane's first
            gp = shfl_down4(f, 1); // Q+1: the next lane's first
        } else {
            const int sl = st / 2;
            fm = shfl_up4(f, sl);
            fp = shfl_down4(f, sl);
            gm = shfl_up4(g, sl);
            gp = shfl_down4(g, sl);
        }
        const int Qf = 2 * l, Qg = 2 * l + 1;
        const Row4<T> nf = pcr_update(f, fm, Qf - st >= 0, fp, Qf + st < 128);
        const Row4<T> ng = pcr_update(g, gm, Qg - st >= 0, gp, Qg + st < 128);
        f = nf;
        g = ng;
    }
    const T xs = quot(f.d, f.b), xe = quot(g.d, g.b);
    // 3. back substitution of the inner unknowns
    asm volatile("" ::: "memory");
    IX[L * l] = xs;
    IX[L * l + L - 1] = xe;
#pragma unroll
    for (int k = 1; k <= L - 2; ++k) {
        const int q = k * 64 + l;
        IX[L * l + k] = IA[q] + IB[q] * xs + ID[q] * xe;
    }
}

// ---------------------------------------------------------------------------
// On-the-fly solve of one row (fine levels, bandwidth-bound): coefficients
// from the metrics, elimination, PCR and back substitution per call.
// LDS: 4 nt (C+1) doubles.
// ---------------------------------------------------------------------------
template <int C, int NT, bool SEG = false, int RG = 1, typename T, typename GM>
__device__ void xline_row(const RowCtx<T> &X, T *__restrict__ e, const T *__restrict__ f, const T *__restrict__ K,
                          const Grid &g, const GM &geo, int j, T *lds) {
    constexpr int CP = C + 1;
    const int tid = threadIdx.x, nt = X.nt;
    const int NT1 = nt * CP;
    T *Ld = lds, *Lb = lds + NT1, *La = lds + 2 * NT1, *Lc = lds + 3 * NT1;
    row_pass<C, NT, SEG, RG>(X, e, f, K, g, geo, j, Ld, La, Lb, Lc);
    __syncthreads();

    // each thread takes its chunk into registers
    T ca[CP], cb[CP], cc[CP], cd[CP];
    const int len = tid < nt ? C + (tid < X.rem ? 1 : 0) : 0;
#pragma unroll
    for (int k = 0; k < CP; ++k)
        if (k < len) {
            cd[k] = Ld[tid * CP + k];
            cb[k] = Lb[tid * CP + k];
            ca[k] = La[tid * CP + k];
            cc[k] = Lc[tid * CP + k];
        }
    __syncthreads();

    // partition elimination in registers
    // start row:  a_0 x_{prev end} + bu x_s + cu x_e = du
    // end row:    a'_L x_s + b'_L x_e + c_L x_{next start} = d'_L
    T ea = 0, eb = 1, ec = 0, ed = 0, oa = 0, ob = 1, oc = 0, od = 0;
    T sg[CP], sh[CP], sm[CP];
    if (tid < nt) {
        const bool longc = len == CP;
        T bu = longc ? cb[C - 1] : cb[C - 2];
        T cu = longc ? cc[C - 1] : cc[C - 2];
        T du = longc ? cd[C - 1] : cd[C - 2];
#pragma unroll
        for (int k = C - 2; k >= 0; --k)
            if (k <= len - 3) {
                const T r = quot(cc[k], bu);
                bu = cb[k] - r * ca[k + 1];
                cu = -r * cu;
                du = cd[k] - r * du;
            }
        ea = ca[0];
        eb = bu;
        ec = cu;
        ed = du;
#pragma unroll
        for (int k = 2; k <= C; ++k)
            if (k <= len - 1) {
                const T r = quot(ca[k], cb[k - 1]);
                ca[k] = -r * ca[k - 1];
                cb[k] = cb[k] - r * cc[k - 1];
                cd[k] = cd[k] - r * cd[k - 1];
            }
        oa = longc ? ca[C] : ca[C - 1];
        ob = longc ? cb[C] : cb[C - 1];
        oc = longc ? cc[C] : cc[C - 1];
        od = longc ? cd[C] : cd[C - 1];
        // inner unknowns as x_k = sg_k + sh_k x_s + sm_k x_e (the back
        // substitution unrolled once here, off the critical path of the
        // interface solve): ca/cb/cc/cd are dead after this
        T G = 0, H = 0, M = 1;
#pragma unroll
        for (int k = C - 1; k >= 1; --k)
            if (k <= len - 2) {
                const T ib = rcp(cb[k]);
                G = (cd[k] - cc[k] * G) * ib;
                H = (-ca[k] - cc[k] * H) * ib;
                M = (-cc[k] * M) * ib;
                sg[k] = G;
                sh[k] = H;
                sm[k] = M;
            }
    }

    // Interface system: the 2 nt chunk-end unknowns (start of chunk t at
    // q = 2t, end at q = 2t+1; tridiagonal in that order), padded to 512 rows
    // with identity rows by the threads past nt.  Staged once through LDS,
    // solved by wave 0 (interface_solve), read back once: two barriers.
    constexpr int IFACE = iface_rows<NT>();
    T *IA = lds + NT1, *IB = IA + IFACE, *IC = IB + IFACE, *ID = IC + IFACE, *IX = ID + IFACE;
    {
        const int q0 = iface_pos<NT>(2 * tid), q1 = iface_pos<NT>(2 * tid + 1);
        IA[q0] = ea;
        IB[q0] = eb;
        IC[q0] = ec;
        ID[q0] = ed;
        IA[q1] = oa;
        IB[q1] = ob;
        IC[q1] = oc;
        ID[q1] = od;
    }
    __syncthreads();
    if (tid < 64) interface_solve<NT>(IA, IB, IC, ID, IX, tid);
    __syncthreads();
    const T x_start = IX[2 * tid], x_end = IX[2 * tid + 1];

    // back substitution, then the coalesced update
    if (tid < nt) {
        const T xs = x_start, xe = x_end;
        const bool longc = len == CP;
#pragma unroll
        for (int k = 1; k <= C - 1; ++k)
            if (k <= len - 2) lds[tid * CP + k] = sg[k] + sh[k] * xs + sm[k] * xe;
        lds[tid * CP] = xs;
        lds[tid * CP + (longc ? C : C - 1)] = xe;
    }
    __syncthreads();
    row_update_add<C, NT, SEG>(X, e, j, lds);
}

// Rows j = row0 + slot*rowstep; the workgroup of slot 0 then also relaxes
// row extra0 and the last slot row extra1 (< 0: none), each after its own
// row.  The boundary rows ride on the odd-row launch this way: row 0 after
// row 1 and row ny-1 after the last odd row, the order of a separate third
// phase, without its launch.
//
// Rows longer than one workgroup's line solve are cut into segments of
// seglen cells, each solved as a line (row_pass).  A launch relaxes segments
// seg0, seg0 + segstep, .. of its rows, nsegl of them per row: launch slot s
// takes row slot s / nsegl, segment seg0 + segstep (s % nsegl).  With
// segstep 2 the even segments of a colour run in one launch and the odd ones
// in the next, so no workgroup reads a cell that another one of the same
// launch writes: a segment's residual reaches one cell beyond each end, into
// a segment of the other parity.  A whole row: seglen = nx, nsegl = 1.
template <int C, int NT, bool SEG, int RG = 3, typename T, typename GM>
__device__ __forceinline__ void xline_rows(T *__restrict__ e, const T *__restrict__ f, const T *__restrict__ K,
                                           const Grid &g, const GM &geo, int row0, int rowstep, int nt, int extra0,
                                           int extra1, int seglen, int seg0, int segstep, int nsegl) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_bytes[];
    T *lds = reinterpret_cast<T *>(lds_bytes);
    RowCtx<T> X0 = row_ctx<T>(g, nt, C);
    int slot = launch_slot(), nslots = (int)gridDim.x;
    if (SEG) {
        X0.nx = seglen;
        X0.rem = seglen - nt * C;
        X0.i0 = (seg0 + segstep * (slot % nsegl)) * seglen;
        slot /= nsegl;
        nslots /= nsegl;
    }
    const int ex0 = slot == 0 ? extra0 : -1, ex1 = slot == nslots - 1 ? extra1 : -1;
#pragma unroll 1
    for (int r = 0; r < 3; ++r) {
        const int jr = r == 0 ? row0 + slot * rowstep : (r == 1 ? ex0 : ex1);
        if (jr < 0) continue;  // uniform over the workgroup
        if (r) __syncthreads();
        // the chunk remainder is made opaque per row, so the per-thread LDS
        // positions derived from it are recomputed in the row (a few integer
        // ops) instead of being hoisted out of the loop and spilled
        RowCtx<T> X = X0;
        asm volatile("" : "+s"(X.rem));
        xline_row<C, NT, SEG, RG>(X, e, f, K, g, geo, jr, lds);
    }
}

// Waves per SIMD of the fine-level kernel: the LDS of one workgroup at
// C = 8 (73.7 KB in FP64, 36.9 KB in FP32) allows two workgroups per CU in
// FP64 and four in FP32, and the register budget is set to match.
template <typename T>
constexpr int xl_waves() {
    return sizeof(T) == 8 ? 2 : 3;
}

// Rows up to 2055 cells: the fine levels are bandwidth-bound and need the
// other workgroups of the CU to cover the elimination and interface phases
// of each one.
template <int C, bool SEG, typename T, typename GM>
__global__ void __launch_bounds__(XL_THREADS) __attribute__((amdgpu_waves_per_eu(xl_waves<T>(), xl_waves<T>())))
k_xline(T *__restrict__ e, const T *__restrict__ f, const T *__restrict__ K, Grid g, GM geo, int row0, int rowstep,
        int nt, int extra0, int extra1, int seglen, int seg0, int segstep, int nsegl) {
    xline_rows<C, XL_THREADS, SEG>(e, f, K, g, geo, row0, rowstep, nt, extra0, extra1, seglen, seg0, segstep, nsegl);
}

// Rows of at most 512 cells: one 64-thread wave per row (as in the batch).
template <int C, bool SEG, typename T, typename GM>
__global__ void __launch_bounds__(XL_SHORT) __attribute__((amdgpu_waves_per_eu(2)))
k_xline_short(T *__restrict__ e, const T *__restrict__ f, const T *__restrict__ K, Grid g, GM geo, int row0,
              int rowstep, int nt, int extra0, int extra1, int seglen, int seg0, int segstep, int nsegl) {
    xline_rows<C, XL_SHORT, SEG>(e, f, K, g, geo, row0, rowstep, nt, extra0, extra1, seglen, seg0, segstep, nsegl);
}

// FP32 rows past 4111 cells (C = 16, 512 threads): one workgroup per CU.
template <int C, typename T, typename GM>
__global__ void __launch_bounds__(XL_WIDE) k_xline_wide(T *__restrict__ e, const T *__restrict__ f,
                                                        const T *__restrict__ K, Grid g, GM geo, int row0,
                                                        int rowstep, int nt, int extra0, int extra1, int seglen,
                                                        int seg0, int segstep, int nsegl) {
    xline_rows<C, XL_WIDE, false>(e, f, K, g, geo, row0, rowstep, nt, extra0, extra1, seglen, seg0, segstep, nsegl);
}

// Longer rows (C = 16): one workgroup per CU in FP64 (139 KB of LDS), all registers.
template <int C, bool SEG, typename T, typename GM>
__global__ void __launch_bounds__(XL_THREADS) k_xline_long(T *__restrict__ e, const T *__restrict__ f,
                                                           const T *__restrict__ K, Grid g, GM geo, int row0,
                                                           int rowstep, int nt, int extra0, int extra1,
                                                           int seglen, int seg0, int segstep, int nsegl) {
    xline_rows<C, XL_THREADS, SEG>(e, f, K, g, geo, row0, rowstep, nt, extra0, extra1, seglen, seg0, segstep, nsegl);
}

// (L e)(i, j) at an interior cell from the face metrics: the stencil of
// cell_lap_interior in the element type T (inverse spacings idx, idy).
template <typename T, typename G>
__device__ __forceinline__ T interior_lap(gconst<T> *e, const G &geo_, int P, size_t c, T idx, T idy) {
    gconst<T> *Jxx = gptr(geo_.Jxx), *Jxy = gptr(geo_.Jxy), *Jyx = gptr(geo_.Jyx), *Jyy = gptr(geo_.Jyy);
    const T fourth = T(0.25);
    gconst<T> *q = e + c;
    const T p = q[0], pe = q[1], pw = q[-1], pn = q[P], ps = q[-P];
    const T ene = q[P + 1], enw = q[P - 1], ese = q[-P + 1], esw = q[-P - 1];
    const T xfe = Jxx[c + 1] * (pe - p) * idx + Jxy[c + 1] * fourth * (ene - ese + pn - ps) * idy;
    const T xfw = Jxx[c] * (p - pw) * idx + Jxy[c] * fourth * (pn - ps + enw - esw) * idy;
    const T yfn = Jyy[c + P] * (pn - p) * idy + Jyx[c + P] * fourth * (ene - enw + pe - pw) * idx;
    const T yfs = Jyy[c] * (p - ps) * idy + Jyx[c] * fourth * (pe - pw + ese - esw) * idx;
    return (xfe - xfw) * idx + (yfn - yfs) * idy;
}

// Preconditioner residual restricted 1x2 in y, with the coarse correction
// cleared:  rc(i, J) = (r(i, 2J) + r(i, 2J+1)) / 2,  r = f - L e  (interior
// stencil from the metrics, ring cells from the table),  ec(i, J) = 0.
template <typename T, typename G>
__device__ __forceinline__ void restrict_cell(T *rc, T *ec, const T *e, const T *f, const T *K, const Grid &gf,
                                              const Grid &gc, const G &geo, int i, int J, int off) {
    const int nring = ring_count(gf.nx, gf.ny);
    const T idx = (T)(1.0 / gf.dx), idy = (T)(1.0 / gf.dy);
    T s = 0;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        const int j = off + 2 * J + b;
        const bool ring = i == 0 || i == gf.nx - 1 || j == 0 || j == gf.ny - 1;
        const T lap = ring ? ring_lap(gptr(e), gptr(K), nring, gf, i, j)
                           : interior_lap(gptr(e), geo, gf.pitch, ix(gf.pitch, i, j), idx, idy);
        s = s + T(0.5) * (f[ix(gf.pitch, i, j)] - lap);
    }
    const size_t c = ix(gc.pitch, i, J);
    rc[c] = s;
    ec[c] = 0;
}

template <typename T, typename G>
__global__ void __launch_bounds__(256) k_residual_restrict_y(T *rc, T *ec, const T *e, const T *f, const T *K,
                                                             Grid gf, Grid gc, G geo, int off) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int J = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= gc.nx || J >= gc.ny) return;
    restrict_cell(rc, ec, e, f, K, gf, gc, geo, i, J, off);
}

// ---------------------------------------------------------------------------
// Batched forms: the workgroup's task or tile names the level (and through
// it the problem); workgroups of inactive problems return at once.
// ---------------------------------------------------------------------------
// The level's fields of the cycle's element type.
__device__ __forceinline__ void level_fields(const BLev *L, int sel, double *&e, const double *&f, const double *&K,
                                             Geo &geo) {
    e = L->e[sel];
    f = L->f[sel];
    K = L->K;
    geo = L->geo;
}
__device__ __forceinline__ void level_fields(const BLev *L, int, float *&e, const float *&f, const float *&K,
                                             GeoF &geo) {
    e = L->e32;
    f = L->f32;
    K = L->K32;
    geo = L->geo32;
}

// Cells whose loads the batched kernels' row pass issues together: as many
// as fit without costing a wave per SIMD (FP32: 2 of the 3 cells of a chunk
// of 2, 3 of 5, 2 of 9; the C = 4 kernel keeps 4 waves with 119 VGPRs, the
// others their 87 and 154-159).  The batched mix's finest FP32 launch set
// went from 0.70 to 0.67 ms, bit-identical.  FP64: 2 cells for every chunk
// length (126, 152, 228 and 240 VGPRs: 4, 3, 2 and 2 waves, as with one).
template <int C, typename T>
constexpr int batch_rg() {
    return sizeof(T) == 8 ? 2 : (C == 4 ? 3 : 2);
}

template <int C, int NT, typename T>
__device__ __forceinline__ void xline_task(const XlTask *tasks, const BLev *levs, const int *active, int sel, int k) {
    const XlTask t = tasks[launch_slot()];
    const BLev *L = levs + t.lev;
    if (!active[L->prob] || (k >= 0 && k >= L->sweeps)) return;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_bytes[];
    T *lds = reinterpret_cast<T *>(lds_bytes);
    const Grid g = L->g;
    T *e;
    const T *f, *K;
    typename std::conditional<sizeof(T) == 8, Geo, GeoF>::type geo;
    level_fields(L, sel, e, f, K, geo);
    const RowCtx<T> X0 = row_ctx<T>(g, L->nt, C);
#pragma unroll 1
    for (int r = 0; r < 2; ++r) {
        const int jr = r == 0 ? t.row : t.extra;
        if (jr < 0) continue;  // uniform over the workgroup
        if (r) __syncthreads();
        RowCtx<T> X = X0;
        asm volatile("" : "+s"(X.rem));
        xline_row<C, NT, false, batch_rg<C, T>()>(X, e, f, K, g, geo, jr, lds);
    }
}

template <int C, typename T>
__global__ void __launch_bounds__(XL_THREADS) __attribute__((amdgpu_waves_per_eu(xl_waves<T>(), xl_waves<T>())))
k_xline_batch(const XlTask *tasks, const BLev *levs, const int *active, int sel, int k) {
    xline_task<C, XL_THREADS, T>(tasks, levs, active, sel, k);
}

template <int C, typename T>
__global__ void __launch_bounds__(XL_THREADS) k_xline_batch_long(const XlTask *tasks, const BLev *levs,
                                                                 const int *active, int sel, int k) {
    xline_task<C, XL_THREADS, T>(tasks, levs, active, sel, k);
}

// Short rows (nx <= 512): one wave per row.
template <int C, typename T>
__global__ void __launch_bounds__(XL_SHORT) __attribute__((amdgpu_waves_per_eu(2)))
k_xline_batch_short(const XlTask *tasks, const BLev *levs, const int *active, int sel, int k) {
    xline_task<C, XL_SHORT, T>(tasks, levs, active, sel, k);
}

// Rows of 513..1024 cells: two waves per row.
template <int C, typename T>
__global__ void __launch_bounds__(XL_MID) __attribute__((amdgpu_waves_per_eu(2)))
k_xline_batch_mid(const XlTask *tasks, const BLev *levs, const int *active, int sel, int k) {
    xline_task<C, XL_MID, T>(tasks, levs, active, sel, k);
}

template <typename T>
__global__ void __launch_bounds__(256) k_restrict_batch(const BTile *tiles, const BLev *levs, const int *active,
                                                        int sel) {
    const BTile t = tiles[blockIdx.x];
    const BLev *F = levs + t.item, *Cl = levs + t.item + 1;
    if (!active[F->prob]) return;
    const int i = t.tx * 64 + (threadIdx.x & 63);
    const Grid gc = Cl->g;
    if (i >= gc.nx) return;
    T *fe, *ce, *cf;
    const T *ff, *K, *cff, *cK;
    typename std::conditional<sizeof(T) == 8, Geo, GeoF>::type geo, cgeo;
    level_fields(F, sel, fe, ff, K, geo);
    level_fields(Cl, 0, ce, cff, cK, cgeo);
    cf = const_cast<T *>(cff);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int J = t.ty * BT_ROWS + (threadIdx.x >> 6) + 4 * r;
        if (J < gc.ny) restrict_cell(cf, ce, (const T *)fe, ff, K, F->g, gc, geo, i, J, F->off);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_prolong_batch(const BTile *tiles, const BLev *levs, const int *active,
                                                       int sel) {
    const BTile t = tiles[blockIdx.x];
    const BLev *F = levs + t.item, *Cl = levs + t.item + 1;
    if (!active[F->prob]) return;
    const int i = t.tx * 64 + (threadIdx.x & 63);
    const Grid gf = F->g, gc = Cl->g;
    if (i >= gf.nx) return;
    T *fe, *ce;
    const T *ff, *K, *cf, *cK;
    typename std::conditional<sizeof(T) == 8, Geo, GeoF>::type geo, cgeo;
    level_fields(F, sel, fe, ff, K, geo);
    level_fields(Cl, 0, ce, cf, cK, cgeo);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int j = t.ty * BT_ROWS + (threadIdx.x >> 6) + 4 * r;
        if (j < gf.ny) prolong_y_cell(fe, (const T *)ce, gf, gc, (T)F->slo, (T)F->shi, i, j, F->off);
    }
}

// Chunk length for a row of nx cells: the smallest C in {2, 4, 8, 16} with
// floor(nx / C) <= 256 chunks.  nx >= C*C then holds, so the remainder is
// spread one cell per chunk (chunks of C or C+1 cells).
int xline_chunk(int nx, int nt_max = XL_THREADS) {
    for (int C = 2; C <= 16; C *= 2)
        if (nx / C <= nt_max) return C;
    return 0;
}

}  // namespace

int ring_table_elems(const Grid &g) { return RING_PLANES * ring_count(g.nx, g.ny); }

void launch_ring_coef(hipStream_t s, double *K, const Grid &g, const Geo &geo, const Bc &bc) {
    const int n = ring_table_elems(g);
    k_ring_coef<<<(n + 255) / 256, 256, 0, s>>>(K, g, geo, bc);
}

size_t xline_lds_bytes(int nx, size_t elem) { return xline_lds_bytes_for(nx, xline_chunk(nx), elem, XL_THREADS); }

size_t xline_lds_bytes_for(int nx, int C, size_t elem, int threads) {
    if (C == 0) return (size_t)-1;
    const size_t nt = (size_t)(nx / C);
    // chunk staging (4 arrays), or one chunk array plus the interface rows
    const size_t tr = 4 * nt * (C + 1), iface = nt * (C + 1) + 5 * 2 * (size_t)threads;
    return (tr > iface ? tr : iface) * elem;
}

bool xline_supported(int nx) { return xline_segments(nx, sizeof(double)) != 0; }

bool xline_single_line(int nx) { return nx >= 4 && xline_chunk(nx) != 0; }

bool xline_whole(int nx, size_t elem) { return xline_segments(nx, elem) == 1; }

int xline_segments(int nx, size_t elem) {
    if (nx < 4) return 0;
    if (xline_chunk(nx) != 0) return 1;
    if (elem == sizeof(float) && xline_chunk(nx, XL_WIDE) == 16) return 1;
    for (int n = 2; n <= 64; n *= 2)
        if (nx % n == 0 && nx / n >= 4 && xline_chunk(nx / n) != 0) return n;
    return 0;
}

int xline_batch_chunk(int nx, int threads, size_t elem) {
    if (nx < 4 || (threads != XL_SHORT && threads != XL_MID && threads != XL_THREADS)) return 0;
    const int C = xline_chunk(nx, threads);
    if (C == 0 || (C == 16 && threads != XL_THREADS)) return 0;
    return xline_lds_bytes_for(nx, C, elem, threads) <= 160 * 1024 ? C : 0;
}

void xline_batch_shape(int nx, int &C, int &threads) {
    threads = nx <= XL_SHORT_MAX_NX ? XL_SHORT : (nx <= XL_MID_MAX_NX ? XL_MID : XL_THREADS);
    C = xline_chunk(nx, threads);
}

template <typename T, typename G>
static void residual_restrict_y(hipStream_t s, T *rc, T *ec, const T *e, const T *f, const T *K, const Grid &gf,
                                const Grid &gc, const G &geo, int off) {
    const dim3 grid((unsigned)((gc.nx + 63) / 64), (unsigned)((gc.ny + 3) / 4));
    k_residual_restrict_y<<<grid, 256, 0, s>>>(rc, ec, e, f, K, gf, gc, geo, off);
}

void launch_residual_restrict_y(hipStream_t s, double *rc, double *ec, const double *e, const double *f,
                                const double *K, const Grid &gf, const Grid &gc, const Geo &geo, int off) {
    residual_restrict_y(s, rc, ec, e, f, K, gf, gc, geo, off);
}

void launch_residual_restrict_y(hipStream_t s, float *rc, float *ec, const float *e, const float *f, const float *K,
                                const Grid &gf, const Grid &gc, const GeoF &geo, int off) {
    residual_restrict_y(s, rc, ec, e, f, K, gf, gc, geo, off);
}

// One launch of the line kernel for rows (SEG = false) or segments of seglen cells.
template <bool SEG, typename T, typename G>
static void xline_launch(hipStream_t s, T *e, const T *f, const T *K, const Grid &g, const G &geo, int row0,
                         int rowstep, int extra0, int extra1, int seglen, int seg0, int segstep, int nsegl,
                         unsigned grid) {
    if (xline_chunk(seglen) == 0) {
        // only whole FP32 rows get here (xline_segments): the 512-thread form
        if constexpr (sizeof(T) == 4 && !SEG) {
            const int nt = seglen / 16;
            const size_t lds = xline_lds_bytes_for(seglen, 16, sizeof(T), XL_WIDE);
            k_xline_wide<16><<<grid, XL_WIDE, lds, s>>>(e, f, K, g, geo, row0, rowstep, nt, extra0, extra1, seglen, seg0, segstep, nsegl);
        }
    } else if (seglen <= XL_SHORT_MAX_NX) {
        const int C = xline_chunk(seglen, XL_SHORT), nt = seglen / C;
        const size_t lds = xline_lds_bytes_for(seglen, C, sizeof(T), XL_SHORT);
        switch (C) {
            case 2: k_xline_short<2, SEG><<<grid, XL_SHORT, lds, s>>>(e, f, K, g, geo, row0, rowstep, nt, extra0, extra1, seglen, seg0, segstep, nsegl); break;
            case 4: k_xline_short<4, SEG><<<grid, XL_SHORT, lds, s>>>(e, f, K, g, geo, row0, rowstep, nt, extra0, extra1, seglen, seg0, segstep, nsegl); break;
            default: k_xline_short<8, SEG><<<grid, XL_SHORT, lds, s>>>(e, f, K, g, geo, row0, rowstep, nt, extra0, extra1, seglen, seg0, segstep, nsegl); break;
        }
    } else {
        const int C = xline_chunk(seglen);
        const size_t lds = xline_lds_bytes(seglen, sizeof(T));
        if (lds > 160 * 1024) {
            g_last_error = "k_xline: row length not supported";
            throw Failure(MGP_ERR_UNSUPPORTED);
        }
        const int nt = seglen / C;
        switch (C) {
            case 2: k_xline<2, SEG><<<grid, XL_THREADS, lds, s>>>(e, f, K, g, geo, row0, rowstep, nt, extra0, extra1, seglen, seg0, segstep, nsegl); break;
            case 4: k_xline<4, SEG><<<grid, XL_THREADS, lds, s>>>(e, f, K, g, geo, row0, rowstep, nt, extra0, extra1, seglen, seg0, segstep, nsegl); break;
            case 8: k_xline<8, SEG><<<grid, XL_THREADS, lds, s>>>(e, f, K, g, geo, row0, rowstep, nt, extra0, extra1, seglen, seg0, segstep, nsegl); break;
            default: k_xline_long<16, SEG><<<grid, XL_THREADS, lds, s>>>(e, f, K, g, geo, row0, rowstep, nt, extra0, extra1, seglen, seg0, segstep, nsegl); break;
        }
    }
}

template <typename T, typename G>
static void xline(hipStream_t s, T *e, const T *f, const T *K, const Grid &g, const G &geo, int row0, int rowstep,
                  int nrows, int extra0, int extra1) {
    if (nrows <= 0) return;
    const int nseg = xline_segments(g.nx, sizeof(T));
    if (nseg == 0) {
        g_last_error = "k_xline: row length not supported";
        throw Failure(MGP_ERR_UNSUPPORTED);
    }
    const int seglen = g.nx / nseg;
    // whole rows in one launch; segmented rows: the even segments, then the odd ones
    const int nlaunch = nseg == 1 ? 1 : 2, segstep = nseg == 1 ? 1 : 2, nsegl = nseg == 1 ? 1 : nseg / 2;
    const unsigned grid = (unsigned)(nrows * nsegl);
    for (int seg0 = 0; seg0 < nlaunch; ++seg0) {
        if (nseg > 1) xline_launch<true>(s, e, f, K, g, geo, row0, rowstep, extra0, extra1, seglen, seg0, segstep, nsegl, grid);
        else xline_launch<false>(s, e, f, K, g, geo, row0, rowstep, extra0, extra1, seglen, seg0, segstep, nsegl, grid);
        const hipError_t err = hipGetLastError();
        if (err != hipSuccess) {
            g_last_error = std::string("k_xline launch: ") + hipGetErrorString(err);
            throw Failure(MGP_ERR_HIP);
        }
    }
}

void launch_xline(hipStream_t s, double *e, const double *f, const double *K, const Grid &g, const Geo &geo,
                  int row0, int rowstep, int nrows, int extra0, int extra1) {
    xline(s, e, f, K, g, geo, row0, rowstep, nrows, extra0, extra1);
}

template <typename T>
static void xline_batch(hipStream_t s, const XlTask *tasks, int ntasks, const BLev *levs, const int *active, int sel,
                        int C, int threads, size_t lds, int k) {
    if (threads == XL_SHORT) {
        switch (C) {
            case 2: k_xline_batch_short<2, T><<<ntasks, XL_SHORT, lds, s>>>(tasks, levs, active, sel, k); break;
            case 4: k_xline_batch_short<4, T><<<ntasks, XL_SHORT, lds, s>>>(tasks, levs, active, sel, k); break;
            case 8: k_xline_batch_short<8, T><<<ntasks, XL_SHORT, lds, s>>>(tasks, levs, active, sel, k); break;
            default:
                g_last_error = "k_xline_batch_short: chunk length must be 2, 4 or 8";
                throw Failure(MGP_ERR_ARG);
        }
    } else if (threads == XL_MID) {
        switch (C) {
            case 2: k_xline_batch_mid<2, T><<<ntasks, XL_MID, lds, s>>>(tasks, levs, active, sel, k); break;
            case 4: k_xline_batch_mid<4, T><<<ntasks, XL_MID, lds, s>>>(tasks, levs, active, sel, k); break;
            case 8: k_xline_batch_mid<8, T><<<ntasks, XL_MID, lds, s>>>(tasks, levs, active, sel, k); break;
            default:
                g_last_error = "k_xline_batch_mid: chunk length must be 2, 4 or 8";
                throw Failure(MGP_ERR_ARG);
        }
    } else switch (C) {
        case 2: k_xline_batch<2, T><<<ntasks, XL_THREADS, lds, s>>>(tasks, levs, active, sel, k); break;
        case 4: k_xline_batch<4, T><<<ntasks, XL_THREADS, lds, s>>>(tasks, levs, active, sel, k); break;
        case 8: k_xline_batch<8, T><<<ntasks, XL_THREADS, lds, s>>>(tasks, levs, active, sel, k); break;
        case 16: k_xline_batch_long<16, T><<<ntasks, XL_THREADS, lds, s>>>(tasks, levs, active, sel, k); break;
        default:
            g_last_error = "k_xline_batch: chunk length must be 2, 4, 8 or 16";
            throw Failure(MGP_ERR_ARG);
    }
}

void launch_xline_batch(hipStream_t s, const XlTask *tasks, int ntasks, const BLev *levs, const int *active, int sel,
                        int C, int threads, size_t lds, int k, bool fp32) {
    if (ntasks <= 0) return;
    if (fp32) xline_batch<float>(s, tasks, ntasks, levs, active, sel, C, threads, lds, k);
    else xline_batch<double>(s, tasks, ntasks, levs, active, sel, C, threads, lds, k);
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        g_last_error = std::string("k_xline_batch launch: ") + hipGetErrorString(err);
        throw Failure(MGP_ERR_HIP);
    }
}

void launch_xline(hipStream_t s, float *e, const float *f, const float *K, const Grid &g, const GeoF &geo, int row0,
                  int rowstep, int nrows, int extra0, int extra1) {
    xline(s, e, f, K, g, geo, row0, rowstep, nrows, extra0, extra1);
}

void launch_restrict_batch(hipStream_t s, const BTile *tiles, int ntiles, const BLev *levs, const int *active,
                           int sel, bool fp32) {
    if (ntiles <= 0) return;
    if (fp32) k_restrict_batch<float><<<ntiles, 256, 0, s>>>(tiles, levs, active, sel);
    else k_restrict_batch<double><<<ntiles, 256, 0, s>>>(tiles, levs, active, sel);
}

void launch_prolong_batch(hipStream_t s, const BTile *tiles, int ntiles, const BLev *levs, const int *active,
                          int sel, bool fp32) {
    if (ntiles <= 0) return;
    if (fp32) k_prolong_batch<float><<<ntiles, 256, 0, s>>>(tiles, levs, active, sel);
    else k_prolong_batch<double><<<ntiles, 256, 0, s>>>(tiles, levs, active, sel);
}

}  // namespace mgp
