// mgp_common.h -- device data layout and parameter blocks shared by the HIP
// kernels and the host driver.
//
// HBM layout of one level.  Every field of a level (cell-centred, x-face,
// y-face) lives in a padded row-major array with one ghost ring:
//   element (i, j) at  base[(j + 1) * pitch + OFFX + i],  i = -1 .. nx+1, j = -1 .. ny+1
// so cell i = 0 of every row starts on a 256-byte boundary (OFFX = 32 doubles)
// and x-face index i = nx and y-face row j = ny fit without a second layout.
// i (xi) is the contiguous direction, as in the reference's column-major
// data(i, j) (arrayutils.f90:107-116).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mgp {

constexpr int OFFX = 32;          // doubles before cell i = 0 in every row
constexpr int PITCH_ALIGN = 32;   // row pitch multiple (256 B)

// BC type / mode codes: arrayutils.f90:27-49
enum BcType : int { BC_NONE = -1, BC_NEUM = 0, BC_DIRI = 1, BC_PERIODIC = 2, BC_CF = 3,
                    BC_EXTRAP0 = 4, BC_EXTRAP1 = 5, BC_EXTRAP2 = 6 };
enum BcMode : int { MODE_UNIFORM = 0, MODE_NONUNIFORM = 1, MODE_LAH = 2 };
enum Side : int { XLO = 0, XHI = 1, YLO = 2, YHI = 3 };

constexpr double BOGUS = 1.2345e300;  // precision.f90:53, corner ghosts

inline int level_pitch(int nx) {
    int p = OFFX + nx + 2;
    return (p + PITCH_ALIGN - 1) / PITCH_ALIGN * PITCH_ALIGN;
}
inline size_t level_elems(int nx, int ny) { return (size_t)level_pitch(nx) * (size_t)(ny + 3); }

struct Grid {
    int nx, ny, pitch;
    double dx, dy;
};

// Metric block of one level (device pointers into padded arrays).
struct Geo {
    const double *J, *Jxx, *Jxy, *Jyx, *Jyy;
};

// FP32 copy of a level's face metrics, same layout (pitch and offsets in
// elements): the mixed-precision preconditioner reads these.
struct GeoF {
    const float *J, *Jxx, *Jxy, *Jyx, *Jyy;
};

// Face-averaged metrics of the fourth-order operator (mgp_fourth.hip), level
// layout: on x faces <J g^11>, <J g^12> and their tangential differences
// (deta/12) d/deta, on y faces <J g^22>, <J g^21> and (dxi/12) d/dxi.
struct Geo4 {
    const double *A4, *Bx4, *dA, *dBx, *C4, *By4, *dC, *dBy;
};

// Boundary description as the kernels see it.  val[] is the UNIFORM value;
// data[] points at device NONUNIFORM face data (ny values for x sides, nx for
// y sides) or is null.
struct Bc {
    int type[4];
    int mode[4];
    double val[4];
    const double *data[4];
};

__host__ __device__ inline size_t ix(int pitch, int i, int j) {
    return (size_t)(j + 1) * (size_t)pitch + (size_t)(OFFX + i);
}

// y-direction prolongation for 1x2 semicoarsening at fine cell (i, j):
// linear in y between the coarse centre J = j/2 (weight 3/4) and its
// neighbour on j's side (J-1 for even j, J+1 for odd j; weight 1/4); beyond a
// y side the neighbour is the ghost slo*c / shi*c (Neumann: sign +1,
// reflected; Dirichlet: sign -1, negated).  ef += the interpolated coarse
// correction.
template <typename T>
__device__ __forceinline__ void prolong_y_cell(T *ef, const T *ec, const Grid &gf, const Grid &gc, T slo, T shi,
                                               int i, int j, int off = 0) {
    const int jc = j - off;  // row within the coarse grid's fine rows
    if (jc < 0 || jc >= 2 * gc.ny) {
        // a row the coarsening left out (at a Neumann side): the correction
        // of the coarse row next to it
        ef[ix(gf.pitch, i, j)] += ec[ix(gc.pitch, i, jc < 0 ? 0 : gc.ny - 1)];
        return;
    }
    const int J = jc >> 1;
    const T c = ec[ix(gc.pitch, i, J)];
    T nb;
    if (jc & 1) nb = J + 1 < gc.ny ? ec[ix(gc.pitch, i, J + 1)] : shi * c;
    else nb = J - 1 >= 0 ? ec[ix(gc.pitch, i, J - 1)] : slo * c;
    ef[ix(gf.pitch, i, j)] += T(0.75) * c + T(0.25) * nb;
}

}  // namespace mgp
