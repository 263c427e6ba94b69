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

// y-direction prolongation for 1x2 semicoarsening at fine cell (i, j), from
// the coarse centre J = j/2 and its neighbour on j's side (J-1 for even j,
// J+1 for odd j); ef += the interpolated coarse correction.  Beyond a y side
// the neighbour is the ghost slo*c / shi*c (Neumann: sign +1, reflected;
// Dirichlet: sign -1, negated).
//
// Jyy == nullptr: linear in the index, weights 3/4 and 1/4.  Else the weights
// are linear in the y "resistance" of the fine level's column i (the
// operator-induced interpolation of semicoarsening): face k between fine
// rows k-1 and k has resistance R(k) = 1/Jyy(i, k) (its y conductance is
// Jyy/dy^2, dy common to the level).  The coarse centre J lies on fine face
// 2J+1, so from fine row j = 2J the distances are R(2J+1)/2 to centre J and
// R(2J) + R(2J-1)/2 to centre J-1 (odd j: R(2J+2) + R(2J+3)/2 to J+1); a
// neighbour beyond the side is the mirror image of centre J across the wall
// face.  On a uniform column this is the linear rule; on strongly stretched
// columns it follows the error that y-dominated smoothing leaves.
template <typename T>
__device__ __forceinline__ void prolong_y_cell(T *ef, const T *ec, const Grid &gf, const Grid &gc, T slo, T shi,
                                               int i, int j, const T *Jyy = nullptr) {
    const int J = j >> 1;
    const T c = ec[ix(gc.pitch, i, J)];
    const bool odd = j & 1;
    const bool inside = odd ? J + 1 < gc.ny : J - 1 >= 0;
    const T nb = inside ? ec[ix(gc.pitch, i, odd ? J + 1 : J - 1)] : (odd ? shi : slo) * c;
    T wc = T(0.75), wn = T(0.25);
    if (Jyy) {
        auto R = [&](int k) { return T(1) / Jyy[ix(gf.pitch, i, k)]; };
        const T ra = T(0.5) * R(2 * J + 1);
        // faces towards the neighbour: 2J, 2J-1 (even j) or 2J+2, 2J+3 (odd j);
        // past the wall face the mirror of face 2J+1
        const int f1 = odd ? 2 * J + 2 : 2 * J, f2 = odd ? 2 * J + 3 : 2 * J - 1;
        const T rb = R(f1) + T(0.5) * (inside ? R(f2) : R(2 * J + 1));
        wc = rb / (ra + rb);
        wn = ra / (ra + rb);
    }
    ef[ix(gf.pitch, i, j)] += wc * c + wn * nb;
}

}  // namespace mgp
