// mgp_fourth.hip -- the fourth-order finite-volume operator of the
// fourth-order deferred correction, and its face-averaged metrics.
//
// Unknowns are cell averages over computational cells; L4 phi is the
// divergence of face-averaged fluxes (Colella, Dorr, Hittinger and Martin,
// "High-order, finite-volume methods in mapped coordinates", J. Comput. Phys.
// 230 (2011) 2952-2976, the product rule for face averages, their eq. (13)):
//   <F^xi>  = <A> <phi_xi> + <B> <phi_eta>
//             + (deta^2/12) (dA/deta d(phi_xi)/deta + dB/deta d(phi_eta)/deta)
//   <F^eta> = <C> <phi_eta> + <B> <phi_xi>
//             + (dxi^2/12) (dC/dxi d(phi_eta)/dxi + dB/dxi d(phi_xi)/dxi)
// with A = J g^11, B = J g^12 = J g^21, C = J g^22.  Face averages of phi and
// of its normal derivative come from cell averages by the fourth-order
// interpolants, centred (-1, 7, 7, -1)/12 and (1, -15, 15, -1)/(12 h), and on
// the first interior face and the boundary face the quartic through five cell
// averages; tangential derivatives of face averages are differences of node
// values interpolated the same way along the face; the correction terms use
// second-order differences (one-sided at the ends).  Boundary faces carry the
// homogeneous Neumann flux, zero.
//
// Every sum is formed left to right in the order the oracle's restatement
// (oracle/mgp_oracle4.c) states it, so on the same metric arrays the two
// agree bit for bit.  One thread per cell evaluates its four face fluxes;
// the field and metric reads go through L1/L2 (a 5 x 5 neighbourhood).
// Algorithmic bytes: phi, out (and rhs) once, the eight metric arrays once,
// about 88 B per cell.
#include <hip/hip_runtime.h>

#include <cmath>

#include "mgp_kernels.h"

namespace mgp {

namespace {

// n cell averages along a line: c(k) = base[k * stride]; m(k) = c(n-1-k).
struct Line4 {
    const double *base;
    long stride;
    int n;
    __device__ __forceinline__ double c(int k) const { return base[(long)k * stride]; }
    __device__ __forceinline__ double m(int k) const { return base[(long)(n - 1 - k) * stride]; }
};

// Point value at node f (0..n) of a line from its cell averages.
__device__ __forceinline__ double face_val4(const Line4 &l, int f) {
    const int n = l.n;
    if (f == 0)
        return ((((137.0 * l.c(0) - 163.0 * l.c(1)) + 137.0 * l.c(2)) - 63.0 * l.c(3)) + 12.0 * l.c(4)) * (1.0 / 60.0);
    if (f == n)
        return ((((137.0 * l.m(0) - 163.0 * l.m(1)) + 137.0 * l.m(2)) - 63.0 * l.m(3)) + 12.0 * l.m(4)) * (1.0 / 60.0);
    if (f == 1)
        return ((((12.0 * l.c(0) + 77.0 * l.c(1)) - 43.0 * l.c(2)) + 17.0 * l.c(3)) - 3.0 * l.c(4)) * (1.0 / 60.0);
    if (f == n - 1)
        return ((((12.0 * l.m(0) + 77.0 * l.m(1)) - 43.0 * l.m(2)) + 17.0 * l.m(3)) - 3.0 * l.m(4)) * (1.0 / 60.0);
    return (7.0 * (l.c(f - 1) + l.c(f)) - (l.c(f - 2) + l.c(f + 1))) * (1.0 / 12.0);
}

// Normal derivative at interior node f (1..n-1); ih12 = 1/(12 h), ih60 = 1/(60 h).
__device__ __forceinline__ double face_der4(const Line4 &l, int f, double ih12, double ih60) {
    const int n = l.n;
    if (f == 1)
        return ((((-50.0 * l.c(0) + 25.0 * l.c(1)) + 45.0 * l.c(2)) - 25.0 * l.c(3)) + 5.0 * l.c(4)) * ih60;
    if (f == n - 1)
        return -(((((-50.0 * l.m(0) + 25.0 * l.m(1)) + 45.0 * l.m(2)) - 25.0 * l.m(3)) + 5.0 * l.m(4)) * ih60);
    return (15.0 * (l.c(f) - l.c(f - 1)) - (l.c(f + 1) - l.c(f - 2))) * ih12;
}

// Face values along the tangential direction of the faces at node `face`
// across: q(t) = face_val4 of line t.  For x faces line t is row t (stride
// 1), for y faces column t (stride pitch).
struct FaceSeq {
    const double *phi;
    int pitch, face, n_across;
    bool xface;
    __device__ __forceinline__ double q(int t) const {
        Line4 l;
        if (xface) {
            l.base = phi + ix(pitch, 0, t);
            l.stride = 1;
        } else {
            l.base = phi + ix(pitch, t, 0);
            l.stride = pitch;
        }
        l.n = n_across;
        return face_val4(l, face);
    }
};

// Node value at tangential node m (0..n) of the face averages q(0..n-1).
__device__ __forceinline__ double tan_node4(const FaceSeq &q, int n, int m) {
    double c[5];
    if (m == 0 || m == 1) {
#pragma unroll
        for (int k = 0; k < 5; ++k) c[k] = q.q(k);
        const Line4 l{c, 1, 5};
        return face_val4(l, m);
    }
    if (m == n || m == n - 1) {
#pragma unroll
        for (int k = 0; k < 5; ++k) c[k] = q.q(n - 5 + k);
        const Line4 l{c, 1, 5};
        return face_val4(l, m - (n - 5));
    }
    const double a = q.q(m - 2), b = q.q(m - 1), d = q.q(m), e = q.q(m + 1);
    return (7.0 * (b + d) - (a + e)) * (1.0 / 12.0);
}

__device__ double flux_x4(const double *phi, const Grid &g, const Geo4 &M, int i, int j) {
    const int nx = g.nx, ny = g.ny, P = g.pitch;
    if (i == 0 || i == nx) return 0.0;
    const double idx = 1.0 / g.dx, idy = 1.0 / g.dy;
    const double ih12 = 1.0 / (12.0 * g.dx), ih60 = 1.0 / (60.0 * g.dx);
    const double hidy = 0.5 * idy, idy2 = idy * idy;
    const Line4 row{phi + ix(P, 0, j), 1, nx};
    const double Dx = face_der4(row, i, ih12, ih60);
    const FaceSeq fs{phi, P, i, nx, true};
    const double T = (tan_node4(fs, ny, j + 1) - tan_node4(fs, ny, j)) * idy;
    double D2[5], S2[5];
    int j0 = j - 2;
    if (j0 < 0) j0 = 0;
    if (j0 > ny - 5) j0 = ny - 5;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double a = phi[ix(P, i - 1, j0 + k)], b = phi[ix(P, i, j0 + k)];
        D2[k] = (b - a) * idx;
        S2[k] = 0.5 * (a + b);
    }
    const int k = j - j0;
    double Cx, Cy;
    if (j == 0) {
        Cx = ((4.0 * D2[k + 1] - 3.0 * D2[k]) - D2[k + 2]) * hidy;
        Cy = (((2.0 * S2[k] - 5.0 * S2[k + 1]) + 4.0 * S2[k + 2]) - S2[k + 3]) * idy2;
    } else if (j == ny - 1) {
        Cx = ((3.0 * D2[k] - 4.0 * D2[k - 1]) + D2[k - 2]) * hidy;
        Cy = (((2.0 * S2[k] - 5.0 * S2[k - 1]) + 4.0 * S2[k - 2]) - S2[k - 3]) * idy2;
    } else {
        Cx = (D2[k + 1] - D2[k - 1]) * hidy;
        Cy = ((S2[k + 1] - 2.0 * S2[k]) + S2[k - 1]) * idy2;
    }
    const size_t c = ix(P, i, j);
    return ((M.A4[c] * Dx + M.Bx4[c] * T) + M.dA[c] * Cx) + M.dBx[c] * Cy;
}

__device__ double flux_y4(const double *phi, const Grid &g, const Geo4 &M, int i, int j) {
    const int nx = g.nx, ny = g.ny, P = g.pitch;
    if (j == 0 || j == ny) return 0.0;
    const double idx = 1.0 / g.dx, idy = 1.0 / g.dy;
    const double ih12 = 1.0 / (12.0 * g.dy), ih60 = 1.0 / (60.0 * g.dy);
    const double hidx = 0.5 * idx, idx2 = idx * idx;
    const Line4 col{phi + ix(P, i, 0), P, ny};
    const double Dy = face_der4(col, j, ih12, ih60);
    const FaceSeq fs{phi, P, j, ny, false};
    const double T = (tan_node4(fs, nx, i + 1) - tan_node4(fs, nx, i)) * idx;
    double E2[5], R2[5];
    int i0 = i - 2;
    if (i0 < 0) i0 = 0;
    if (i0 > nx - 5) i0 = nx - 5;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double a = phi[ix(P, i0 + k, j - 1)], b = phi[ix(P, i0 + k, j)];
        E2[k] = (b - a) * idy;
        R2[k] = 0.5 * (a + b);
    }
    const int k = i - i0;
    double Ey, Ry;
    if (i == 0) {
        Ey = ((4.0 * E2[k + 1] - 3.0 * E2[k]) - E2[k + 2]) * hidx;
        Ry = (((2.0 * R2[k] - 5.0 * R2[k + 1]) + 4.0 * R2[k + 2]) - R2[k + 3]) * idx2;
    } else if (i == nx - 1) {
        Ey = ((3.0 * E2[k] - 4.0 * E2[k - 1]) + E2[k - 2]) * hidx;
        Ry = (((2.0 * R2[k] - 5.0 * R2[k - 1]) + 4.0 * R2[k - 2]) - R2[k - 3]) * idx2;
    } else {
        Ey = (E2[k + 1] - E2[k - 1]) * hidx;
        Ry = ((R2[k + 1] - 2.0 * R2[k]) + R2[k - 1]) * idx2;
    }
    const size_t c = ix(P, i, j);
    return ((M.C4[c] * Dy + M.By4[c] * T) + M.dC[c] * Ey) + M.dBy[c] * Ry;
}

__global__ void __launch_bounds__(256) k_lap4(double *__restrict__ out, const double *__restrict__ phi,
                                              const double *__restrict__ rhs, Grid g, Geo4 M) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    const double idx = 1.0 / g.dx, idy = 1.0 / g.dy;
    const double fe = flux_x4(phi, g, M, i + 1, j);
    const double fw = flux_x4(phi, g, M, i, j);
    const double fn = flux_y4(phi, g, M, i, j + 1);
    const double fs = flux_y4(phi, g, M, i, j);
    const double lap = (fe - fw) * idx + (fn - fs) * idy;
    const size_t c = ix(g.pitch, i, j);
    out[c] = rhs ? rhs[c] - lap : lap;
}

// J g^ij of x = s + amp sin(2 pi s) sin(2 pi t), y = (exp(beta t) - 1)/(exp(beta) - 1),
// s = xi/L, t = eta/H, with the chain rule for L and H.
__device__ __forceinline__ void warped_agc(double amp, double beta, double L, double H, double em1, double xi,
                                           double eta, double &A, double &B, double &C) {
    const double two_pi = 6.283185307179586;
    const double s = xi / L, t = eta / H;
    const double x_xi = 1.0 + amp * two_pi * cos(two_pi * s) * sin(two_pi * t);
    const double x_eta = amp * two_pi * sin(two_pi * s) * cos(two_pi * t);
    const double y_eta = beta == 0.0 ? 1.0 : beta * exp(beta * t) / em1;
    const double J = x_xi * y_eta;
    A = (x_eta * x_eta + y_eta * y_eta) / J * (L / H);
    B = -(x_xi * x_eta) / J;
    C = (x_xi * x_xi) / J * (H / L);
}

constexpr double GL_T0 = 0.1127016653792583, GL_T2 = 0.8872983346207417;  // 1/2 -+ sqrt(15)/10

// One thread per node (i, j): the x face (i, j) (j < ny) and the y face (i, j) (i < nx).
__global__ void __launch_bounds__(256) k_metric4(Geo4W M, Grid g, double amp, double beta, double L, double H,
                                                 double em1) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i > g.nx || j > g.ny) return;
    const double dx = g.dx, dy = g.dy;
    const double tg[3] = {GL_T0, 0.5, GL_T2};
    const size_t c = ix(g.pitch, i, j);
    if (j < g.ny) {
        const double xi = i * dx, w12x = dy * (1.0 / 12.0);
        double a[3], b[3], cc, a0, b0, a1, b1;
        for (int k = 0; k < 3; ++k) warped_agc(amp, beta, L, H, em1, xi, (j + tg[k]) * dy, a[k], b[k], cc);
        warped_agc(amp, beta, L, H, em1, xi, j * dy, a0, b0, cc);
        warped_agc(amp, beta, L, H, em1, xi, (j + 1) * dy, a1, b1, cc);
        M.A4[c] = ((5.0 * a[0] + 8.0 * a[1]) + 5.0 * a[2]) * (1.0 / 18.0);
        M.Bx4[c] = ((5.0 * b[0] + 8.0 * b[1]) + 5.0 * b[2]) * (1.0 / 18.0);
        M.dA[c] = (a1 - a0) * w12x;
        M.dBx[c] = (b1 - b0) * w12x;
    }
    if (i < g.nx) {
        const double eta = j * dy, w12y = dx * (1.0 / 12.0);
        double aa, b[3], cc[3], b0, c0, b1, c1;
        for (int k = 0; k < 3; ++k) warped_agc(amp, beta, L, H, em1, (i + tg[k]) * dx, eta, aa, b[k], cc[k]);
        warped_agc(amp, beta, L, H, em1, i * dx, eta, aa, b0, c0);
        warped_agc(amp, beta, L, H, em1, (i + 1) * dx, eta, aa, b1, c1);
        M.C4[c] = ((5.0 * cc[0] + 8.0 * cc[1]) + 5.0 * cc[2]) * (1.0 / 18.0);
        M.By4[c] = ((5.0 * b[0] + 8.0 * b[1]) + 5.0 * b[2]) * (1.0 / 18.0);
        M.dC[c] = (c1 - c0) * w12y;
        M.dBy[c] = (b1 - b0) * w12y;
    }
}

}  // namespace

void launch_lap4(hipStream_t s, double *out, const double *phi, const double *rhs, const Grid &g, const Geo4 &m) {
    const dim3 grid((unsigned)((g.nx + 63) / 64), (unsigned)((g.ny + 3) / 4));
    k_lap4<<<grid, 256, 0, s>>>(out, phi, rhs, g, m);
}

void launch_metric4(hipStream_t s, const mgp_mapping &m, const Grid &g, const Geo4W &out) {
    const double L = m.L > 0.0 ? m.L : g.nx * g.dx, H = m.H > 0.0 ? m.H : g.ny * g.dy;
    const dim3 grid((unsigned)((g.nx + 1 + 63) / 64), (unsigned)((g.ny + 1 + 3) / 4));
    k_metric4<<<grid, 256, 0, s>>>(out, g, m.amp, m.beta, L, H, std::expm1(m.beta));
}

}  // namespace mgp
