// mgp_geometry.hip -- metric assembly on the GPU: J and J g^ij of a mapping,
// written straight into the level layout (no host metric arrays).
//
// MAP_TERRAIN: the reference's test geometry (test.f90:320-716).  x = xi,
//   y = el(x) + (H - el(x)) eta / H over el(x) = amp sin^2(pi x / L)
//   (amp = 1/4 in the reference); L = nx dx and H = ny dy unless given.  J and J g^ij
//   are differences of y at staggered points (fill_dxdXi, :566-646), so the
//   discrete metrics satisfy the geometric conservation law exactly:
//     J      = (y(cell, node j+1) - y(cell, node j)) / deta
//     J g^11 = (y(node, node j+1) - y(node, node j)) / deta      (x-faces)
//     J g^12 = -(y(cell i, cell) - y(cell i-1, cell)) / dxi       (x-faces)
//     J g^21 = -(y(node i+1, node) - y(node i, node)) / dxi       (y-faces)
//     J g^22 = (1 + y_xi^2) / y_eta, y_eta from y(cell, cell)     (y-faces)
// MAP_WARPED: the analytic warped/stretched mapping of the benchmark,
//   x = s + amp sin(2 pi s) sin(2 pi t), y = (exp(beta t) - 1)/(exp(beta) - 1)
//   with s = xi / L, t = eta / H; exact derivatives at cells and faces.
#include <hip/hip_runtime.h>

#include <cmath>

#include "mgp2d.h"
#include "mgp_geometry.h"
#include "mgp_host.h"

namespace mgp {
namespace {

struct MapParams {
    int kind;
    double amp, beta, L, H, expm1_beta;
};

__device__ double terrain_y(const Grid &g, const MapParams &m, int i, double offx, int j, double offy) {
    const double pi = 3.141592653589793;  // 4 atan(1) in double
    const double xx = (i + offx) * g.dx;
    const double sn = sin(pi * xx / m.L);
    const double el = m.amp * (sn * sn);
    const double yfrac = (j + offy) * g.dy / m.H;
    return el + (m.H - el) * yfrac;
}

struct Warped {
    double J, g11, g12, g22;
};

// J and J g^ij of the warped/stretched mapping at (xi, eta), chain rule for
// the domain extents L and H.
__device__ Warped warped(const MapParams &m, double xi, double eta) {
    const double two_pi = 6.283185307179586;
    const double s = xi / m.L, t = eta / m.H;
    const double x_xi = 1.0 + m.amp * two_pi * cos(two_pi * s) * sin(two_pi * t);
    const double x_eta = m.amp * two_pi * sin(two_pi * s) * cos(two_pi * t);
    const double y_xi = 0.0;
    const double y_eta = m.beta == 0.0 ? 1.0 : m.beta * exp(m.beta * t) / m.expm1_beta;
    const double J = x_xi * y_eta - x_eta * y_xi;
    Warped w;
    w.J = J / (m.L * m.H);
    w.g11 = (x_eta * x_eta + y_eta * y_eta) / J * (m.L / m.H);
    w.g12 = -(x_xi * x_eta + y_xi * y_eta) / J;
    w.g22 = (x_xi * x_xi + y_xi * y_xi) / J * (m.H / m.L);
    return w;
}

// One thread per node (i, j), i <= nx, j <= ny: the cell, x-face and y-face
// whose lower-left corner it is.
__global__ void __launch_bounds__(256) k_metric(double *__restrict__ J, double *__restrict__ Jxx,
                                                double *__restrict__ Jxy, double *__restrict__ Jyx,
                                                double *__restrict__ Jyy, Grid g, MapParams m) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i > g.nx || j > g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    const double sx = 1.0 / g.dx, sy = 1.0 / g.dy;
    if (m.kind == MGP_MAP_TERRAIN) {
        if (i < g.nx && j < g.ny) J[c] = (terrain_y(g, m, i, 0.5, j + 1, 0.0) - terrain_y(g, m, i, 0.5, j, 0.0)) * sy;
        if (j < g.ny) {
            Jxx[c] = (terrain_y(g, m, i, 0.0, j + 1, 0.0) - terrain_y(g, m, i, 0.0, j, 0.0)) * sy;
            Jxy[c] = -((terrain_y(g, m, i, 0.5, j, 0.5) - terrain_y(g, m, i - 1, 0.5, j, 0.5)) * sx);
        }
        if (i < g.nx) {
            const double yxi = (terrain_y(g, m, i + 1, 0.0, j, 0.0) - terrain_y(g, m, i, 0.0, j, 0.0)) * sx;
            const double yeta = (terrain_y(g, m, i, 0.5, j, 0.5) - terrain_y(g, m, i, 0.5, j - 1, 0.5)) * sy;
            Jyx[c] = -yxi;
            Jyy[c] = (1.0 + yxi * yxi) / yeta;
        }
        return;
    }
    const double xc = (i + 0.5) * g.dx, yc = (j + 0.5) * g.dy, xn = i * g.dx, yn = j * g.dy;
    if (i < g.nx && j < g.ny) J[c] = warped(m, xc, yc).J;
    if (j < g.ny) {
        const Warped w = warped(m, xn, yc);
        Jxx[c] = w.g11;
        Jxy[c] = w.g12;
    }
    if (i < g.nx) {
        const Warped w = warped(m, xc, yn);
        Jyx[c] = w.g12;
        Jyy[c] = w.g22;
    }
}

}  // namespace

void check_mapping(const mgp_mapping *m) {
    if (!m || (m->kind != MGP_MAP_TERRAIN && m->kind != MGP_MAP_WARPED) || !std::isfinite(m->amp) ||
        !std::isfinite(m->beta) || !std::isfinite(m->L) || !std::isfinite(m->H)) {
        g_last_error = "mapping: kind must be MGP_MAP_TERRAIN or MGP_MAP_WARPED, parameters finite";
        throw Failure(MGP_ERR_ARG);
    }
}

void launch_metric(hipStream_t s, const mgp_mapping &mp, const Grid &g, double *J, double *Jxx, double *Jxy,
                   double *Jyx, double *Jyy) {
    MapParams m;
    m.kind = mp.kind;
    m.amp = mp.amp;
    m.beta = mp.beta;
    m.L = mp.L > 0.0 ? mp.L : g.nx * g.dx;
    m.H = mp.H > 0.0 ? mp.H : g.ny * g.dy;
    m.expm1_beta = std::expm1(mp.beta);
    const dim3 grid((g.nx + 1 + 63) / 64, (g.ny + 1 + 3) / 4);
    k_metric<<<grid, 256, 0, s>>>(J, Jxx, Jxy, Jyx, Jyy, g, m);
}

}  // namespace mgp
