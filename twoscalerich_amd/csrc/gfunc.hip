// gfunc.hip - the two-scale RHS G(u) on CDNA4 (gfx950), FP64.
//
// One G evaluation is two launches on one stream:
//   k_micro       leading workgroups, one thread per macro element: the Keff-weighted
//                 sub-face fluxes and the BC half-edge inflows of the macro mesh;
//                 then one workgroup per micro cell: node heads (the cell's Dirichlet
//                 heads interpolated from the macro heads), element soils, the
//                 node-centric van Genuchten evaluation in LDS, a deterministic
//                 node-centric gather of sub-face fluxes, and the exchange term (net
//                 inflow into the cell's non-Dirichlet CVs) by a wavefront shuffle
//                 reduction;
//   k_macro_node  one thread per macro node: gather of sub-face fluxes, BC inflows
//                 and the exchange, divided by the macro storage.
// The discrete model (geometry conventions, unknown layout, micro Dirichlet heads,
// macro storage) follows the reference's data structures: main.cpp:260,
// save_solution.cpp:40-101, micro_mesh_properties.cpp:259-351,
// macro_mesh_properties.cpp:289-363, read_effective_conductivity.cpp:44-58, Keff.h:3.
// Every kernel reads the state as u + veps*v (the finite-difference perturbation of
// the Krylov iteration, never materialised) and can write (G - g0) * ieps directly.
#include "device.h"
#include "wave.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace tsr {

// ----------------------------------------------------------------------------------
// van Genuchten (1980) / Mualem (1976, l = 1/2) with specific storage Ss.
// C = (ts-tr) m n x Se / (|h| (1+x)) + Ss,  K = Ks sqrt(Se) (1 - (x/(1+x))^m)^2,
// x = (alpha|h|)^n, Se = (1+x)^-m; 1 - (x/(1+x))^m = -expm1(-m log1p(1/x)).
// ----------------------------------------------------------------------------------
__device__ __forceinline__ void vg_ck(const DSoil& s, double h, double Ss, double& C, double& K)
{
    if (h < 0.0) {
        const double L = log(s.al * (-h));
        const double x = exp(s.n * L);
        const double P = log1p(x);
        const double Se = exp(-s.m * P);
        C = (s.ts - s.tr) * s.m * s.n * x * Se / ((-h) * (1.0 + x)) + Ss;
        const double B = -expm1(-s.m * log1p(1.0 / x));
        K = s.Ks * sqrt(Se) * B * B;
    } else {
        C = Ss;
        K = s.Ks;
    }
}

// vg_ck plus dC/dh and dK/dh (same C and K, bit for bit).  With x = (alpha|h|)^n,
// Se = (1+x)^-m, B = 1 - (x/(1+x))^m, K = Ks sqrt(Se) B^2, C - Ss = Cw:
//   dC/dh = Cw / h * (n - 1 - n (1+m) x / (1+x))
//   dK/dh = -K m n / ((1+x) h) * (x/2 + 2 (1-B)/B)
// Non-finite derivatives (h -> 0-) are recorded as 0.
__device__ __forceinline__ void vg_ckd(const DSoil& s, double h, double Ss, double& C, double& K, double& dC,
                                       double& dK)
{
    if (h < 0.0) {
        const double L = log(s.al * (-h));
        const double x = exp(s.n * L);
        const double P = log1p(x);
        const double Se = exp(-s.m * P);
        const double Cw = (s.ts - s.tr) * s.m * s.n * x * Se / ((-h) * (1.0 + x));
        C = Cw + Ss;
        const double B = -expm1(-s.m * log1p(1.0 / x));
        K = s.Ks * sqrt(Se) * B * B;
        const double ih = 1.0 / h, ix1 = 1.0 / (1.0 + x);
        dC = Cw * ih * (s.n - 1.0 - s.n * (1.0 + s.m) * x * ix1);
        dK = -K * s.m * s.n * ix1 * ih * (0.5 * x + 2.0 * (1.0 - B) / B);
        if (!isfinite(dC)) dC = 0.0;
        if (!isfinite(dK)) dK = 0.0;
    } else {
        C = Ss;
        K = s.Ks;
        dC = 0.0;
        dK = 0.0;
    }
}

__device__ __forceinline__ void keff_dev(const DScen& S, double h, double K[4])
{
    const int n = S.keff_n;
    const double lh = (h < 0.0) ? log10(-h) : -INFINITY;
    double tt = (lh - S.keff_first) / S.keff_logr;
    tt = fmin(fmax(tt, 0.0), (double)(n - 1));
    int i = (int)tt;
    if (i > n - 2) i = n - 2;
    const double fr = tt - (double)i;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double a = as_global(S.keff)[c * n + i], b = as_global(S.keff)[c * n + i + 1];
        K[c] = a + fr * (b - a);
    }
}

__device__ __forceinline__ double ld_state(const GArgs& a, double ve, long long i)
{
    double x = a.u[i];
    if (a.v) x += ve * a.v[i];
    return x;
}

__device__ __forceinline__ double macro_head(const DScen& S, const GArgs& a, double ve, int j)
{
    return as_global(S.dir)[j] ? S.dirichlet_head : ld_state(a, ve, S.u_off + j);
}

// Head at state index gi (-1: a Dirichlet macro node, head dh).
__device__ __forceinline__ double state_or(const GArgs& a, double ve, int gi, double dh)
{
    return gi < 0 ? dh : ld_state(a, ve, gi);
}

__device__ __forceinline__ double wave_sum(double x) { return wave_sum_dpp(x); }

__device__ __forceinline__ void store_out(const DBatch& B, const GArgs& a, int s, long long i, double g)
{
    if (a.g0) g = (g - a.g0[i]) * a.ieps[s];
    if (!isfinite(g)) atomicOr(B.bad + s, 1);
    a.out[i] = g;
}

// ---------------------------------------------------------------- macro elements
// Vertex heads hv and interpolant coefficients al of macro element e (linear: al =
// (a, b, c) of a x + b y + c; bilinear: (a, b, c, d) of a x + b y + c x y + d).
__device__ __forceinline__ void macro_interp(const DScen& S, const GArgs& a, double ve, int e, double hv[4],
                                             double al[4])
{
    const int nv = S.nv, ne = S.ne;
    for (int i = 0; i < 4; ++i) { hv[i] = 0.0; al[i] = 0.0; }
    for (int i = 0; i < nv; ++i) hv[i] = macro_head(S, a, ve, as_global(S.el)[i * ne + e]);
    for (int k = 0; k < nv; ++k) {
        double acc = 0.0;
        for (int i = 0; i < nv; ++i) acc += hv[i] * as_global(S.sf)[(k * nv + i) * ne + e];
        al[k] = acc;
    }
}

// Workgroups of bs threads per scenario that the macro elements take in a k_micro launch.
__host__ __device__ __forceinline__ int macro_wgs(const DBatch& B, int bs) { return (B.macro_ne_max + bs - 1) / bs; }

// Macro element e of scenario s: sub-face fluxes qe (Keff at the sub-face midpoints) and
// boundary-edge fluxes bq, gathered per node by k_macro_node.  Runs in the leading
// workgroups of the k_micro launch (the micro cells do not depend on it).
__device__ __forceinline__ void macro_elem(const DBatch& B, const GArgs& a, int s, int e)
{
    const DScen& S = B.scen[s];
    if (e >= S.ne) return;
    const int nv = S.nv, ne = S.ne;
    const double ve = a.v ? a.veps[s] : 0.0;
    double hv[4], al[4];
    macro_interp(S, a, ve, e, hv, al);
    // one sub-face and one half-edge at a time (few registers: this code shares its
    // kernel's register budget with the micro cells)
#pragma unroll 1
    for (int j = 0; j < nv; ++j) {
        const double x = as_global(S.midx)[j * ne + e], y = as_global(S.midy)[j * ne + e];
        double gx, gy, hip;
        if (nv == 3) { gx = al[0]; gy = al[1]; hip = al[0] * x + al[1] * y + al[2]; }
        else { gx = al[0] + al[2] * y; gy = al[1] + al[2] * x; hip = al[0] * x + al[1] * y + al[2] * x * y + al[3]; }
        double K[4];
        keff_dev(S, hip, K);
        const double nx = as_global(S.nrm)[(2 * j) * ne + e], ny = as_global(S.nrm)[(2 * j + 1) * ne + e];
        as_global_w(S.qe)[e * nv + j] = -((K[0] * gx + K[1] * (gy + 1.0)) * nx + (K[2] * gx + K[3] * (gy + 1.0)) * ny);
    }
#pragma unroll 1
    for (int p = 0; p < nv; ++p) {
        const int type = as_global(S.bce)[p * ne + e];
#pragma unroll 1
        for (int hk = 0; hk < 2; ++hk) {
            double f = 0.0;
            if (type == TSR_BC_FLUX) {
                f = S.boundary_flux * as_global(S.be_len)[(2 * p + hk) * ne + e];
            } else if (type == TSR_BC_DHDX0 || type == TSR_BC_DHDY0) {
                const double bx = as_global(S.be_mid)[(4 * p + 2 * hk) * ne + e], by = as_global(S.be_mid)[(4 * p + 2 * hk + 1) * ne + e];
                const double bnx = as_global(S.be_nrm)[(4 * p + 2 * hk) * ne + e], bny = as_global(S.be_nrm)[(4 * p + 2 * hk + 1) * ne + e];
                double gx, gy, hip;
                if (nv == 3) { gx = al[0]; gy = al[1]; hip = al[0] * bx + al[1] * by + al[2]; }
                else { gx = al[0] + al[2] * by; gy = al[1] + al[2] * bx; hip = al[0] * bx + al[1] * by + al[2] * bx * by + al[3]; }
                if (type == TSR_BC_DHDX0) gx = 0.0; else gy = 0.0;
                double K[4];
                keff_dev(S, hip, K);
                f = (K[0] * gx + K[1] * (gy + 1.0)) * bnx + (K[2] * gx + K[3] * (gy + 1.0)) * bny;
            }
            as_global_w(S.bq)[e * 2 * nv + 2 * p + hk] = f;
        }
    }
}

// Barycentric coordinates of (px, py) in an incident macro element whose shape-function
// coefficients are m[0..8] (gradient x, gradient y, constant for each vertex).
__device__ __forceinline__ void macro_bary(const double* m, double px, double py, double l[3])
{
#pragma unroll
    for (int v = 0; v < 3; ++v) l[v] = m[v] * px + m[3 + v] * py + m[6 + v];
}

// Micro Dirichlet node i of a node-hosted cell: its position, clamped to the macro
// domain's bounding box.
__device__ __forceinline__ void dir_point(const DScen& S, const DTmpl& T, double hx, double hy, int i, double& px,
                                          double& py)
{
    px = fmin(fmax(hx + as_global(T.vec)[i], as_global(S.box)[0]), as_global(S.box)[1]);
    py = fmin(fmax(hy + as_global(T.vec)[T.nn + i], as_global(S.box)[2]), as_global(S.box)[3]);
}

// Upload-time choice of the macro element that interpolates each micro Dirichlet node
// of a node-hosted cell: among the host node's incident elements, the first with the
// largest smallest barycentric coordinate (ties within 1e-12 keep the earlier one), so
// a point on a shared edge takes the lower-numbered element.  One thread per
// (cell, node); the same expressions as the interpolation in k_micro.
__global__ __launch_bounds__(256) void k_dir_elem(DBatch B)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)B.n_cells * B.nn_max) return;
    const int cell = (int)(id / B.nn_max), i = (int)(id % B.nn_max);
    const DScen& S = B.scen[B.cell_scen[cell]];
    const DTmpl& T = B.tmpl[B.cell_tmpl[cell]];
    int best = 0;
    if (i < T.nn && as_global(T.bnd)[i] == 1 && S.host == TSR_HOST_NODE && S.model != 1) {
        const int host = cell - S.first_cell;
        const int q0 = as_global(S.ptr)[host], ninc = as_global(S.ptr)[host + 1] - q0;
        double px, py;
        dir_point(S, T, as_global(S.nodes)[host], as_global(S.nodes)[S.nn + host], i, px, py);
        double bmin = -INFINITY;
        for (int k = 0; k < ninc; ++k) {
            const int e = as_global(S.inc)[q0 + k] >> 2;
            double m[9], l[3];
#pragma unroll
            for (int j = 0; j < 9; ++j) m[j] = as_global(S.sf)[j * S.ne + e];
            macro_bary(m, px, py, l);
            const double mn = fmin(fmin(l[0], l[1]), l[2]);
            if (k == 0 || mn > bmin + 1e-12) { bmin = mn; best = k; }
        }
    }
    B.dir_elem[id] = (unsigned char)best;
}

// Upload-time copy of each node-hosted cell's macro interpolation data (one thread
// per (cell, incident element slot)).
__global__ __launch_bounds__(256) void k_cell_macro(DBatch B)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)B.n_cells * kMaxNodeInc) return;
    const int cell = (int)(id / kMaxNodeInc), k = (int)(id % kMaxNodeInc);
    const DScen& S = B.scen[B.cell_scen[cell]];
    double* sf = B.cm_sf + id * 9;
    int* vi = B.cm_vi + id * 3;
    for (int j = 0; j < 9; ++j) sf[j] = 0.0;
    for (int v = 0; v < 3; ++v) vi[v] = -1;
    if (S.host != TSR_HOST_NODE) {
        if (k == 0) { B.cm_host[cell] = -1; B.cm_ninc[cell] = 0; }
        return;
    }
    const int host = cell - S.first_cell;
    const int q0 = as_global(S.ptr)[host], ninc = as_global(S.ptr)[host + 1] - q0;
    if (k == 0) {
        B.cm_host[cell] = as_global(S.dir)[host] ? -1 : (int)(S.u_off + host);
        B.cm_ninc[cell] = (unsigned char)ninc;
    }
    if (k >= ninc) return;
    const int e = as_global(S.inc)[q0 + k] >> 2;
    for (int j = 0; j < 9; ++j) sf[j] = as_global(S.sf)[j * S.ne + e];
    for (int v = 0; v < 3; ++v) {
        const int n = as_global(S.el)[v * S.ne + e];
        vi[v] = as_global(S.dir)[n] ? -1 : (int)(S.u_off + n);
    }
}

// ---------------------------------------------------------------- micro cells
__device__ __forceinline__ int elem_soil(const DBatch& B, const DTmpl& T, int cell, int e)
{
    if (B.max_discs > 0) {
        const double cx = as_global(T.cen)[e], cy = as_global(T.cen)[T.ne + e];
        const double* d = B.discs + (size_t)cell * B.max_discs * 3;
        for (int k = 0; k < B.max_discs; ++k) {
            const double r = d[3 * k + 2];
            if (r > 0.0) {
                const double dx = cx - d[3 * k], dy = cy - d[3 * k + 1];
                if (dx * dx + dy * dy < r * r) return TSR_SOIL_B;
            }
        }
        return TSR_SOIL_A;
    }
    return T.soil ? (int)as_global(T.soil)[e] : TSR_SOIL_B;
}

// Upload-time soil map: one thread per (cell, 32-element word).
__global__ __launch_bounds__(256) void k_soil_bits(DBatch B)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)B.n_cells * B.soil_words) return;
    const int cell = (int)(id / B.soil_words), w = (int)(id % B.soil_words);
    const DTmpl& T = B.tmpl[B.cell_tmpl[cell]];
    unsigned bits = 0u;
    for (int k = 0; k < 32; ++k) {
        const int e = 32 * w + k;
        if (e < T.ne && elem_soil(B, T, cell, e) == TSR_SOIL_A) bits |= 1u << k;
    }
    B.soil_bits[id] = bits;
}

// Visit the incident (element, local vertex) pairs of node i in ascending element
// order (the order the oracle's CSR gives).  Generic templates read the CSR; the
// structured grid calls f on the (at most six) pairs of node (ix, iy) in straight-line
// code, so nothing is indexed dynamically and everything stays in registers.
__device__ __forceinline__ int div_r(const DTmpl& T, int n) { return (int)__umulhi((unsigned)n, T.mul_r); }
__device__ __forceinline__ int div_rm(const DTmpl& T, int n) { return (int)__umulhi((unsigned)n, T.mul_rm); }

// Node of the b-th Dirichlet node: structured grids walk the perimeter (bottom row,
// right column, top row, left column); generic templates use their list.
template <bool STRUCT>
__device__ __forceinline__ int dirichlet_node(const DTmpl& T, int b)
{
    if constexpr (STRUCT) {
        const int rm = T.r - 1;
        if (b < rm) return b;                                   // (b, 0)
        if (b < 2 * rm) return (b - rm) * T.r + rm;             // (r-1, b-rm)
        if (b < 3 * rm) return rm * T.r + (3 * rm - b);         // (3rm-b, r-1)
        return (4 * rm - b) * T.r;                              // (0, 4rm-b)
    } else {
        return as_global(T.bnd_list)[b];
    }
}

template <bool STRUCT, class F>
__device__ __forceinline__ void for_incident(const DTmpl& T, int i, F&& f)
{
    if constexpr (STRUCT) {
        const int r = T.r, rm = r - 1, iy = div_r(T, i), ix = i - iy * r;
        const bool L = ix >= 1, R = ix < rm, D = iy >= 1, U = iy < rm;
        const int sLD = 2 * ((iy - 1) * rm + ix - 1), sRD = 2 * ((iy - 1) * rm + ix);
        const int sLU = 2 * (iy * rm + ix - 1), sRU = 2 * (iy * rm + ix);
        if (L && D) { f(sLD, 2); f(sLD + 1, 1); }
        if (R && D) f(sRD + 1, 2);
        if (L && U) f(sLU, 1);
        if (R && U) { f(sRU, 0); f(sRU + 1, 0); }
    } else {
        const int q1 = as_global(T.ptr)[i + 1];
        for (int q = as_global(T.ptr)[i]; q < q1; ++q) {
            const int v = as_global(T.inc)[q];
            f(v >> 2, v & 3);
        }
    }
}

template <bool STRUCT, int R = 0>
__device__ __forceinline__ bool is_dirichlet(const DTmpl& T, int i)
{
    if constexpr (STRUCT) {
        const int r = R > 0 ? R : T.r;
        const int iy = R > 0 ? (int)((unsigned)i / (unsigned)R) : div_r(T, i), ix = i - iy * r;
        return ix == 0 || iy == 0 || ix == r - 1 || iy == r - 1;
    } else {
        return as_global(T.bnd)[i] == 1;
    }
}

// Element soil of a micro cell from its soil words (staged in LDS): bit e set = soil A.
constexpr int kMaxSoilWords = 64;   // micro templates of at most 2048 elements
__device__ __forceinline__ bool soil_is_a(const unsigned* sw, int e) { return (sw[e >> 5] >> (e & 31)) & 1u; }

// Net inflow of interior node i = (ix, iy) of the uniform right-triangle grid with
// spacing d (1 <= ix, iy <= r-2: all six incident elements exist).  On this grid the
// element geometry is exact rationals times d (checked against the template's own
// geometry by detect_structured):
//   A = (a, a+1, a+r+1): d grad h = (h1 - h0, h2 - h1); sub-face normals / d =
//       (1/3, -1/6), (-1/6, 1/3), (-1/6, -1/6)
//   B = (a, a+r+1, a+r): d grad h = (h1 - h2, h2 - h0); normals / d =
//       (1/6, 1/6), (-1/3, 1/6), (1/6, -1/3)
//   sub-face j's K weights (the P1 functions at its midpoint): 5/12 on v_j and v_{j+1},
//   1/6 on v_{j+2};  q_j = -K_j ((grad h + e_y) . n_j), leaving v_j, entering v_{j+1}.
// The incident pairs are summed in ascending element order, net += q_in; net -= q_out,
// as the generic visitor does:
//   LD.A (2), LD.B (1), RD.B (2), LU.A (1), RU.A (0), RU.B (0)   (local vertex of i)
template <int TYPE, int J>
__device__ __forceinline__ double q_struct(double h0, double h1, double h2, double k0, double k1, double k2, double d)
{
    constexpr double c5 = 5.0 / 12.0, c6 = 1.0 / 6.0, c3 = 1.0 / 3.0;
    double Kf;
    if constexpr (J == 0) Kf = c5 * k0 + c5 * k1 + c6 * k2;
    else if constexpr (J == 1) Kf = c6 * k0 + c5 * k1 + c5 * k2;
    else Kf = c5 * k0 + c6 * k1 + c5 * k2;
    // (grad h) . n_j = ax * (d gx) + ay * (d gy); the gravity term (e_y . n_j) = ay * d
    double gxd, gyd, ax, ay;
    if constexpr (TYPE == 0) {
        gxd = h1 - h0; gyd = h2 - h1;
        ax = J == 0 ? c3 : -c6;
        ay = J == 1 ? c3 : -c6;
    } else {
        gxd = h1 - h2; gyd = h2 - h0;
        ax = J == 1 ? -c3 : c6;
        ay = J == 2 ? -c3 : c6;
    }
    return -Kf * (gxd * ax + (gyd + d) * ay);
}

template <int TYPE, int LV>
__device__ __forceinline__ void pair_struct(double& net, int e, int v0, int v1, int v2, const double* h,
                                            const double* KA, const double* KB, const unsigned* sw, double d)
{
    const double* K = soil_is_a(sw, e) ? KA : KB;
    const double h0 = h[v0], h1 = h[v1], h2 = h[v2];
    const double k0 = K[v0], k1 = K[v1], k2 = K[v2];
    constexpr int JIN = LV == 0 ? 2 : LV - 1;
    net += q_struct<TYPE, JIN>(h0, h1, h2, k0, k1, k2, d);
    net -= q_struct<TYPE, LV>(h0, h1, h2, k0, k1, k2, d);
    __builtin_amdgcn_sched_barrier(0);   // one pair at a time (register pressure)
}

// R > 0: the grid size is a compile-time constant (neighbour offsets become immediates);
// R = 0: the template's own r.  The arithmetic is the same either way.
template <int R = 0>
__device__ __forceinline__ double struct_net(const DTmpl& T, int i, const double* h, const double* KA,
                                             const double* KB, const unsigned* sw)
{
    const int r = R > 0 ? R : T.r, rm = r - 1;
    const int iy = R > 0 ? (int)((unsigned)i / (unsigned)R) : div_r(T, i), ix = i - iy * r;
    const double d = T.dx;
    const int iLD = i - r - 1, iD = i - r, iL = i - 1, iR = i + 1, iU = i + r, iRU = i + r + 1;
    const int eLD = 2 * ((iy - 1) * rm + ix - 1), eRD = 2 * ((iy - 1) * rm + ix);
    const int eLU = 2 * (iy * rm + ix - 1), eRU = 2 * (iy * rm + ix);
    double net = 0.0;
    pair_struct<0, 2>(net, eLD, iLD, iD, i, h, KA, KB, sw, d);
    pair_struct<1, 1>(net, eLD + 1, iLD, i, iL, h, KA, KB, sw, d);
    pair_struct<1, 2>(net, eRD + 1, iD, iR, i, h, KA, KB, sw, d);
    pair_struct<0, 1>(net, eLU, iL, i, iU, h, KA, KB, sw, d);
    pair_struct<0, 0>(net, eRU, i, iR, iRU, h, KA, KB, sw, d);
    pair_struct<1, 0>(net, eRU + 1, i, iRU, iU, h, KA, KB, sw, d);
    return net;
}

// Net inflow of node i of a generic template: its incident (element, local vertex)
// pairs from the CSR in ascending element order, the two sub-face fluxes each pair
// needs, net += q_in; net -= q_out.
__device__ __forceinline__ double generic_net(const DTmpl& T, int i, const double* h, const double* KA,
                                              const double* KB, const unsigned* sw)
{
    const int ne = T.ne;
    double net = 0.0;
    for (int qq = as_global(T.ptr)[i]; qq < as_global(T.ptr)[i + 1]; ++qq) {
        const int e = as_global(T.inc)[qq] >> 2, lv = as_global(T.inc)[qq] & 3;
        const int v0 = as_global(T.el)[e], v1 = as_global(T.el)[ne + e], v2 = as_global(T.el)[2 * ne + e];
        const double h0 = h[v0], h1 = h[v1], h2 = h[v2];
        const double gx = as_global(T.sf)[e] * h0 + as_global(T.sf)[ne + e] * h1 + as_global(T.sf)[2 * ne + e] * h2;
        const double gy = as_global(T.sf)[3 * ne + e] * h0 + as_global(T.sf)[4 * ne + e] * h1 + as_global(T.sf)[5 * ne + e] * h2;
        const double* Ks = soil_is_a(sw, e) ? KA : KB;
        const double k0 = Ks[v0], k1 = Ks[v1], k2 = Ks[v2];
        const int jin = lv == 0 ? 2 : lv - 1;   // (lv + 2) % 3: the sub-face entering v_lv
        double qj[2];
#pragma unroll
        for (int t2 = 0; t2 < 2; ++t2) {
            const int j = t2 == 0 ? jin : lv;
            const double Kf = as_global(T.w)[(j * 3) * ne + e] * k0 + as_global(T.w)[(j * 3 + 1) * ne + e] * k1 + as_global(T.w)[(j * 3 + 2) * ne + e] * k2;
            qj[t2] = -Kf * (gx * as_global(T.nrm)[(2 * j) * ne + e] + (gy + 1.0) * as_global(T.nrm)[(2 * j + 1) * ne + e]);
        }
        net += qj[0];
        net -= qj[1];
    }
    return net;
}

// Linearised FD products keep each thread's nodes (tid + k bs, k < kLinNPT) in
// registers between the load phase and the gather (g0 is prefetched with the state).
constexpr int kLinNPT = 4;

// What a linearised FD product reads of one micro node: g0 at every node; at interior
// nodes the state, the direction and the recorded expansion c = (storage, its derivative,
// K_A, K_A', K_B, K_B'), values from the FP64 components and derivatives from the FP32
// ones.  Loading into this register block first and applying later lets a thread issue
// every load of its nodes before it waits on any of them.
struct LinNode {
    double u, v, c[3];   // state, direction, recorded storage / K_A / K_B (FP64)
    float d[3];          // their recorded h-derivatives (FP32, widened where used)
};

template <bool STRUCT>
__device__ __forceinline__ void lin_load(const DTmpl& T, const GArgs& a, size_t NN, long long off, int nn, int i,
                                         LinNode& x)
{
    if (i >= nn) return;
    const long long gi = off + i;
    if (is_dirichlet<STRUCT>(T, i)) return;
    x.u = a.u[gi];
    x.v = a.v ? a.v[gi] : 0.0;
    const double* L = a.lin + gi;
    const float* Ld = reinterpret_cast<const float*>(a.lin + kLinComp64 * NN) + gi;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        x.c[c] = L[(c + 1) * NN];
        x.d[c] = Ld[c * NN];
    }
}

// Recorded head (element-hosted cells only: node-hosted cells take the head's offset
// from the macro perturbation) and the four conductivity components at a Dirichlet node.
struct LinDir {
    double hb, c[2];     // recorded head, K_A, K_B (FP64)
    float d[2];          // h-derivatives of K_A, K_B (FP32)
};

__device__ __forceinline__ void lin_load_dir(const GArgs& a, size_t NN, long long gi, bool need_hb, LinDir& x)
{
    const double* L = a.lin + gi;
    const float* Ld = reinterpret_cast<const float*>(a.lin + kLinComp64 * NN) + gi;
    x.hb = need_hb ? L[0] : 0.0;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        x.c[c] = L[(c + 2) * NN];
        x.d[c] = Ld[(c + 1) * NN];
    }
}

__device__ __forceinline__ void store_fd(const DBatch& B, const GArgs& a, int s, long long i, double g, double g0v)
{
    g = (g - g0v) * a.ieps[s];
    if (!isfinite(g)) atomicOr(B.bad + s, 1);
    __builtin_nontemporal_store(g, &a.out[i]);
}

// Launch shape (tuned per problem by tune_micro): workgroup size bs (64..1024, a
// multiple of 64, each thread owning nodes tid, tid + bs, ...; at most 8 of them) and a
// register budget WPE (amdgpu_waves_per_eu: 1 = compiler default, 6, 8).  Results do
// not depend on the shape: every per-node computation is the same whichever thread does
// it, and the exchange sum uses a fixed 256-slot order.
//
// MODE 0 evaluates G; MODE 1 also records the expansion point (GArgs::rec); MODE 2 takes
// capacities and conductivities from the recorded first-order expansion (GArgs::lin).
template <bool STRUCT, int WPE, int MODE>
__global__ __launch_bounds__(kMicroMaxBS) __attribute__((amdgpu_waves_per_eu(WPE))) void k_micro(DBatch B, GArgs a)
{
    extern __shared__ __align__(16) double smem[];
    __shared__ double s_bc[6];
    // per incident macro element of a node-hosted cell: shape-function coefficients (9),
    // vertex heads (3) and, for MODE 2, the vertex heads' perturbations (3)
    __shared__ double s_mel[kMaxNodeInc][16];
    __shared__ double s_q[4];
    __shared__ double s_geo[2][22];
    __shared__ unsigned s_sw[kMaxSoilWords];
    // the leading workgroups of the launch evaluate the macro elements (one thread each,
    // every scenario); the rest are the micro cells
    const int mwg = macro_wgs(B, blockDim.x);
    if ((int)blockIdx.x < mwg * B.n_scen) {
        const int sm = blockIdx.x / mwg;
        if (!a.active || a.active[sm]) macro_elem(B, a, sm, (blockIdx.x % mwg) * blockDim.x + threadIdx.x);
        return;
    }
    const int cell = blockIdx.x - mwg * B.n_scen;
    const int s = B.cell_scen[cell];
    if (a.active && !a.active[s]) return;
    const int t = B.cell_tmpl[cell];
    const DTmpl& T = B.tmpl[t];
    const DScen& S = B.scen[s];
    const int nn = T.nn, ne = T.ne, tid = threadIdx.x, bs = blockDim.x;
    const DSoil soilA = S.A, soilB = S.B;   // uniform: scalar registers
    double* h = smem;
    double* KA = h + B.nn_max;
    double* KB = KA + B.nn_max;
    // MODE 0/1: node storage capacity, then the node's net inflow.  MODE 2 keeps both in
    // registers (each thread owns its nodes from load to gather) and has no such array.
    double* stl = KB + B.nn_max;
    const unsigned* sw = s_sw;
    const double ve = a.v ? a.veps[s] : 0.0;
    const long long off = B.cell_off[cell];
    const int host = cell - S.first_cell;
    const int model = S.model;
    const bool node_host = (S.host == TSR_HOST_NODE);
    const size_t NN = (size_t)B.N;
    const int nbnd = STRUCT ? 4 * (T.r - 1) : T.nbnd;
    const int nsw = B.soil_words;

    // ---- Loads.  Everything that depends only on the cell is issued before any of it is
    // used, so one memory round trip (two for the macro heads around a node-hosted cell:
    // vertex index, then head) precedes the first barrier: the cell's soil words; this
    // thread's first two nodes (their state; MODE 2: the recorded expansion, and the
    // recorded conductivities of the thread's first Dirichlet node); for node-hosted
    // cells the host's head and position and, for models 2/3, the shape-function
    // coefficients, vertex heads and vertex perturbations of the host's incident macro
    // elements (unused slots hold zero coefficients and the Dirichlet head).
    const unsigned swv = tid < nsw ? B.soil_bits[(size_t)cell * nsw + tid] : 0u;
    int best = 0;   // node-hosted cells: macro element of this thread's first Dirichlet node
    if (node_host && model != 1 && tid < nbnd) best = B.dir_elem[(size_t)cell * B.nn_max + dirichlet_node<STRUCT>(T, tid)];
    LinNode xn[2];
    LinDir xd;
    double hs[2] = {0.0, 0.0};   // MODE 0/1: state at this thread's first two nodes
    if constexpr (MODE == 2) {
        lin_load<STRUCT>(T, a, NN, off, nn, tid, xn[0]);
        lin_load<STRUCT>(T, a, NN, off, nn, tid + bs, xn[1]);
        if (tid < nbnd) lin_load_dir(a, NN, off + dirichlet_node<STRUCT>(T, tid), !node_host, xd);
    } else {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int i = tid + k * bs;
            if (i < nn && !is_dirichlet<STRUCT>(T, i)) hs[k] = ld_state(a, ve, off + i);
        }
    }
    // the host's incident-element data are spread over the threads (one coefficient, or
    // one vertex head and its perturbation, per thread: few registers each); both
    // tables are contiguous per cell, so the loads coalesce
    const bool has_mel = node_host && model != 1;
    constexpr int kMelSf = kMaxNodeInc * 9, kMelV = kMaxNodeInc * 3;
    double msf[2] = {0.0, 0.0}, mhv = 0.0, mpv = 0.0;
    if (has_mel) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int q = tid + r * bs;
            if (q < kMelSf) msf[r] = B.cm_sf[(size_t)cell * kMelSf + q];
        }
        if (tid < kMelV) {
            const int gi = B.cm_vi[(size_t)cell * kMelV + tid];
            mhv = state_or(a, ve, gi, S.dirichlet_head);
            mpv = (MODE == 2 && a.v && gi >= 0) ? ve * a.v[gi] : 0.0;
        }
    }
    double hb0 = 0.0, hb1 = 0.0;   // tid 0 of a node-hosted cell: host head and its perturbation
    if (node_host && tid == 0) {
        // the host's state and its Dirichlet flag are loaded side by side
        const long long gh = S.u_off + host;
        const double hu = ld_state(a, ve, gh);
        const double pu = (MODE == 2 && a.v) ? ve * a.v[gh] : 0.0;
        const bool hdir = as_global(S.dir)[host] != 0;
        hb0 = hdir ? S.dirichlet_head : hu;
        hb1 = hdir ? 0.0 : pu;
    }

    // ---- LDS stores of the loaded data (the element-hosted interpolant is formed here)
    if (tid < nsw) s_sw[tid] = swv;
    if (has_mel) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int q = tid + r * bs;
            if (q < kMelSf) s_mel[q / 9][q % 9] = msf[r];
        }
        if (tid < kMelV) {
            s_mel[tid / 3][9 + tid % 3] = mhv;
            s_mel[tid / 3][12 + tid % 3] = mpv;
        }
    }
    if (tid == 0) {
        if (!node_host) {
            double hv[4], al[4];
            macro_interp(S, a, ve, host, hv, al);
            if (S.nv == 3) { s_bc[0] = al[0]; s_bc[1] = al[1]; s_bc[2] = 0.0; s_bc[3] = al[2]; }
            else { s_bc[0] = al[0]; s_bc[1] = al[1]; s_bc[2] = al[2]; s_bc[3] = al[3]; }
            s_bc[4] = as_global(S.cen)[host];
            s_bc[5] = as_global(S.cen)[S.ne + host];
        } else {
            s_bc[2] = hb1;   // host perturbation (MODE 2, model 1)
            s_bc[3] = hb0;
            s_bc[4] = as_global(S.nodes)[host];
            s_bc[5] = as_global(S.nodes)[S.nn + host];
        }
    }
    if constexpr (STRUCT) {
        if (tid < 44) s_geo[tid / 22][tid % 22] = as_global(T.sg)[tid / 22][tid % 22];
    }
    // MODE 2: interior heads, storage capacities (registers) and conductivities (LDS) of
    // this thread's first two nodes from the recorded expansion; MODE 0/1: their heads
    double g0r[kLinNPT], str[kLinNPT];   // MODE 2: G(u) and storage of this thread's nodes
    auto lin_node = [&](int k, const LinNode& x) {
        const int i = tid + k * bs;
        str[k] = 1.0;
        if (i >= nn) return;
        if (is_dirichlet<STRUCT>(T, i)) return;
        // the perturbation itself is the offset from the expansion point (the recorded
        // head is u), so component 0 is not read here
        const double dv = a.v ? ve * x.v : 0.0;
        h[i] = x.u + dv;
        str[k] = x.c[0] + (double)x.d[0] * dv;
        KA[i] = x.c[1] + (double)x.d[1] * dv;
        KB[i] = x.c[2] + (double)x.d[2] * dv;
    };
    if constexpr (MODE == 2) {
        lin_node(0, xn[0]);
        lin_node(1, xn[1]);
    } else {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int i = tid + k * bs;
            if (i < nn && !is_dirichlet<STRUCT>(T, i)) h[i] = hs[k];
        }
    }
    __syncthreads();

    // ---- the rest of the nodes (templates with more than two nodes per thread)
    if constexpr (MODE == 2) {
#pragma unroll
        for (int k = 2; k < kLinNPT; ++k) {
            LinNode& x = xn[k & 1];
            lin_load<STRUCT>(T, a, NN, off, nn, tid + k * bs, x);
            lin_node(k, x);
        }
    } else {
        for (int i = tid + 2 * bs; i < nn; i += bs)
            if (!is_dirichlet<STRUCT>(T, i)) h[i] = ld_state(a, ve, off + i);
    }
    // Dirichlet heads from the macro field (first nbnd threads); MODE 2 also expands the
    // conductivities there about the recorded head, by the head's perturbation d
    for (int b = tid; b < nbnd; b += bs) {
        const int i = dirichlet_node<STRUCT>(T, b);
        double hi, d = 0.0;
        if (node_host) {
            if (model == 1) {
                hi = s_bc[3];
                d = s_bc[2];
            } else {
                // P1 interpolation in the incident macro element chosen at upload
                // (k_dir_elem): weights max(lambda, 0), normalised
                double px, py, lam[3];
                dir_point(S, T, s_bc[4], s_bc[5], i, px, py);
                if (b != tid) best = B.dir_elem[(size_t)cell * B.nn_max + i];
                macro_bary(s_mel[best], px, py, lam);
                double ws = 0.0, hsum = 0.0, psum = 0.0;
#pragma unroll
                for (int v = 0; v < 3; ++v) {
                    const double w = fmax(lam[v], 0.0);
                    ws += w;
                    hsum += w * s_mel[best][9 + v];
                    psum += w * s_mel[best][12 + v];
                }
                hi = hsum / ws;
                d = psum / ws;
            }
        } else {
            const double X = (model == 1) ? s_bc[4] : s_bc[4] + as_global(T.vec)[i];
            const double Y = (model == 1) ? s_bc[5] : s_bc[5] + as_global(T.vec)[nn + i];
            hi = s_bc[0] * X + s_bc[1] * Y + s_bc[2] * X * Y + s_bc[3];
        }
        h[i] = hi;
        if constexpr (MODE == 2) {
            if (b != tid) lin_load_dir(a, NN, off + i, !node_host, xd);
            if (!node_host) d = hi - xd.hb;
            KA[i] = xd.c[0] + (double)xd.d[0] * d;
            KB[i] = xd.c[1] + (double)xd.d[1] * d;
        }
    }
    __syncthreads();

    // node-centric constitutive evaluation (MODE 0/1)
    if constexpr (MODE != 2) {
        for (int i = tid; i < nn; i += bs) {
            {
                double cva = 0.0, cvb = 0.0;
                for_incident<STRUCT>(T, i, [&](int e, int) {
                    const double sc = STRUCT ? s_geo[e & 1][21] : as_global(T.scv)[e];
                    if (soil_is_a(sw, e)) cva += sc; else cvb += sc;
                });
                double cA = 0.0, cB = 0.0, kA = 0.0, kB = 0.0, dcA = 0.0, dcB = 0.0, dkA = 0.0, dkB = 0.0;
                // first evaluation with the lane's own soil (no divergence), second only at interfaces
                const bool hasA = cva > 0.0, hasB = cvb > 0.0;
                {
                    DSoil sl;
                    sl.Ks = hasA ? soilA.Ks : soilB.Ks; sl.tr = hasA ? soilA.tr : soilB.tr;
                    sl.ts = hasA ? soilA.ts : soilB.ts; sl.al = hasA ? soilA.al : soilB.al;
                    sl.n = hasA ? soilA.n : soilB.n; sl.m = hasA ? soilA.m : soilB.m;
                    double C, K, dC = 0.0, dK = 0.0;
                    if constexpr (MODE == 1) vg_ckd(sl, h[i], B.Ss, C, K, dC, dK);
                    else vg_ck(sl, h[i], B.Ss, C, K);
                    if (hasA) { cA = C; kA = K; dcA = dC; dkA = dK; } else { cB = C; kB = K; dcB = dC; dkB = dK; }
                }
                if (hasA && hasB) {
                    double C, K, dC = 0.0, dK = 0.0;
                    if constexpr (MODE == 1) vg_ckd(soilB, h[i], B.Ss, C, K, dC, dK);
                    else vg_ck(soilB, h[i], B.Ss, C, K);
                    cB = C; kB = K; dcB = dC; dkB = dK;
                }
                KA[i] = kA;
                KB[i] = kB;
                const double st = cA * cva + cB * cvb;
                stl[i] = st;
                if constexpr (MODE == 1) {
                    double* R = a.rec + off + i;
                    R[0] = h[i];
                    R[NN] = st;
                    R[2 * NN] = kA;
                    R[3 * NN] = kB;
                    float* Rd = reinterpret_cast<float*>(a.rec + kLinComp64 * NN) + off + i;
                    Rd[0] = (float)(dcA * cva + dcB * cvb);
                    Rd[NN] = (float)dkA;
                    Rd[2 * NN] = (float)dkB;
                }
            }
        }
        __syncthreads();
    }

    // Exchange Q = sum of the interior net inflows in a fixed order, independent of bs:
    // 256 slots t take nodes t, t + 256, ... in increasing order; the 64 slots of group g
    // are added by the wavefront shuffle tree and the four group sums in a fixed tree.
    // bs is a multiple of 64, so the slots tid + j bs a thread holds for j < 256 / bs
    // put group g = w + j bs / 64 in one wavefront w = tid / 64, lane = slot - 64 g:
    // that wavefront sums it (slot_value(j): the thread's j-th slot).
    auto exchange_groups = [&](auto slot_value) {
        const int wv = tid >> 6;
#pragma unroll
        for (int j2 = 0; j2 < 4; ++j2) {
            if (j2 * bs >= 256) break;
            const int g = wv + j2 * (bs >> 6);
            if (g < 4) {   // uniform across the wavefront: every lane takes part
                const double r = wave_sum(slot_value(j2));
                if ((tid & 63) == 0) s_q[g] = r;
            }
        }
    };

    // node-centric fluxes and gather: each interior node visits its incident (element,
    // local vertex) pairs in ascending element order and evaluates the two sub-face
    // fluxes it needs, q_j = -K_j (grad h . n_j + n_j,y) (q_j leaves v_j, enters
    // v_{j+1}): net = sum q_{(lv+2)%3} - q_{lv}.  Each sub-face flux is the same
    // expression whichever node evaluates it, so the result does not depend on who does.
    auto node_net = [&](int i) -> double {
        if constexpr (STRUCT) return struct_net(T, i, h, KA, KB, sw);
        else return generic_net(T, i, h, KA, KB, sw);
    };
    if constexpr (MODE == 2) {
        // G(u) at this thread's interior nodes, loaded now so the loads are in flight
        // during the gather (G(u) is 0 at micro Dirichlet nodes: nothing to load there)
#pragma unroll
        for (int k = 0; k < kLinNPT; ++k) {
            const int i = tid + k * bs;
            g0r[k] = (i < nn && !is_dirichlet<STRUCT>(T, i)) ? a.g0[off + i] : 0.0;
        }
        // the exchange slots (node i -> slot i % 256) this thread owns: with bs dividing
        // 256, node tid + k bs falls in slot tid + (k % (256 / bs)) bs, and each slot's
        // nodes are added in increasing order, as in the LDS version below
        double qs[kLinNPT] = {0.0, 0.0, 0.0, 0.0};
        const int nslot = 256 / bs;
#pragma unroll
        for (int k = 0; k < kLinNPT; ++k) {
            const int i = tid + k * bs;
            if (i >= nn) continue;
            double gi = 0.0, net = 0.0;
            if (!is_dirichlet<STRUCT>(T, i)) {
                net = node_net(i);
                gi = net / str[k];
            }
#pragma unroll
            for (int j2 = 0; j2 < kLinNPT; ++j2)
                if (j2 == k % nslot) qs[j2] += net;
            store_fd(B, a, s, off + i, gi, g0r[k]);
        }
        exchange_groups([&](int j2) { return qs[j2]; });
        __syncthreads();
        if (tid == 0) B.Q[cell] = (s_q[0] + s_q[1]) + (s_q[2] + s_q[3]);
        return;
    } else {
        for (int i = tid; i < nn; i += bs) {
            double gi = 0.0, net = 0.0;
            if (!is_dirichlet<STRUCT>(T, i)) {
                net = node_net(i);
                gi = net / stl[i];
            }
            stl[i] = net;   // only this thread reads or writes entry i in this phase
            store_out(B, a, s, off + i, gi);
        }
    }
    __syncthreads();
    exchange_groups([&](int j2) {
        double Qacc = 0.0;
        for (int i = tid + j2 * bs; i < nn; i += 256) Qacc += stl[i];
        return Qacc;
    });
    __syncthreads();
    if (tid == 0) B.Q[cell] = (s_q[0] + s_q[1]) + (s_q[2] + s_q[3]);
}

// ---------------------------------------------------------------- linearised FD products on uniform grids
// Perimeter position p (0 <= p < 4 (r - 1)) of an r x r grid -> node index: the bottom
// row, the top row, then the left and the right column without their corners.
__device__ __forceinline__ int perim_node(int p, int r)
{
    if (p < r) return p;
    if (p < 2 * r) return (r - 1) * r + p - r;
    const int q = p - 2 * r;
    return q < r - 2 ? (q + 1) * r : (q - r + 3) * r + r - 1;
}

// Largest grid size of a k_micro_lin problem (templates of at most 1024 nodes).
constexpr int kLinMaxR = 32;

// k_micro_lin<R, W, NPT, WPE>: the linearised FD product (k_micro mode 2) for problems
// whose micro templates are all uniform r x r grids, one workgroup of W wavefronts per
// cell, thread tid owning nodes tid + k 64W (k < NPT) and perimeter positions
// tid + q 64W.  R > 0 is the grid size as a compile-time constant (every template of
// the problem has r = R): neighbour offsets and the node -> (ix, iy) split are then
// immediates; R = 0 reads r per cell.  WPE is the register budget (waves per SIMD).
//
// Every load a cell needs of HBM is issued before the first wait: per owned node the
// recorded conductivities and their derivatives, and at interior nodes the state, the
// direction, the recorded storage and G(u); per perimeter position the Dirichlet node's
// interpolation data; the soil words; the macro interpolation data of the host
// (node-hosted cells) or the host element's interpolant (element-hosted).  One memory
// round trip (two for the macro heads: vertex index, then head) therefore precedes the
// gather, and with W = 1 the workgroup barriers are a single wavefront's.  The Dirichlet
// nodes are evaluated densely by perimeter position (4 (r - 1) of them: 2 of 8
// wavefront passes at r = 21, 64 W = 256), their owners parking the recorded
// conductivities in the nodes' LDS slots.  Per node the arithmetic is that of k_micro
// mode 2 (same expressions, same order), and the exchange slots are summed in the same
// order, so the results are the same bit for bit.
template <int R, int W, int NPT, int WPE>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(WPE))) void k_micro_lin(DBatch B, GArgs a)
{
    constexpr int BS = 64 * W;
    extern __shared__ __align__(16) double smem[];
    __shared__ double s_bc[6];
    __shared__ double s_mel[kMaxNodeInc][16];
    __shared__ double s_q[4];
    __shared__ unsigned s_sw[kMaxSoilWords];
    const int mwg = macro_wgs(B, BS);
    if ((int)blockIdx.x < mwg * B.n_scen) {
        const int sm = blockIdx.x / mwg;
        if (!a.active || a.active[sm]) macro_elem(B, a, sm, (blockIdx.x % mwg) * BS + threadIdx.x);
        return;
    }
    const int cell = blockIdx.x - mwg * B.n_scen;
    const int tid = threadIdx.x;
    // ---- loads.  First everything indexed by the cell alone (one memory level): the
    // host's macro interpolation data (every cell has its slots: -1 vertex indices and
    // zero coefficients where unused), the soil words, the cell's scenario, template
    // and state offset.  Nothing here waits on the scenario's description.
    constexpr int kMelSf = kMaxNodeInc * 9, kMelV = kMaxNodeInc * 3;
    int gi = -1;
    double msf[2] = {0.0, 0.0};
    if (tid < kMelV) gi = B.cm_vi[(size_t)cell * kMelV + tid];
#pragma unroll
    for (int q2 = 0; q2 < 2; ++q2) {
        const int q = tid + q2 * BS;
        if (q < kMelSf) msf[q2] = B.cm_sf[(size_t)cell * kMelSf + q];
    }
    const int nsw = B.soil_words;
    const unsigned swv = tid < nsw ? B.soil_bits[(size_t)cell * nsw + tid] : 0u;
    const bool uni = B.cell_stride > 0;   // one scenario, one template, cells in order
    const int s = uni ? 0 : B.cell_scen[cell];
    const long long off = uni ? B.cell_off0 + (long long)cell * B.cell_stride : B.cell_off[cell];
    const DTmpl& T = B.tmpl[uni ? 0 : B.cell_tmpl[cell]];
    if (a.active && !a.active[s]) return;
    const DScen& S = B.scen[s];
    const int r = R > 0 ? R : T.r;
    const int nn = R > 0 ? R * R : T.nn;
    const int ld = R > 0 ? R * R : B.nn_max;   // LDS stride of the three node arrays
    double* h = smem;
    double* KA = h + ld;
    double* KB = KA + ld;
    const size_t NN = (size_t)B.N;
    // per-cell stream bases (uniform): the loads below index them by node only
    const double* pu = a.u + off;
    const double* pv = a.v ? a.v + off : nullptr;
    const double* pg = a.g0 + off;
    const double* L = a.lin + off;
    const float* Ld = reinterpret_cast<const float*>(a.lin + kLinComp64 * NN) + off;

    // then the nodes (two levels: offset, then data): every owned node's recorded
    // conductivities and derivatives, and interior nodes' state, direction, recorded
    // storage and G(u) (n0..n3, nf: zero at Dirichlet nodes).  The Dirichlet nodes are
    // also taken densely by perimeter slots, thread tid having perimeter positions
    // tid + q BS: their template position, interpolating macro element and (for
    // element-hosted cells) recorded head.
    double n0[NPT], n1[NPT], n2[NPT], n3[NPT], xa[NPT], xb[NPT];
    float nf[NPT], da[NPT], db[NPT];
    bool dir[NPT];
    const int model = S.model;
    const bool node_host = (S.host == TSR_HOST_NODE);
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
        const int i = tid + k * BS;
        n0[k] = n1[k] = n2[k] = n3[k] = xa[k] = xb[k] = 0.0;
        nf[k] = da[k] = db[k] = 0.0f;
        dir[k] = false;
        if (i >= nn) continue;
        dir[k] = is_dirichlet<true, R>(T, i);
        xa[k] = L[2 * NN + i];
        xb[k] = L[3 * NN + i];
        da[k] = Ld[NN + i];
        db[k] = Ld[2 * NN + i];
        if (!dir[k]) {
            n0[k] = pu[i];
            n1[k] = pv ? pv[i] : 0.0;
            n2[k] = L[NN + i];
            nf[k] = Ld[i];
            n3[k] = pg[i];
        }
    }
    constexpr int NPER = (4 * ((R > 0 ? R : kLinMaxR) - 1) + BS - 1) / BS;
    const int nper = 4 * (r - 1);
    double bx[NPER], by[NPER], bh[NPER];
    int be[NPER];
#pragma unroll
    for (int q = 0; q < NPER; ++q) {
        const int p = tid + q * BS;
        bx[q] = by[q] = bh[q] = 0.0;
        be[q] = 0;
        if (p >= nper) continue;
        const int i = perim_node(p, r);
        be[q] = B.dir_elem[(size_t)cell * B.nn_max + i];
        bx[q] = as_global(T.vec)[i];
        by[q] = as_global(T.vec)[nn + i];
        if (!node_host) bh[q] = L[i];
    }
    // last, what waits on a loaded index: the macro vertex heads of a node-hosted cell,
    // the host's head
    const bool has_mel = node_host && model != 1;
    const int host = cell - S.first_cell;
    double mu = 0.0, mv = 0.0;
    if (has_mel && gi >= 0) {
        mu = a.u[gi];
        if (a.v) mv = a.v[gi];
    }
    double hu0 = 0.0, hv0 = 0.0;
    bool hdir = false;
    if (node_host && tid == 0) {
        const long long gh = S.u_off + host;
        hu0 = a.u[gh];
        if (a.v) hv0 = a.v[gh];
        hdir = as_global(S.dir)[host] != 0;
    }
    const double ve = a.v ? a.veps[s] : 0.0;
    // the same expressions as state_or / ld_state: u + ve v, or the Dirichlet head
    const double mhv = gi < 0 ? S.dirichlet_head : (a.v ? mu + ve * mv : mu);
    const double mpv = (a.v && gi >= 0) ? ve * mv : 0.0;
    double hb0 = 0.0, hb1 = 0.0;
    if (node_host && tid == 0) {
        const double hu = a.v ? hu0 + ve * hv0 : hu0;
        const double pu0 = a.v ? ve * hv0 : 0.0;
        hb0 = hdir ? S.dirichlet_head : hu;
        hb1 = hdir ? 0.0 : pu0;
    }

    // ---- LDS: soil words, macro data, interior nodes' heads and conductivities
    if (tid < nsw) s_sw[tid] = swv;
    if (has_mel) {
#pragma unroll
        for (int q2 = 0; q2 < 2; ++q2) {
            const int q = tid + q2 * BS;
            if (q < kMelSf) s_mel[q / 9][q % 9] = msf[q2];
        }
        if (tid < kMelV) {
            s_mel[tid / 3][9 + tid % 3] = mhv;
            s_mel[tid / 3][12 + tid % 3] = mpv;
        }
    }
    if (tid == 0) {
        if (!node_host) {
            double hv[4], al[4];
            macro_interp(S, a, ve, host, hv, al);
            if (S.nv == 3) { s_bc[0] = al[0]; s_bc[1] = al[1]; s_bc[2] = 0.0; s_bc[3] = al[2]; }
            else { s_bc[0] = al[0]; s_bc[1] = al[1]; s_bc[2] = al[2]; s_bc[3] = al[3]; }
            s_bc[4] = as_global(S.cen)[host];
            s_bc[5] = as_global(S.cen)[S.ne + host];
        } else {
            s_bc[2] = hb1;
            s_bc[3] = hb0;
            s_bc[4] = as_global(S.nodes)[host];
            s_bc[5] = as_global(S.nodes)[S.nn + host];
        }
    }
    // (a Dirichlet node parks its recorded conductivities and derivatives in its own
    // LDS slots until its head is known, so they hold no registers across the barrier)
    double str[NPT];
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
        const int i = tid + k * BS;
        str[k] = 1.0;
        if (i >= nn) continue;
        if (dir[k]) {
            h[i] = __hiloint2double(__float_as_int(db[k]), __float_as_int(da[k]));
            KA[i] = xa[k];
            KB[i] = xb[k];
            continue;
        }
        const double dv = a.v ? ve * n1[k] : 0.0;
        h[i] = n0[k] + dv;
        str[k] = n2[k] + (double)nf[k] * dv;
        KA[i] = xa[k] + (double)da[k] * dv;
        KB[i] = xb[k] + (double)db[k] * dv;
    }
    __syncthreads();

    // ---- Dirichlet nodes: heads from the macro field, conductivities expanded about the
    // recorded head by the head's perturbation d
#pragma unroll
    for (int q = 0; q < NPER; ++q) {
        const int p = tid + q * BS;
        if (p >= nper) continue;
        const int i = perim_node(p, r);
        double hi, d = 0.0;
        if (node_host) {
            if (model == 1) {
                hi = s_bc[3];
                d = s_bc[2];
            } else {
                // dir_point's expressions with the prefetched template position
                const double px = fmin(fmax(s_bc[4] + bx[q], as_global(S.box)[0]), as_global(S.box)[1]);
                const double py = fmin(fmax(s_bc[5] + by[q], as_global(S.box)[2]), as_global(S.box)[3]);
                double lam[3];
                macro_bary(s_mel[be[q]], px, py, lam);
                double ws = 0.0, hsum = 0.0, psum = 0.0;
#pragma unroll
                for (int v = 0; v < 3; ++v) {
                    const double w = fmax(lam[v], 0.0);
                    ws += w;
                    hsum += w * s_mel[be[q]][9 + v];
                    psum += w * s_mel[be[q]][12 + v];
                }
                hi = hsum / ws;
                d = psum / ws;
            }
        } else {
            const double X = (model == 1) ? s_bc[4] : s_bc[4] + bx[q];
            const double Y = (model == 1) ? s_bc[5] : s_bc[5] + by[q];
            hi = s_bc[0] * X + s_bc[1] * Y + s_bc[2] * X * Y + s_bc[3];
            d = hi - bh[q];
        }
        const double pd = h[i], ka = KA[i], kb = KB[i];
        h[i] = hi;
        KA[i] = ka + (double)__int_as_float(__double2loint(pd)) * d;
        KB[i] = kb + (double)__int_as_float(__double2hiint(pd)) * d;
    }
    __syncthreads();

    // ---- gather, the FD quotient, and the exchange slots (node i -> slot i % 256; the
    // slots of a group of 64 lie in one wavefront since W divides 4)
    constexpr int NSLOT = 4 / W;
    double qs[NSLOT];
#pragma unroll
    for (int j2 = 0; j2 < NSLOT; ++j2) qs[j2] = 0.0;
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
        const int i = tid + k * BS;
        if (i >= nn) continue;
        double gi = 0.0, net = 0.0;
        if (!dir[k]) {
            net = struct_net<R>(T, i, h, KA, KB, s_sw);
            gi = net / str[k];
        }
        qs[k % NSLOT] += net;
        store_fd(B, a, s, off + i, gi, n3[k]);
    }
    const int wv = tid >> 6;
#pragma unroll
    for (int j2 = 0; j2 < NSLOT; ++j2) {
        const double q = wave_sum(qs[j2]);
        if ((tid & 63) == 0) s_q[wv + j2 * W] = q;
    }
    __syncthreads();
    if (tid == 0) B.Q[cell] = (s_q[0] + s_q[1]) + (s_q[2] + s_q[3]);
}

// ---------------------------------------------------------------- macro nodes
__global__ __launch_bounds__(256) void k_macro_node(DBatch B, GArgs a)
{
    const int s = blockIdx.y;
    if (a.active && !a.active[s]) return;
    const DScen& S = B.scen[s];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= S.nn) return;
    const int nv = S.nv, ne = S.ne;
    const int q0 = as_global(S.ptr)[j], q1 = as_global(S.ptr)[j + 1];
    // what does not wait on the incidence list is loaded first: the node's state and
    // Dirichlet flag, and (node-hosted model 3) its cell's exchange term
    const bool dj = as_global(S.dir)[j] != 0;
    const double hj = ld_state(a, a.v ? a.veps[s] : 0.0, S.u_off + j);
    const bool qnode = S.model == 3 && S.host != TSR_HOST_ELEMENT;
    double qc = 0.0, qw = 0.0;
    if (qnode) {
        const int cell = S.first_cell + j;
        qc = B.Q[cell];
        qw = as_global(S.cv)[j] / B.tmpl[B.cell_tmpl[cell]].cell_area;
    }
    // the incident (element, vertex) pairs in CSR order, kMacroNodeBatch at a time: every
    // pair's index, then every flux of the batch, is loaded before any is summed (two
    // memory round trips per batch instead of two per pair); the sum keeps the CSR order
    constexpr int kMacroNodeBatch = 8;
    double net = 0.0;
    for (int qb = q0; qb < q1; qb += kMacroNodeBatch) {
        int ev[kMacroNodeBatch];
#pragma unroll
        for (int q = 0; q < kMacroNodeBatch; ++q) ev[q] = qb + q < q1 ? as_global(S.inc)[qb + q] : 0;
        double f[kMacroNodeBatch][4];
#pragma unroll
        for (int q = 0; q < kMacroNodeBatch; ++q) {
            f[q][0] = f[q][1] = f[q][2] = f[q][3] = 0.0;
            if (qb + q >= q1) continue;
            const int e = ev[q] >> 2, lv = ev[q] & 3;
            const int prev = (lv + nv - 1) % nv;
            f[q][0] = as_global(S.qe)[e * nv + prev];
            f[q][1] = as_global(S.qe)[e * nv + lv];
            f[q][2] = as_global(S.bq)[e * 2 * nv + 2 * lv];
            f[q][3] = as_global(S.bq)[e * 2 * nv + 2 * prev + 1];
        }
#pragma unroll
        for (int q = 0; q < kMacroNodeBatch; ++q) {
            if (qb + q >= q1) break;
            net += f[q][0];
            net -= f[q][1];
            net += f[q][2];
            net += f[q][3];
        }
    }
    if (qnode) {
        net -= qc * qw;
    } else if (S.model == 3) {
        for (int qq = q0; qq < q1; ++qq) {
            const int e = as_global(S.inc)[qq] >> 2, lv = as_global(S.inc)[qq] & 3;
            const int cell = S.first_cell + e;
            net -= B.Q[cell] * as_global(S.scv)[lv * ne + e] / B.tmpl[B.cell_tmpl[cell]].cell_area;
        }
    }
    double g = 0.0;
    if (!dj) {
        double C, K;
        vg_ck(S.A, hj, B.Ss, C, K);
        g = net / (S.epsA * C * as_global(S.cv)[j]);
    }
    store_out(B, a, s, S.u_off + j, g);
}

// ---------------------------------------------------------------- upload
struct DevProblem {
    void* buf = nullptr;              // one HBM allocation
    size_t bytes = 0;
    DBatch B{};
    size_t lds = 0, lds2 = 0;           // dynamic LDS of k_micro (modes 0/1, mode 2)
    bool all_structured = false;      // every template is a uniform grid: analytic k_micro
    int micro_bs[3] = {0, 0, 0};      // tuned launch shape per k_micro mode (0 = not tuned yet)
    int micro_wpe[3] = {1, 1, 1};
    // linearised FD products by k_micro_lin: -1 not used (generic templates), 0 with the
    // grid size read per cell, R > 0 with R compiled in (every template is R x R)
    int lin_r = -1;
    double tune_ms[64];               // measured candidates (for reports)
    int tune_n = 0;
    long long n_macro = 0;            // macro nodes of all scenarios (the rest of N are micro nodes)
};

namespace {

struct Arena {
    std::vector<unsigned char> host;
    template <class T> size_t put(const T* src, size_t n)
    {
        size_t off = (host.size() + 255) & ~size_t(255);
        host.resize(off + sizeof(T) * n);
        if (n) std::memcpy(host.data() + off, src, sizeof(T) * n);
        return off;
    }
    template <class T> size_t reserve(size_t n)
    {
        size_t off = (host.size() + 255) & ~size_t(255);
        host.resize(off + sizeof(T) * n, 0);
        return off;
    }
};

template <class T> T* at(void* base, size_t off) { return reinterpret_cast<T*>(static_cast<unsigned char*>(base) + off); }

void csr(int nn, int ne, int nv, const int* el, std::vector<int>& ptr, std::vector<int>& inc)
{
    std::vector<int> cnt(nn + 1, 0);
    for (int e = 0; e < ne; ++e)
        for (int lv = 0; lv < nv; ++lv) cnt[el[lv * ne + e] + 1]++;
    ptr.assign(nn + 1, 0);
    for (int i = 0; i < nn; ++i) ptr[i + 1] = ptr[i] + cnt[i + 1];
    inc.assign(ptr[nn], 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int e = 0; e < ne; ++e)
        for (int lv = 0; lv < nv; ++lv) inc[fill[el[lv * ne + e]]++] = e * 4 + lv;
}

}  // namespace

static std::vector<const void*> micro_kernels();

// A template is the generator's uniform r x r right-triangle grid when its
// connectivity, Dirichlet set and every element's geometry (to 1e-12) match the
// pattern; then sg[0], sg[1] hold the A and B element geometry for the analytic kernel.
static bool detect_structured(const micro_mesh_struct& m, const double* w, const double* scv, double sg[2][22])
{
    const int nn = m.no_nodes, ne = m.no_elements;
    const int r = (int)std::lround(std::sqrt((double)nn));
    bool st = r >= 2 && r * r == nn && ne == 2 * (r - 1) * (r - 1);
    for (int sq = 0; st && sq < (r - 1) * (r - 1); ++sq) {
        const int i = sq % (r - 1), jj = sq / (r - 1), a0 = jj * r + i;
        const int eA = 2 * sq, eB = eA + 1;
        st = m.elements[eA] == a0 && m.elements[ne + eA] == a0 + 1 && m.elements[2 * ne + eA] == a0 + r + 1 &&
             m.elements[eB] == a0 && m.elements[ne + eB] == a0 + r + 1 && m.elements[2 * ne + eB] == a0 + r;
    }
    for (int i = 0; st && i < nn; ++i) {
        const int ix = i % r, iy = i / r;
        const bool per = ix == 0 || ix == r - 1 || iy == 0 || iy == r - 1;
        st = (m.boundary_nodes[i] == 1) == per;
    }
    auto geo = [&](int e, double* g) {
        for (int k = 0; k < 6; ++k) g[k] = m.shape_funcs[k * ne + e];
        for (int k = 0; k < 6; ++k) g[6 + k] = m.normals[k * ne + e];
        for (int k = 0; k < 9; ++k) g[12 + k] = w[(size_t)k * ne + e];
        g[21] = scv[e];
    };
    if (st) {
        geo(0, sg[0]);
        geo(1, sg[1]);
        // the analytic right-triangle geometry (k_micro's structured stencil): square
        // cells of side d in both directions
        const double d = m.nodes[1] - m.nodes[0];
        const double an[2][21] = {
            {-1 / d, 1 / d, 0, 0, -1 / d, 1 / d, d / 3, -d / 6, -d / 6, d / 3, -d / 6, -d / 6,
             5. / 12, 5. / 12, 1. / 6, 1. / 6, 5. / 12, 5. / 12, 5. / 12, 1. / 6, 5. / 12},
            {0, 1 / d, -1 / d, -1 / d, 0, 1 / d, d / 6, d / 6, -d / 3, d / 6, d / 6, -d / 3,
             5. / 12, 5. / 12, 1. / 6, 1. / 6, 5. / 12, 5. / 12, 5. / 12, 1. / 6, 5. / 12}};
        st = d > 0.0 && std::fabs((m.nodes[nn + r] - m.nodes[nn]) - d) <= 1e-12 * d;
        for (int t = 0; t < 2 && st; ++t)
            for (int k = 0; k < 21 && st; ++k) {
                const double scale = k < 6 ? 1 / d : k < 12 ? d : 1.0;
                st = std::fabs(sg[t][k] - an[t][k]) <= 1e-12 * scale;
            }
        double g[22];
        for (int e = 0; st && e < ne; ++e) {
            geo(e, g);
            const double* ref = sg[e & 1];
            for (int k = 0; k < 22 && st; ++k) st = std::fabs(g[k] - ref[k]) <= 1e-12 * std::max(1.0, std::fabs(ref[k]));
        }
    }
    return st;
}

// Sub-face K weights and SCV areas of a template (the arrays k_micro<false> reads).
static void template_weights(const micro_mesh_struct& m, std::vector<double>& w, std::vector<double>& scv)
{
    const int nn = m.no_nodes, ne = m.no_elements;
    w.assign((size_t)9 * ne, 0.0);
    scv.assign(ne, 0.0);
    for (int e = 0; e < ne; ++e) {
        const int a0 = m.elements[e], b0 = m.elements[ne + e], c0 = m.elements[2 * ne + e];
        const double xa = m.nodes[a0], ya = m.nodes[nn + a0];
        const double xb = m.nodes[b0], yb = m.nodes[nn + b0];
        const double xc = m.nodes[c0], yc = m.nodes[nn + c0];
        const double det = (xb - xa) * (yc - ya) - (xc - xa) * (yb - ya);
        scv[e] = std::fabs(det) / 6.0;
        for (int j = 0; j < 3; ++j) {
            const double mx = m.midx[j * ne + e], my = m.midy[j * ne + e];
            for (int vv = 0; vv < 3; ++vv)
                w[(size_t)(j * 3 + vv) * ne + e] = m.shape_funcs[(0 * 3 + vv) * ne + e] * mx +
                                                   m.shape_funcs[(1 * 3 + vv) * ne + e] * my +
                                                   m.shape_funcs[(2 * 3 + vv) * ne + e];
        }
    }
}

DevProblem* upload_problem(const tsr_batch_view& v, std::string& err)
{
    Arena A;
    struct TOff { size_t ptr, inc, el, sf, nrm, w, scv, vec, cen, bnd, soil, blist; int nn, ne; double ca; bool has_soil;
                  int structured, r, nbnd; unsigned mul_r, mul_rm; double dx = 0.0; double sg[2][22]; };
    struct SOff { size_t keff, el, nodes, sf, nrm, midx, midy, cv, scv, bce, blen, bmid, bnrm, cen, ptr, inc, dir, qe, bq; };
    std::vector<TOff> to(v.n_tmpl);
    std::vector<SOff> so(v.n_scen);
    int nn_max = 0, ne_max = 0, mnn = 0, mne = 0;
    for (int t = 0; t < v.n_tmpl; ++t) {
        const micro_mesh_struct& m = v.tmpl[t];
        if (m.no_verts != 3) { err = "micro templates must be triangles"; return nullptr; }
        const int nn = m.no_nodes, ne = m.no_elements;
        if (nn > kMicroMaxBS * 8) { err = "micro template has more than 8192 nodes"; return nullptr; }
        nn_max = std::max(nn_max, nn);
        ne_max = std::max(ne_max, ne);
        std::vector<int> ptr, inc;
        csr(nn, ne, 3, m.elements, ptr, inc);
        std::vector<double> w, scv;
        template_weights(m, w, scv);
        TOff& o = to[t];
        o.nn = nn; o.ne = ne; o.ca = m.cell_area;
        o.structured = 0; o.r = 0; o.mul_r = 0; o.mul_rm = 0;
        if (detect_structured(m, w.data(), scv.data(), o.sg)) {
            const int r = (int)std::lround(std::sqrt((double)nn));
            const unsigned mr = (unsigned)(0xFFFFFFFFull / (unsigned)r) + 1u;
            const unsigned mrm = (unsigned)(0xFFFFFFFFull / (unsigned)(r - 1)) + 1u;
            bool exact = r >= 3;
            for (unsigned n = 0; exact && n < 4u * (unsigned)nn; ++n)
                exact = (unsigned)(((unsigned long long)n * mr) >> 32) == n / (unsigned)r &&
                        (unsigned)(((unsigned long long)n * mrm) >> 32) == n / (unsigned)(r - 1);
            if (exact) { o.structured = 1; o.r = r; o.mul_r = mr; o.mul_rm = mrm; o.dx = m.nodes[1] - m.nodes[0]; }
        }
        std::vector<int> blist;
        for (int i = 0; i < nn; ++i) if (m.boundary_nodes[i] == 1) blist.push_back(i);
        o.nbnd = (int)blist.size();
        o.blist = A.put(blist.data(), blist.size());
        o.ptr = A.put(ptr.data(), ptr.size());
        o.inc = A.put(inc.data(), inc.size());
        o.el = A.put(m.elements, (size_t)3 * ne);
        o.sf = A.put(m.shape_funcs, (size_t)9 * ne);
        o.nrm = A.put(m.normals, (size_t)6 * ne);
        o.w = A.put(w.data(), w.size());
        o.scv = A.put(scv.data(), scv.size());
        o.vec = A.put(m.vectors, (size_t)2 * nn);
        o.cen = A.put(m.centroid_elements, (size_t)2 * ne);
        o.bnd = A.put(m.boundary_nodes, (size_t)nn);
        o.has_soil = v.tmpl_soil && v.tmpl_soil[t];
        o.soil = o.has_soil ? A.put(v.tmpl_soil[t], (size_t)ne) : 0;
    }
    // scenarios sharing one macro mesh share its device arrays
    std::vector<int> share(v.n_scen, -1);
    for (int s = 0; s < v.n_scen; ++s) {
        const tsr_scenario& sc = v.scen[s];
        const macro_mesh_struct& M = sc.ud.macro_mesh;
        const Keff_struct& K = sc.ud.effective_conductivity;
        const int nv = M.no_verts, nn = M.no_nodes, ne = M.no_elements;
        mnn = std::max(mnn, nn);
        mne = std::max(mne, ne);
        SOff& o = so[s];
        o.keff = A.put(K.Keff_values, (size_t)4 * K.no_hvalues);
        for (int r = 0; r < s; ++r)
            if (v.scen[r].ud.macro_mesh.elements == M.elements && v.scen[r].ud.macro_mesh.no_nodes == nn) { share[s] = r; break; }
        if (sc.host_kind == TSR_HOST_NODE && nv != 3) { err = "cells hosted by macro nodes need a triangle macro mesh"; return nullptr; }
        if (share[s] < 0) {
            std::vector<int> ptr, inc;
            csr(nn, ne, nv, M.elements, ptr, inc);
            for (int j = 0; j < nn; ++j)
                if (sc.host_kind == TSR_HOST_NODE && (ptr[j + 1] - ptr[j] > kMaxNodeInc || ptr[j + 1] == ptr[j])) {
                    err = "a macro node hosting a cell needs 1 to 12 incident elements";
                    return nullptr;
                }
            std::vector<unsigned char> dir(nn, 0);
            for (int e = 0; e < ne; ++e)
                for (int p = 0; p < nv; ++p)
                    if (M.boundary_condition_edges[p * ne + e] == TSR_BC_DIRICHLET) {
                        dir[M.elements[p * ne + e]] = 1;
                        dir[M.elements[((p + 1) % nv) * ne + e]] = 1;
                    }
            o.el = A.put(M.elements, (size_t)nv * ne);
            o.nodes = A.put(M.nodes, (size_t)2 * nn);
            o.sf = A.put(M.shape_funcs, (size_t)nv * nv * ne);
            o.nrm = A.put(M.normals, (size_t)2 * nv * ne);
            o.midx = A.put(M.midx, (size_t)nv * ne);
            o.midy = A.put(M.midy, (size_t)nv * ne);
            o.cv = A.put(M.CV_area, (size_t)nn);
            o.scv = A.put(M.SCV_area, (size_t)nv * ne);
            o.bce = A.put(M.boundary_condition_edges, (size_t)nv * ne);
            o.blen = A.put(M.boundary_edge_lens, (size_t)2 * nv * ne);
            o.bmid = A.put(M.boundary_edge_midpts, (size_t)4 * nv * ne);
            o.bnrm = A.put(M.boundary_edge_normals, (size_t)4 * nv * ne);
            o.cen = A.put(M.centroid_elements, (size_t)2 * ne);
            o.ptr = A.put(ptr.data(), ptr.size());
            o.inc = A.put(inc.data(), inc.size());
            o.dir = A.put(dir.data(), dir.size());
        } else {
            const SOff& r = so[share[s]];
            const size_t keep = o.keff;
            o = r;
            o.keff = keep;
        }
        o.qe = A.reserve<double>((size_t)nv * ne);
        o.bq = A.reserve<double>((size_t)2 * nv * ne);
    }
    // cells
    std::vector<int> cell_scen(v.n_cells);
    for (int s = 0; s < v.n_scen; ++s)
        for (int c = 0; c < v.scen[s].n_cells; ++c) cell_scen[v.scen[s].first_cell + c] = s;
    const size_t o_ct = A.put(v.cell_tmpl, (size_t)v.n_cells);
    const size_t o_cs = A.put(cell_scen.data(), cell_scen.size());
    const size_t o_co = A.put(v.cell_offset, (size_t)v.n_cells);
    const size_t o_disc = v.max_discs > 0 ? A.put(v.cell_discs, (size_t)v.n_cells * v.max_discs * 3) : 0;
    const size_t o_Q = A.reserve<double>((size_t)v.n_cells);
    const size_t o_de = A.reserve<unsigned char>((size_t)v.n_cells * nn_max);
    const int soil_words = (ne_max + 31) / 32;
    if (soil_words > kMaxSoilWords) {
        err = "micro templates of more than " + std::to_string(32 * kMaxSoilWords) + " elements are not supported";
        return nullptr;
    }
    const size_t o_sb = A.reserve<unsigned>((size_t)v.n_cells * soil_words);
    const size_t o_cmsf = A.reserve<double>((size_t)v.n_cells * kMaxNodeInc * 9);
    const size_t o_cmvi = A.reserve<int>((size_t)v.n_cells * kMaxNodeInc * 3);
    const size_t o_cmh = A.reserve<int>((size_t)v.n_cells);
    const size_t o_cmn = A.reserve<unsigned char>((size_t)v.n_cells);
    const size_t o_bad = A.reserve<int>((size_t)v.n_scen);
    const size_t o_T = A.reserve<DTmpl>(v.n_tmpl);
    const size_t o_S = A.reserve<DScen>(v.n_scen);

    if (v.N >= 0x7fffffffLL) { err = "more than 2^31 - 1 unknowns"; return nullptr; }
    DevProblem* P = new DevProblem();
    P->bytes = A.host.size();
    if (hipMalloc(&P->buf, P->bytes) != hipSuccess) { err = "hipMalloc failed for the problem"; delete P; return nullptr; }
    void* base = P->buf;
    // fill the descriptor tables with device addresses, then copy everything once
    for (int t = 0; t < v.n_tmpl; ++t) {
        const TOff& o = to[t];
        DTmpl T;
        T.nn = o.nn; T.ne = o.ne; T.cell_area = o.ca;
        T.ptr = at<int>(base, o.ptr); T.inc = at<int>(base, o.inc); T.el = at<int>(base, o.el);
        T.sf = at<double>(base, o.sf); T.nrm = at<double>(base, o.nrm); T.w = at<double>(base, o.w);
        T.scv = at<double>(base, o.scv); T.vec = at<double>(base, o.vec); T.cen = at<double>(base, o.cen);
        T.bnd = at<int>(base, o.bnd);
        T.soil = o.has_soil ? at<unsigned char>(base, o.soil) : nullptr;
        T.structured = o.structured;
        T.r = o.r;
        T.mul_r = o.mul_r;
        T.mul_rm = o.mul_rm;
        T.dx = o.dx;
        T.nbnd = o.nbnd;
        T.bnd_list = at<int>(base, o.blist);
        std::memcpy(T.sg, o.sg, sizeof(T.sg));
        std::memcpy(A.host.data() + o_T + t * sizeof(DTmpl), &T, sizeof(DTmpl));
    }
    for (int s = 0; s < v.n_scen; ++s) {
        const tsr_scenario& sc = v.scen[s];
        const macro_mesh_struct& M = sc.ud.macro_mesh;
        const Keff_struct& K = sc.ud.effective_conductivity;
        const SOff& o = so[s];
        DScen S;
        auto soil = [](const soil_struct& x) { return DSoil{x.Ksat, x.thetar, x.thetas, x.alpha, x.n, x.m}; };
        S.A = soil(sc.ud.soilA); S.B = soil(sc.ud.soilB);
        S.epsA = K.epsilonA; S.dirichlet_head = sc.dirichlet_head; S.boundary_flux = sc.ud.boundary_flux;
        S.keff_logr = K.logr; S.keff_first = K.logh_first; S.keff_n = K.no_hvalues;
        S.box[0] = S.box[2] = INFINITY;
        S.box[1] = S.box[3] = -INFINITY;
        for (int i = 0; i < M.no_nodes; ++i) {
            S.box[0] = std::fmin(S.box[0], M.nodes[i]); S.box[1] = std::fmax(S.box[1], M.nodes[i]);
            S.box[2] = std::fmin(S.box[2], M.nodes[M.no_nodes + i]); S.box[3] = std::fmax(S.box[3], M.nodes[M.no_nodes + i]);
        }
        S.model = sc.ud.model_type; S.host = sc.host_kind; S.nv = M.no_verts; S.nn = M.no_nodes; S.ne = M.no_elements;
        S.n_cells = sc.n_cells; S.first_cell = sc.first_cell; S.u_off = sc.u_offset;
        S.keff = at<double>(base, o.keff);
        S.el = at<int>(base, o.el); S.nodes = at<double>(base, o.nodes); S.sf = at<double>(base, o.sf);
        S.nrm = at<double>(base, o.nrm); S.midx = at<double>(base, o.midx); S.midy = at<double>(base, o.midy);
        S.cv = at<double>(base, o.cv); S.scv = at<double>(base, o.scv); S.bce = at<int>(base, o.bce);
        S.be_len = at<double>(base, o.blen); S.be_mid = at<double>(base, o.bmid); S.be_nrm = at<double>(base, o.bnrm);
        S.cen = at<double>(base, o.cen); S.ptr = at<int>(base, o.ptr); S.inc = at<int>(base, o.inc);
        S.dir = at<unsigned char>(base, o.dir);
        S.qe = at<double>(base, o.qe); S.bq = at<double>(base, o.bq);
        std::memcpy(A.host.data() + o_S + s * sizeof(DScen), &S, sizeof(DScen));
    }
    if (hipMemcpy(P->buf, A.host.data(), P->bytes, hipMemcpyHostToDevice) != hipSuccess) {
        err = "hipMemcpy failed for the problem";
        hipFree(P->buf);
        delete P;
        return nullptr;
    }
    DBatch& B = P->B;
    B.n_scen = v.n_scen; B.n_tmpl = v.n_tmpl; B.n_cells = v.n_cells; B.max_discs = v.max_discs;
    B.nn_max = nn_max; B.ne_max = ne_max; B.macro_nn_max = mnn; B.macro_ne_max = mne;
    B.scen = at<DScen>(base, o_S); B.tmpl = at<DTmpl>(base, o_T);
    B.cell_tmpl = at<int>(base, o_ct); B.cell_scen = at<int>(base, o_cs); B.cell_off = at<long long>(base, o_co);
    B.cell_off0 = v.n_cells > 0 ? v.cell_offset[0] : 0;
    B.cell_stride = 0;
    if (v.n_scen == 1 && v.n_tmpl == 1 && v.n_cells > 0) {
        const long long st = v.tmpl[0].no_nodes;
        bool lin = true;
        for (int c = 0; c < v.n_cells && lin; ++c) lin = v.cell_offset[c] == B.cell_off0 + c * st && v.cell_tmpl[c] == 0;
        if (lin) B.cell_stride = (int)st;
    }
    B.discs = v.max_discs > 0 ? at<double>(base, o_disc) : nullptr;
    B.Q = at<double>(base, o_Q); B.bad = at<int>(base, o_bad);
    B.dir_elem = at<unsigned char>(base, o_de);
    B.soil_bits = at<unsigned>(base, o_sb);
    B.soil_words = soil_words;
    B.cm_sf = at<double>(base, o_cmsf); B.cm_vi = at<int>(base, o_cmvi);
    B.cm_host = at<int>(base, o_cmh); B.cm_ninc = at<unsigned char>(base, o_cmn);
    B.Ss = v.Ss; B.N = v.N;
    for (int q = 0; q < v.n_scen; ++q) P->n_macro += v.scen[q].ud.macro_mesh.no_nodes;
    {
        const long long n = (long long)v.n_cells * nn_max, nw = (long long)v.n_cells * soil_words;
        if (n > 0) hipLaunchKernelGGL(k_dir_elem, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, B);
        if (nw > 0) hipLaunchKernelGGL(k_soil_bits, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, 0, B);
        const long long nk = (long long)v.n_cells * kMaxNodeInc;
        if (nk > 0) hipLaunchKernelGGL(k_cell_macro, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, 0, B);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(0) != hipSuccess) {
            err = "set-up kernels (Dirichlet interpolation, soil map, cell macro data) failed";
            hipFree(P->buf);
            delete P;
            return nullptr;
        }
    }
    P->all_structured = !std::getenv("TSR_GENERIC_MICRO");
    for (int t = 0; t < v.n_tmpl; ++t) P->all_structured = P->all_structured && to[t].structured;
    if (P->all_structured && nn_max <= 1024 && !std::getenv("TSR_MICRO_LIN_OFF")) {
        P->lin_r = 0;
        bool same = true;
        for (int t = 1; t < v.n_tmpl; ++t) same = same && to[t].r == to[0].r;
        if (same && v.n_tmpl > 0)
            for (int rr : {11, 16, 21, 31})
                if (to[0].r == rr) P->lin_r = rr;
    }
    P->lds = (size_t)4 * nn_max * sizeof(double);
    P->lds2 = (size_t)3 * nn_max * sizeof(double);   // linearised mode: no storage array
    if (P->lds > 64 * 1024) {
        bool ok = true;
        for (const void* f : micro_kernels()) ok = ok && hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)P->lds) == hipSuccess;
        if (!ok) {
            err = "micro template needs more LDS than a workgroup may allocate";
            hipFree(P->buf);
            delete P;
            return nullptr;
        }
    }
    return P;
}

void free_problem(DevProblem* p)
{
    if (!p) return;
    if (p->buf) hipFree(p->buf);
    delete p;
}

const DBatch& batch_of(const DevProblem* p) { return p->B; }
size_t micro_lds_bytes(const DevProblem* p) { return p->lds; }

static const int kWpe[3] = {1, 6, 8};

using MicroFn = void (*)(DBatch, GArgs);

template <bool ST, int M>
static MicroFn micro_fn(int wpe)
{
    return wpe == 8 ? (MicroFn)k_micro<ST, 8, M> : wpe == 6 ? (MicroFn)k_micro<ST, 6, M> : (MicroFn)k_micro<ST, 1, M>;
}

static MicroFn micro_fn(bool structured, int mode, int wpe)
{
    if (structured) return mode == 2 ? micro_fn<true, 2>(wpe) : mode == 1 ? micro_fn<true, 1>(wpe) : micro_fn<true, 0>(wpe);
    return mode == 2 ? micro_fn<false, 2>(wpe) : mode == 1 ? micro_fn<false, 1>(wpe) : micro_fn<false, 0>(wpe);
}

// k_micro_lin instances: grid size R (0 = read per cell), W wavefronts per cell, NPT
// nodes per thread (64 W NPT >= R^2; R = 0 serves templates of up to 1024 nodes), each
// with the compiler's register budget (f) and, for W >= 2, a budget of 8 waves per SIMD
// (f8: at most 64 VGPRs)
struct LinVariant { int R, W, NPT; MicroFn f, f8; };
#define TSR_LIN1(R, W, NPT) {R, W, NPT, k_micro_lin<R, W, NPT, 1>, nullptr}
#define TSR_LIN2(R, W, NPT) {R, W, NPT, k_micro_lin<R, W, NPT, 1>, k_micro_lin<R, W, NPT, 8>}
static const std::vector<LinVariant>& lin_variants()
{
    static const std::vector<LinVariant> v = {
        TSR_LIN1(11, 1, 2), TSR_LIN2(11, 2, 1),
        TSR_LIN1(16, 1, 4), TSR_LIN2(16, 2, 2), TSR_LIN2(16, 4, 1),
        TSR_LIN1(21, 1, 7), TSR_LIN2(21, 2, 4), TSR_LIN2(21, 4, 2),
        TSR_LIN2(31, 2, 8), TSR_LIN2(31, 4, 4),
        TSR_LIN2(0, 2, 8), TSR_LIN2(0, 4, 4),
    };
    return v;
}
#undef TSR_LIN1
#undef TSR_LIN2

static std::vector<const void*> micro_kernels()
{
    std::vector<const void*> v;
    for (int st = 0; st < 2; ++st)
        for (int mode = 0; mode < 3; ++mode)
            for (int w : kWpe) v.push_back((const void*)micro_fn(st == 1, mode, w));
    for (const LinVariant& x : lin_variants()) {
        v.push_back((const void*)x.f);
        if (x.f8) v.push_back((const void*)x.f8);
    }
    return v;
}

static int micro_mode(const GArgs& a) { return a.lin ? 2 : a.rec ? 1 : 0; }

// the k_micro_lin instance of this problem with bs = 64 W threads per cell and register
// budget wpe (1 or 8 waves per SIMD; nullptr: none)
static MicroFn lin_fn(const DevProblem* p, int bs, int wpe)
{
    for (const LinVariant& x : lin_variants())
        if (x.R == p->lin_r && 64 * x.W == bs && 64 * x.W * x.NPT >= p->B.nn_max)
            return wpe == 1 ? x.f : wpe == 8 ? x.f8 : nullptr;
    return nullptr;
}

static void launch_micro(const DevProblem* p, const GArgs& a, int bs, int wpe, hipStream_t st)
{
    const DBatch& B = p->B;
    MicroFn f = nullptr;
    if (micro_mode(a) == 2 && p->lin_r >= 0) f = lin_fn(p, bs, wpe);
    if (!f) f = micro_fn(p->all_structured, micro_mode(a), wpe);
    const int grid = B.n_cells + macro_wgs(B, bs) * B.n_scen;
    hipLaunchKernelGGL(f, dim3(grid), dim3(bs), micro_mode(a) == 2 ? p->lds2 : p->lds, st, B, a);
}

// Autotuner: on the first G of a problem (per k_micro mode), time every admissible
// launch shape on the real inputs (one warm-up and three timed launches each, HIP
// events) and keep the fastest.  Every candidate computes the same result, so the
// output of the final launch (the chosen shape) is the G the caller asked for.
// Workgroup sizes the micro-cell kernel admits for a mode: at most 8 strided nodes per
// thread; the linearised mode also needs bs to divide 256 and at most kLinNPT nodes per
// thread.
static std::vector<int> micro_bs_candidates(const DevProblem* p, int mode)
{
    const DBatch& B = p->B;
    std::vector<int> cand_bs;
    if (mode == 2 && p->lin_r >= 0) {
        // TSR_LIN_BS pins the shape (tests of each shape); otherwise every compiled one
        const char* pin = std::getenv("TSR_LIN_BS");
        for (int bs : {64, 128, 256})
            if (lin_fn(p, bs, 1) && (!pin || std::atoi(pin) == bs)) cand_bs.push_back(bs);
        if (!cand_bs.empty()) return cand_bs;
    }
    for (int bs : {64, 128, 192, 256, 384, 512, 768, 1024}) {
        const bool ok2 = mode != 2 || (256 % bs == 0 && bs * kLinNPT >= B.nn_max);
        if (bs <= kMicroMaxBS && bs * 8 >= B.nn_max && ok2) cand_bs.push_back(bs);
    }
    if (cand_bs.empty()) cand_bs.push_back(kMicroMaxBS);
    return cand_bs;
}

static hipError_t tune_micro(DevProblem* p, const GArgs& a, hipStream_t st)
{
    const DBatch& B = p->B;
    const int mode = micro_mode(a);
    const std::vector<int> cand_bs = micro_bs_candidates(p, mode);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return hipErrorUnknown;
    // register budgets of k_micro whose code spills to scratch are not candidates:
    // besides being slow, a spilling build of this kernel (mode 1, 8 waves per SIMD) has
    // given wrong results.  k_micro_lin's budgets (1 and, for W >= 2, 8 waves per SIMD)
    // are all timed: its 8-wave builds of the larger shapes spill a few registers and
    // can still be the fastest, and every budget gives the same products bit for bit
    // (tested).  TSR_MICRO_WPE pins the budget (tests of each budget; a pinned budget is
    // taken even when its code spills).
    std::vector<int> cand_w;
    const char* pin_w = std::getenv("TSR_MICRO_WPE");
    const bool lin = mode == 2 && p->lin_r >= 0 && lin_fn(p, cand_bs[0], 1);
    if (lin) {
        for (int w : {1, 8})
            if (!pin_w || std::atoi(pin_w) == w) cand_w.push_back(w);
    } else for (int w : kWpe) {
        hipFuncAttributes fa;
        if (pin_w ? std::atoi(pin_w) == w
                  : hipFuncGetAttributes(&fa, (const void*)micro_fn(p->all_structured, mode, w)) == hipSuccess &&
                        fa.localSizeBytes == 0)
            cand_w.push_back(w);
    }
    if (cand_w.empty()) cand_w.push_back(1);
    double best = 1e300;
    int best_bs = cand_bs[0], best_w = cand_w[0];
    if (mode == 0) p->tune_n = 0;
    for (int bs : cand_bs)
        for (int w : cand_w) {
            if (lin && !lin_fn(p, bs, w)) continue;   // no 8-wave build of one-wavefront shapes
            launch_micro(p, a, bs, w, st);
            hipEventRecord(e0, st);
            for (int r = 0; r < 3; ++r) launch_micro(p, a, bs, w, st);
            hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) != hipSuccess) { hipEventDestroy(e0); hipEventDestroy(e1); return hipErrorUnknown; }
            float ms = 0.0f;
            hipEventElapsedTime(&ms, e0, e1);
            const double avg = ms / 3.0;
            if (std::getenv("TSR_TUNE_LOG"))
                std::fprintf(stderr, "[tsr] k_micro mode %d: bs %d, waves/EU %d: %.1f us\n", mode, bs, w, avg * 1e3);
            if (mode == 0 && p->tune_n < 64) p->tune_ms[p->tune_n++] = avg;
            if (avg < best) { best = avg; best_bs = bs; best_w = w; }
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (best == 1e300) best_w = 1;   // a pinned budget no shape of this problem has
    p->micro_bs[mode] = best_bs;
    p->micro_wpe[mode] = best_w;
    launch_micro(p, a, best_bs, best_w, st);
    return hipGetLastError();
}

hipError_t launch_gfunc_part(const DevProblem* p, const GArgs& a, int part, hipStream_t st, bool tune)
{
    const DBatch& B = p->B;
    if (part == 0) {
        // nothing: the macro elements run in the leading workgroups of the micro launch
    } else if (part == 1) {
        const int mode = micro_mode(a);
        if (mode == 2 && !lin_mode_ok(p)) return hipErrorInvalidValue;
        if (p->micro_bs[mode] == 0) {
            if (tune) return tune_micro(const_cast<DevProblem*>(p), a, st);
            // not tuned yet and tuning not wanted (it creates events): the smallest
            // admissible shape with the compiler's register budget (results do not
            // depend on the shape)
            launch_micro(p, a, micro_bs_candidates(p, mode)[0], 1, st);
            return hipGetLastError();
        }
        // algorithmic bytes: u (and v, G(u) for an FD product) read, the result written,
        // per micro node
        const double bpu = 8.0 * (2 + (a.v ? 1 : 0) + (a.g0 ? 1 : 0));
        const double fr = a.active ? kt_active_fraction() : 1.0;
        const bool kt = kt_begin(mode == 2 ? KT_MICRO_LIN : KT_MICRO_FULL, fr * bpu * (double)(B.N - p->n_macro), st);
        launch_micro(p, a, p->micro_bs[mode], p->micro_wpe[mode], st);
        if (kt) kt_end(st);
    } else {
        dim3 gn((B.macro_nn_max + 63) / 64, B.n_scen);
        const double bpu = 8.0 * (2 + (a.v ? 1 : 0) + (a.g0 ? 1 : 0));
        const double fr = a.active ? kt_active_fraction() : 1.0;
        const bool kt = kt_begin(KT_MACRO_NODE, fr * bpu * (double)p->n_macro, st);
        hipLaunchKernelGGL(k_macro_node, gn, dim3(64), 0, st, B, a);
        if (kt) kt_end(st);
    }
    return hipGetLastError();
}

bool lin_mode_ok(const DevProblem* p) { return p->B.nn_max <= 256 * kLinNPT; }

bool gfunc_tuned(const DevProblem* p, bool lin) { return p->micro_bs[lin ? 2 : 0] != 0; }

void micro_shape(const DevProblem* p, int* bs, int* wpe)
{
    *bs = p->micro_bs[0];
    *wpe = p->micro_wpe[0];
}

hipError_t launch_gfunc(const DevProblem* p, const GArgs& a, hipStream_t st, bool tune)
{
    for (int part = 0; part < 3; ++part) {
        const hipError_t e = launch_gfunc_part(p, a, part, st, tune);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace tsr

// 1 when the micro-cell kernel can use the analytic uniform-grid path for this template.
extern "C" int tsr_template_structured(const micro_mesh_struct* m)
{
    std::vector<double> w, scv;
    if (m->no_verts != 3) return 0;
    tsr::template_weights(*m, w, scv);
    double sg[2][22];
    return tsr::detect_structured(*m, w.data(), scv.data(), sg) ? 1 : 0;
}
