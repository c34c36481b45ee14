// mesh.cpp - host-side set-up of the two-scale problem: mesh properties, the
// reference's text formats, and structured generators.
//
// The arrays written here are the reference's data format (element-major SoA,
// 0-based indices), consumed unchanged by the HIP kernels:
//   macro: macro_mesh_properties.cpp:15-482   micro: micro_mesh_properties.cpp:15-459
// Conventions kept because the format depends on them:
//   * sub-face j runs from the midpoint of edge (v_j, v_{j+1}) to the centroid; its
//     normal is the sub-face vector rotated by -90 deg, times -sign(ordering) (micro,
//     macro triangles) or unsigned (macro quads), so it points from v_j's SCV into
//     v_{j+1}'s and its length is the sub-face length (micro_mesh_properties.cpp:282-296,
//     macro_mesh_properties.cpp:309-321);
//   * (midx, midy) = midpoint of the sub-face;
//   * shape_funcs[(m*nv + i)*ne + e] = coefficient m of basis function i, with
//     P1 h = a0 x + a1 y + a2 and Q1 h = a0 x + a1 y + a2 x y + a3 (save_solution.cpp:57-99);
//   * boundary half-edges k = 0 (v_p side) and 1 (v_{p+1} side) of edge p: length,
//     midpoint and outward length-scaled normal (macro_mesh_properties.cpp:326-363);
//   * macro edge types from node labels -(j+1) -> macro_bc[j], later labels win
//     (macro_mesh_properties.cpp:95-102,170-190); micro Dirichlet nodes are the nodes
//     of edges labelled -1 (micro_mesh_properties.cpp:93-106).
// The geometry itself is computed by this file's own routines.  Triangle SCV areas are
// the true median-dual areas (element area / 3, see scv_area).  element_area is the
// sum of the element's SCV areas (the reference zeroes it over no_nodes entries,
// micro_mesh_properties.cpp:208-211, which leaves some entries unset).
#include "tsr_internal.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace tsr {

namespace {

inline int sgn(double v) { return (0.0 < v) - (v < 0.0); }

template <class T> T* alloc(size_t n) { T* p = new T[n > 0 ? n : 1]; std::memset(p, 0, sizeof(T) * (n > 0 ? n : 1)); return p; }

// Inverse of the 4x4 matrix [x y xy 1] at the quad vertices: column i holds the
// coefficients of basis i.  Gauss-Jordan with partial pivoting.
bool q1_coeffs(const double xv[4], const double yv[4], double C[4][4])
{
    double A[4][8];
    for (int k = 0; k < 4; ++k) {
        A[k][0] = xv[k]; A[k][1] = yv[k]; A[k][2] = xv[k] * yv[k]; A[k][3] = 1.0;
        for (int c = 0; c < 4; ++c) A[k][4 + c] = (k == c) ? 1.0 : 0.0;
    }
    for (int col = 0; col < 4; ++col) {
        int p = col;
        for (int r = col + 1; r < 4; ++r) if (std::fabs(A[r][col]) > std::fabs(A[p][col])) p = r;
        if (A[p][col] == 0.0) return false;
        if (p != col) for (int c = 0; c < 8; ++c) std::swap(A[p][c], A[col][c]);
        const double d = A[col][col];
        for (int c = 0; c < 8; ++c) A[col][c] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == col) continue;
            const double f = A[r][col];
            if (f == 0.0) continue;
            for (int c = 0; c < 8; ++c) A[r][c] -= f * A[col][c];
        }
    }
    // A[:,4:8] = M^{-1}; coefficient m of basis i = Minv[m][i]
    for (int m = 0; m < 4; ++m)
        for (int i = 0; i < 4; ++i) C[m][i] = A[m][4 + i];
    return true;
}

// Per-element geometry shared by the macro and micro properties.
struct ElemGeom {
    int nv;
    double xv[4], yv[4], cx, cy, ordering;
    double mid[4][2], f[4][2];
};

void elem_geom(int nv, const int* vert, const double* nodes, int nn, ElemGeom& G)
{
    G.nv = nv;
    G.cx = G.cy = 0.0;
    for (int j = 0; j < nv; ++j) {
        G.xv[j] = nodes[vert[j]];
        G.yv[j] = nodes[nn + vert[j]];
        G.cx += G.xv[j];
        G.cy += G.yv[j];
    }
    G.cx /= (double)nv;
    G.cy /= (double)nv;
    G.ordering = 0.0;
    for (int j = 0; j < nv; ++j) {
        const int k = (j + 1) % nv;
        G.ordering += (G.xv[k] - G.xv[j]) * (G.yv[k] + G.yv[j]);
        G.mid[j][0] = 0.5 * (G.xv[j] + G.xv[k]);
        G.mid[j][1] = 0.5 * (G.yv[j] + G.yv[k]);
    }
    for (int j = 0; j < nv; ++j) { G.f[j][0] = G.cx - G.mid[j][0]; G.f[j][1] = G.cy - G.mid[j][1]; }
}

// SCV area of vertex j.  Triangles: the quadrilateral (v_j, m_j, c, m_{j-1}) has
// diagonals c - v_j and m_j - m_{j-1}, so its area is |(c - v_j) x (m_j - m_{j-1})| / 2
// = element area / 3.  (The reference crosses with m_j - m_{j+1},
// micro_mesh_properties.cpp:250-270 and macro_mesh_properties.cpp:272-294, which gives
// element area / 6 and control volumes summing to half the domain.)  Quads: a quarter
// of the bounding rectangle, as the reference (macro_mesh_properties.cpp:302).
double scv_area(const ElemGeom& G, int j)
{
    if (G.nv == 3) {
        const int k = (j + 2) % 3;
        const double px = G.cx - G.xv[j], py = G.cy - G.yv[j];
        const double qx = G.mid[j][0] - G.mid[k][0], qy = G.mid[j][1] - G.mid[k][1];
        return 0.5 * std::fabs(px * qy - py * qx);
    }
    return 0.25 * std::fabs((G.xv[2] - G.xv[0]) * (G.yv[2] - G.yv[0]));
}

void shape_coeffs(const ElemGeom& G, int e, int ne, double* sf)
{
    if (G.nv == 3) {
        const double* x = G.xv; const double* y = G.yv;
        const double det = (x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]);
        for (int i = 0; i < 3; ++i) {
            const int a = (i + 1) % 3, b = (i + 2) % 3;
            sf[(0 * 3 + i) * ne + e] = (y[a] - y[b]) / det;
            sf[(1 * 3 + i) * ne + e] = (x[b] - x[a]) / det;
            sf[(2 * 3 + i) * ne + e] = (x[a] * y[b] - x[b] * y[a]) / det;
        }
    } else {
        double C[4][4];
        if (!q1_coeffs(G.xv, G.yv, C)) { std::fprintf(stderr, "tsr: degenerate quad %d\n", e); return; }
        for (int m = 0; m < 4; ++m)
            for (int i = 0; i < 4; ++i) sf[(m * 4 + i) * ne + e] = C[m][i];
    }
}

void boundary_half_edges(const ElemGeom& G, int j, int e, int ne, double* lens, double* mids, double* nrm)
{
    const int k = (j + 1) % G.nv;
    const double s = -(double)sgn(G.ordering);
    const double mx = G.mid[j][0], my = G.mid[j][1];
    lens[(2 * j) * ne + e] = std::hypot(G.xv[j] - mx, G.yv[j] - my);
    lens[(2 * j + 1) * ne + e] = std::hypot(G.xv[k] - mx, G.yv[k] - my);
    mids[(4 * j + 0) * ne + e] = 0.5 * (G.xv[j] + mx);
    mids[(4 * j + 1) * ne + e] = 0.5 * (G.yv[j] + my);
    mids[(4 * j + 2) * ne + e] = 0.5 * (G.xv[k] + mx);
    mids[(4 * j + 3) * ne + e] = 0.5 * (G.yv[k] + my);
    const double b1x = mx - G.xv[j], b1y = my - G.yv[j];
    const double b2x = G.xv[k] - mx, b2y = G.yv[k] - my;
    nrm[(4 * j + 0) * ne + e] = s * b1y;
    nrm[(4 * j + 1) * ne + e] = -s * b1x;
    nrm[(4 * j + 2) * ne + e] = s * b2y;
    nrm[(4 * j + 3) * ne + e] = -s * b2x;
}

}  // namespace

int build_macro(int nv, int nn, const double* nodes_xy, int nbe, const int* bedges, int ne, const int* elems,
                const int* macro_bc, macro_mesh_struct& M)
{
    if (nv != 3 && nv != 4) return 2;
    M.no_verts = nv; M.no_nodes = nn; M.no_elements = ne;
    M.nodes = alloc<double>(2 * (size_t)nn);
    std::memcpy(M.nodes, nodes_xy, sizeof(double) * 2 * nn);
    M.elements = alloc<int>((size_t)nv * ne);
    std::memcpy(M.elements, elems, sizeof(int) * nv * ne);
    // node label masks: bit j set if the node lies on an edge labelled -(j+1)
    std::vector<unsigned> mask(nn, 0u);
    for (int b = 0; b < nbe; ++b) {
        const int lab = bedges[3 * b + 2];
        if (lab <= -1 && lab >= -5) {
            mask[bedges[3 * b]] |= 1u << (-lab - 1);
            mask[bedges[3 * b + 1]] |= 1u << (-lab - 1);
        }
    }
    M.boundary_condition_edges = alloc<int>((size_t)nv * ne);
    std::vector<int> bedge((size_t)nv * ne, 0);
    for (int e = 0; e < ne; ++e)
        for (int p = 0; p < nv; ++p) {
            const unsigned common = mask[elems[p * ne + e]] & mask[elems[((p + 1) % nv) * ne + e]];
            if (common) {
                int j = 31 - __builtin_clz(common);
                M.boundary_condition_edges[p * ne + e] = macro_bc[j];
                bedge[(size_t)p * ne + e] = 1;
            }
        }
    M.boundary_edge_lens = alloc<double>(2 * (size_t)nv * ne);
    M.boundary_edge_normals = alloc<double>(4 * (size_t)nv * ne);
    M.boundary_edge_midpts = alloc<double>(4 * (size_t)nv * ne);
    M.normals = alloc<double>(2 * (size_t)nv * ne);
    M.CV_area = alloc<double>(nn);
    M.midx = alloc<double>((size_t)nv * ne);
    M.midy = alloc<double>((size_t)nv * ne);
    M.shape_funcs = alloc<double>((size_t)nv * nv * ne);
    M.SCV_area = alloc<double>((size_t)nv * ne);
    M.centroid_elements = alloc<double>(2 * (size_t)ne);
    for (int e = 0; e < ne; ++e) {
        int vert[4];
        for (int j = 0; j < nv; ++j) vert[j] = elems[j * ne + e];
        ElemGeom G;
        elem_geom(nv, vert, M.nodes, nn, G);
        M.centroid_elements[e] = G.cx;
        M.centroid_elements[ne + e] = G.cy;
        double s;
        if (nv == 3) {
            // orientation from (c - v0) x (m0 - m1), macro_mesh_properties.cpp:295-298
            const double px = G.cx - G.xv[0], py = G.cy - G.yv[0];
            const double qx = G.mid[0][0] - G.mid[1][0], qy = G.mid[0][1] - G.mid[1][1];
            s = -(double)sgn(px * qy - py * qx);
        } else {
            s = 1.0;
        }
        for (int j = 0; j < nv; ++j) {
            const double a = scv_area(G, j);
            M.SCV_area[j * ne + e] = a;
            M.CV_area[vert[j]] += a;
            M.normals[(2 * j) * ne + e] = s * G.f[j][1];
            M.normals[(2 * j + 1) * ne + e] = -s * G.f[j][0];
            M.midx[j * ne + e] = 0.5 * (G.mid[j][0] + G.cx);
            M.midy[j * ne + e] = 0.5 * (G.mid[j][1] + G.cy);
            if (bedge[(size_t)j * ne + e])
                boundary_half_edges(G, j, e, ne, M.boundary_edge_lens, M.boundary_edge_midpts, M.boundary_edge_normals);
        }
        shape_coeffs(G, e, ne, M.shape_funcs);
    }
    return 0;
}

int build_micro(double cw, double chh, int nv, int nn, const double* nodes_xy, int nbe, const int* bedges, int ne,
                const int* elems, micro_mesh_struct& m)
{
    if (nv != 3 && nv != 4) return 2;
    m.no_verts = nv; m.no_nodes = nn; m.no_elements = ne;
    m.cell_area = cw * chh;
    m.nodes = alloc<double>(2 * (size_t)nn);
    std::memcpy(m.nodes, nodes_xy, sizeof(double) * 2 * nn);
    m.elements = alloc<int>((size_t)nv * ne);
    std::memcpy(m.elements, elems, sizeof(int) * nv * ne);
    m.boundary_nodes = alloc<int>(nn);
    for (int b = 0; b < nbe; ++b)
        if (bedges[3 * b + 2] == -1) { m.boundary_nodes[bedges[3 * b]] = 1; m.boundary_nodes[bedges[3 * b + 1]] = 1; }
    m.boundary_edges = alloc<int>((size_t)nv * ne);
    m.boundary_edge_lens = alloc<double>(2 * (size_t)nv * ne);
    m.boundary_edge_midpts = alloc<double>(4 * (size_t)nv * ne);
    m.boundary_edge_normals = alloc<double>(4 * (size_t)nv * ne);
    m.normals = alloc<double>(2 * (size_t)nv * ne);
    m.CV_area = alloc<double>(nn);
    m.midx = alloc<double>((size_t)nv * ne);
    m.midy = alloc<double>((size_t)nv * ne);
    m.shape_funcs = alloc<double>((size_t)nv * nv * ne);
    m.vectors = alloc<double>(2 * (size_t)nn);
    m.centroid_elements = alloc<double>(2 * (size_t)ne);
    m.element_area = alloc<double>(ne);
    for (int i = 0; i < nn; ++i) {
        m.vectors[i] = m.nodes[i] - 0.5 * cw;
        m.vectors[nn + i] = m.nodes[nn + i] - 0.5 * chh;
    }
    for (int e = 0; e < ne; ++e) {
        int vert[4];
        for (int j = 0; j < nv; ++j) vert[j] = elems[j * ne + e];
        for (int j = 0; j < nv; ++j)
            if (m.boundary_nodes[vert[j]] == 1 && m.boundary_nodes[vert[(j + 1) % nv]] == 1) m.boundary_edges[j * ne + e] = 1;
        ElemGeom G;
        elem_geom(nv, vert, m.nodes, nn, G);
        m.centroid_elements[e] = G.cx;
        m.centroid_elements[ne + e] = G.cy;
        const double s = -(double)sgn(G.ordering);
        for (int j = 0; j < nv; ++j) {
            const double a = scv_area(G, j);
            m.CV_area[vert[j]] += a;
            m.element_area[e] += a;
            m.normals[(2 * j) * ne + e] = s * G.f[j][1];
            m.normals[(2 * j + 1) * ne + e] = -s * G.f[j][0];
            m.midx[j * ne + e] = 0.5 * (G.mid[j][0] + G.cx);
            m.midy[j * ne + e] = 0.5 * (G.mid[j][1] + G.cy);
            if (m.boundary_edges[j * ne + e])
                boundary_half_edges(G, j, e, ne, m.boundary_edge_lens, m.boundary_edge_midpts, m.boundary_edge_normals);
        }
        shape_coeffs(G, e, ne, m.shape_funcs);
    }
    return 0;
}

void free_macro(macro_mesh_struct& M)
{
    delete[] M.elements; delete[] M.nodes; delete[] M.shape_funcs; delete[] M.normals;
    delete[] M.boundary_edge_midpts; delete[] M.boundary_edge_normals; delete[] M.boundary_edge_lens;
    delete[] M.boundary_condition_edges; delete[] M.CV_area; delete[] M.SCV_area; delete[] M.midx;
    delete[] M.midy; delete[] M.centroid_elements;
    std::memset(&M, 0, sizeof(M));
}

void free_micro(micro_mesh_struct& m)
{
    delete[] m.elements; delete[] m.boundary_nodes; delete[] m.nodes; delete[] m.shape_funcs; delete[] m.normals;
    delete[] m.boundary_edges; delete[] m.boundary_edge_midpts; delete[] m.boundary_edge_normals;
    delete[] m.boundary_edge_lens; delete[] m.centroid_elements; delete[] m.CV_area; delete[] m.midx;
    delete[] m.midy; delete[] m.vectors; delete[] m.element_area;
    std::memset(&m, 0, sizeof(m));
}

// ---- structured generators --------------------------------------------------------
// [x0, x0+w] x [y0, y0+w], r x r nodes row-major (node = j*r + i), each square split
// along its (i,j)-(i+1,j+1) diagonal into two counter-clockwise right triangles.
void structured_grid(int r, double x0, double y0, double w, std::vector<double>& xy, std::vector<int>& el,
                     std::vector<int>& be, const int labels[4])
{
    const int nn = r * r, nq = (r - 1) * (r - 1), ne = 2 * nq;
    xy.assign(2 * (size_t)nn, 0.0);
    const double h = w / (double)(r - 1);
    for (int j = 0; j < r; ++j)
        for (int i = 0; i < r; ++i) {
            xy[j * r + i] = (i == r - 1) ? x0 + w : x0 + h * i;
            xy[nn + j * r + i] = (j == r - 1) ? y0 + w : y0 + h * j;
        }
    el.assign(3 * (size_t)ne, 0);
    int e = 0;
    for (int j = 0; j < r - 1; ++j)
        for (int i = 0; i < r - 1; ++i) {
            const int a = j * r + i, b = a + 1, c = a + r + 1, d = a + r;
            el[e] = a; el[ne + e] = b; el[2 * ne + e] = c; ++e;
            el[e] = a; el[ne + e] = c; el[2 * ne + e] = d; ++e;
        }
    // labels: top, bottom, right, left
    be.clear();
    for (int i = 0; i < r - 1; ++i) {
        be.push_back((r - 1) * r + i); be.push_back((r - 1) * r + i + 1); be.push_back(labels[0]);
        be.push_back(i); be.push_back(i + 1); be.push_back(labels[1]);
        be.push_back(i * r + r - 1); be.push_back((i + 1) * r + r - 1); be.push_back(labels[2]);
        be.push_back(i * r); be.push_back((i + 1) * r); be.push_back(labels[3]);
    }
}

// ---- text formats -----------------------------------------------------------------
// Mesh file (tests/Richards_Equation/*.msh): [micro: cell_width cell_height] no_verts
// no_nodes, x y per node, no_boundary_edges, (n1 n2 label) per edge, no_elements,
// no_verts 1-based vertex ids per element.
thread_local double cell_area_ld = 0.0;

int read_msh(const char* path, bool micro, double& cw, double& chh, int& nv, std::vector<double>& xy,
             std::vector<int>& be, std::vector<int>& el)
{
    FILE* f = std::fopen(path, "r");
    if (!f) return 1;
    int nn = 0, nb = 0, ne = 0;
    bool ok = true;
    if (micro) {
        // cell sizes are read and multiplied in long double, as the reference does
        // (micro_mesh_properties.cpp:54-57,194); cw*chh is then formed by build_micro
        long double lw = 0, lh = 0;
        ok = std::fscanf(f, "%Lf %Lf", &lw, &lh) == 2;
        cw = (double)lw;
        chh = (double)lh;
        cell_area_ld = (double)(lw * lh);
    }
    ok = ok && std::fscanf(f, "%d %d", &nv, &nn) == 2 && nn > 0 && (nv == 3 || nv == 4);
    if (ok) {
        xy.assign(2 * (size_t)nn, 0.0);
        for (int i = 0; i < nn && ok; ++i) ok = std::fscanf(f, "%lf %lf", &xy[i], &xy[nn + i]) == 2;
    }
    ok = ok && std::fscanf(f, "%d", &nb) == 1 && nb >= 0;
    if (ok) {
        be.assign(3 * (size_t)nb, 0);
        for (int b = 0; b < nb && ok; ++b) {
            ok = std::fscanf(f, "%d %d %d", &be[3 * b], &be[3 * b + 1], &be[3 * b + 2]) == 3;
            be[3 * b] -= 1; be[3 * b + 1] -= 1;
        }
    }
    ok = ok && std::fscanf(f, "%d", &ne) == 1 && ne > 0;
    if (ok) {
        el.assign((size_t)nv * ne, 0);
        for (int e = 0; e < ne && ok; ++e)
            for (int j = 0; j < nv && ok; ++j) { ok = std::fscanf(f, "%d", &el[(size_t)j * ne + e]) == 1; el[(size_t)j * ne + e] -= 1; }
    }
    std::fclose(f);
    if (!ok) return 3;
    for (int v : el) if (v < 0 || v >= nn) return 3;
    for (size_t b = 0; b < be.size(); b += 3) if (be[b] < 0 || be[b] >= nn || be[b + 1] < 0 || be[b + 1] >= nn) return 3;
    return 0;
}

// Keff.txt: epsilonA, no_hvalues, logr, logh_first, then rows (log10|h|, K11, K12, K21,
// K22) stored as Keff_values[c*n + i] (read_effective_conductivity.cpp:36-58).
int read_keff(const char* path, Keff_struct& K)
{
    FILE* f = std::fopen(path, "r");
    if (!f) return 1;
    int n = 0;
    bool ok = std::fscanf(f, "%lf %d %lf %lf", &K.epsilonA, &n, &K.logr, &K.logh_first) == 4 && n >= 2;
    if (ok) {
        K.no_hvalues = n;
        K.logh_values = alloc<double>(n);
        K.Keff_values = alloc<double>(4 * (size_t)n);
        for (int i = 0; i < n && ok; ++i)
            ok = std::fscanf(f, "%lf %lf %lf %lf %lf", &K.logh_values[i], &K.Keff_values[i], &K.Keff_values[n + i],
                             &K.Keff_values[2 * n + i], &K.Keff_values[3 * n + i]) == 5;
    }
    std::fclose(f);
    return ok ? 0 : 3;
}

}  // namespace tsr

// ---- reference-signature loaders (include/main.h:23-28), C++ linkage --------------
int macro_mesh_properties(const char* mesh_file, macro_mesh_struct& macro_mesh, int* macro_bc)
{
    double cw = 0, ch = 0;
    int nv = 0;
    std::vector<double> xy;
    std::vector<int> be, el;
    const int rc = tsr::read_msh(mesh_file, false, cw, ch, nv, xy, be, el);
    if (rc) return 1;
    const int nn = (int)xy.size() / 2, ne = (int)el.size() / nv;
    return tsr::build_macro(nv, nn, xy.data(), (int)be.size() / 3, be.data(), ne, el.data(), macro_bc, macro_mesh);
}

int micro_mesh_properties(const char* mesh_file, micro_mesh_struct& micro_mesh)
{
    double cw = 0, ch = 0;
    int nv = 0;
    std::vector<double> xy;
    std::vector<int> be, el;
    const int rc = tsr::read_msh(mesh_file, true, cw, ch, nv, xy, be, el);
    if (rc) return 1;
    const int nn = (int)xy.size() / 2, ne = (int)el.size() / nv;
    const int r = tsr::build_micro(cw, ch, nv, nn, xy.data(), (int)be.size() / 3, be.data(), ne, el.data(), micro_mesh);
    micro_mesh.cell_area = tsr::cell_area_ld;
    return r;
}

int read_effective_conductivity(const char* Keff_file, Keff_struct& effective_conductivity)
{
    return tsr::read_keff(Keff_file, effective_conductivity);
}

// Defaults of set_exprem_options.cpp:16-26.
int set_exprem_options(exprem_options_struct& o)
{
    o.AbsTol = 1.0e-5; o.RelTol = 1.0e-5; o.MinKrylov = 2; o.MaxKrylov = 100; o.MinStep = 1.0e-10;
    o.MaxStep = 1000.0; o.MaxNumSteps = 100000; o.SafetyFac = 0.8; o.MinStepDecFac = 0.1;
    o.MaxStepIncFac = 1.1; o.InitStep = 1.0e-3;
    return 0;
}
