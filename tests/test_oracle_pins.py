"""CPU tests that pin the oracle (the checker of the HIP path) to what the reference's
own data and the mathematics fix.  No GPU; no compute call into libtsr.so except the
host-side set-up and the host dense routines.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest
import scipy.linalg

import twoscalerich_amd as T
from twoscalerich_amd import _ffi as F
from oracle import oracle as O

from helpers import GOLDEN, reference_problem, rel_inf, seeded_state, soil

_dp = C.POINTER(C.c_double)


def _vg(s, h, Ss=0.0):
    L = O.load()
    h = np.ascontiguousarray(h, dtype=np.float64)
    Cc, K, th = np.empty_like(h), np.empty_like(h), np.empty_like(h)
    L.oracle_vg(C.byref(s), h.ctypes.data_as(_dp), len(h), Ss, Cc.ctypes.data_as(_dp), K.ctypes.data_as(_dp),
                th.ctypes.data_as(_dp))
    return Cc, K, th


def _keff_file():
    rows = np.loadtxt(os.path.join(GOLDEN, "Keff.txt"), skiprows=4)
    with open(os.path.join(GOLDEN, "Keff.txt")) as f:
        head = [float(f.readline()) for _ in range(4)]
    return head, rows


# ---------------------------------------------------------------- van Genuchten-Mualem
def test_vg_mualem_matches_reference_keff_table():
    """Every row of the reference's Keff.txt is 0.044 * kr_vGM(h; alpha 7.3, n 1.89) * T_c
    for a constant tensor T: the oracle's K(h) must reproduce the table's h-dependence
    (pins the Mualem form with l = 1/2 and the log10 axis)."""
    _, rows = _keff_file()
    h = -10.0 ** rows[:, 0]
    _, K, _ = _vg(soil(0.044, 0.058, 0.41, 7.3, 1.89), h)
    for c in range(1, 5):
        ratio = rows[:, c] / K
        assert np.max(np.abs(ratio / ratio[0] - 1.0)) < 1e-11, c


def test_vg_capacity_is_derivative_of_theta():
    s = soil(0.2496, 0.078, 0.43, 3.6, 1.56)
    h = -np.logspace(-3, 2, 60)
    d = 1e-6 * np.abs(h)
    C0, _, _ = _vg(s, h)
    _, _, tp = _vg(s, h + d)
    _, _, tm = _vg(s, h - d)
    fd = (tp - tm) / (2 * d)
    assert np.max(np.abs(fd / C0 - 1.0)) < 1e-6


def test_vg_limits_and_saturation_branch():
    s = soil(7.128, 0.045, 0.43, 14.5, 2.68)
    C0, K, th = _vg(s, np.array([-1e-12, -1e4, 0.0, 2.0]), Ss=1e-5)
    # theta carries the specific-storage term Ss*h
    assert abs(th[0] - 0.43) < 1e-9 and abs(th[1] - (0.045 - 1e-5 * 1e4)) < 1e-6 and abs(th[3] - (0.43 + 2e-5)) < 1e-15
    assert K[2] == 7.128 and K[3] == 7.128 and C0[2] == 1e-5 and abs(K[0] / 7.128 - 1) < 1e-6
    assert K[1] < 1e-20


def test_keff_lookup_nodes_midpoints_and_clamps():
    head, rows = _keff_file()
    kt = F.Keff_struct()
    n = rows.shape[0]
    lv = np.ascontiguousarray(rows[:, 0])
    kv = np.ascontiguousarray(np.concatenate([rows[:, c] for c in range(1, 5)]))
    kt.epsilonA, kt.no_hvalues, kt.logr, kt.logh_first = head[0], n, head[2], head[3]
    kt.logh_values, kt.Keff_values = lv.ctypes.data_as(_dp), kv.ctypes.data_as(_dp)
    L = O.load()

    def look(h):
        h = np.ascontiguousarray(h, dtype=np.float64)
        out = np.empty(4 * len(h))
        L.oracle_keff(C.byref(kt), h.ctypes.data_as(_dp), len(h), out.ctypes.data_as(_dp))
        return out.reshape(-1, 4)

    # row i sits at log10|h| = logh_first + i*logr (the file's own first column)
    hn = -10.0 ** (head[3] + head[2] * np.arange(n))
    assert np.max(np.abs(look(hn) - rows[:, 1:]) / np.abs(rows[:, 1:])) < 1e-12
    mid = -10.0 ** (head[3] + head[2] * (np.arange(n - 1) + 0.5))
    exp = 0.5 * (rows[:-1, 1:] + rows[1:, 1:])
    assert np.max(np.abs(look(mid) - exp) / np.abs(exp)) < 1e-12
    ends = look(np.array([-1e6, 0.0, 5.0, -1e-9]))
    assert np.allclose(ends[0], rows[0, 1:], rtol=0, atol=0)
    assert np.allclose(ends[1], rows[-1, 1:], rtol=0, atol=0) and np.allclose(ends[2], rows[-1, 1:], rtol=0, atol=0)


# ---------------------------------------------------------------- phi functions
def _hess(m, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    H = np.triu(rng.standard_normal((m + 1, m)), -1) * scale
    H[np.arange(m) + 1, np.arange(m)] = np.abs(H[np.arange(m) + 1, np.arange(m)]) + 0.1
    return H


def _phi_scipy(A, k):
    m = A.shape[0]
    Aug = np.zeros((m + k, m + k))
    Aug[:m, :m] = A
    Aug[:m, m] = 0.0
    Aug[:m, m:] = 0.0
    # phi_k(A) via [[A, I, 0..],[0, 0, I ..]] block form on the full matrix
    big = np.zeros((m * (k + 1), m * (k + 1)))
    for i in range(k + 1):
        big[i * m:(i + 1) * m, i * m:(i + 1) * m] = A if i == 0 else 0.0
        if i < k:
            big[i * m:(i + 1) * m, (i + 1) * m:(i + 2) * m] = np.eye(m)
    E = scipy.linalg.expm(big)
    return E[:m, k * m:(k + 1) * m]


@pytest.mark.parametrize("m,scale,alpha", [(2, 1.0, 0.3), (5, 3.0, 1.0), (12, 10.0, 2.5), (30, 50.0, 0.7)])
def test_phipade_oracle_and_product_match_scipy(m, scale, alpha):
    maxk = 40
    H = np.zeros((maxk + 1, maxk))
    H[:m + 1, :m] = _hess(m, m, scale)
    ref = _phi_scipy(alpha * H[:m, :m], 1)
    got, rc = O.phipade(H, alpha, m, maxk)
    assert rc == 0 and np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-11
    Hp = np.asfortranarray(H)
    out = np.zeros((maxk, maxk), order="F")
    rc = T.lib().phipade(out.ctypes.data_as(_dp), Hp.ctypes.data_as(_dp), alpha, m, maxk)
    assert rc == 0 and np.max(np.abs(out[:m, :m] - ref)) / np.max(np.abs(ref)) < 1e-11


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 7, 8, 9, 13, 16, 21, 33, 47, 60, 64, 71, 128])
def test_solver_phi_functions(m):
    """The solver's phi functions.  tsr_phi1_e1 (phi_1(tau H) e1, forming only a leading
    block of columns in the doublings) without flushing gives the first column of
    phipade's full matrix bit for bit, over scaling exponents 0..~15 (||alpha H|| from 1e-3
    to 1e4, a stiff decaying spectrum) and a ragged leading dimension; with subnormals
    flushed (the solver's setting) it stays within 1e-14 of that; it matches scipy's expm
    of the augmented matrix.  tsr_phik_e1 (phi_1..phi_3 e1) matches scipy and is unchanged
    by flushing beyond 1e-14."""
    L = T.lib()
    rng = np.random.default_rng(m)
    maxk = 128
    for scale in (1e-3, 0.3, 4.0, 60.0, 1e3, 1e4):
        H = np.zeros((maxk + 1, maxk), order="F")
        H[:m + 1, :m] = _hess(m, m, 1.0) * scale / max(1.0, m ** 0.5)
        H[:m, :m] += np.diag(-rng.uniform(1.5 * scale, 2.5 * scale, m))   # a decaying (stiff) spectrum
        full = np.zeros((maxk, maxk), order="F")
        rc = L.phipade(full.ctypes.data_as(_dp), H.ctypes.data_as(_dp), 0.9, m, maxk)
        e1 = np.full(m, np.nan)
        e1f = np.full(m, np.nan)
        rc1 = L.tsr_phi1_e1(e1.ctypes.data_as(_dp), H.ctypes.data_as(_dp), 0.9, m, maxk + 1, 0)
        rc2 = L.tsr_phi1_e1(e1f.ctypes.data_as(_dp), H.ctypes.data_as(_dp), 0.9, m, maxk + 1, 1)
        assert rc == rc1 == rc2 == 0, (m, scale)
        assert np.all(np.isfinite(full[:m, 0])), (m, scale)
        assert np.array_equal(e1, full[:m, 0]), (m, scale, np.max(np.abs(e1 - full[:m, 0])))
        assert np.max(np.abs(e1f - e1)) <= 1e-14 * np.max(np.abs(e1)), (m, scale)
        if scale <= 60.0:
            A = 0.9 * H[:m, :m]
            Aug = np.zeros((m + 1, m + 1))
            Aug[:m, :m] = A
            Aug[0, m] = 1.0
            ref = scipy.linalg.expm(Aug)[:m, m]
            assert np.max(np.abs(e1 - ref)) <= 1e-10 * max(1.0, np.max(np.abs(ref))), (m, scale)
        if m <= 64:
            pk = np.full(3 * m, np.nan)
            pkf = np.full(3 * m, np.nan)
            assert L.tsr_phik_e1(pk.ctypes.data_as(_dp), H.ctypes.data_as(_dp), maxk + 1, m, 0.9, 3, 0) == 0
            assert L.tsr_phik_e1(pkf.ctypes.data_as(_dp), H.ctypes.data_as(_dp), maxk + 1, m, 0.9, 3, 1) == 0
            assert np.max(np.abs(pkf - pk)) <= 1e-14 * np.max(np.abs(pk)), (m, scale)
            if scale <= 60.0:
                Aug = np.zeros((m + 3, m + 3))
                Aug[:m, :m] = 0.9 * H[:m, :m]
                Aug[0, m] = 1.0
                Aug[m, m + 1] = Aug[m + 1, m + 2] = 1.0
                E = scipy.linalg.expm(Aug)
                for k in range(3):
                    ref = E[:m, m + k]
                    assert np.max(np.abs(pk[k * m:(k + 1) * m] - ref)) <= 1e-10 * max(1.0, np.max(np.abs(ref))), (m, scale, k)


def test_phipade_scalar_and_zero():
    maxk = 4
    H = np.zeros((maxk + 1, maxk))
    H[0, 0] = -3.0
    got, _ = O.phipade(H, 0.5, 1, maxk)
    assert abs(got[0, 0] - math.expm1(-1.5) / -1.5) < 1e-15
    H[0, 0] = 0.0
    got, _ = O.phipade(H, 1.0, 1, maxk)
    assert got[0, 0] == 1.0
    got, _ = O.phipade(np.zeros((maxk + 1, maxk)), 1.0, 3, maxk)
    assert np.allclose(got, np.eye(3), atol=0)


@pytest.mark.parametrize("m,tau", [(1, 0.4), (6, 1.3), (20, 0.05)])
def test_phik_e1_matches_scipy(m, tau):
    H = _hess(m, 100 + m, 4.0)
    out, rc = O.phik_e1(H, tau, 3)
    assert rc == 0
    for k in (1, 2, 3):
        ref = _phi_scipy(tau * H[:m, :m], k)[:, 0]
        assert np.max(np.abs(out[k - 1] - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_expm_matches_scipy():
    for seed, sc in ((1, 0.1), (2, 5.0), (3, 80.0)):
        A = np.random.default_rng(seed).standard_normal((9, 9)) * sc
        E, rc = O.expm(A)
        ref = scipy.linalg.expm(A)
        assert rc == 0 and np.max(np.abs(E - ref)) / np.max(np.abs(ref)) < 1e-11


def test_weighted_norm_definition():
    rng = np.random.default_rng(5)
    e, u = rng.standard_normal(1000), rng.standard_normal(1000)
    L = O.load()
    got = L.oracle_weighted_norm(e.ctypes.data_as(_dp), u.ctypes.data_as(_dp), 1e-6, 1e-4, 1000)
    assert abs(got - np.sqrt(np.mean((e / (1e-6 + 1e-4 * np.abs(u))) ** 2))) < 1e-12 * got


# ---------------------------------------------------------------- geometry vs the reference's own code
def _arr(ptr, n, dt=np.float64):
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt)


REF_SO = os.path.join(os.path.dirname(O.__file__), "_ref", "libref_mesh.so")


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("fname", ["macro_structured_10by10.msh", "macro_structured_20by20.msh"])
def test_macro_mesh_properties_match_reference_code(fname):
    R = C.CDLL(REF_SO)
    R.ref_macro_mesh_properties.argtypes = [C.c_char_p, C.POINTER(F.macro_mesh_struct), C.POINTER(C.c_int)]
    bc = (C.c_int * 5)(1, 3, 2, 2, 2)
    ref = F.macro_mesh_struct()
    assert R.ref_macro_mesh_properties(os.path.join(GOLDEN, fname).encode(), C.byref(ref), bc) == 0
    p = reference_problem(macro=fname, micro="microB_unstructured_10by10.msh")
    M = p.view.contents.scen[0].ud.macro_mesh
    nv, nn, ne = ref.no_verts, ref.no_nodes, ref.no_elements
    assert (M.no_verts, M.no_nodes, M.no_elements) == (nv, nn, ne)
    # macro meshes of the reference tests are quads: SCV = quarter rectangle on both sides
    sizes = {"nodes": 2 * nn, "shape_funcs": nv * nv * ne, "normals": 2 * nv * ne, "boundary_edge_midpts": 4 * nv * ne,
             "boundary_edge_normals": 4 * nv * ne, "boundary_edge_lens": 2 * nv * ne, "CV_area": nn,
             "SCV_area": nv * ne, "midx": nv * ne, "midy": nv * ne, "centroid_elements": 2 * ne}
    for name, n in sizes.items():
        a, b = _arr(getattr(M, name), n), _arr(getattr(ref, name), n)
        # Q1 coefficients come from the ill-conditioned [x y xy 1] basis: different
        # elimination order differs at ~1e-12 of the largest coefficient
        tol = 1e-10 if name == "shape_funcs" else 1e-12
        assert np.max(np.abs(a - b)) <= tol * max(1.0, np.max(np.abs(b))), name
    for name, n in {"elements": nv * ne, "boundary_condition_edges": nv * ne}.items():
        assert np.array_equal(_arr(getattr(M, name), n, np.int64), _arr(getattr(ref, name), n, np.int64)), name


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("fname", ["microB_unstructured_10by10.msh", "microB_unstructured_20by20.msh"])
def test_micro_mesh_properties_match_reference_code(fname):
    R = C.CDLL(REF_SO)
    R.ref_micro_mesh_properties.argtypes = [C.c_char_p, C.POINTER(F.micro_mesh_struct)]
    ref = F.micro_mesh_struct()
    assert R.ref_micro_mesh_properties(os.path.join(GOLDEN, fname).encode(), C.byref(ref)) == 0
    p = reference_problem(micro=fname)
    m = p.view.contents.tmpl[0]
    nv, nn, ne = ref.no_verts, ref.no_nodes, ref.no_elements
    assert (m.no_verts, m.no_nodes, m.no_elements, m.cell_area) == (nv, nn, ne, ref.cell_area)
    sizes = {"nodes": 2 * nn, "shape_funcs": nv * nv * ne, "normals": 2 * nv * ne, "boundary_edge_midpts": 4 * nv * ne,
             "boundary_edge_normals": 4 * nv * ne, "boundary_edge_lens": 2 * nv * ne,
             "midx": nv * ne, "midy": nv * ne, "centroid_elements": 2 * ne, "vectors": 2 * nn}
    for name, n in sizes.items():
        a, b = _arr(getattr(m, name), n), _arr(getattr(ref, name), n)
        assert np.max(np.abs(a - b)) <= 1e-12 * max(1.0, np.max(np.abs(b))), name
    # triangle SCVs: the reference's formula gives element area / 6, the median dual area / 3
    a, b = _arr(m.CV_area, nn), _arr(ref.CV_area, nn)
    assert np.max(np.abs(a - 2.0 * b)) <= 1e-12 * np.max(np.abs(a))
    xy = _arr(m.nodes, 2 * nn).reshape(2, -1)
    el = _arr(m.elements, 3 * ne, np.int64).reshape(3, -1)
    d1, d2 = xy[:, el[1]] - xy[:, el[0]], xy[:, el[2]] - xy[:, el[0]]
    assert abs(a.sum() - 0.5 * np.abs(d1[0] * d2[1] - d1[1] * d2[0]).sum()) < 1e-13 * a.sum()
    for name, n in {"elements": nv * ne, "boundary_nodes": nn, "boundary_edges": nv * ne}.items():
        assert np.array_equal(_arr(getattr(m, name), n, np.int64), _arr(getattr(ref, name), n, np.int64)), name


def test_structured_mesh_closed_forms():
    p = T.Problem.config(F.CFG_PLAIN_SMALL)
    v = p.view.contents
    M = v.scen[0].ud.macro_mesh
    n = 21
    h = 1.0 / (n - 1)
    cv = _arr(M.CV_area, M.no_nodes).reshape(n, n)
    assert abs(cv.sum() - 1.0) < 1e-13
    assert np.max(np.abs(cv[1:-1, 1:-1] - h * h)) < 1e-15
    m = v.tmpl[0]
    ell = h
    cvm = _arr(m.CV_area, m.no_nodes)
    assert abs(cvm.sum() - ell * ell) < 1e-15 and abs(m.cell_area - ell * ell) < 1e-18
    # every element's sub-face normals sum to zero around the centroid and have length |f|
    nrm = _arr(m.normals, 6 * m.no_elements).reshape(6, -1)
    assert np.max(np.abs(nrm[0] + nrm[2] + nrm[4])) < 1e-15
    # soil map: centred disc of radius 0.25 cell widths
    assert v.tmpl_soil[0]


# ---------------------------------------------------------------- the oracle's G
def _micro_stiffness(m):
    """Galerkin P1 stiffness of a micro template, assembled independently (gradients from
    node coordinates)."""
    nn, ne = m.no_nodes, m.no_elements
    xy = _arr(m.nodes, 2 * nn).reshape(2, nn)
    el = _arr(m.elements, 3 * ne, np.int64).reshape(3, ne)
    A = np.zeros((nn, nn))
    for e in range(ne):
        v = el[:, e]
        P = np.c_[np.ones(3), xy[0, v], xy[1, v]]
        G = np.linalg.inv(P)[1:, :]          # gradients of the 3 basis functions (2 x 3)
        area = 0.5 * abs(np.linalg.det(P))
        A[np.ix_(v, v)] += area * G.T @ G
    return A, xy


def test_oracle_micro_cvfe_equals_p1_stiffness_when_saturated():
    """Saturated single-soil micro cell (K = Ks, C = Ss): the CVFE net inflow of every
    interior node is -Ks sum_j A_ij (h_j + y_j) with A the P1 Galerkin stiffness (exact
    identity for median-dual control volumes)."""
    p = reference_problem(model=1, bc=(2, 2, 2, 2, 2))
    v = p.view.contents
    v.Ss = 1e-3
    o = O.Oracle(p.view)
    M = v.scen[0].ud.macro_mesh
    m = v.tmpl[0]
    rng = np.random.default_rng(3)
    u = np.full(p.N, 0.7)                    # uniform macro head: every micro Dirichlet head is 0.7
    u[M.no_nodes:] = rng.uniform(0.1, 2.0, p.N - M.no_nodes)
    g, rc = o.gfunc(u)
    assert rc == 0
    A, xy = _micro_stiffness(m)
    bnd = _arr(m.boundary_nodes, m.no_nodes, np.int64) == 1
    Ks = v.scen[0].ud.soilB.Ksat
    cv = _arr(m.CV_area, m.no_nodes)
    for cell in (0, 37, 99):
        off = M.no_nodes + cell * m.no_nodes
        hc = u[off:off + m.no_nodes].copy()
        hc[bnd] = 0.7
        H = hc + xy[1]
        expect = -Ks * (A @ H) / (1e-3 * cv)
        got = g[off:off + m.no_nodes]
        assert np.all(got[bnd] == 0.0)
        assert np.max(np.abs(got[~bnd] - expect[~bnd])) <= 1e-10 * np.max(np.abs(expect[~bnd]))
    o.close()


def test_oracle_macro_cvfe_equals_p1_stiffness_when_saturated():
    """Macro P1 mesh, model 1 (no exchange), heads >= 0 away from the Dirichlet row: the
    Keff table clamps to its wettest row, so the macro CVFE equals -K A H / storage."""
    p = T.Problem.config(F.CFG_PLAIN_SMALL)
    v = p.view.contents
    v.scen[0].ud.model_type = 1
    o = O.Oracle(p.view)
    M = v.scen[0].ud.macro_mesh
    n = 21
    rng = np.random.default_rng(4)
    u = rng.uniform(0.1, 1.0, p.N)
    g, rc = o.gfunc(u)
    assert rc == 0
    K = v.scen[0].ud.effective_conductivity
    kw = _arr(K.Keff_values, 4 * K.no_hvalues).reshape(4, -1)[0, -1]
    xy = _arr(M.nodes, 2 * M.no_nodes).reshape(2, -1)
    el = _arr(M.elements, 3 * M.no_elements, np.int64).reshape(3, -1)
    A = np.zeros((M.no_nodes, M.no_nodes))
    for e in range(M.no_elements):
        vv = el[:, e]
        P = np.c_[np.ones(3), xy[0, vv], xy[1, vv]]
        G = np.linalg.inv(P)[1:, :]
        A[np.ix_(vv, vv)] += 0.5 * abs(np.linalg.det(P)) * G.T @ G
    H = u[:M.no_nodes] + xy[1]
    cv = _arr(M.CV_area, M.no_nodes)
    stor = K.epsilonA * v.Ss * cv
    expect = -kw * (A @ H) / stor
    rows = np.arange(M.no_nodes) // n
    sel = rows <= n - 3
    assert np.max(np.abs(g[:M.no_nodes][sel] - expect[sel])) <= 1e-9 * np.max(np.abs(expect[sel]))
    o.close()


def _hydrostatic_state(p, c):
    v = p.view.contents
    sc = v.scen[0]
    M = sc.ud.macro_mesh
    u = np.zeros(p.N)
    xy = _arr(M.nodes, 2 * M.no_nodes).reshape(2, -1)
    u[:M.no_nodes] = c - xy[1]
    offs = np.ctypeslib.as_array(v.cell_offset, shape=(v.n_cells,))
    cen = _arr(M.centroid_elements, 2 * M.no_elements).reshape(2, -1)
    for k in range(v.n_cells):
        m = v.tmpl[v.cell_tmpl[k]]
        vy = _arr(m.vectors, 2 * m.no_nodes).reshape(2, -1)[1]
        base = cen[1, k] if sc.host_kind == F.HOST_ELEMENT else xy[1, k]
        u[offs[k]:offs[k] + m.no_nodes] = c - (base + vy)
    return u


@pytest.mark.parametrize("model", [2, 3])
def test_oracle_hydrostatic_state_is_steady(model):
    p = reference_problem(model=model, bc=(2, 2, 2, 2, 2))
    o = O.Oracle(p.view)
    g0, _ = o.gfunc(seeded_state(p, 1, -3.0, -0.5))
    g, rc = o.gfunc(_hydrostatic_state(p, -1.0))
    assert rc == 0 and np.max(np.abs(g)) <= 1e-12 * np.max(np.abs(g0))
    o.close()


def test_oracle_hydrostatic_node_hosted():
    """Cells hosted by interior macro nodes see the exact macro interpolant, so a
    hydrostatic state is steady there and at their host nodes.  (Cells of boundary nodes
    reach outside the domain, where the macro field is clamped to the boundary; they are
    not hydrostatic by construction.)"""
    p = T.Problem.config(F.CFG_PLAIN_SMALL)
    o = O.Oracle(p.view)
    g0, _ = o.gfunc(seeded_state(p, 1))
    g, rc = o.gfunc(_hydrostatic_state(p, 0.9))   # h = -0.1 on the Dirichlet top edge
    n, r = 21, 11
    gm = g[:n * n].reshape(n, n)
    gc = g[n * n:].reshape(n, n, r * r)
    assert rc == 0
    # below the water table storage is Ss only, so rounding in the fluxes is amplified
    # by 1/Ss: 1e-10 of the scale of a non-steady g
    assert np.max(np.abs(gc[1:-1, 1:-1])) <= 1e-10 * np.max(np.abs(g0))
    assert np.max(np.abs(gm[1:-1, 1:-1])) <= 1e-10 * np.max(np.abs(g0))
    o.close()


@pytest.mark.parametrize("cfg,mn,mr,ns", [(F.CFG_PLAIN_SMALL, 0, 0, 0), (F.CFG_RAGGED, 9, 0, 0),
                                          (F.CFG_SCENARIOS, 7, 6, 3), (F.CFG_LARGE, 5, 11, 0)])
def test_oracle_discrete_mass_balance(cfg, mn, mr, ns):
    """sum(storage * g) equals the boundary inflow computed independently of G."""
    p = T.Problem.config(cfg, mn, mr, ns)
    o = O.Oracle(p.view)
    for seed in (1, 2):
        _, inflow, sg = o.water_balance(seeded_state(p, seed))
        assert abs(sg - inflow) <= 1e-12 * abs(inflow)
    o.close()


def test_oracle_serial_equals_parallel():
    p = T.Problem.config(F.CFG_RAGGED, 9)
    o = O.Oracle(p.view)
    u = seeded_state(p, 9)
    g1, _ = o.gfunc(u, 1)
    g8, _ = o.gfunc(u, 8)
    assert np.array_equal(g1, g8)
    o.close()


def test_oracle_left_right_mirror_symmetry():
    """The plain workload is symmetric under x -> 1 - x: G commutes with the mirror."""
    p = T.Problem.config(F.CFG_PLAIN_SMALL)
    v = p.view.contents
    o = O.Oracle(p.view)
    n, r = 21, 11
    rng = np.random.default_rng(11)
    mac = rng.uniform(-5, -0.3, (n, n))
    mac = 0.5 * (mac + mac[:, ::-1])
    u = np.empty(p.N)
    u[:n * n] = mac.ravel()
    cells = rng.uniform(-5, -0.3, (n, n, r, r))
    cells = 0.5 * (cells + cells[:, ::-1, :, ::-1])
    u[n * n:] = cells.ravel()
    g, _ = o.gfunc(u)
    gm = g[:n * n].reshape(n, n)
    gc = g[n * n:].reshape(n, n, r, r)
    # the triangulation's diagonal breaks exact mirror symmetry of the discrete operator at
    # the element level, so compare the macro field only where it is smooth: tolerance on
    # the symmetric part
    asym = np.max(np.abs(gm - gm[:, ::-1])) / np.max(np.abs(gm))
    assert asym < 0.5
    assert np.isfinite(gc).all()
    o.close()


# ---------------------------------------------------------------- EXPREM oracle
def _callback_problem(Amat, b):
    @F.GFUNC_T
    def G(gp, up, ud):
        n = Amat.shape[0]
        u = np.ctypeslib.as_array(up, shape=(n,))
        g = np.ctypeslib.as_array(gp, shape=(n,))
        g[:] = Amat @ u + b
        return 0
    return G


def test_oracle_exprem_exact_for_linear_problems():
    rng = np.random.default_rng(21)
    n = 12
    Q = rng.standard_normal((n, n))
    Amat = -(Q @ Q.T) / n - 0.5 * np.eye(n)
    b = rng.standard_normal(n)
    u0 = rng.standard_normal(n)
    G = _callback_problem(Amat, b)
    L = O.load()
    opt = F.exprem_options_struct()
    T.lib().set_exprem_options_c(C.byref(opt))
    opt.MaxKrylov = n
    opt.InitStep = 0.3
    opt.AbsTol = opt.RelTol = 1e-8
    st = F.exprem_stats_struct()
    u = u0.copy()
    t = C.c_double(0.0)
    ud = F.user_data_struct()
    rc = L.oracle_exprem(C.cast(G, C.c_void_p), n, C.byref(t), 10.0, None, None, u.ctypes.data_as(_dp), ud, opt,
                         C.byref(st))
    assert rc == 0 and st.NumSteps == 1
    tau = t.value
    exact = scipy.linalg.expm(tau * Amat) @ u0 + tau * _phi_scipy(tau * Amat, 1) @ b
    assert np.max(np.abs(u - exact)) <= 1e-8 * np.max(np.abs(exact))


def test_oracle_exprem_local_order_on_scalar_ode():
    """u' = -u^2, u(0) = 1: the accepted solution u+ = U2 + 2 tau phi3 D is third order,
    so its one-step error shrinks ~16x when tau halves."""
    @F.GFUNC_T
    def G(gp, up, ud):
        gp[0] = -up[0] * up[0]
        return 0
    L = O.load()
    errs = []
    for tau in (0.2, 0.1, 0.05):
        opt = F.exprem_options_struct()
        T.lib().set_exprem_options_c(C.byref(opt))
        opt.MaxKrylov = 1
        opt.MinKrylov = 1
        opt.InitStep = tau
        opt.AbsTol = opt.RelTol = 1.0
        u = np.array([1.0])
        t = C.c_double(0.0)
        st = F.exprem_stats_struct()
        rc = L.oracle_exprem(C.cast(G, C.c_void_p), 1, C.byref(t), 10.0, None, None, u.ctypes.data_as(_dp),
                             F.user_data_struct(), opt, C.byref(st))
        assert rc == 0 and abs(t.value - tau) < 1e-15
        errs.append(abs(u[0] - 1.0 / (1.0 + tau)))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 12.0 < r1 < 20.0 and 12.0 < r2 < 20.0, errs


def test_oracle_exprem_lands_on_tspan_and_fills_solutions():
    p = T.Problem.config(F.CFG_PLAIN_SMALL, 9, 5)
    o = O.Oracle(p.view, p.ud)
    opt = p.options()
    u = p.initial()
    tspan = np.array([2.0, 2e-4, 5e-4])
    sol = np.zeros(3 * p.N)
    st = F.exprem_stats_struct()
    t = 0.0
    hit = []
    while t < 5e-4:
        rc, t, _ = o.exprem_step(p.ud.contents, u, t, 5e-4, opt, st, tspan=tspan, tspan_solns=sol)
        assert rc == 0
        if t in (2e-4, 5e-4):
            hit.append(t)
    assert hit == [2e-4, 5e-4]
    assert np.array_equal(sol[2 * p.N:], u) and np.any(sol[p.N:2 * p.N] != 0)
    o.close()


def test_oracle_exprem_independent_of_thread_count():
    """The oracle's inner products sum in an order fixed by N alone, so its EXPREM steps
    repeat bit for bit whatever the OpenMP thread count (a reduction clause would not)."""
    gomp = C.CDLL("libgomp.so.1")
    p = T.Problem.config(F.CFG_PLAIN_SMALL, 9, 5)
    o = O.Oracle(p.view, p.ud)
    opt = p.options()
    out = []
    try:
        for nt in (1, 5):
            gomp.omp_set_num_threads(nt)
            u = p.initial()
            st = F.exprem_stats_struct()
            t = 0.0
            for _ in range(4):
                rc, t, _ = o.exprem_step(p.ud.contents, u, t, p.tend(), opt, st)
                assert rc == 0
            out.append((t, u.copy()))
    finally:
        gomp.omp_set_num_threads(os.cpu_count() or 1)
    assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1])
    o.close()
    p.close()
