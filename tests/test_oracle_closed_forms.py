"""CPU tests that pin branches of the oracle (the checker of the HIP path) which the
reference's data do not reach, by closed forms the mathematics fixes:

  * macro CVFE on the reference's quad (Q1) meshes and on the configurations' P1
    triangles, with every macro BC type of src/main.cpp:175-184 (1 flux, 2 no flux,
    3 dh/dx = 0, 4 dh/dy = 0): a linear total-head field with a constant anisotropic
    conductivity tensor makes the Darcy flux q = -K grad(H) uniform, so every interior
    control volume balances to zero and every boundary control volume's inflow is a sum
    over its boundary half-edges of closed-form terms in the half-edge length and the
    domain's outward normal (taken from the node coordinates, not from the mesh arrays);
  * EXPREM's error estimate and step-size law on u' = -u^2 (Krylov dimension 1, where
    phi_1, phi_3 and the exprb32 estimate have closed forms): the rejected first attempt,
    the retried step and the next proposal follow
    tau * min(MaxStepIncFac, max(MinStepDecFac, SafetyFac * err^(-1/3)));
  * the weighted RMS norm's special cases.

Each check is shown to fail for a planted mistake of the kind it guards against (a
flipped boundary normal, a transposed tensor, a dropped gravity term, a wrong step-law
exponent): the planted variants are computed here on the CPU.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import twoscalerich_amd as T
from twoscalerich_amd import _ffi as F
from oracle import oracle as O

from helpers import GOLDEN, soil

_dp = C.POINTER(C.c_double)
KTEN = np.array([[2.0, 0.3], [0.5, 1.5]])      # constant, anisotropic, not symmetric


def _arr(ptr, n, dt=np.float64):
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt)


def _write_const_keff(path, K, eps_a=0.6, rows=5):
    """A Keff table in the reference's format (read_effective_conductivity.cpp:15-69)
    whose every row is the tensor K."""
    logr, first = -1.0, 1.0
    with open(path, "w") as f:
        f.write(f"{eps_a!r}\n{rows}\n{logr!r}\n{first!r}\n")
        for i in range(rows):
            f.write(f"{first + i * logr!r} {float(K[0, 0])!r} {float(K[0, 1])!r} {float(K[1, 0])!r} {float(K[1, 1])!r}\n")


def _outward(bx, by, box, tol=1e-12):
    """Outward unit normal of the domain side a boundary half-edge midpoint lies on."""
    x0, x1, y0, y1 = box
    if abs(bx - x0) <= tol:
        return np.array([-1.0, 0.0])
    if abs(bx - x1) <= tol:
        return np.array([1.0, 0.0])
    if abs(by - y0) <= tol:
        return np.array([0.0, -1.0])
    if abs(by - y1) <= tol:
        return np.array([0.0, 1.0])
    raise AssertionError(f"half-edge midpoint ({bx}, {by}) is not on the domain boundary")


def _expected_macro_inflow(M, K, a, b, bflux, plant=None):
    """Net inflow into every macro CV for total head H = a x + b y + c (head h = H - y)
    and constant tensor K: interior CVs 0; a boundary node collects, per boundary
    half-edge of length l on a side with outward normal n,
        q . (l n)                      (the interior sub-faces close the CV: -q . sum n_int)
      + type 1: bflux l
        type 3: (K (0, b)) . (l n)     (dh/dx = 0, gravity kept)
        type 4: (K (a, 1)) . (l n)     (dh/dy = 0: d(h + y)/dy = 1)
    with q = -K (a, b).  Dirichlet (type 5) nodes are returned as NaN (their g is 0).
    `plant` computes a deliberately wrong variant (for the power checks)."""
    nv, nn, ne = M.no_verts, M.no_nodes, M.no_elements
    xy = _arr(M.nodes, 2 * nn).reshape(2, nn)
    box = (xy[0].min(), xy[0].max(), xy[1].min(), xy[1].max())
    el = _arr(M.elements, nv * ne, np.int64).reshape(nv, ne)
    bce = _arr(M.boundary_condition_edges, nv * ne, np.int64).reshape(nv, ne)
    lens = _arr(M.boundary_edge_lens, 2 * nv * ne).reshape(2 * nv, ne)
    mids = _arr(M.boundary_edge_midpts, 4 * nv * ne).reshape(4 * nv, ne)
    Kq = K.T if plant == "transpose" else K
    q = -Kq @ np.array([a, b if plant != "no_gravity" else b - 1.0])
    out = np.zeros(nn)
    dirichlet = np.zeros(nn, bool)
    for e in range(ne):
        for p in range(nv):
            typ = bce[p, e]
            if typ == 0:
                continue
            for hk in range(2):
                j = el[(p + hk) % nv, e]
                if typ == 5:
                    dirichlet[j] = True
                    continue
                ln = lens[2 * p + hk, e]
                n = _outward(mids[4 * p + 2 * hk, e], mids[4 * p + 2 * hk + 1, e], box)
                if plant == "flip_normal":
                    n = -n
                f = q @ (ln * n)
                if typ == 1:
                    f += bflux * ln
                elif typ == 3:
                    f += (Kq @ np.array([0.0, b])) @ (ln * n)
                elif typ == 4:
                    f += (Kq @ np.array([a, 1.0])) @ (ln * n)
                out[j] += f
    out[dirichlet] = np.nan
    return out


def _check_linear_field(p, o, a, b, c, K, bflux):
    """Oracle G on h = a x + (b - 1) y + c (all heads >= 0, so every storage is
    eps_A * Ss * CV) against the closed form; returns the relative mismatch of each
    planted variant too."""
    v = p.view.contents
    M = v.scen[0].ud.macro_mesh
    nn = M.no_nodes
    xy = _arr(M.nodes, 2 * nn).reshape(2, nn)
    u = np.full(p.N, 0.5)
    u[:nn] = a * xy[0] + (b - 1.0) * xy[1] + c
    assert u[:nn].min() >= 0.0
    g, rc = o.gfunc(u)
    assert rc == 0
    eps_a = v.scen[0].ud.effective_conductivity.epsilonA
    stor = eps_a * v.Ss * _arr(M.CV_area, nn)
    net = g[:nn] * stor
    exp = _expected_macro_inflow(M, K, a, b, bflux)
    fin = np.isfinite(exp)
    scale = np.max(np.abs(exp[fin]))
    assert scale > 1e-6
    assert np.all(g[:nn][~fin] == 0.0)
    err = np.max(np.abs(net[fin] - exp[fin])) / scale
    planted = {}
    for plant in ("flip_normal", "transpose", "no_gravity"):
        e2 = _expected_macro_inflow(M, K, a, b, bflux, plant)
        planted[plant] = np.max(np.abs(net[fin] - e2[fin])) / scale
    return err, planted


@pytest.mark.parametrize("bc", [(1, 3, 2, 2, 2), (1, 4, 3, 3, 2), (3, 4, 4, 3, 1), (4, 3, 1, 2, 4)])
@pytest.mark.parametrize("mesh", ["macro_structured_10by10.msh", "macro_structured_20by20.msh"])
def test_oracle_q1_macro_linear_field_every_bc_type(tmp_path, mesh, bc):
    """Reference quad meshes (Q1 macro CVFE, the path the reference's own inputs take),
    model 2 (no exchange term at the macro scale), constant anisotropic Keff."""
    keff = os.path.join(tmp_path, "Keff_const.txt")
    _write_const_keff(keff, KTEN)
    A = soil(0.044, 0.058, 0.41, 7.3, 1.89)
    B = soil(0.044e-3, 0.058, 0.41, 7.3, 1.89)
    bflux = 0.37
    p = T.Problem.reference(os.path.join(GOLDEN, mesh), os.path.join(GOLDEN, "microB_unstructured_10by10.msh"),
                            keff, A, B, bflux, list(bc), 2, -2.0, F.HOST_ELEMENT)
    p.view.contents.Ss = 1e-3
    o = O.Oracle(p.view)
    M = p.view.contents.scen[0].ud.macro_mesh
    assert M.no_verts == 4
    types = set(_arr(M.boundary_condition_edges, 4 * M.no_elements, np.int64)) - {0}
    assert types == set(bc)
    for a, b, c in ((0.3, 0.7, 2.0), (-0.8, 1.9, 3.0), (0.0, 0.0, 1.5)):
        err, planted = _check_linear_field(p, o, a, b, c, KTEN, bflux)
        # the Q1 coefficients come from the ill-conditioned [x y xy 1] basis (~1e-12)
        assert err <= 1e-10, (a, b, err)
        if a != 0.0:
            assert min(planted.values()) > 1e-3, planted
    o.close()
    p.close()


@pytest.mark.parametrize("top", [1, 3, 4])
def test_oracle_p1_macro_linear_field(top):
    """The configurations' right-triangle macro mesh (P1 macro CVFE), model 1, with the
    Dirichlet top edge relabelled to BC type `top` and the Keff table overwritten with
    a constant anisotropic tensor."""
    p = T.Problem.config(F.CFG_PLAIN_SMALL, 9, 5)
    v = p.view.contents
    sc = v.scen[0]
    sc.ud.model_type = 1
    sc.ud.boundary_flux = -0.21
    v.Ss = 1e-3
    M = sc.ud.macro_mesh
    nv, ne = M.no_verts, M.no_elements
    bce = np.ctypeslib.as_array(M.boundary_condition_edges, shape=(nv * ne,))
    bce[bce == 5] = top
    Kt = sc.ud.effective_conductivity
    kv = np.ctypeslib.as_array(Kt.Keff_values, shape=(4 * Kt.no_hvalues,)).reshape(4, -1)
    kv[:] = KTEN.reshape(4, 1)
    o = O.Oracle(p.view)
    for a, b, c in ((0.3, 0.7, 2.0), (-0.8, 1.9, 3.0)):
        err, planted = _check_linear_field(p, o, a, b, c, KTEN, sc.ud.boundary_flux)
        assert err <= 1e-12, (a, b, err)
        assert min(planted.values()) > 1e-3, planted
    o.close()
    p.close()


# ---------------------------------------------------------------- EXPREM step law
def _phi(z, k):
    """phi_k(z) = (e^z - sum_{i<k} z^i / i!) / z^k for scalar z, by its series where the
    quotient would cancel."""
    if abs(z) < 0.1:
        return sum(z ** i / math.factorial(i + k) for i in range(30))
    s = math.exp(z) - sum(z ** i / math.factorial(i) for i in range(k))
    return s / z ** k


def _scalar_step_closed_form(u0, tau0, tol, opt, law_exp=-1.0 / 3.0, safety=None):
    """One EXPREM step on u' = -u^2 with Krylov dimension 1 (exact Jacobian J = -2 u0):
    U2 = u0 + tau phi1(tau J) g0, D = G(U2) - g0 - J (U2 - u0), e = 2 tau phi3(tau J) D,
    err = |e| / (AbsTol + RelTol |u0|).  Rejected attempts retry with
    tau * max(MinStepDecFac, SafetyFac err^exp); the accepted step proposes
    tau * min(MaxStepIncFac, max(MinStepDecFac, SafetyFac err^exp))."""
    sf = opt.SafetyFac if safety is None else safety
    G = lambda u: -u * u
    J = -2.0 * u0
    g0 = G(u0)
    tau = tau0
    attempts = []
    while True:
        u2 = u0 + tau * _phi(tau * J, 1) * g0
        D = G(u2) - g0 - J * (u2 - u0)
        e = 2.0 * tau * _phi(tau * J, 3) * D
        err = abs(e) / (tol + tol * abs(u0))
        attempts.append((tau, err))
        if err <= 1.0:
            nxt = tau * min(opt.MaxStepIncFac, max(opt.MinStepDecFac, sf * err ** law_exp))
            return u2 + e, tau, nxt, attempts
        tau *= max(opt.MinStepDecFac, sf * err ** law_exp)


def test_oracle_step_law_on_scalar_ode():
    @F.GFUNC_T
    def G(gp, up, ud):
        gp[0] = -up[0] * up[0]
        return 0
    L = O.load()
    opt = F.exprem_options_struct()
    T.lib().set_exprem_options_c(C.byref(opt))
    opt.MaxKrylov = 1
    opt.MinKrylov = 1
    opt.InitStep = 0.5
    tol = 1.3e-3
    opt.AbsTol = opt.RelTol = tol
    u = np.array([1.0])
    t = C.c_double(0.0)
    st = F.exprem_stats_struct()
    rc = L.oracle_exprem(C.cast(G, C.c_void_p), 1, C.byref(t), 10.0, None, None, u.ctypes.data_as(_dp),
                         F.user_data_struct(), opt, C.byref(st))
    assert rc == 0
    last = np.zeros(8)
    L.oracle_exprem_last(last.ctypes.data_as(_dp))
    u_cf, tau_cf, next_cf, att = _scalar_step_closed_form(1.0, 0.5, tol, opt)
    assert len(att) >= 2 and att[0][1] > 1.0 >= att[-1][1]          # rejected, then accepted
    assert last[2] == len(att) - 1 and abs(last[3] / att[-1][1] - 1) < 1e-6   # rejects, accepted err
    # both branches of the law are active: the retry factor and the proposal are inside
    # the clamps [MinStepDecFac, MaxStepIncFac]
    assert opt.MinStepDecFac < att[1][0] / att[0][0] and next_cf / tau_cf < opt.MaxStepIncFac
    # the FD Jacobian differs from J by O(sqrt(eps)): relative 1e-7 in tau and the proposal
    assert abs(t.value / tau_cf - 1) < 1e-7 and abs(st.CurrentStep / next_cf - 1) < 1e-7
    assert abs(u[0] - u_cf) < 2e-8
    # the check has power: a wrong exponent or safety factor moves tau by far more
    _, tau_b, next_b, _ = _scalar_step_closed_form(1.0, 0.5, tol, opt, law_exp=-0.5)
    _, tau_c, next_c, _ = _scalar_step_closed_form(1.0, 0.5, tol, opt, safety=0.9)
    for tb, nb in ((tau_b, next_b), (tau_c, next_c)):
        assert abs(tb / t.value - 1) > 1e-3 or abs(nb / st.CurrentStep - 1) > 1e-3


def test_oracle_step_growth_is_capped():
    """A step far inside the tolerance proposes MaxStepIncFac * tau (the cap of the law)."""
    @F.GFUNC_T
    def G(gp, up, ud):
        gp[0] = -up[0] * up[0]
        return 0
    L = O.load()
    opt = F.exprem_options_struct()
    T.lib().set_exprem_options_c(C.byref(opt))
    opt.MaxKrylov = 1
    opt.MinKrylov = 1
    opt.InitStep = 1e-3
    opt.AbsTol = opt.RelTol = 1e-2
    u = np.array([1.0])
    t = C.c_double(0.0)
    st = F.exprem_stats_struct()
    rc = L.oracle_exprem(C.cast(G, C.c_void_p), 1, C.byref(t), 10.0, None, None, u.ctypes.data_as(_dp),
                         F.user_data_struct(), opt, C.byref(st))
    _, tau_cf, next_cf, att = _scalar_step_closed_form(1.0, 1e-3, 1e-2, opt)
    assert rc == 0 and len(att) == 1 and att[0][1] < 1e-3
    assert abs(st.CurrentStep / (opt.MaxStepIncFac * t.value) - 1) < 1e-14 and abs(next_cf / st.CurrentStep - 1) < 1e-12


# ---------------------------------------------------------------- weighted norm
def test_weighted_norm_special_cases():
    """The error norm of EXPREM (include/exprem.h:27) is the weighted RMS
    sqrt(mean((e_i / (AbsTol + RelTol |u_i|))^2)): e_i = c w_i gives |c|; a single
    nonzero entry gives its weighted size / sqrt(N) (a max norm would not divide); u = 0
    leaves AbsTol alone as the weight.  The product's host export agrees."""
    L = O.load()
    P = T.lib()
    rng = np.random.default_rng(8)
    n = 1000
    u = rng.standard_normal(n)
    at, rt = 1e-6, 1e-4
    w = at + rt * np.abs(u)

    def both(e, uu):
        e = np.ascontiguousarray(e); uu = np.ascontiguousarray(uu)
        a = L.oracle_weighted_norm(e.ctypes.data_as(_dp), uu.ctypes.data_as(_dp), at, rt, n)
        b = P.weighted_norm(e.ctypes.data_as(_dp), uu.ctypes.data_as(_dp), at, rt, n)
        return a, b
    for c in (3.5, -0.25):
        a, b = both(c * w, u)
        assert abs(a - abs(c)) < 1e-13 * abs(c) and abs(b - a) < 1e-13 * abs(c)
    e = np.zeros(n)
    e[17] = 4.0 * w[17]
    a, b = both(e, u)
    assert abs(a - 4.0 / math.sqrt(n)) < 1e-13 and abs(b - a) < 1e-13
    e = rng.standard_normal(n)
    a, b = both(e, np.zeros(n))
    assert abs(a - np.linalg.norm(e) / (at * math.sqrt(n))) < 1e-12 * a and abs(b - a) < 1e-12 * a


# ---------------------------------------------------------------- water balance, reference meshes
@pytest.mark.parametrize("model", [1, 2, 3])
@pytest.mark.parametrize("mesh,micro", [("macro_structured_10by10.msh", "microB_unstructured_10by10.msh"),
                                        ("macro_structured_20by20.msh", "microB_unstructured_20by20.msh")])
def test_oracle_discrete_mass_balance_reference_meshes(mesh, micro, model):
    """The reference's own inputs (quad macro mesh, BC (1, 3, 2, 2, 2) of main.cpp:180-184,
    models 1-3 of main.cpp:187-192): sum(storage * g) equals the inflow computed apart
    from G (flux, free-drainage and Dirichlet-neighbour half-edges, exchange)."""
    from helpers import reference_problem, seeded_state
    p = reference_problem(macro=mesh, micro=micro, model=model)
    o = O.Oracle(p.view)
    for seed in (1, 2):
        _, inflow, sg = o.water_balance(seeded_state(p, seed, -3.0, -0.2))
        assert abs(sg - inflow) <= 1e-12 * max(abs(inflow), 1e-300), (seed, sg, inflow)
    o.close()
    p.close()
