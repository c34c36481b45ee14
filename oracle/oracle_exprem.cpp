// oracle_exprem.cpp - CPU oracle (TEST INFRASTRUCTURE ONLY) of one EXPREM step.
//
// The reference's exprem (include/main.h:29-32) is absent from the tree
// (compile.sh:18-21); the call contract is fixed by main.cpp:311-341 (one accepted
// step per call, t and u0 updated in place, return -1 on failure), the option and
// statistics structs (exprem_options.h, exprem_stats.h, set_exprem_options.cpp:16-26),
// tspan storage (main.cpp:87-94,270-274) and phipade (src/phipade.cpp).  The method:
//
//   exponential Rosenbrock-Euler with the exprb32 embedded error estimate
//   (Hochbruck, Ostermann & Schweitzer, SINUM 47 (2009)):
//     U2 = u + tau*phi1(tau J) G(u)
//     D  = G(U2) - G(u) - J (U2 - u)
//     u+ = U2 + 2 tau phi3(tau J) D,      err = WRMS(2 tau phi3(tau J) D)
//   phi products by Arnoldi with finite-difference J v = (G(u + eps v) - G(u)) / eps,
//   eps = sqrt(DBL_EPSILON) (1 + ||u||_2) (Brown & Saad 1990), classical Gram-Schmidt
//   with one reorthogonalisation (CGS2).  J(U2 - u) = V_{m+1} Hbar_m y by the Arnoldi
//   relation (no extra RHS call).  Krylov stopping: tau * (tau beta h_{m+1,m}
//   |e_m^T phi1(tau H_m) e1|) * WRMS(v_{m+1}) <= 0.1 (generalised residual,
//   Hochbruck-Lubich-Selhofer 1998).  Step law
//     tau_new = tau * min(MaxStepIncFac, max(MinStepDecFac, SafetyFac * err^(-1/3))).
//   A rejected step reuses the Krylov basis of G(u) (it does not depend on tau).
// Every choice above is the build's own [the reference gives no exprem source]; the HIP
// driver implements the same steps and constants.
#include "oracle.h"

#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

constexpr double kKrylovTol = 0.1;

double dot(const double* a, const double* b, long long n)
{
    return oracle_fixed_sum(n, [&](long long i) { return a[i] * b[i]; });
}

double wrms(const double* e, const double* u, double at, double rt, long long n)
{
    return oracle_weighted_norm(e, u, at, rt, n);
}

struct Ws {
    long long N = 0;
    int mmax = 0;
    std::vector<double> V, W, g0, w, up, u2, g2, D, e, H, HD;
};

Ws& ws(long long N, int mmax)
{
    static Ws W;
    if (W.N != N || W.mmax != mmax) {
        W.N = N; W.mmax = mmax;
        W.V.assign((size_t)N * (mmax + 1), 0.0);
        W.W.assign((size_t)N * (mmax + 1), 0.0);
        W.g0.assign(N, 0.0); W.w.assign(N, 0.0); W.up.assign(N, 0.0); W.u2.assign(N, 0.0);
        W.g2.assign(N, 0.0); W.D.assign(N, 0.0); W.e.assign(N, 0.0);
        W.H.assign((size_t)(mmax + 1) * mmax, 0.0);
        W.HD.assign((size_t)(mmax + 1) * mmax, 0.0);
    }
    return W;
}

struct Last { double krylov_m, krylov_mD, rejects, err, tau, beta, hnorm, fevals; } g_last;

// Arnoldi for the FD Jacobian at u, starting vector b/beta.  Basis columns in B
// (length N, ld N), Hessenberg in H (ld mmax+1).  stop(m, hn) returns true when the
// m-dimensional space is enough.  Returns m (>= 1), or -1 if G failed.
template <class Stop>
int arnoldi(tsr_gfunc_t G, const user_data_struct& ud, const double* u, const double* g0, const double* b,
            double beta, double* B, double* H, int mmax, double epsfd, Ws& W, int& fevals, bool& breakdown,
            double& hn_out, Stop stop)
{
    const long long N = W.N;
    const int ld = W.mmax + 1;
    breakdown = false;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < N; ++i) B[i] = b[i] / beta;
    int m = 0;
    for (int j = 0; j < mmax; ++j) {
        const double* vj = B + (size_t)j * N;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < N; ++i) W.up[i] = u[i] + epsfd * vj[i];
        const int flag = G(W.w.data(), W.up.data(), ud);
        ++fevals;
        if (flag != 0) return -1;
        const double ie = 1.0 / epsfd;
        double* w = W.w.data();
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < N; ++i) w[i] = (w[i] - g0[i]) * ie;
        const double wnorm0 = std::sqrt(dot(w, w, N));
        for (int k = 0; k <= j; ++k) H[k + (size_t)j * ld] = 0.0;
        for (int pass = 0; pass < 2; ++pass) {
            std::vector<double> h(j + 1);
            for (int k = 0; k <= j; ++k) h[k] = dot(B + (size_t)k * N, w, N);
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < N; ++i) {
                double s = w[i];
                for (int k = 0; k <= j; ++k) s -= B[(size_t)k * N + i] * h[k];
                w[i] = s;
            }
            for (int k = 0; k <= j; ++k) H[k + (size_t)j * ld] += h[k];
        }
        const double hn = std::sqrt(dot(w, w, N));
        H[(j + 1) + (size_t)j * ld] = hn;
        m = j + 1;
        hn_out = hn;
        if (!(hn > 1e-12 * wnorm0) || hn == 0.0) { breakdown = true; break; }
        double* vn = B + (size_t)(j + 1) * N;
        const double ih = 1.0 / hn;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < N; ++i) vn[i] = w[i] * ih;
        if (stop(m, hn)) break;
    }
    return m;
}

}  // namespace

extern "C" {

void oracle_exprem_last(double* out8)
{
    std::memcpy(out8, &g_last, sizeof(g_last));
}

int oracle_exprem(int (*Gfunc)(double*, double*, user_data_struct), int N, double* tp, double tend,
                  double* tspan, double* tspan_solns, double* u0, user_data_struct ud,
                  exprem_options_struct opt, exprem_stats_struct* st)
{
    double& t = *tp;
    if (t == 0.0) {
        st->CurrentStep = opt.InitStep; st->CurrentTime = 0.0; st->CurrentKryDim = 0;
        st->CurrentSoln = u0; st->NumSteps = 0; st->FuncEvals = 0;
    }
    if (st->NumSteps >= opt.MaxNumSteps) return -1;
    const int mmax = opt.MaxKrylov;
    Ws& W = ws(N, mmax);
    const int ld = mmax + 1;
    int fe = 0;
    const double at = opt.AbsTol, rt = opt.RelTol;

    // step size proposal, output landing
    double proposal = std::fmin(std::fmax(st->CurrentStep, opt.MinStep), opt.MaxStep);
    const int nout = tspan ? (int)tspan[0] : 0;
    int out_k = -1;
    double tstop = tend;
    // next output time after t (tspan[0] = count, main.cpp:87-94); land on it if <= tend
    for (int k = 1; k <= nout; ++k)
        if (tspan[k] > t * (1.0 + 1e-14) + 1e-300 && tspan[k] <= tstop) { tstop = tspan[k]; out_k = k; }
    double tau = proposal;
    bool landing = false;
    if (t + tau >= tstop - 1e-13 * std::fabs(tstop)) { tau = tstop - t; landing = true; }

    double* g0 = W.g0.data();
    if (Gfunc(g0, u0, ud) != 0) return -1;
    ++fe;
    const double beta = std::sqrt(dot(g0, g0, N));
    const double unorm = std::sqrt(dot(u0, u0, N));
    const double epsfd = std::sqrt(DBL_EPSILON) * (1.0 + unorm);

    int m = 0, mD = 0, rejects = 0;
    double err = 0.0, hn = 0.0;
    bool breakdown = false;
    std::vector<double> phi((size_t)mmax * mmax), y(mmax + 1), Hy(mmax + 2), z(3 * (mmax + 1));

    auto wrms_col = [&](const double* v) { return wrms(v, u0, at, rt, N); };
    auto krylov_err = [&](int mm, double hnn, double tt, const double* vnext) {
        if (oracle_phipade(phi.data(), W.H.data(), tt, mm, mmax)) return (double)INFINITY;
        return tt * (tt * beta * hnn * std::fabs(phi[mm - 1])) * wrms_col(vnext);
    };

    if (beta > 0.0) {
        m = arnoldi(Gfunc, ud, u0, g0, g0, beta, W.V.data(), W.H.data(), mmax, epsfd, W, fe, breakdown, hn,
                    [&](int mm, double hnn) {
                        if (mm < opt.MinKrylov) return false;
                        return krylov_err(mm, hnn, tau, W.V.data() + (size_t)mm * N) <= kKrylovTol;
                    });
        if (m < 0) return -1;
        if (!breakdown) {
            // shrink tau until the Krylov estimate passes at this dimension
            while (krylov_err(m, hn, tau, W.V.data() + (size_t)m * N) > kKrylovTol) {
                tau *= 0.5; landing = false;
                if (tau < opt.MinStep) return -1;
            }
        }
    }

    for (;;) {
        double* u2 = W.u2.data();
        if (beta > 0.0) {
            if (oracle_phipade(phi.data(), W.H.data(), tau, m, mmax)) return -1;
            for (int k = 0; k < m; ++k) y[k] = tau * beta * phi[k];
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < N; ++i) {
                double s = u0[i];
                for (int k = 0; k < m; ++k) s += W.V[(size_t)k * N + i] * y[k];
                u2[i] = s;
            }
        } else {
            std::memcpy(u2, u0, sizeof(double) * N);
        }
        const int flag = Gfunc(W.g2.data(), u2, ud);
        ++fe;
        bool ok = (flag == 0);
        if (ok) {
            const int mv = breakdown ? m : m + 1;   // columns of V_{m+1} available
            for (int i = 0; i < mv; ++i) {
                double s = 0.0;
                for (int k = 0; k < m; ++k) s += W.H[i + (size_t)k * ld] * y[k];
                Hy[i] = s;
            }
            double* D = W.D.data();
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < N; ++i) {
                double s = W.g2[i] - g0[i];
                if (beta > 0.0) for (int k = 0; k < mv; ++k) s -= W.V[(size_t)k * N + i] * Hy[k];
                D[i] = s;
            }
            const double betaD = std::sqrt(dot(D, D, N));
            double* e = W.e.data();
            if (betaD > 0.0 && std::isfinite(betaD)) {
                bool bdD = false;
                double hnD = 0.0;
                std::vector<double> ph(3 * (size_t)mmax);
                mD = arnoldi(Gfunc, ud, u0, g0, D, betaD, W.W.data(), W.HD.data(), mmax, epsfd, W, fe, bdD, hnD,
                             [&](int mm, double hnn) {
                                 if (mm < opt.MinKrylov) return false;
                                 if (oracle_phik_e1(ph.data(), W.HD.data(), ld, mm, tau, 3)) return false;
                                 const double est = 2.0 * tau * betaD * (tau * hnn * std::fabs(ph[2 * mm + mm - 1]))
                                                    * wrms_col(W.W.data() + (size_t)mm * N);
                                 return est <= kKrylovTol;
                             });
                if (mD < 0) return -1;
                if (oracle_phik_e1(ph.data(), W.HD.data(), ld, mD, tau, 3)) return -1;
                for (int k = 0; k < mD; ++k) z[k] = 2.0 * tau * betaD * ph[2 * mD + k];
#pragma omp parallel for schedule(static)
                for (long long i = 0; i < N; ++i) {
                    double s = 0.0;
                    for (int k = 0; k < mD; ++k) s += W.W[(size_t)k * N + i] * z[k];
                    e[i] = s;
                }
                err = wrms(e, u0, at, rt, N);
            } else if (betaD == 0.0) {
                std::memset(e, 0, sizeof(double) * N);
                err = 0.0;
            } else {
                ok = false;
            }
        }
        if (ok && std::isfinite(err) && err <= 1.0) {
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < N; ++i) u0[i] = u2[i] + W.e[i];
            break;
        }
        ++rejects;
        double fac = (ok && std::isfinite(err)) ? std::fmax(opt.MinStepDecFac, opt.SafetyFac * std::pow(err, -1.0 / 3.0))
                                                : 0.5;
        tau *= fac;
        landing = false;
        if (tau < opt.MinStep) return -1;
    }

    // accepted
    const double tau_taken = tau;
    t = landing ? tstop : t + tau;
    double fac = (err > 0.0) ? std::fmin(opt.MaxStepIncFac, std::fmax(opt.MinStepDecFac, opt.SafetyFac * std::pow(err, -1.0 / 3.0)))
                             : opt.MaxStepIncFac;
    double next = tau_taken * fac;
    if (landing && tau_taken < proposal) next = std::fmax(next, proposal);
    next = std::fmin(std::fmax(next, opt.MinStep), opt.MaxStep);
    if (landing && out_k > 0 && tspan_solns) std::memcpy(tspan_solns + (size_t)out_k * N, u0, sizeof(double) * N);
    st->CurrentStep = next;
    st->CurrentTime = t;
    st->CurrentKryDim = m;
    st->CurrentSoln = u0;
    st->NumSteps += 1;
    st->FuncEvals += fe;
    g_last = {(double)m, (double)mD, (double)rejects, err, tau_taken, beta, hn, (double)fe};
    return 0;
}

}  // extern "C"
