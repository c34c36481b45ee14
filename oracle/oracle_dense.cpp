// oracle_dense.cpp - small dense FP64 kit of the CPU oracle (TEST INFRASTRUCTURE ONLY).
//
// oracle_phipade restates the reference's phipade (src/phipade.cpp:15-217): the
// (6,6) Pade approximant of phi_1 (coefficients from Hochbruck, Lubich & Selhofer,
// SISC 19(5) 1998, p.1567, as cited at phipade.cpp:59-60) after scaling by 2^-s with
// s = max(ceil(1 + log2 ||A||_inf), 0) (:99-110), an LU solve (:154-167), and s
// doublings phi(2A) = (e^A + I) phi(A) / 2, e^{2A} = (e^A)^2 (:170-195).  Column-major,
// H with leading dimension MaxKrylov+1, phiH with MaxKrylov (:93, :201).  The m = 1
// case returns expm1(a)/a (and 1 at a = 0, where the reference's formula gives 0/0).
// The reference's BLAS/LAPACK calls are replaced by the textbook loops below; its
// exit() on a singular denominator becomes a return code of 1.
//
// oracle_expm is the (6,6) Pade exponential with scaling and squaring (Moler & Van
// Loan, SIAM Review 45 (2003), "method 3"); oracle_phik_e1 evaluates phi_1..phi_k(tau H)
// e1 from the exponential of the augmented matrix [[tau H, e1, 0],[0, 0, I],[0, 0, 0]]
// (Saad, SINUM 29 (1992); Sidje, ACM TOMS 24 (1998), Theorem 1).
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

// C = alpha*A*B + beta*C, n x n column-major
void gemm(int n, double alpha, const double* A, const double* B, double beta, double* C)
{
    std::vector<double> T((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) {
            const double b = B[k + (size_t)j * n];
            if (b == 0.0) continue;
            for (int i = 0; i < n; ++i) T[i + (size_t)j * n] += A[i + (size_t)k * n] * b;
        }
    for (size_t i = 0; i < (size_t)n * n; ++i) C[i] = alpha * T[i] + (beta == 0.0 ? 0.0 : beta * C[i]);
}

double norm_inf(int n, const double* A)
{
    double best = 0.0;
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += std::fabs(A[i + (size_t)j * n]);
        if (s > best || std::isnan(s)) best = s;
    }
    return best;
}

// Solve A X = B in place (B <- X), LU with partial pivoting.  Returns 0 or 1 (singular).
int lu_solve(int n, double* A, double* B, int nrhs)
{
    std::vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::fabs(A[k + (size_t)k * n]);
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[i + (size_t)k * n]) > best) { best = std::fabs(A[i + (size_t)k * n]); p = i; }
        if (best == 0.0) return 1;
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(A[k + (size_t)j * n], A[p + (size_t)j * n]);
        const double d = A[k + (size_t)k * n];
        for (int i = k + 1; i < n; ++i) A[i + (size_t)k * n] /= d;
        for (int j = k + 1; j < n; ++j) {
            const double a = A[k + (size_t)j * n];
            if (a == 0.0) continue;
            for (int i = k + 1; i < n; ++i) A[i + (size_t)j * n] -= A[i + (size_t)k * n] * a;
        }
    }
    for (int r = 0; r < nrhs; ++r) {
        double* b = B + (size_t)r * n;
        for (int k = 0; k < n; ++k) if (piv[k] != k) std::swap(b[k], b[piv[k]]);
        for (int k = 0; k < n; ++k) { const double x = b[k]; for (int i = k + 1; i < n; ++i) b[i] -= A[i + (size_t)k * n] * x; }
        for (int k = n - 1; k >= 0; --k) {
            b[k] /= A[k + (size_t)k * n];
            const double x = b[k];
            for (int i = 0; i < k; ++i) b[i] -= A[i + (size_t)k * n] * x;
        }
    }
    return 0;
}

}  // namespace

extern "C" {

int oracle_phipade(double* phiH, const double* H, double alpha, int m, int MaxKrylov)
{
    if (m == 1) {
        const double a = alpha * H[0];
        phiH[0] = (a == 0.0) ? 1.0 : std::expm1(a) / a;
        return 0;
    }
    static const double an[6] = {1.0 / 26.0, 5.0 / 156.0, 1.0 / 858.0, 1.0 / 5720.0, 1.0 / 205920.0, 1.0 / 8648640.0};
    static const double bd[6] = {-6.0 / 13.0, 5.0 / 52.0, -5.0 / 429.0, 1.0 / 1144.0, -1.0 / 25740.0, 1.0 / 1235520.0};
    const size_t mm = (size_t)m * m;
    std::vector<double> A(mm), P(mm), Num(mm, 0.0), Den(mm, 0.0), E(mm, 0.0), T(mm);
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) A[i + (size_t)j * m] = alpha * H[i + (size_t)j * (MaxKrylov + 1)];
    const double nA = norm_inf(m, A.data());
    int s = (nA > 0.0) ? (int)std::ceil(1.0 + std::log(nA) / std::log(2.0)) : 0;
    if (s < 0) s = 0;
    const double sc = 1.0 / std::pow(2.0, (double)s);
    for (auto& x : A) x *= sc;
    for (int i = 0; i < m; ++i) { Num[i + (size_t)i * m] = 1.0; Den[i + (size_t)i * m] = 1.0; E[i + (size_t)i * m] = 1.0; }
    P = A;
    for (int k = 0; k < 6; ++k) {
        if (k > 0) { gemm(m, 1.0, P.data(), A.data(), 0.0, T.data()); P = T; }
        for (size_t i = 0; i < mm; ++i) { Num[i] += an[k] * P[i]; Den[i] += bd[k] * P[i]; }
    }
    if (lu_solve(m, Den.data(), Num.data(), m)) return 1;   // Num <- phi_1(A)
    if (s > 0) {
        gemm(m, 1.0, A.data(), Num.data(), 1.0, E.data());   // E = A*phi + I = e^A
        T = Num;
        gemm(m, 0.5, E.data(), Num.data(), 0.5, T.data());
        Num = T;
        for (int k = 1; k < s; ++k) {
            gemm(m, 1.0, E.data(), E.data(), 0.0, T.data());
            E = T;
            T = Num;
            gemm(m, 0.5, E.data(), Num.data(), 0.5, T.data());
            Num = T;
        }
    }
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) phiH[i + (size_t)j * MaxKrylov] = Num[i + (size_t)j * m];
    return 0;
}

int oracle_expm(double* E, const double* A0, int n)
{
    // Pade (6,6): c_k = (12-k)! 6! / (12! k! (6-k)!)
    double c[7];
    c[0] = 1.0;
    for (int k = 1; k <= 6; ++k) c[k] = c[k - 1] * (double)(6 - k + 1) / (double)(k * (12 - k + 1));
    const size_t nn = (size_t)n * n;
    std::vector<double> A(A0, A0 + nn), P(nn), Nm(nn, 0.0), Dm(nn, 0.0), T(nn);
    const double nA = norm_inf(n, A.data());
    int s = 0;
    if (nA > 0.5) s = (int)std::ceil(std::log2(nA / 0.5));
    const double sc = std::ldexp(1.0, -s);
    for (auto& x : A) x *= sc;
    for (int i = 0; i < n; ++i) { Nm[i + (size_t)i * n] = c[0]; Dm[i + (size_t)i * n] = c[0]; }
    P = A;
    for (int k = 1; k <= 6; ++k) {
        if (k > 1) { gemm(n, 1.0, P.data(), A.data(), 0.0, T.data()); P = T; }
        const double sg = (k & 1) ? -1.0 : 1.0;
        for (size_t i = 0; i < nn; ++i) { Nm[i] += c[k] * P[i]; Dm[i] += sg * c[k] * P[i]; }
    }
    if (lu_solve(n, Dm.data(), Nm.data(), n)) return 1;
    for (int k = 0; k < s; ++k) { gemm(n, 1.0, Nm.data(), Nm.data(), 0.0, T.data()); Nm = T; }
    std::memcpy(E, Nm.data(), nn * sizeof(double));
    return 0;
}

// out[(k-1)*m + i] = [phi_k(tau*H) e1]_i, k = 1..kmax; H column-major with leading dim ldh.
int oracle_phik_e1(double* out, const double* H, int ldh, int m, double tau, int kmax)
{
    const int n = m + kmax;
    std::vector<double> A((size_t)n * n, 0.0), E((size_t)n * n);
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) A[i + (size_t)j * n] = tau * H[i + (size_t)j * ldh];
    A[0 + (size_t)m * n] = 1.0;
    for (int k = 1; k < kmax; ++k) A[(m + k - 1) + (size_t)(m + k) * n] = 1.0;
    if (oracle_expm(E.data(), A.data(), n)) return 1;
    for (int k = 1; k <= kmax; ++k)
        for (int i = 0; i < m; ++i) out[(size_t)(k - 1) * m + i] = E[i + (size_t)(m + k - 1) * n];
    return 0;
}

// sqrt( (1/N) sum (e_i / (AbsTol + RelTol |u_i|))^2 )   (include/exprem.h:27)
double oracle_weighted_norm(const double* e, const double* u, double AbsTol, double RelTol, long long N)
{
    const double s = oracle_fixed_sum(N, [&](long long i) {
        const double r = e[i] / (AbsTol + RelTol * std::fabs(u[i]));
        return r * r;
    });
    return std::sqrt(s / (double)N);
}

}  // extern "C"
