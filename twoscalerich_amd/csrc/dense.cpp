// dense.cpp - small FP64 dense routines on the m x m Hessenberg matrix of the Krylov
// process (m <= 64): phi_1 by the reference's (6,6) Pade scheme, exp by (6,6) Pade with
// scaling and squaring, and phi_k(tau H) e1 from the exponential of an augmented matrix.
//
// phipade follows src/phipade.cpp:15-217: scale A = alpha*H by 2^-s with
// s = max(ceil(1 + log2 ||A||_inf), 0), (6,6) Pade approximant of phi_1 with the
// coefficients of Hochbruck, Lubich & Selhofer (SISC 19(5) 1998, p.1567), LU solve, then
// s doublings phi(2A) = (e^A + I) phi(A) / 2, e^{2A} = (e^A)^2.  The coefficient table
// must equal the published one.  Differences from the reference: a singular
// denominator, or a non-finite entry of alpha H, returns 1 (the reference calls exit on
// the first and carries the second through), and m = 1 with H = 0 gives phi_1 = 1 (the
// reference's (e^a - 1) a / log(e^a) / a is 0/0 there).
#include "tsr_internal.h"

#include <xmmintrin.h>

#include <cmath>
#include <cstring>
#include <vector>

namespace tsr {

namespace {

// The hot loops are built twice: for x86-64-v3 (AVX2, FMA; picked at load time where
// the CPU has it) and for baseline x86-64.
#define TSR_HOST_CLONES __attribute__((target_clones("arch=x86-64-v3", "default")))

// C <- alpha * A * B + beta * C, n x n column-major.  Each element of A B is summed over k
// in increasing order (a zero entry of B adds a zero product, which changes no finite
// sum); blocks of 8 rows x 4 columns keep their 32 sums in registers, so every loaded
// element of A serves four columns.  With jc < n only the columns of the leading blocks of
// four that cover columns 0..jc-1 are formed and updated, each by the same code as in
// the full product (so bit for bit the same values); the other columns of C are left.
typedef double v4d __attribute__((vector_size(32)));
#define TSR_LOAD4(v, p) std::memcpy(&(v), (p), sizeof(v4d))
#define TSR_STORE4(p, v) std::memcpy((p), &(v), sizeof(v4d))

TSR_HOST_CLONES void mm(int n, double alpha, const double* A, const double* B, double beta, double* C, double* tmp,
                        int jc = -1)
{
    const int n8 = n & ~7, n4 = n & ~3;
    // columns formed: whole blocks of four up to jc, or every column
    const int nc = (jc < 0 || jc >= n) ? n : std::min(n, (jc + 3) & ~3);
    const int nb4 = std::min(n4, nc);
    for (int j0 = 0; j0 < nb4; j0 += 4) {
        const double* b0 = B + (size_t)j0 * n;
        const double* b1 = b0 + n;
        const double* b2 = b1 + n;
        const double* b3 = b2 + n;
        for (int i0 = 0; i0 < n8; i0 += 8) {
            v4d c00 = {0, 0, 0, 0}, c10 = c00, c01 = c00, c11 = c00, c02 = c00, c12 = c00, c03 = c00, c13 = c00;
            for (int k = 0; k < n; ++k) {
                const double* a = A + (size_t)k * n + i0;
                v4d a0, a1;
                TSR_LOAD4(a0, a);
                TSR_LOAD4(a1, a + 4);
                const double s0 = b0[k], s1 = b1[k], s2 = b2[k], s3 = b3[k];
                c00 = c00 + a0 * s0; c10 = c10 + a1 * s0;
                c01 = c01 + a0 * s1; c11 = c11 + a1 * s1;
                c02 = c02 + a0 * s2; c12 = c12 + a1 * s2;
                c03 = c03 + a0 * s3; c13 = c13 + a1 * s3;
            }
            double* t = tmp + (size_t)j0 * n + i0;
            TSR_STORE4(t, c00); TSR_STORE4(t + 4, c10);
            TSR_STORE4(t + n, c01); TSR_STORE4(t + n + 4, c11);
            TSR_STORE4(t + 2 * n, c02); TSR_STORE4(t + 2 * n + 4, c12);
            TSR_STORE4(t + 3 * n, c03); TSR_STORE4(t + 3 * n + 4, c13);
        }
        for (int i = n8; i < n; ++i)
            for (int c = 0; c < 4; ++c) {
                const double* bc = b0 + (size_t)c * n;
                double acc = 0.0;
                for (int k = 0; k < n; ++k) acc = acc + A[i + (size_t)k * n] * bc[k];
                tmp[i + (size_t)(j0 + c) * n] = acc;
            }
    }
    for (int j = n4; j < nc; ++j) {
        const double* bj = B + (size_t)j * n;
        for (int i = 0; i < n; ++i) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc = acc + A[i + (size_t)k * n] * bj[k];
            tmp[i + (size_t)j * n] = acc;
        }
    }
    const size_t nn = (size_t)n * nc;
    if (beta == 0.0) for (size_t i = 0; i < nn; ++i) C[i] = alpha * tmp[i];
    else for (size_t i = 0; i < nn; ++i) C[i] = alpha * tmp[i] + beta * C[i];
}

double inf_norm(int n, const double* A)
{
    double r = 0.0;
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += std::fabs(A[i + (size_t)j * n]);
        if (!(s <= r)) r = s;
    }
    return r;
}

// Gaussian elimination with partial pivoting; X overwrites B (n x nrhs); R: n * nrhs
// doubles of workspace.
TSR_HOST_CLONES bool solve(int n, double* A, double* B, int nrhs, double* R)
{
    int p[256];   // n <= 2 * kMaxKrylovCap + 3 (augmented matrices)
    if (n > 256) return false;
    for (int k = 0; k < n; ++k) {
        int piv = k;
        double big = std::fabs(A[k + (size_t)k * n]);
        for (int i = k + 1; i < n; ++i) {
            const double v = std::fabs(A[i + (size_t)k * n]);
            if (v > big) { big = v; piv = i; }
        }
        if (!(big > 0.0)) return false;
        p[k] = piv;
        if (piv != k)
            for (int j = 0; j < n; ++j) std::swap(A[k + (size_t)j * n], A[piv + (size_t)j * n]);
        const double inv = 1.0 / A[k + (size_t)k * n];
        for (int i = k + 1; i < n; ++i) A[i + (size_t)k * n] *= inv;
        for (int j = k + 1; j < n; ++j) {
            const double f = A[k + (size_t)j * n];
            if (f == 0.0) continue;
            for (int i = k + 1; i < n; ++i) A[i + (size_t)j * n] -= A[i + (size_t)k * n] * f;
        }
    }
    // the right-hand sides row by row (row i of every column contiguous), so the
    // substitutions run across the columns; each element sees the same operations in the
    // same order as in a column-by-column solve
    for (int r = 0; r < nrhs; ++r)
        for (int i = 0; i < n; ++i) R[(size_t)i * nrhs + r] = B[i + (size_t)r * n];
    for (int k = 0; k < n; ++k)
        if (p[k] != k)
            for (int r = 0; r < nrhs; ++r) std::swap(R[(size_t)k * nrhs + r], R[(size_t)p[k] * nrhs + r]);
    for (int k = 0; k < n; ++k) {
        const double* rk = &R[(size_t)k * nrhs];
        for (int i = k + 1; i < n; ++i) {
            const double a = A[i + (size_t)k * n];
            double* ri = &R[(size_t)i * nrhs];
            for (int r = 0; r < nrhs; ++r) ri[r] -= a * rk[r];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        double* rk = &R[(size_t)k * nrhs];
        const double d = A[k + (size_t)k * n];
        for (int r = 0; r < nrhs; ++r) rk[r] /= d;
        for (int i = 0; i < k; ++i) {
            const double a = A[i + (size_t)k * n];
            double* ri = &R[(size_t)i * nrhs];
            for (int r = 0; r < nrhs; ++r) ri[r] -= a * rk[r];
        }
    }
    for (int r = 0; r < nrhs; ++r)
        for (int i = 0; i < n; ++i) B[i + (size_t)r * n] = R[(size_t)i * nrhs + r];
    return true;
}

// Flush-to-zero and denormals-are-zero (MXCSR bits 15 and 6) on this thread within the
// scope.  The squarings of a stiff tau*H (eigenvalues far in the left half-plane) drive
// entries of e^A into the subnormal range (below 2.2e-308), where every x86 operation on
// them takes a microcode assist; taken as zero they change only entries of that size,
// and a stiff phi_1 evaluation runs 3-7x faster.
struct FlushDenormals {
    unsigned saved = 0;
    bool on;
    explicit FlushDenormals(bool f) : on(f)
    {
        if (on) { saved = _mm_getcsr(); _mm_setcsr(saved | 0x8040u); }
    }
    ~FlushDenormals()
    {
        if (on) _mm_setcsr(saved);
    }
};

}  // namespace

int phi1_pade(double* phiH, const double* H, double alpha, int m, int ldh, int ldphi, int ncol, bool flush)
{
    const FlushDenormals ftz(flush);
    // ncol < m: only the leading columns of phi are wanted.  The Pade stage and the
    // exponential's squarings need whole matrices; the doublings of phi itself then form
    // only the leading block of four columns (one product in two of each doubling), with
    // the same operations as the whole product, so those columns are the same bit for bit.
    const int jc = (ncol > 0 && ncol < m) ? ncol : m;
    // a non-finite entry leaves the Pade denominator undefined: reported as singular
    // (the scaling exponent below would otherwise convert an infinite logarithm to int)
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i)
            if (!std::isfinite(alpha * H[i + (size_t)j * ldh])) return 1;
    if (m == 1) {
        const double a = alpha * H[0];
        phiH[0] = (a == 0.0) ? 1.0 : std::expm1(a) / a;
        return 0;
    }
    static const double num[6] = {1.0 / 26.0, 5.0 / 156.0, 1.0 / 858.0, 1.0 / 5720.0, 1.0 / 205920.0, 1.0 / 8648640.0};
    static const double den[6] = {-6.0 / 13.0, 5.0 / 52.0, -5.0 / 429.0, 1.0 / 1144.0, -1.0 / 25740.0, 1.0 / 1235520.0};
    const size_t n2 = (size_t)m * m;
    std::vector<double> buf(6 * n2);
    double *A = buf.data(), *P = A + n2, *N = P + n2, *D = N + n2, *E = D + n2, *T = E + n2;
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) A[i + (size_t)j * m] = alpha * H[i + (size_t)j * ldh];
    const double na = inf_norm(m, A);
    int s = na > 0.0 ? (int)std::ceil(1.0 + std::log(na) / std::log(2.0)) : 0;
    if (s < 0) s = 0;
    const double sc = 1.0 / std::pow(2.0, (double)s);
    for (size_t i = 0; i < n2; ++i) { A[i] *= sc; P[i] = A[i]; N[i] = 0.0; D[i] = 0.0; E[i] = 0.0; }
    for (int i = 0; i < m; ++i) { N[i + (size_t)i * m] = 1.0; D[i + (size_t)i * m] = 1.0; E[i + (size_t)i * m] = 1.0; }
    std::vector<double> tmp(n2);
    for (int k = 0; k < 6; ++k) {
        if (k > 0) { mm(m, 1.0, P, A, 0.0, T, tmp.data()); std::memcpy(P, T, n2 * sizeof(double)); }
        for (size_t i = 0; i < n2; ++i) { N[i] += num[k] * P[i]; D[i] += den[k] * P[i]; }
    }
    if (!solve(m, D, N, m, tmp.data())) return 1;
    if (s > 0) {
        mm(m, 1.0, A, N, 1.0, E, tmp.data());
        std::memcpy(T, N, n2 * sizeof(double));
        mm(m, 0.5, E, N, 0.5, T, tmp.data(), jc);
        std::memcpy(N, T, n2 * sizeof(double));
        for (int k = 1; k < s; ++k) {
            mm(m, 1.0, E, E, 0.0, T, tmp.data());
            std::memcpy(E, T, n2 * sizeof(double));
            std::memcpy(T, N, n2 * sizeof(double));
            mm(m, 0.5, E, N, 0.5, T, tmp.data(), jc);
            std::memcpy(N, T, n2 * sizeof(double));
        }
    }
    for (int j = 0; j < jc; ++j)
        for (int i = 0; i < m; ++i) phiH[i + (size_t)j * ldphi] = N[i + (size_t)j * m];
    return 0;
}

int expm_pade6(double* E, const double* A0, int n)
{
    double c[7];
    c[0] = 1.0;
    for (int k = 1; k <= 6; ++k) c[k] = c[k - 1] * (double)(7 - k) / (double)(k * (13 - k));
    const size_t n2 = (size_t)n * n;
    std::vector<double> buf(5 * n2);
    double *A = buf.data(), *P = A + n2, *N = P + n2, *D = N + n2, *T = D + n2;
    std::vector<double> tmp(n2);
    std::memcpy(A, A0, n2 * sizeof(double));
    const double na = inf_norm(n, A);
    int s = 0;
    if (na > 0.5) s = (int)std::ceil(std::log2(na / 0.5));
    const double sc = std::ldexp(1.0, -s);
    for (size_t i = 0; i < n2; ++i) { A[i] *= sc; P[i] = A[i]; N[i] = 0.0; D[i] = 0.0; }
    for (int i = 0; i < n; ++i) { N[i + (size_t)i * n] = 1.0; D[i + (size_t)i * n] = 1.0; }
    for (int k = 1; k <= 6; ++k) {
        if (k > 1) { mm(n, 1.0, P, A, 0.0, T, tmp.data()); std::memcpy(P, T, n2 * sizeof(double)); }
        const double sg = (k & 1) ? -1.0 : 1.0;
        for (size_t i = 0; i < n2; ++i) { N[i] += c[k] * P[i]; D[i] += sg * c[k] * P[i]; }
    }
    if (!solve(n, D, N, n, tmp.data())) return 1;
    for (int k = 0; k < s; ++k) { mm(n, 1.0, N, N, 0.0, T, tmp.data()); std::memcpy(N, T, n2 * sizeof(double)); }
    std::memcpy(E, N, n2 * sizeof(double));
    return 0;
}

// out[(k-1)*m + i] = [phi_k(tau H) e1]_i for k = 1..kmax (Saad 1992; Sidje 1998, Thm 1).
int phik_e1(double* out, const double* H, int ldh, int m, double tau, int kmax, bool flush)
{
    const FlushDenormals ftz(flush);
    const int n = m + kmax;
    std::vector<double> A((size_t)n * n, 0.0), E((size_t)n * n);
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) A[i + (size_t)j * n] = tau * H[i + (size_t)j * ldh];
    A[(size_t)m * n] = 1.0;
    for (int k = 1; k < kmax; ++k) A[(m + k - 1) + (size_t)(m + k) * n] = 1.0;
    if (expm_pade6(E.data(), A.data(), n)) return 1;
    for (int k = 1; k <= kmax; ++k)
        for (int i = 0; i < m; ++i) out[(size_t)(k - 1) * m + i] = E[i + (size_t)(m + k - 1) * n];
    return 0;
}

}  // namespace tsr

extern "C" int phipade(double* phiH, double* H, double alpha, int m, int MaxKrylov)
{
    return tsr::phi1_pade(phiH, H, alpha, m, MaxKrylov + 1, MaxKrylov);
}

extern "C" int tsr_phi1_e1(double* out, const double* H, double alpha, int m, int ldh, int flush_denormals)
{
    if (m < 1 || ldh < m) return 1;
    return tsr::phi1_pade(out, H, alpha, m, ldh, m, 1, flush_denormals != 0);
}

extern "C" int tsr_phik_e1(double* out, const double* H, int ldh, int m, double tau, int kmax, int flush_denormals)
{
    if (m < 1 || ldh < m || kmax < 1) return 1;
    return tsr::phik_e1(out, H, ldh, m, tau, kmax, flush_denormals != 0);
}
