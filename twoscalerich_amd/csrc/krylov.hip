// krylov.hip - bandwidth-bound FP64 vector kernels of the Arnoldi process (CGS2) and
// the EXPREM update, segmented by scenario so one launch serves a whole batch.
//
// Every kernel streams rows; workgroup b owns rows [r0[b], r1[b]) of one scenario
// (kVecR rows per thread, strided by the workgroup size so each load instruction is
// one contiguous 2 KiB segment).  Reductions are two-stage and deterministic: per
// workgroup a wavefront shuffle tree and a fixed-order sum over its 4 wavefronts into
// part[k * nblk + b] (column-major, so the second stage reads contiguously), then
// k_reduce gives each (scenario, column) its own workgroup, which sums that column's
// partials with a fixed strided assignment and a fixed tree.  No FP64 atomics
// anywhere, so every rerun is bitwise identical.
//
// Basis columns are stored unnormalised; v_k = cs[k] * raw_k.  The scale is folded
// into the coefficients by k_reduce, so no pass ever rescales a column.
#include "device.h"
#include "wave.h"

#include <cfloat>
#include <cmath>
#include <mutex>

namespace tsr {

// wave_sum_v: the DPP wavefront sum of device.h (wave_sum_dpp)
__device__ __forceinline__ double wave_sum_v(double x) { return wave_sum_dpp(x); }

// Sum a per-thread value over the workgroup into lds slot; result valid for thread 0.
__device__ __forceinline__ double block_sum(double x, double* red)
{
    x = wave_sum_v(x);
    const int tid = threadIdx.x;
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = x;
    __syncthreads();
    double s = 0.0;
    if (tid == 0)
        for (int w = 0; w < kVecBS / 64; ++w) s += red[w];
    return s;
}

// Four basis columns k..k+3 of a CGS2 pass for a thread's kVecR rows, every load issued
// before any is used (with FULL, every row of the block exists): pass_axpy_group
// w -= a coef column by column, pass_dot_group p[c] += a w.
template <bool FULL>
__device__ __forceinline__ void pass_axpy_group(const double* const* cols, int k, const long long* idx, const bool* ok,
                                                const double* cf, double* wr)
{
    double av[kVecR][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const gptr<double> col = as_global(cols[k + c]);
#pragma unroll
        for (int r = 0; r < kVecR; ++r) av[r][c] = (FULL || ok[r]) ? __builtin_nontemporal_load(&col[idx[r]]) : 0.0;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < kVecR; ++r)
            if (FULL || ok[r]) wr[r] -= av[r][c] * cf[k + c];
}

template <bool FULL>
__device__ __forceinline__ void pass_dot_group(const double* const* cols, int k, const long long* idx, const bool* ok,
                                               const double* wr, double* pp)
{
    double av[kVecR][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const gptr<double> col = as_global(cols[k + c]);
#pragma unroll
        for (int r = 0; r < kVecR; ++r) av[r][c] = (FULL || ok[r]) ? __builtin_nontemporal_load(&col[idx[r]]) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < kVecR; ++r)
        if (FULL || ok[r]) {
#pragma unroll
            for (int c = 0; c < 4; ++c) pp[c] += av[r][c] * wr[r];
        }
}

// part[k][b] below is part[k * ldp + b] (ldp = number of workgroups).
// mode PASS_DOTS:       part[k][b] = sum_rows col_k * w                     (k < nk), part[nk][b] = sum w^2
// mode PASS_AXPY_DOTS:  w -= sum_k col_k * coef[s][k]; then as PASS_DOTS (w written back)
// mode PASS_AXPY_NORM:  w -= sum_k col_k * coef[s][k]; part[0][b] = sum w^2, part[1][b] = sum (w/wt)^2,
//                       wt = at + rt |u|
__global__ __launch_bounds__(kVecBS) void k_pass(int mode, const double* const* cols, int nk, double* w,
                                                const double* u, const double* coef, int ldc, BlkMap M,
                                                double* part, int ldp, const int* active, double at, double rt)
{
    __shared__ double lds[kVecBS / 64][kMaxCols + 1];
    __shared__ double red[kVecBS / 64];
    const int b = blockIdx.x;
    const int s = M.scen[b];
    if (active && !active[s]) return;
    const long long r0 = M.r0[b], r1 = M.r1[b];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double wr[kVecR];
    long long idx[kVecR];
    bool ok[kVecR];
#pragma unroll
    for (int r = 0; r < kVecR; ++r) {
        idx[r] = r0 + tid + (long long)r * kVecBS;
        ok[r] = idx[r] < r1;
        wr[r] = ok[r] ? w[idx[r]] : 0.0;
    }
    const bool full = r1 - r0 == (long long)kVecBS * kVecR;   // every row of the block exists
    if (mode != PASS_DOTS) {
        const double* cf = coef + (size_t)s * ldc;
        int k = 0;
        for (; k + 4 <= nk; k += 4) {   // four columns at a time, the same order
            if (full) pass_axpy_group<true>(cols, k, idx, ok, cf, wr);
            else pass_axpy_group<false>(cols, k, idx, ok, cf, wr);
        }
        for (; k < nk; ++k) {
            const double c = cf[k];
            const gptr<double> col = as_global(cols[k]);
#pragma unroll
            for (int r = 0; r < kVecR; ++r)
                if (ok[r]) wr[r] -= col[idx[r]] * c;
        }
#pragma unroll
        for (int r = 0; r < kVecR; ++r)
            if (ok[r]) w[idx[r]] = wr[r];
    }
    if (mode == PASS_AXPY_NORM) {
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int r = 0; r < kVecR; ++r)
            if (ok[r]) {
                s0 += wr[r] * wr[r];
                const double q = wr[r] / (at + rt * fabs(u[idx[r]]));
                s1 += q * q;
            }
        s0 = block_sum(s0, red);
        s1 = block_sum(s1, red);
        if (tid == 0) { part[b] = s0; part[(size_t)ldp + b] = s1; }
        return;
    }
    // four columns at a time: every load of the group in flight together, four
    // independent shuffle trees
    int k = 0;
    for (; k + 4 <= nk; k += 4) {
        double pp[4] = {0.0, 0.0, 0.0, 0.0};
        if (full) pass_dot_group<true>(cols, k, idx, ok, wr, pp);
        else pass_dot_group<false>(cols, k, idx, ok, wr, pp);
        double p0 = pp[0], p1 = pp[1], p2 = pp[2], p3 = pp[3];
        p0 = wave_sum_v(p0);
        p1 = wave_sum_v(p1);
        p2 = wave_sum_v(p2);
        p3 = wave_sum_v(p3);
        if (lane == 0) { lds[wv][k] = p0; lds[wv][k + 1] = p1; lds[wv][k + 2] = p2; lds[wv][k + 3] = p3; }
    }
    for (; k < nk; ++k) {
        const gptr<double> col = as_global(cols[k]);
        double p = 0.0;
#pragma unroll
        for (int r = 0; r < kVecR; ++r)
            if (ok[r]) p += col[idx[r]] * wr[r];
        p = wave_sum_v(p);
        if (lane == 0) lds[wv][k] = p;
    }
    {
        double p = 0.0;
#pragma unroll
        for (int r = 0; r < kVecR; ++r) p += wr[r] * wr[r];
        p = wave_sum_v(p);
        if (lane == 0) lds[wv][nk] = p;
    }
    __syncthreads();
    for (int kk = tid; kk <= nk; kk += kVecBS) {
        double acc = 0.0;
        for (int x = 0; x < kVecBS / 64; ++x) acc += lds[x][kk];
        part[(size_t)kk * ldp + b] = acc;
    }
}

// ---------------------------------------------------------------- DCGS2
// Four basis columns k..k+3 of a DCGS2 pass, for a thread's kVecR rows: every load of the
// group is issued before any is used (with FULL, every row of the block exists, so no
// row test stands between the loads), then the per-row arithmetic in the passes' fixed
// order.  The passes stream the basis, w, R_j and u with non-temporal loads (and the
// update pass writes w and R_j with non-temporal stores): each is touched once per
// pass, and keeping them out of the caches' working set measured 5-10 % faster passes
// and FD products.  dca_group: column sums sx[c] += a w, px[c] += a R_j; dcb_group: w -= a coef,
// R_j -= a coefr, column by column.
template <bool FULL, class T>
__device__ __forceinline__ void dca_group(const T* const* cols, int k, const long long* idx, const bool* ok,
                                          const double* wr, const double* rj, double* sx, double* px)
{
    double av[kVecR][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const gptr<T> col = as_global(cols[k + c]);
#pragma unroll
        for (int r = 0; r < kVecR; ++r) av[r][c] = (FULL || ok[r]) ? (double)__builtin_nontemporal_load(&col[idx[r]]) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < kVecR; ++r)
        if (FULL || ok[r]) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                sx[c] += av[r][c] * wr[r];
                px[c] += av[r][c] * rj[r];
            }
        }
}

template <bool FULL, class T>
__device__ __forceinline__ void dcb_group(const T* const* cols, int k, const long long* idx, const bool* ok,
                                          const double* cf, const double* cr, double* wr, double* rn)
{
    double av[kVecR][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const gptr<T> col = as_global(cols[k + c]);
#pragma unroll
        for (int r = 0; r < kVecR; ++r) av[r][c] = (FULL || ok[r]) ? (double)__builtin_nontemporal_load(&col[idx[r]]) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < kVecR; ++r)
        if (FULL || ok[r]) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                wr[r] -= av[r][c] * cf[k + c];
                rn[r] -= av[r][c] * cr[k + c];
            }
        }
}

// Classical Gram-Schmidt with the reorthogonalisation of each column delayed by one
// iteration (the "DCGS2" scheme of Swirydowicz, Langou, Ananthan, Yang and Thomas,
// Numer. Linear Algebra Appl. 2021), so one Arnoldi iteration reads the basis twice
// instead of three times.  At iteration j the raw columns R_0..R_{j-1} are final and
// R_j is the once-projected previous vector with provisional scale cs_j.
//
// k_pass_dca, per row (nk = j + 1 columns):
//   part[k][b]        = sum R_k w      (k <= j)
//   part[nk + k][b]   = sum R_k R_j    (k < j)
//   part[2 nk - 1][b] = sum R_j R_j
//   part[2 nk][b]     = sum w w
// Columns k < j are read as T (FP64, or the FP32 copies of TSR_PATH_F32BASIS); R_j, the
// once-projected vector of the previous iteration, is always read in FP64.
template <class T>
__global__ __launch_bounds__(kVecBS) void k_pass_dca(const T* const* cols, const double* Rj, int nk, const double* w,
                                                    BlkMap M, double* part, int ldp, const int* active)
{
    __shared__ double lds[kVecBS / 64][2 * kMaxCols + 2];
    const int b = blockIdx.x;
    const int s = M.scen[b];
    if (active && !active[s]) return;
    const long long r0 = M.r0[b], r1 = M.r1[b];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int j = nk - 1;
    double wr[kVecR], rj[kVecR];
    long long idx[kVecR];
    bool ok[kVecR];
#pragma unroll
    for (int r = 0; r < kVecR; ++r) {
        idx[r] = r0 + tid + (long long)r * kVecBS;
        ok[r] = idx[r] < r1;
        wr[r] = ok[r] ? __builtin_nontemporal_load(&w[idx[r]]) : 0.0;
        rj[r] = ok[r] ? __builtin_nontemporal_load(&Rj[idx[r]]) : 0.0;
    }
    int k = 0;
    const bool full = r1 - r0 == (long long)kVecBS * kVecR;   // every row of the block exists
    for (; k + 4 <= j; k += 4) {   // four columns at a time
        double sx[4] = {0.0, 0.0, 0.0, 0.0}, px[4] = {0.0, 0.0, 0.0, 0.0};
        if (full) dca_group<true>(cols, k, idx, ok, wr, rj, sx, px);
        else dca_group<false>(cols, k, idx, ok, wr, rj, sx, px);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            sx[c] = wave_sum_v(sx[c]);
            px[c] = wave_sum_v(px[c]);
        }
        if (lane == 0) {
#pragma unroll
            for (int c = 0; c < 4; ++c) { lds[wv][k + c] = sx[c]; lds[wv][nk + k + c] = px[c]; }
        }
    }
    for (; k < j; ++k) {
        const gptr<T> c0 = as_global(cols[k]);
        double s0 = 0.0, p0 = 0.0;
#pragma unroll
        for (int r = 0; r < kVecR; ++r)
            if (ok[r]) { const double a0 = (double)c0[idx[r]]; s0 += a0 * wr[r]; p0 += a0 * rj[r]; }
        s0 = wave_sum_v(s0); p0 = wave_sum_v(p0);
        if (lane == 0) { lds[wv][k] = s0; lds[wv][nk + k] = p0; }
    }
    {
        double sj = 0.0, rr = 0.0, ww = 0.0;
#pragma unroll
        for (int r = 0; r < kVecR; ++r) { sj += rj[r] * wr[r]; rr += rj[r] * rj[r]; ww += wr[r] * wr[r]; }
        sj = wave_sum_v(sj); rr = wave_sum_v(rr); ww = wave_sum_v(ww);
        if (lane == 0) { lds[wv][j] = sj; lds[wv][2 * nk - 1] = rr; lds[wv][2 * nk] = ww; }
    }
    __syncthreads();
    for (int kk = tid; kk <= 2 * nk; kk += kVecBS) {
        double acc = 0.0;
        for (int x = 0; x < kVecBS / 64; ++x) acc += lds[x][kk];
        part[(size_t)kk * ldp + b] = acc;
    }
}

// Coefficients of iteration j from the reduced sums d (layout of k_pass_dca), one
// thread per scenario.  With a_k = <v_k, v~_j> (v~_j = cs_j R_j) and
// gamma = ||v~_j - sum a_k v_k||:
//   v_j = (v~_j - sum_k a_k v_k) / gamma          (R_j -= sum coefr_k R_k, cs_j /= gamma)
//   column j-1 of H: H[k][j-1] += a_k H[j][j-1], H[j][j-1] *= gamma
//   column j:  h_k = cs_k d_k (k < j), h_j = (cs_j d_j - sum a_k h_k) / gamma
//   w' = w - sum_{k<j} (h_k - h_j a_k / gamma) cs_k R_k - (h_j cs_j / gamma) R_j
// res[1] and nwj[j] receive ||w||^2 (before projection).
__global__ void k_dc_finish(KrylovState K, int j, int n_scen, const int* active)
{
    // one workgroup per scenario: the inputs are staged in LDS by all threads, then
    // thread 0 does the (order-dependent, O(j)) arithmetic
    __shared__ double sd[2 * kMaxCols + 2], scs[kMaxCols + 1];
    const int s = blockIdx.x;
    if (active && !active[s]) return;
    const int nk = j + 1;
    for (int i = threadIdx.x; i <= 2 * nk; i += blockDim.x) sd[i] = K.dots[(size_t)s * K.ld2 + i];
    for (int i = threadIdx.x; i <= j; i += blockDim.x) scs[i] = K.cs[(size_t)s * K.ld + i];
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double* d = sd;
    double* coef = K.coef + (size_t)s * K.ld;
    double* coefr = K.coefr + (size_t)s * K.ld;
    double* H = K.H + (size_t)s * K.hstride;
    double* hj_col = H + (size_t)j * K.ldh;
    K.res[(size_t)s * 8 + 1] = d[2 * nk];
    K.nwj[(size_t)s * K.ld + j] = d[2 * nk];
    if (j == 0) {
        const double h0 = scs[0] * d[0];
        hj_col[0] = h0;
        coef[0] = scs[0] * h0;
        return;
    }
    __shared__ double sa[kMaxCols], sh[kMaxCols];   // a_k, h_k
    const double csj = scs[j];
    double suma2 = 0.0;
    for (int k = 0; k < j; ++k) {
        const double a = scs[k] * csj * d[nk + k];
        sa[k] = a;
        suma2 += a * a;
    }
    const double g2 = csj * csj * d[2 * nk - 1] - suma2;
    const double gam = (csj > 0.0 && g2 > 0.0) ? sqrt(g2) : 1.0;
    double* hp = H + (size_t)(j - 1) * K.ldh;
    const double hprev = hp[j];
    for (int k = 0; k < j; ++k) hp[k] += sa[k] * hprev;
    hp[j] = hprev * gam;
    double ah = 0.0;
    for (int k = 0; k < j; ++k) {
        const double hk = scs[k] * d[k];
        sh[k] = hk;
        hj_col[k] = hk;
        ah += sa[k] * hk;
    }
    const double hjj = (csj * d[j] - ah) / gam;
    hj_col[j] = hjj;
    for (int k = 0; k < j; ++k) {
        coef[k] = (sh[k] - hjj * sa[k] / gam) * scs[k];
        coefr[k] = csj > 0.0 ? sa[k] * scs[k] / csj : 0.0;
    }
    coef[j] = hjj * csj / gam;
    K.cs[(size_t)s * K.ld + j] = csj / gam;
}

// w -= sum_{k<=j} coef_k R_k;  R_j -= sum_{k<j} coefr_k R_k (both from the old R_j);
// part[0][b] = sum w'^2, part[1][b] = sum (w'/wt(u))^2.  The old R_j is read from Rj_in
// (FP64); the corrected one goes to Rj_out (for j > 0, or for every j when all_j: the
// FP32 copy of column 0 is written here too).
template <class T>
__global__ __launch_bounds__(kVecBS) void k_pass_dcb(const T* const* cols, const double* Rj_in, T* Rj_out, int all_j,
                                                    int nk, double* w, const double* u, const double* coef,
                                                    const double* coefr, int ldc, BlkMap M, double* part, int ldp,
                                                    const int* active, double at, double rt)
{
    __shared__ double red[kVecBS / 64];
    // rows in the reverse order of k_pass_dca's: the blocks that start first find the
    // rows k_pass_dca read last still in the memory-side cache (each block's rows and
    // partials are its own, so the order changes no result)
    const int b = gridDim.x - 1 - blockIdx.x;
    const int s = M.scen[b];
    if (active && !active[s]) return;
    const long long r0 = M.r0[b], r1 = M.r1[b];
    const int tid = threadIdx.x;
    const int j = nk - 1;
    const double* cf = coef + (size_t)s * ldc;
    const double* cr = coefr + (size_t)s * ldc;
    double wr[kVecR], rj[kVecR], rn[kVecR];
    long long idx[kVecR];
    bool ok[kVecR];
    const double cj = cf[j];
#pragma unroll
    for (int r = 0; r < kVecR; ++r) {
        idx[r] = r0 + tid + (long long)r * kVecBS;
        ok[r] = idx[r] < r1;
        rj[r] = ok[r] ? __builtin_nontemporal_load(&Rj_in[idx[r]]) : 0.0;
        wr[r] = ok[r] ? __builtin_nontemporal_load(&w[idx[r]]) - cj * rj[r] : 0.0;
        rn[r] = rj[r];
    }
    int k = 0;
    const bool full = r1 - r0 == (long long)kVecBS * kVecR;   // every row of the block exists
    for (; k + 4 <= j; k += 4) {   // four columns at a time
        if (full) dcb_group<true>(cols, k, idx, ok, cf, cr, wr, rn);
        else dcb_group<false>(cols, k, idx, ok, cf, cr, wr, rn);
    }
    for (; k < j; ++k) {
        const double c = cf[k], c2 = cr[k];
        const gptr<T> col = as_global(cols[k]);
#pragma unroll
        for (int r = 0; r < kVecR; ++r)
            if (ok[r]) {
                const double a = (double)col[idx[r]];
                wr[r] -= a * c;
                rn[r] -= a * c2;
            }
    }
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int r = 0; r < kVecR; ++r)
        if (ok[r]) {
            __builtin_nontemporal_store(wr[r], &w[idx[r]]);
            if (all_j || j > 0) __builtin_nontemporal_store((T)rn[r], &Rj_out[idx[r]]);
            s0 += wr[r] * wr[r];
            const double q = wr[r] / (at + rt * fabs(__builtin_nontemporal_load(&u[idx[r]])));
            s1 += q * q;
        }
    s0 = block_sum(s0, red);
    s1 = block_sum(s1, red);
    if (tid == 0) { part[b] = s0; part[(size_t)ldp + b] = s1; }
}

// part[0..2][b] = sum a^2, sum (a/wt(u))^2, sum u^2
__global__ __launch_bounds__(kVecBS) void k_norms(const double* a, const double* u, BlkMap M, double* part, int ldp,
                                                 const int* active, double at, double rt)
{
    __shared__ double red[kVecBS / 64];
    const int b = blockIdx.x;
    const int s = M.scen[b];
    if (active && !active[s]) return;
    const long long r0 = M.r0[b], r1 = M.r1[b];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int r = 0; r < kVecR; ++r) {
        const long long i = r0 + threadIdx.x + (long long)r * kVecBS;
        if (i < r1) {
            const double x = a[i], y = u[i];
            s0 += x * x;
            const double q = x / (at + rt * fabs(y));
            s1 += q * q;
            s2 += y * y;
        }
    }
    s0 = block_sum(s0, red);
    s1 = block_sum(s1, red);
    s2 = block_sum(s2, red);
    if (threadIdx.x == 0) { part[b] = s0; part[(size_t)ldp + b] = s1; part[(size_t)2 * ldp + b] = s2; }
}

// Four basis columns k..k+3 at a thread's kVecR rows, every load issued before any is
// used (with FULL, every row of the block exists; rows past the end read as 0).
template <bool FULL, class T>
__device__ __forceinline__ void load_group(const T* const* cols, int k, const long long* idx, const bool* ok,
                                           double av[kVecR][4])
{
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const gptr<T> col = as_global(cols[k + c]);
#pragma unroll
        for (int r = 0; r < kVecR; ++r) av[r][c] = (FULL || ok[r]) ? (double)__builtin_nontemporal_load(&col[idx[r]]) : 0.0;
    }
}

// out = base + sgn * sum_{k < nk_s[s]} col_k * coef[s][k] - extra_sub
// partmode 0: nothing; 1: part[0][b] = sum out^2; 2: part[0][b] = sum (c/wt(u))^2 of the
// combination c = sum col_k coef_k alone (the error vector) with out = base + c.
template <class T>
__global__ __launch_bounds__(kVecBS) void k_lincomb(double* out, const double* base, const T* const* cols,
                                                   const int* nk_s, const double* coef, int ldc, double sgn, BlkMap M,
                                                   double* part, int ldp, int partmode, const double* u, double at,
                                                   double rt, const double* extra_sub, const int* active)
{
    __shared__ double red[kVecBS / 64];
    const int b = blockIdx.x;
    const int s = M.scen[b];
    if (active && !active[s]) return;
    const long long r0 = M.r0[b], r1 = M.r1[b];
    const int nk = nk_s[s];
    const double* cf = coef + (size_t)s * ldc;
    double acc[kVecR];
    long long idx[kVecR];
    bool ok[kVecR];
#pragma unroll
    for (int r = 0; r < kVecR; ++r) {
        idx[r] = r0 + threadIdx.x + (long long)r * kVecBS;
        ok[r] = idx[r] < r1;
        acc[r] = 0.0;
    }
    const bool full = r1 - r0 == (long long)kVecBS * kVecR;   // every row of the block exists
    int k = 0;
    for (; k + 4 <= nk; k += 4) {   // four columns at a time, the same order
        double av[kVecR][4];
        if (full) load_group<true>(cols, k, idx, ok, av);
        else load_group<false>(cols, k, idx, ok, av);
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int r = 0; r < kVecR; ++r)
                if (ok[r]) acc[r] += av[r][c] * cf[k + c];
    }
    for (; k < nk; ++k) {
        const double c = cf[k];
        const gptr<T> col = as_global(cols[k]);
#pragma unroll
        for (int r = 0; r < kVecR; ++r)
            if (ok[r]) acc[r] += (double)col[idx[r]] * c;
    }
    double s0 = 0.0;
#pragma unroll
    for (int r = 0; r < kVecR; ++r) {
        if (!ok[r]) continue;
        double o = base ? base[idx[r]] : 0.0;
        o += sgn * acc[r];
        if (extra_sub) o -= extra_sub[idx[r]];
        out[idx[r]] = o;
        if (partmode == 1) s0 += o * o;
        else if (partmode == 2) { const double q = acc[r] / (at + rt * fabs(u[idx[r]])); s0 += q * q; }
    }
    if (partmode) {
        s0 = block_sum(s0, red);
        if (threadIdx.x == 0) part[b] = s0;
    }
}

// out1 = base1 + sum_{k < nk1} coef1_k col_k and out2 = sum_{k < nk2} coef2_k col_k in one
// pass over the columns: the two combinations of an EXPREM step that share the basis
// (U2 and the Arnoldi-relation term of D).  Each sum runs over increasing k as in
// k_lincomb, so both outputs equal two separate k_lincomb calls bit for bit.
template <class T>
__global__ __launch_bounds__(kVecBS) void k_lincomb2(double* out1, const double* base1, double* out2,
                                                    const T* const* cols, const int* nk1_s, const double* coef1,
                                                    const int* nk2_s, const double* coef2, int ldc, BlkMap M,
                                                    const int* active)
{
    const int b = blockIdx.x;
    const int s = M.scen[b];
    if (active && !active[s]) return;
    const long long r0 = M.r0[b], r1 = M.r1[b];
    const int nk1 = nk1_s[s], nk2 = nk2_s[s];
    const int nk = nk1 > nk2 ? nk1 : nk2;
    const double* cf1 = coef1 + (size_t)s * ldc;
    const double* cf2 = coef2 + (size_t)s * ldc;
    double a1[kVecR], a2[kVecR];
    long long idx[kVecR];
    bool ok[kVecR];
#pragma unroll
    for (int r = 0; r < kVecR; ++r) {
        idx[r] = r0 + threadIdx.x + (long long)r * kVecBS;
        ok[r] = idx[r] < r1;
        a1[r] = 0.0;
        a2[r] = 0.0;
    }
    const bool full = r1 - r0 == (long long)kVecBS * kVecR;   // every row of the block exists
    int k = 0;
    for (; k + 4 <= nk; k += 4) {   // four columns at a time, the same order
        double av[kVecR][4];
        if (full) load_group<true>(cols, k, idx, ok, av);
        else load_group<false>(cols, k, idx, ok, av);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bool u1 = k + c < nk1, u2 = k + c < nk2;
            const double c1 = u1 ? cf1[k + c] : 0.0, c2 = u2 ? cf2[k + c] : 0.0;
#pragma unroll
            for (int r = 0; r < kVecR; ++r)
                if (ok[r]) {
                    if (u1) a1[r] += av[r][c] * c1;
                    if (u2) a2[r] += av[r][c] * c2;
                }
        }
    }
    for (; k < nk; ++k) {
        const gptr<T> col = as_global(cols[k]);
        const bool u1 = k < nk1, u2 = k < nk2;
        const double c1 = u1 ? cf1[k] : 0.0, c2 = u2 ? cf2[k] : 0.0;
#pragma unroll
        for (int r = 0; r < kVecR; ++r)
            if (ok[r]) {
                const double x = (double)col[idx[r]];
                if (u1) a1[r] += x * c1;
                if (u2) a2[r] += x * c2;
            }
    }
#pragma unroll
    for (int r = 0; r < kVecR; ++r) {
        if (!ok[r]) continue;
        double o = base1[idx[r]];
        o += a1[r];
        out1[idx[r]] = o;
        out2[idx[r]] = a2[r];
    }
}

// FP32 copy of a column's rows (scenarios with mask set)
__global__ __launch_bounds__(kVecBS) void k_to_f32(float* dst, const double* src, BlkMap M, const int* mask)
{
    const int b = blockIdx.x;
    const int s = M.scen[b];
    if (mask && !mask[s]) return;
    const long long r0 = M.r0[b], r1 = M.r1[b];
#pragma unroll
    for (int r = 0; r < kVecR; ++r) {
        const long long i = r0 + threadIdx.x + (long long)r * kVecBS;
        if (i < r1) dst[i] = (float)src[i];
    }
}

__global__ __launch_bounds__(kVecBS) void k_masked_copy(double* dst, const double* src, BlkMap M, const int* mask)
{
    const int b = blockIdx.x;
    const int s = M.scen[b];
    if (mask && !mask[s]) return;
    const long long r0 = M.r0[b], r1 = M.r1[b];
#pragma unroll
    for (int r = 0; r < kVecR; ++r) {
        const long long i = r0 + threadIdx.x + (long long)r * kVecBS;
        if (i < r1) dst[i] = src[i];
    }
}

// Second stage: workgroup (s, k) sums column k of scenario s's workgroup partials
// (thread t takes partials t, t+256, ... in order, then a fixed tree), then finishes
// the Gram-Schmidt bookkeeping for that column:
//   RED_H1:   h1_k = cs_k sum_k; H[k][j] = h1_k; coef_k = cs_k h1_k; (k = nk) res[1] = ||w||^2 before CGS
//   RED_H2:   h2_k = cs_k sum_k; H[k][j] += h2_k; coef_k = cs_k h2_k
//   RED_NORM: (k = 0) n = sqrt(sum_0); H[j+1][j] = n; cs[j+1] = 1/n; veps = eps cs[j+1]; res[0] = n;
//             (k = 1) res[2] = sqrt(sum_1 / N_s) (WRMS of w; divided by n on the host)
//   RED_START (column 0 = start vector): (k = 0) beta = sqrt(sum_0); cs[0] = 1/beta; res[0] = beta;
//             (k = 2) res[3] = ||u||_2; eps = sqrt(DBL_EPSILON)(1 + ||u||_2) unless j < 0; ieps = 1/eps;
//             then k_start_finish sets veps = eps cs[0]
//   RED_SUM:  res[k] = sum_k
constexpr int kRedBS = 1024;   // threads of a k_reduce workgroup

__global__ __launch_bounds__(kRedBS) void k_reduce(int mode, const double* part, int ldp, int nk, const int* blk_first,
                                                const long long* scen_rows, KrylovState K, int j, const int* active)
{
    __shared__ double red[kRedBS / 64];
    const int s = blockIdx.x;
    const int k = blockIdx.y;
    if (active && !active[s]) return;
    const int nsum = (mode == RED_H1 || mode == RED_H2) ? nk + 1 : nk;
    if (k >= nsum) return;
    const int b0 = blk_first[s], b1 = blk_first[s + 1];
    const double* col = part + (size_t)k * ldp;
    // thread t takes partials t, t + kRedBS, ...: eight running sums (partials 8i + r go
    // to sum r), so eight loads are in flight per thread, combined in a fixed tree
    double xs[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int b = b0 + threadIdx.x;
    for (; b + 7 * kRedBS < b1; b += 8 * kRedBS) {
#pragma unroll
        for (int r = 0; r < 8; ++r) xs[r] += col[b + r * kRedBS];
    }
    for (int r = 0; b < b1; b += kRedBS, ++r) xs[r] += col[b];
    double x = ((xs[0] + xs[1]) + (xs[2] + xs[3])) + ((xs[4] + xs[5]) + (xs[6] + xs[7]));
    x = wave_sum_v(x);
    const int tid = threadIdx.x;
    if ((tid & 63) == 0) red[tid >> 6] = x;
    __syncthreads();
    if (tid != 0) return;
    double acc = 0.0;
    for (int w = 0; w < kRedBS / 64; ++w) acc += red[w];
    double* cs = K.cs + (size_t)s * K.ld;
    double* coef = K.coef + (size_t)s * K.ld;
    double* H = K.H + (size_t)s * K.hstride;
    double* res = K.res + (size_t)s * 8;
    if (mode == RED_COLS) {
        K.dots[(size_t)s * K.ld2 + k] = acc;
        return;
    }
    if (mode == RED_H1 || mode == RED_H2) {
        if (k == nk) {
            if (mode == RED_H1) { res[1] = acc; K.nwj[(size_t)s * K.ld + j] = acc; }
            return;
        }
        const double hk = cs[k] * acc;
        if (mode == RED_H1) H[(size_t)j * K.ldh + k] = hk;
        else H[(size_t)j * K.ldh + k] += hk;
        coef[k] = cs[k] * hk;
    } else if (mode == RED_NORM) {
        if (k == 0) {
            // breakdown (the host's test): the next column is zero and so is its FD step
            const double n = sqrt(acc);
            const bool bd = !(n > 1e-12 * sqrt(K.nwj[(size_t)s * K.ld + j]));
            H[(size_t)j * K.ldh + j + 1] = n;
            cs[j + 1] = bd ? 0.0 : 1.0 / n;
            K.veps[s] = K.eps[s] * cs[j + 1];
            res[0] = n;
        } else if (k == 1) {
            res[2] = sqrt(acc / (double)scen_rows[s]);
            K.wrj[(size_t)s * K.ld + j] = res[2];
        }
    } else if (mode == RED_START) {
        if (k == 0) {
            const double beta = sqrt(acc);
            cs[0] = 1.0 / beta;
            res[0] = beta;
        } else if (k == 2) {
            res[3] = sqrt(acc);
            if (j >= 0) {
                K.eps[s] = sqrt(DBL_EPSILON) * (1.0 + res[3]);
                K.ieps[s] = 1.0 / K.eps[s];
            }
        }
    } else {
        res[k] = acc;
    }
}

// veps = eps cs[0] once both RED_START columns are done (one thread per scenario)
__global__ void k_start_finish(KrylovState K, int n_scen, const int* active)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_scen || (active && !active[s])) return;
    K.veps[s] = K.eps[s] * K.cs[(size_t)s * K.ld];
}

// Per-scenario state summary for the step controller: out[s*6 + 0..5] = mean, min, max of
// the macro heads and of the matrix soil's effective saturation
// Se(h) = (1 + (alpha |h|)^n)^(-m) (1 for h >= 0).  One 1024-thread workgroup per
// scenario; the sums run in a fixed order (thread-strided, then the wavefront tree, then
// the sixteen wavefronts in order).
constexpr int kStatsBS = 1024;
__global__ __launch_bounds__(kStatsBS) void k_macro_stats(const DScen* scen, const double* u, double* out)
{
    __shared__ double red[6][kStatsBS / 64];
    const int s = blockIdx.x, tid = threadIdx.x;
    const DScen& S = scen[s];
    const double* h = u + S.u_off;
    double sh = 0.0, mnh = INFINITY, mxh = -INFINITY, ss = 0.0, mns = INFINITY, mxs = -INFINITY;
    for (int j = tid; j < S.nn; j += kStatsBS) {
        const double hj = h[j];
        const double se = hj < 0.0 ? exp(-S.A.m * log1p(exp(S.A.n * log(S.A.al * (-hj))))) : 1.0;
        sh += hj; mnh = fmin(mnh, hj); mxh = fmax(mxh, hj);
        ss += se; mns = fmin(mns, se); mxs = fmax(mxs, se);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sh += __shfl_xor(sh, o, 64);
        ss += __shfl_xor(ss, o, 64);
        mnh = fmin(mnh, __shfl_xor(mnh, o, 64));
        mxh = fmax(mxh, __shfl_xor(mxh, o, 64));
        mns = fmin(mns, __shfl_xor(mns, o, 64));
        mxs = fmax(mxs, __shfl_xor(mxs, o, 64));
    }
    if ((tid & 63) == 0) {
        const int w = tid >> 6;
        red[0][w] = sh; red[1][w] = mnh; red[2][w] = mxh; red[3][w] = ss; red[4][w] = mns; red[5][w] = mxs;
    }
    __syncthreads();
    if (tid == 0) {
        double a0 = 0.0, a3 = 0.0, a1 = INFINITY, a2 = -INFINITY, a4 = INFINITY, a5 = -INFINITY;
        for (int w = 0; w < kStatsBS / 64; ++w) {
            a0 += red[0][w]; a3 += red[3][w];
            a1 = fmin(a1, red[1][w]); a2 = fmax(a2, red[2][w]);
            a4 = fmin(a4, red[4][w]); a5 = fmax(a5, red[5][w]);
        }
        double* o = out + (size_t)s * 6;
        const double n = (double)S.nn;
        o[0] = a0 / n; o[1] = a1; o[2] = a2; o[3] = a3 / n; o[4] = a4; o[5] = a5;
    }
}

// Readbacks in one launch.  The host syncs on the stream a few times per Arnoldi
// window; instead of one copy per array, one kernel writes everything the host needs
// into a mapped pinned host buffer (the host reads it after the sync).
//
// Window (scenario s, at out + s * stride): Hessenberg columns c0..jt-1, rows 0..col+1
// each, packed in column order; cs[0..jt]; nwj[0..jt-1]; wrj[0..jt-1]; bad[s].
__global__ void k_pack_window(KrylovState K, int c0, int jt, int* bad, const int* active, double* out,
                              int stride)
{
    const int s = blockIdx.x;
    if (active && !active[s]) return;
    double* o = out + (size_t)s * stride;
    const double* H = K.H + (size_t)s * K.hstride;
    int base = 0;
    for (int col = c0; col < jt; ++col) {
        for (int i = threadIdx.x; i <= col + 1; i += blockDim.x) o[base + i] = H[(size_t)col * K.ldh + i];
        base += col + 2;
    }
    const double* cs = K.cs + (size_t)s * K.ld;
    for (int i = threadIdx.x; i <= jt; i += blockDim.x) o[base + i] = cs[i];
    base += jt + 1;
    const double* nw = K.nwj + (size_t)s * K.ld;
    const double* wr = K.wrj + (size_t)s * K.ld;
    for (int i = threadIdx.x; i < jt; i += blockDim.x) {
        o[base + i] = nw[i];
        o[base + jt + i] = wr[i];
    }
    base += 2 * jt;
    if (threadIdx.x == 0) {
        o[base] = (double)bad[s];
        bad[s] = 0;   // read and cleared: the next window starts clean
    }
}

// Per-scenario scalars (scenario s at out + s * 16): res[0..7], bad, eps, stats[0..5]
// (eps and stats only when given).
__global__ void k_pack_small(const double* res, int* bad, const double* eps, const double* stats, const float* ctl,
                             int n_scen, double* out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_scen) return;
    double* o = out + (size_t)s * kPackSmall;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = res[(size_t)s * 8 + k];
    o[8] = (double)bad[s];
    bad[s] = 0;
    if (eps) o[9] = eps[s];
    if (stats) {
#pragma unroll
        for (int k = 0; k < 6; ++k) o[10 + k] = stats[(size_t)s * 6 + k];
    }
    if (ctl) {
        o[16] = ctl[2 * s];
        o[17] = ctl[2 * s + 1];
    }
}

// ---------------------------------------------------------------- launchers
hipError_t launch_pack_window(const KrylovState& K, int c0, int jt, int* bad, const int* active, int n_scen,
                              double* out, int stride, hipStream_t st)
{
    hipLaunchKernelGGL(k_pack_window, dim3(n_scen), dim3(64), 0, st, K, c0, jt, bad, active, out, stride);
    return hipGetLastError();
}

hipError_t launch_pack_small(const double* res, int* bad, const double* eps, const double* stats, const float* ctl,
                             int n_scen, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_pack_small, dim3((n_scen + 63) / 64), dim3(64), 0, st, res, bad, eps, stats, ctl, n_scen, out);
    return hipGetLastError();
}

hipError_t launch_macro_stats(const DScen* scen, int n_scen, const double* u, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_macro_stats, dim3(n_scen), dim3(kStatsBS), 0, st, scen, u, out);
    return hipGetLastError();
}

hipError_t launch_pass(int mode, const double* const* cols, int nk, double* w, const double* u, const double* coef,
                       int ldc, const BlkMap& M, int nblk, double* part, int ldp, const int* active, double at,
                       double rt, hipStream_t st)
{
    // algorithmic bytes: nk columns and w read; w written back by the two axpy passes;
    // u read by the last
    const double rows = (double)nblk * kVecBS * kVecR * (active ? kt_active_fraction() : 1.0);
    const double bytes = 8.0 * rows * (nk + 1 + (mode != PASS_DOTS ? 1 : 0) + (mode == PASS_AXPY_NORM ? 1 : 0));
    const bool kt = kt_begin(KT_PASS, bytes, st);
    hipLaunchKernelGGL(k_pass, dim3(nblk), dim3(kVecBS), 0, st, mode, cols, nk, w, u, coef, ldc, M, part, ldp, active,
                       at, rt);
    if (kt) kt_end(st);
    return hipGetLastError();
}

hipError_t launch_norms(const double* a, const double* u, const BlkMap& M, int nblk, double* part, int ldp,
                        const int* active, double at, double rt, hipStream_t st)
{
    hipLaunchKernelGGL(k_norms, dim3(nblk), dim3(kVecBS), 0, st, a, u, M, part, ldp, active, at, rt);
    return hipGetLastError();
}

hipError_t launch_lincomb(double* out, const double* base, const double* const* cols, const int* nk_s,
                          const double* coef, int ldc, double sgn, const BlkMap& M, int nblk, double* part, int ldp,
                          int partmode, const double* u, double at, double rt, const double* extra_sub,
                          const int* active, hipStream_t st)
{
    hipLaunchKernelGGL(k_lincomb<double>, dim3(nblk), dim3(kVecBS), 0, st, out, base, cols, nk_s, coef, ldc, sgn, M,
                       part, ldp, partmode, u, at, rt, extra_sub, active);
    return hipGetLastError();
}

hipError_t launch_lincomb(double* out, const double* base, const float* const* cols, const int* nk_s,
                          const double* coef, int ldc, double sgn, const BlkMap& M, int nblk, double* part, int ldp,
                          int partmode, const double* u, double at, double rt, const double* extra_sub,
                          const int* active, hipStream_t st)
{
    hipLaunchKernelGGL(k_lincomb<float>, dim3(nblk), dim3(kVecBS), 0, st, out, base, cols, nk_s, coef, ldc, sgn, M,
                       part, ldp, partmode, u, at, rt, extra_sub, active);
    return hipGetLastError();
}

hipError_t launch_lincomb2(double* out1, const double* base1, double* out2, const double* const* cols,
                           const int* nk1_s, const double* coef1, const int* nk2_s, const double* coef2, int ldc,
                           const BlkMap& M, int nblk, const int* active, hipStream_t st)
{
    hipLaunchKernelGGL(k_lincomb2<double>, dim3(nblk), dim3(kVecBS), 0, st, out1, base1, out2, cols, nk1_s, coef1, nk2_s,
                       coef2, ldc, M, active);
    return hipGetLastError();
}

hipError_t launch_lincomb2(double* out1, const double* base1, double* out2, const float* const* cols,
                           const int* nk1_s, const double* coef1, const int* nk2_s, const double* coef2, int ldc,
                           const BlkMap& M, int nblk, const int* active, hipStream_t st)
{
    hipLaunchKernelGGL(k_lincomb2<float>, dim3(nblk), dim3(kVecBS), 0, st, out1, base1, out2, cols, nk1_s, coef1, nk2_s,
                       coef2, ldc, M, active);
    return hipGetLastError();
}

hipError_t launch_reduce(int mode, const double* part, int ldp, int nk, const int* blk_first, int n_scen,
                         const long long* scen_rows, KrylovState K, int j, const int* active, hipStream_t st)
{
    const int ncol = (mode == RED_H1 || mode == RED_H2) ? nk + 1 : nk;
    hipLaunchKernelGGL(k_reduce, dim3(n_scen, ncol), dim3(kRedBS), 0, st, mode, part, ldp, nk, blk_first, scen_rows, K, j,
                       active);
    if (mode == RED_START)
        hipLaunchKernelGGL(k_start_finish, dim3((n_scen + 255) / 256), dim3(256), 0, st, K, n_scen, active);
    return hipGetLastError();
}

template <class T>
static hipError_t dcgs2(int nk, const T* const* cols, const double* Rj_in, T* Rj_out, int all_j, double* w,
                        const double* u, KrylovState K, const BlkMap& M, int nblk, double* part, int ldp,
                        const int* blk_first, int n_scen, const long long* scen_rows, const int* active, double at,
                        double rt, hipStream_t st)
{
    const int j = nk - 1;
    // algorithmic bytes per row: k_pass_dca reads columns k < j (as T), R_j and w;
    // k_pass_dcb reads them and u, writes w and (j > 0 or all_j) the corrected R_j (as T)
    const double rows = (double)nblk * kVecBS * kVecR * (active ? kt_active_fraction() : 1.0);
    const double bca = rows * (sizeof(T) * (double)j + 16.0);
    const double bcb = rows * (sizeof(T) * (double)j + 32.0 + ((all_j || j > 0) ? sizeof(T) : 0));
    bool kt = kt_begin(KT_DCA, bca, st);
    hipLaunchKernelGGL(k_pass_dca<T>, dim3(nblk), dim3(kVecBS), 0, st, cols, Rj_in, nk, w, M, part, ldp, active);
    if (kt) kt_end(st);
    hipLaunchKernelGGL(k_reduce, dim3(n_scen, 2 * nk + 1), dim3(kRedBS), 0, st, (int)RED_COLS, part, ldp, 2 * nk + 1,
                       blk_first, scen_rows, K, j, active);
    hipLaunchKernelGGL(k_dc_finish, dim3(n_scen), dim3(64), 0, st, K, j, n_scen, active);
    kt = kt_begin(KT_DCB, bcb, st);
    hipLaunchKernelGGL(k_pass_dcb<T>, dim3(nblk), dim3(kVecBS), 0, st, cols, Rj_in, Rj_out, all_j, nk, w, u, K.coef,
                       K.coefr, K.ld, M, part, ldp, active, at, rt);
    if (kt) kt_end(st);
    hipLaunchKernelGGL(k_reduce, dim3(n_scen, 2), dim3(kRedBS), 0, st, (int)RED_NORM, part, ldp, 2, blk_first, scen_rows, K,
                       j, active);
    return hipGetLastError();
}

hipError_t launch_dcgs2(int nk, const double* const* cols, double* Rj, double* w, const double* u, KrylovState K,
                        const BlkMap& M, int nblk, double* part, int ldp, const int* blk_first, int n_scen,
                        const long long* scen_rows, const int* active, double at, double rt, hipStream_t st)
{
    return dcgs2<double>(nk, cols, Rj, Rj, 0, w, u, K, M, nblk, part, ldp, blk_first, n_scen, scen_rows, active, at, rt,
                         st);
}

hipError_t launch_dcgs2_f32(int nk, const float* const* cols32, const double* Rj, float* Rj32, double* w,
                            const double* u, KrylovState K, const BlkMap& M, int nblk, double* part, int ldp,
                            const int* blk_first, int n_scen, const long long* scen_rows, const int* active, double at,
                            double rt, hipStream_t st)
{
    return dcgs2<float>(nk, cols32, Rj, Rj32, 1, w, u, K, M, nblk, part, ldp, blk_first, n_scen, scen_rows, active, at,
                        rt, st);
}

hipError_t launch_to_f32(float* dst, const double* src, const BlkMap& M, int nblk, const int* mask, hipStream_t st)
{
    hipLaunchKernelGGL(k_to_f32, dim3(nblk), dim3(kVecBS), 0, st, dst, src, M, mask);
    return hipGetLastError();
}

// ---------------------------------------------------------------- allocation-free scratch
// The FD steps of tsr_solver_jv_noalloc: one array per GPU, held until the process
// ends, shared by every call on that GPU (calls that overlap on one GPU race).  Host
// code reaches it only through the address hipGetSymbolAddress gives for each device.
static __device__ __align__(256) double scratch[2 * kNoAllocMaxScen];

double* noalloc_scratch()
{
    constexpr int kMaxDev = 64;
    static std::mutex mu;
    static double* addr[kMaxDev] = {};
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        addr[dev] = static_cast<double*>(p);
    }
    return addr[dev];
}

__global__ void k_fill_eps(double* de, int n_scen, double eps)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_scen) return;
    de[q] = eps;
    de[n_scen + q] = 1.0 / eps;
}

hipError_t launch_fill_eps(double* de, int n_scen, double eps, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_eps, dim3((n_scen + 255) / 256), dim3(256), 0, st, de, n_scen, eps);
    return hipGetLastError();
}

hipError_t launch_masked_copy(double* dst, const double* src, const BlkMap& M, int nblk, const int* mask,
                              hipStream_t st)
{
    hipLaunchKernelGGL(k_masked_copy, dim3(nblk), dim3(kVecBS), 0, st, dst, src, M, mask);
    return hipGetLastError();
}

}  // namespace tsr
