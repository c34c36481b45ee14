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

#include <cfloat>
#include <cmath>

namespace tsr {

__device__ __forceinline__ double wave_sum_v(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

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
    if (mode != PASS_DOTS) {
        const double* cf = coef + (size_t)s * ldc;
        for (int k = 0; k < nk; ++k) {
            const double c = cf[k];
            const double* col = cols[k];
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
    // four columns at a time: 16 independent loads in flight per thread, four
    // independent shuffle trees
    int k = 0;
    for (; k + 4 <= nk; k += 4) {
        const double* c0 = cols[k];
        const double* c1 = cols[k + 1];
        const double* c2 = cols[k + 2];
        const double* c3 = cols[k + 3];
        double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
#pragma unroll
        for (int r = 0; r < kVecR; ++r)
            if (ok[r]) {
                const long long i = idx[r];
                p0 += c0[i] * wr[r];
                p1 += c1[i] * wr[r];
                p2 += c2[i] * wr[r];
                p3 += c3[i] * wr[r];
            }
        p0 = wave_sum_v(p0);
        p1 = wave_sum_v(p1);
        p2 = wave_sum_v(p2);
        p3 = wave_sum_v(p3);
        if (lane == 0) { lds[wv][k] = p0; lds[wv][k + 1] = p1; lds[wv][k + 2] = p2; lds[wv][k + 3] = p3; }
    }
    for (; k < nk; ++k) {
        const double* col = cols[k];
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

// out = base + sgn * sum_{k < nk_s[s]} col_k * coef[s][k] - extra_sub
// partmode 0: nothing; 1: part[0][b] = sum out^2; 2: part[0][b] = sum (c/wt(u))^2 of the
// combination c = sum col_k coef_k alone (the error vector) with out = base + c.
__global__ __launch_bounds__(kVecBS) void k_lincomb(double* out, const double* base, const double* const* cols,
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
    for (int k = 0; k < nk; ++k) {
        const double c = cf[k];
        const double* col = cols[k];
#pragma unroll
        for (int r = 0; r < kVecR; ++r)
            if (ok[r]) acc[r] += col[idx[r]] * c;
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
__global__ __launch_bounds__(256) void k_reduce(int mode, const double* part, int ldp, int nk, const int* blk_first,
                                                const long long* scen_rows, KrylovState K, int j, const int* active)
{
    __shared__ double red[256 / 64];
    const int s = blockIdx.x;
    const int k = blockIdx.y;
    if (active && !active[s]) return;
    const int nsum = (mode == RED_H1 || mode == RED_H2) ? nk + 1 : nk;
    if (k >= nsum) return;
    const int b0 = blk_first[s], b1 = blk_first[s + 1];
    const double* col = part + (size_t)k * ldp;
    double x = 0.0;
    for (int b = b0 + threadIdx.x; b < b1; b += 256) x += col[b];
    x = wave_sum_v(x);
    const int tid = threadIdx.x;
    if ((tid & 63) == 0) red[tid >> 6] = x;
    __syncthreads();
    if (tid != 0) return;
    double acc = 0.0;
    for (int w = 0; w < 4; ++w) acc += red[w];
    double* cs = K.cs + (size_t)s * K.ld;
    double* coef = K.coef + (size_t)s * K.ld;
    double* H = K.H + (size_t)s * K.hstride;
    double* res = K.res + (size_t)s * 8;
    if (mode == RED_H1 || mode == RED_H2) {
        if (k == nk) {
            if (mode == RED_H1) res[1] = acc;
            return;
        }
        const double hk = cs[k] * acc;
        if (mode == RED_H1) H[(size_t)j * K.ldh + k] = hk;
        else H[(size_t)j * K.ldh + k] += hk;
        coef[k] = cs[k] * hk;
    } else if (mode == RED_NORM) {
        if (k == 0) {
            const double n = sqrt(acc);
            H[(size_t)j * K.ldh + j + 1] = n;
            cs[j + 1] = 1.0 / n;
            K.veps[s] = K.eps[s] * cs[j + 1];
            res[0] = n;
        } else if (k == 1) {
            res[2] = sqrt(acc / (double)scen_rows[s]);
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

// ---------------------------------------------------------------- launchers
hipError_t launch_pass(int mode, const double* const* cols, int nk, double* w, const double* u, const double* coef,
                       int ldc, const BlkMap& M, int nblk, double* part, int ldp, const int* active, double at,
                       double rt, hipStream_t st)
{
    hipLaunchKernelGGL(k_pass, dim3(nblk), dim3(kVecBS), 0, st, mode, cols, nk, w, u, coef, ldc, M, part, ldp, active,
                       at, rt);
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
    hipLaunchKernelGGL(k_lincomb, dim3(nblk), dim3(kVecBS), 0, st, out, base, cols, nk_s, coef, ldc, sgn, M, part, ldp,
                       partmode, u, at, rt, extra_sub, active);
    return hipGetLastError();
}

hipError_t launch_reduce(int mode, const double* part, int ldp, int nk, const int* blk_first, int n_scen,
                         const long long* scen_rows, KrylovState K, int j, const int* active, hipStream_t st)
{
    const int ncol = (mode == RED_H1 || mode == RED_H2) ? nk + 1 : nk;
    hipLaunchKernelGGL(k_reduce, dim3(n_scen, ncol), dim3(256), 0, st, mode, part, ldp, nk, blk_first, scen_rows, K, j,
                       active);
    if (mode == RED_START)
        hipLaunchKernelGGL(k_start_finish, dim3((n_scen + 255) / 256), dim3(256), 0, st, K, n_scen, active);
    return hipGetLastError();
}

hipError_t launch_masked_copy(double* dst, const double* src, const BlkMap& M, int nblk, const int* mask,
                              hipStream_t st)
{
    hipLaunchKernelGGL(k_masked_copy, dim3(nblk), dim3(kVecBS), 0, st, dst, src, M, mask);
    return hipGetLastError();
}

}  // namespace tsr
