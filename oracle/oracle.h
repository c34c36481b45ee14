// oracle.h - C interface of the CPU oracle (TEST INFRASTRUCTURE ONLY).
//
// Loaded by tests/, the smoke entry and the bench's cpu_baseline leg through ctypes;
// never linked into or called by the product library.
#ifndef TSR_ORACLE_H
#define TSR_ORACLE_H

#include <cstdio>
#include "../include/tsr/tsr_types.h"

extern "C" {

// ---- RHS (oracle_model.cpp) ----
void* oracle_prepare(const tsr_batch_view* view);
void  oracle_release(void* h);
int   oracle_gfunc(void* h, double* g, const double* u, int nthreads);
void  oracle_water_balance(void* h, const double* u, const double* g, double* water, double* inflow,
                           double* stor_g);
void  oracle_vg(const soil_struct* s, const double* h, int n, double Ss, double* C, double* K, double* theta);
void  oracle_keff(const Keff_struct* T, const double* h, int n, double* K4);
void  oracle_register(const user_data_struct* ud, void* handle);
int   Gfunc_serial(double* g, double* u, user_data_struct user_data);
int   Gfunc_parallel(double* g, double* u, user_data_struct user_data);

// ---- dense kit and phi functions (oracle_dense.cpp) ----
int    oracle_phipade(double* phiH, const double* H, double alpha, int m, int MaxKrylov);
int    oracle_expm(double* E, const double* A, int n);
int    oracle_phik_e1(double* out, const double* H, int ldh, int m, double tau, int kmax);
double oracle_weighted_norm(const double* e, const double* u, double AbsTol, double RelTol, long long N);

// ---- integrator (oracle_exprem.cpp) ----
int oracle_exprem(int (*Gfunc)(double*, double*, user_data_struct), int N, double* t, double tend,
                  double* tspan, double* tspan_solns, double* u0, user_data_struct user_data,
                  exprem_options_struct options, exprem_stats_struct* stats);
void oracle_exprem_last(double* out8);
}

#include <vector>

// sum_{i < n} f(i) in an order fixed by n alone: 4096-term chunks summed in parallel, the
// chunk sums added in chunk order, so the result does not depend on the thread count or
// on which thread finishes first (an OpenMP reduction clause combines in arrival order).
template <class F>
inline double oracle_fixed_sum(long long n, F f)
{
    const long long kChunk = 4096;
    const long long nc = (n + kChunk - 1) / kChunk;
    std::vector<double> part((size_t)(nc > 0 ? nc : 1), 0.0);
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < nc; ++c) {
        double s = 0.0;
        const long long e = (c + 1) * kChunk < n ? (c + 1) * kChunk : n;
        for (long long i = c * kChunk; i < e; ++i) s += f(i);
        part[(size_t)c] = s;
    }
    double s = 0.0;
    for (long long c = 0; c < nc; ++c) s += part[(size_t)c];
    return s;
}

#endif
