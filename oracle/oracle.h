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

#endif
