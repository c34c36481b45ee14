// tsr_internal.h - declarations shared by the library's host sources.
#ifndef TSR_INTERNAL_H
#define TSR_INTERNAL_H

#include <string>
#include <vector>
#include "../../include/tsr/tsr.h"

namespace tsr {

int  build_macro(int nv, int nn, const double* nodes_xy, int nbe, const int* bedges, int ne, const int* elems,
                 const int* macro_bc, macro_mesh_struct& M);
int  build_micro(double cw, double ch, int nv, int nn, const double* nodes_xy, int nbe, const int* bedges, int ne,
                 const int* elems, micro_mesh_struct& m);
void free_macro(macro_mesh_struct& M);
void free_micro(micro_mesh_struct& m);
void structured_grid(int r, double x0, double y0, double w, std::vector<double>& xy, std::vector<int>& el,
                     std::vector<int>& be, const int labels[4]);
int  read_msh(const char* path, bool micro, double& cw, double& ch, int& nv, std::vector<double>& xy,
              std::vector<int>& be, std::vector<int>& el);
int  read_keff(const char* path, Keff_struct& K);

// Small dense routines on the Krylov Hessenberg matrix (dense.cpp).
// phi_1(alpha H) of the leading m x m block of H (leading dimension ldh) into phiH
// (leading dimension ldphi); ncol > 0 writes only its first ncol columns (cheaper); flush
// evaluates with subnormals flushed to zero (faster for stiff alpha H).  Returns 1 on a
// singular Pade denominator or a non-finite alpha H.
int phi1_pade(double* phiH, const double* H, double alpha, int m, int ldh, int ldphi, int ncol = 0, bool flush = false);
int expm_pade6(double* E, const double* A, int n);
int phik_e1(double* out, const double* H, int ldh, int m, double tau, int kmax, bool flush = false);

void set_error(const std::string& s);

// Host van Genuchten-Mualem conductivity (used to tabulate Keff).
double vg_K_host(const soil_struct& s, double h);

// Problem registry: reference-signature entry points find the batch view of a
// user_data_struct by (macro_mesh.elements, N).
void registry_put(const user_data_struct& ud, const tsr_batch_view* v);
const tsr_batch_view* registry_get(const user_data_struct& ud);
void registry_drop(const tsr_batch_view* v);
// registry_get, else the reference's own problem around the caller's structs (err set
// when N fits neither hosting rule)
const tsr_batch_view* registry_find(const user_data_struct& ud, std::string& err);
// free the problem adopted for a caller's user_data (0), or 1 if none
int release_adopted(const user_data_struct& ud);
// free the solvers that Gfunc_hip / exprem made for a problem's view (solver.cpp)
void solver_cache_drop(const tsr_batch_view* v);

}  // namespace tsr

#endif
