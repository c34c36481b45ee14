// refabi.cpp - the reference's C++-linkage entry points, so a caller compiled against
// the reference's own headers links against libtsr.so without source changes.
//
// The reference declares these with C++ linkage:
//   include/main.h:21-22     Gfunc_serial, Gfunc_parallel (sources absent, compile.sh:18-21)
//   include/main.h:33-34     save_solution (src/save_solution.cpp:14-140)
//   include/exprem.h:27-30   weighted_norm, phipade, check_step, check_flag_Gfunc
// libtsr.so also exports the last four with C linkage (include/tsr/tsr.h), for C and
// ctypes callers.  C++ cannot declare a C-linkage and a C++-linkage function of the same
// name and parameters in one translation unit, so this file does not include tsr.h: it
// reaches the C-linkage implementations through asm labels.
//
// Both RHS names resolve to the device RHS: the library has one G, on the GPU, and no
// CPU path.  exprem keeps the whole step on the device for all three names
// (Gfunc_hip, Gfunc_parallel, Gfunc_serial).
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "../../include/tsr/tsr_types.h"

extern "C" int Gfunc_hip(double* g, double* u, user_data_struct user_data);
extern "C" int c_phipade(double* phiH, double* H, double alpha, int m, int MaxKrylov) __asm__("phipade");
extern "C" double c_weighted_norm(double* e, double* u, double AbsTol, double RelTol, int N) __asm__("weighted_norm");
extern "C" void c_check_step(double dt, double MinStep) __asm__("check_step");
extern "C" double c_check_flag_Gfunc(int flag, double dt) __asm__("check_flag_Gfunc");
namespace tsr {
// micro Dirichlet calibration of save_solution for cells hosted by macro elements
// (problem.cpp); returns 0, or 1 when the problem is not element-hosted
int calibrate_dirichlet(const user_data_struct& ud, double* u);
}

// include/main.h:21-22 (selected at src/main.cpp:198-213)
int Gfunc_serial(double* g, double* u, user_data_struct user_data) { return Gfunc_hip(g, u, user_data); }
int Gfunc_parallel(double* g, double* u, user_data_struct user_data) { return Gfunc_hip(g, u, user_data); }

// include/exprem.h:27-30
double weighted_norm(double* PhiError, double* u, double AbsTol, double RelTol, int N)
{
    return c_weighted_norm(PhiError, u, AbsTol, RelTol, N);
}
int phipade(double* phiH, double* H, double alpha, int m, int MaxKrylov)
{
    return c_phipade(phiH, H, alpha, m, MaxKrylov);
}
void check_step(double dt, double MinStep) { c_check_step(dt, MinStep); }
double check_flag_Gfunc(int flag, double dt) { return c_check_flag_Gfunc(flag, dt); }

// include/main.h:33-34, src/save_solution.cpp:14-140: set every stored solution's micro
// Dirichlet entries to the macro interpolant (those unknowns are not evolved; their
// heads are what G uses), then write the file: the count of stored solutions, then one
// line per solution, "t u_0 ... u_{N-1}" (t = 0 for the initial one), 20 significant
// digits in scientific notation.
void save_solution(std::string solution_file, double* tspan, double* tspan_solns, int no_solns,
                   user_data_struct user_data)
{
    const long long N = user_data.N;
    for (int j = 0; j <= no_solns; ++j) tsr::calibrate_dirichlet(user_data, tspan_solns + (size_t)j * N);
    std::ofstream f(solution_file.c_str());
    if (!f.is_open()) {
        std::fprintf(stderr, "save_solution: cannot open %s\n", solution_file.c_str());
        return;
    }
    f << no_solns + 1 << "\n" << std::scientific << std::setprecision(20);
    for (int j = 0; j <= no_solns; ++j) {
        f << (j == 0 ? 0.0 : tspan[j]);
        const double* u = tspan_solns + (size_t)j * N;
        for (long long k = 0; k < N; ++k) f << " " << u[k];
        f << "\n";
    }
}
