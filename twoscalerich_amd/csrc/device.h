// device.h - device-side problem layout shared by the HIP sources.
//
// HBM layout: one allocation per problem holding every template's geometry (SoA,
// the reference's element-major order), every scenario's macro mesh and Keff table,
// the per-cell tables, and G's scratch (macro interpolant coefficients, macro
// sub-face fluxes, BC half-edge inflows, per-cell exchange).  The unknown vector
// itself is [scenario 0: macro nodes | cells ...][scenario 1: ...] in FP64.
#ifndef TSR_DEVICE_H
#define TSR_DEVICE_H

#include <hip/hip_runtime.h>
#include "../../include/tsr/tsr_types.h"

namespace tsr {

constexpr int kMicroMaxBS = 1024;     // micro-cell workgroup size limit
constexpr int kMaxNodeInc = 12;       // macro elements around a node hosting a cell
constexpr int kVecBS = 256;           // threads per workgroup of the vector kernels
constexpr int kVecR = 2;              // rows per thread of the vector kernels
constexpr int kMaxCols = 128;         // Krylov basis columns a pass can read

struct DSoil { double Ks, tr, ts, al, n, m; };

struct DTmpl {
    int nn, ne;
    const int* ptr;                   // [nn+1] node -> incident (e*4 + lv), e ascending
    const int* inc;
    const int* el;                    // [3*ne]
    const double* sf;                 // [9*ne] shape-function coefficients (rows 0..5 used)
    const double* nrm;                // [6*ne]
    const double* w;                  // [9*ne] (j*3+v)*ne + e: nodal-K weight of vertex v at sub-face j
    const double* scv;                // [ne] SCV area of each vertex (|det|/6)
    const double* vec;                // [2*nn]
    const double* cen;                // [2*ne]
    const int* bnd;                   // [nn]
    const unsigned char* soil;        // [ne] or nullptr
    double cell_area;
    // Uniform r x r right-triangle grid (the generator's layout: node j*r+i, square
    // (i,j) split into A = (a, a+1, a+r+1) and B = (a, a+r+1, a+r), element 2*(j*(r-1)+i)
    // + {0,1}, Dirichlet nodes = perimeter): geometry of one A and one B element serves
    // every element.  sg[t] = sf x-coeffs (3), y-coeffs (3), normals (6), sub-face
    // K weights (9, j*3+v), SCV area (1).
    int structured;
    int r;
    double dx;                        // grid spacing (structured templates)
    unsigned mul_r, mul_rm;           // n / r == __umulhi(n, mul_r), n / (r-1) == __umulhi(n, mul_rm), n < 2^16
    int nbnd;                         // Dirichlet nodes (generic templates: listed in bnd_list)
    const int* bnd_list;              // [nbnd] (generic templates)
    double sg[2][22];
};

struct DScen {
    DSoil A, B;
    double epsA, dirichlet_head, boundary_flux;
    double keff_logr, keff_first;
    double box[4];                    // bounding box of the macro nodes (xmin, xmax, ymin, ymax)
    int model, host, nv, nn, ne, n_cells, first_cell, keff_n;
    long long u_off;
    const double* keff;               // [4*n]
    const int* el;
    const double* nodes;
    const double* sf;
    const double* nrm;
    const double* midx;
    const double* midy;
    const double* cv;
    const double* scv;
    const int* bce;
    const double* be_len;
    const double* be_mid;
    const double* be_nrm;
    const double* cen;
    const int* ptr;
    const int* inc;
    const unsigned char* dir;
    double* qe;                       // [nv*ne] scratch
    double* bq;                       // [2*nv*ne] scratch
};

struct DBatch {
    int n_scen, n_tmpl, n_cells, max_discs;
    int nn_max, ne_max, macro_nn_max, macro_ne_max;
    const DScen* scen;
    const DTmpl* tmpl;
    const int* cell_tmpl;
    const int* cell_scen;
    const long long* cell_off;
    // cell_stride > 0: every cell c starts at cell_off0 + c * cell_stride (one scenario,
    // one template), so kernels can place a cell without loading its table entries
    long long cell_off0;
    int cell_stride;
    const double* discs;
    double* Q;                        // [n_cells]
    int* bad;                         // [n_scen] set when G produced a non-finite value
    // node-hosted cells: for micro Dirichlet node i of cell c, entry c * nn_max + i is the
    // index (into the host node's incident macro elements) of the element whose P1
    // interpolant gives the node's head; fixed by the geometry, filled once at upload
    unsigned char* dir_elem;
    // element soils of every cell (bit e of the cell's soil_words words: 1 = soil A),
    // from the template or the cell's inclusion discs; fixed, filled once at upload
    unsigned* soil_bits;
    int soil_words;
    // node-hosted cells, per incident macro element k < kMaxNodeInc of the host node:
    // shape-function coefficients (cm_sf[(c * kMaxNodeInc + k) * 9 + j], the macro
    // mesh's shape_funcs of that element) and the state index of each vertex's head
    // (cm_vi[(c * kMaxNodeInc + k) * 3 + v], -1 at Dirichlet macro nodes); cm_host[c]
    // the host node's state index (-1: Dirichlet), cm_ninc[c] its number of elements.
    // Geometry only, filled once at upload, so a cell's set-up is one load deep.
    double* cm_sf;
    int* cm_vi;
    int* cm_host;
    unsigned char* cm_ninc;
    double Ss;
    long long N;
};

// Per-scenario perturbation / finite-difference arguments of one G launch:
//   state read  = u + veps[s] * v        (v may be null)
//   value written = (G - g0) * ieps[s]   (g0 may be null)
//
// Linearised Jacobian products: a G(u) launch with rec set also writes, per micro node
// (SoA over the global index): FP64 components 0 the head used, 1 the node's storage
// capacity sum(scv C), 2 K of soil A, 3 K of soil B (rec[c * N + i]); then, from
// rec + 4 N viewed as floats, FP32 components 0 d(storage)/dh, 1 dK_A/dh, 2 dK_B/dh
// ((float*)(rec + 4 N))[c * N + i]; K and dK/dh are 0 where the soil is absent.  A launch
// with lin set (an FD product around that same u) takes the micro nodes' capacities and
// conductivities from the first-order expansion X(h) = X0 + X0' (h - h0) instead of the
// van Genuchten functions.  X0' multiplies h - h0 = eps v, so its FP32 rounding (2^-24
// relative) changes the product by that factor of a term of order eps; X0 itself, which
// the product differences against G(u), stays FP64.
constexpr int kLinComp64 = 4;
constexpr int kLinComp32 = 3;
constexpr size_t kLinBytesPerNode = 8 * kLinComp64 + 4 * kLinComp32;
struct GArgs {
    const double* u;
    const double* v;
    const double* veps;
    const double* g0;
    const double* ieps;
    const int* active;                // scenario mask (nullptr = all)
    double* out;
    double* rec = nullptr;            // record the expansion point (G(u) launches)
    const double* lin = nullptr;      // use the recorded expansion (FD launches)
};

struct DevProblem;
DevProblem* upload_problem(const tsr_batch_view& v, std::string& err);
void        free_problem(DevProblem* p);
const DBatch& batch_of(const DevProblem* p);
size_t      micro_lds_bytes(const DevProblem* p);
// tune = false: a micro-cell kernel not tuned yet runs with a fixed admissible shape
// instead of timing the candidates (no events are created)
hipError_t  launch_gfunc(const DevProblem* p, const GArgs& a, hipStream_t st, bool tune = true);
// one of the three G kernels: 0 macro elements, 1 micro cells, 2 macro nodes
hipError_t  launch_gfunc_part(const DevProblem* p, const GArgs& a, int part, hipStream_t st, bool tune = true);
// the linearised mode needs micro templates of at most 1024 nodes
bool        lin_mode_ok(const DevProblem* p);
// the micro-cell kernel of FD products (the linearised mode when lin) is tuned already
bool        gfunc_tuned(const DevProblem* p, bool lin);
// tuned micro-cell launch shape (bs = 0 before the first G)
void        micro_shape(const DevProblem* p, int* bs, int* wpe);

// Sampled device timing of kernel launches (the solver's measurement mode, HIP events
// on the launch's stream).  Launchers bracket the kernels of these classes with
// kt_begin / kt_end; kt_begin returns false, and records nothing, unless timing is on
// for the calling thread and this launch of the class is sampled.  `bytes` is the
// launch's algorithmic HBM traffic with every scenario active.
enum KtClass {
    KT_MICRO_FULL = 0,   // k_micro modes 0/1 (G, or an FD product of the full RHS)
    KT_MICRO_LIN = 1,    // k_micro mode 2 (linearised FD product)
    KT_MACRO_NODE = 2,   // k_macro_node
    KT_DCA = 3,          // k_pass_dca (DCGS2 projections)
    KT_DCB = 4,          // k_pass_dcb (DCGS2 update)
    KT_PASS = 5,         // k_pass (CGS2 passes of the plain path)
    KT_COUNT = 6
};
bool kt_begin(int cls, double bytes, hipStream_t st);
// share of the unknowns the current scenario mask keeps active (1 without a mask): the
// launchers scale a masked launch's bytes by it
double kt_active_fraction();
void kt_end(hipStream_t st);

// Row blocks of the vector kernels: block b covers rows [r0, r1) of scenario scen.
struct BlkMap {
    const int* scen;
    const long long* r0;
    const long long* r1;
};

enum PassMode { PASS_DOTS = 0, PASS_AXPY_DOTS = 1, PASS_AXPY_NORM = 2 };
enum RedMode { RED_H1 = 0, RED_H2 = 1, RED_NORM = 2, RED_START = 3, RED_SUM = 4, RED_COLS = 5 };

struct KrylovState {              // per-scenario device arrays, stride ld
    double* cs;                   // column scales: v_k = cs[k] * raw column k
    double* coef;                 // coefficients of the current axpy
    double* H;                    // Hessenberg, column-major (ldh = mmax+1) per scenario
    double* eps;                  // FD step
    double* veps;                 // eps * cs[j] for the next G call
    double* ieps;                 // 1 / eps
    double* res;                  // misc results per scenario (8 slots)
    double* coefr;                // DCGS2: coefficients of the delayed correction of column j
    double* dots;                 // DCGS2: reduced sums of k_pass_dca, stride ld2
    double* nwj;                  // ||w||^2 before projection, per iteration j (stride ld)
    double* wrj;                  // WRMS of the new column, per iteration j (stride ld)
    int ld, ldh, hstride, ld2;
};

hipError_t launch_pass(int mode, const double* const* cols, int nk, double* w, const double* u, const double* coef,
                       int ldc, const BlkMap& M, int nblk, double* part, int ldp, const int* active, double at,
                       double rt, hipStream_t st);
hipError_t launch_norms(const double* a, const double* u, const BlkMap& M, int nblk, double* part, int ldp,
                        const int* active, double at, double rt, hipStream_t st);
hipError_t launch_lincomb(double* out, const double* base, const double* const* cols, const int* nk_s,
                          const double* coef, int ldc, double sgn, const BlkMap& M, int nblk, double* part, int ldp,
                          int partmode, const double* u, double at, double rt, const double* extra_sub,
                          const int* active, hipStream_t st);
hipError_t launch_lincomb(double* out, const double* base, const float* const* cols, const int* nk_s,
                          const double* coef, int ldc, double sgn, const BlkMap& M, int nblk, double* part, int ldp,
                          int partmode, const double* u, double at, double rt, const double* extra_sub,
                          const int* active, hipStream_t st);
// out1 = base1 + sum_{k < nk1[s]} coef1 col_k, out2 = sum_{k < nk2[s]} coef2 col_k, one pass
hipError_t launch_lincomb2(double* out1, const double* base1, double* out2, const double* const* cols,
                           const int* nk1_s, const double* coef1, const int* nk2_s, const double* coef2, int ldc,
                           const BlkMap& M, int nblk, const int* active, hipStream_t st);
hipError_t launch_lincomb2(double* out1, const double* base1, double* out2, const float* const* cols,
                           const int* nk1_s, const double* coef1, const int* nk2_s, const double* coef2, int ldc,
                           const BlkMap& M, int nblk, const int* active, hipStream_t st);
hipError_t launch_reduce(int mode, const double* part, int ldp, int nk, const int* blk_first, int n_scen,
                         const long long* scen_rows, KrylovState K, int j, const int* active, hipStream_t st);
// One Arnoldi orthogonalisation by DCGS2 (nk = j + 1 columns; w becomes raw column j+1
// and column j, Rj = cols[j], is corrected in place): k_pass_dca, column reduction,
// k_dc_finish, k_pass_dcb, then the RED_NORM bookkeeping.
hipError_t launch_dcgs2(int nk, const double* const* cols, double* Rj, double* w, const double* u, KrylovState K,
                        const BlkMap& M, int nblk, double* part, int ldp, const int* blk_first, int n_scen,
                        const long long* scen_rows, const int* active, double at, double rt, hipStream_t st);
// The same with the basis read from FP32 copies: columns k < j from cols32, R_j from its
// FP64 column Rj (left as it is), the corrected R_j written to its FP32 copy Rj32 (for
// j = 0 too, so column 0 gets its copy).
hipError_t launch_dcgs2_f32(int nk, const float* const* cols32, const double* Rj, float* Rj32, double* w,
                            const double* u, KrylovState K, const BlkMap& M, int nblk, double* part, int ldp,
                            const int* blk_first, int n_scen, const long long* scen_rows, const int* active, double at,
                            double rt, hipStream_t st);
hipError_t launch_to_f32(float* dst, const double* src, const BlkMap& M, int nblk, const int* mask, hipStream_t st);
hipError_t launch_masked_copy(double* dst, const double* src, const BlkMap& M, int nblk, const int* mask,
                              hipStream_t st);

// Allocation-free FD products (tsr_solver_jv_noalloc): the per-scenario FD step and its
// inverse, [eps x S | 1/eps x S], live in a file-scope device array of
// 2 * kNoAllocMaxScen doubles (krylov.hip).  noalloc_scratch() gives the calling thread's
// current device's address of it (resolved once per device), or nullptr on error;
// launch_fill_eps writes the 2 S entries.
constexpr int kNoAllocMaxScen = 4096;
double*    noalloc_scratch();
hipError_t launch_fill_eps(double* de, int n_scen, double eps, hipStream_t st);
// Readbacks into a mapped pinned host buffer (layouts in krylov.hip); both clear the
// bad flags they read
hipError_t launch_pack_window(const KrylovState& K, int c0, int jt, int* bad, const int* active, int n_scen,
                              double* out, int stride, hipStream_t st);
// One small readback per scenario, kPackSmall doubles: res[0..7], the bad flag (then
// cleared), eps, the macro-state summary (stats, 6), the controller's output (ctl, 2)
constexpr int kPackSmall = 18;
hipError_t launch_pack_small(const double* res, int* bad, const double* eps, const double* stats, const float* ctl,
                             int n_scen, double* out, hipStream_t st);
// out[s*6 + 0..5]: mean, min, max macro head; mean, min, max matrix-soil saturation
hipError_t launch_macro_stats(const DScen* scen, int n_scen, const double* u, double* out, hipStream_t st);

}  // namespace tsr

#endif
