/*
 * tsr_types.h - data types of the TwoScaleRich time-stepping boundary.
 *
 * The first seven structs are byte-for-byte the reference's plain-C carriers, so a
 * caller written against the reference (its main.cpp builds a user_data_struct and
 * passes it by value) links against this library unchanged.  Field order, types and
 * array layouts must match; each struct cites the reference header it mirrors.
 *
 * Array layout convention (reference): element-major SoA, a[j*no_elements + e].
 */
#ifndef TSR_TYPES_H
#define TSR_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

/* A caller that already has the reference's own definitions of these seven structs (its
 * main.h, user_data.h, ...: they carry no include guards) defines
 * TSR_USE_REFERENCE_HEADERS before including this header, which then uses those. */
#ifndef TSR_USE_REFERENCE_HEADERS

/* include/macro_mesh.h:1-19 */
struct macro_mesh_struct {
    int    no_verts;                 /* 3 = P1 triangles, 4 = Q1 quads */
    int    no_nodes;
    int    no_elements;
    int    *elements;                /* [no_verts*ne], 0-based */
    double *nodes;                   /* [2*no_nodes]: x then y */
    double *shape_funcs;             /* [no_verts*no_verts*ne]: coeff m of basis i at (m*nv+i)*ne+e */
    double *normals;                 /* [2*nv*ne]: sub-face j normal (x,y) at (2j)*ne+e, (2j+1)*ne+e */
    double *boundary_edge_midpts;    /* [4*nv*ne] */
    double *boundary_edge_normals;   /* [4*nv*ne] */
    double *boundary_edge_lens;      /* [2*nv*ne] */
    int    *boundary_condition_edges;/* [nv*ne]: 0 interior, else BC type 1..5 */
    double *CV_area;                 /* [no_nodes] */
    double *SCV_area;                /* [nv*ne] */
    double *midx;                    /* [nv*ne] sub-face midpoints */
    double *midy;
    double *centroid_elements;       /* [2*ne] */
};

/* include/micro_mesh.h:1-22 */
struct micro_mesh_struct {
    int    no_nodes;
    int    no_elements;
    int    *elements;
    int    *boundary_nodes;          /* [no_nodes]: 1 = Dirichlet node of the micro problem */
    double *nodes;
    double *shape_funcs;
    double *normals;
    int    *boundary_edges;
    double *boundary_edge_midpts;
    double *boundary_edge_normals;
    double *boundary_edge_lens;
    double *centroid_elements;
    double *CV_area;
    double *midx;
    double *midy;
    double cell_area;
    double *vectors;                 /* [2*no_nodes]: node minus cell centre */
    int    no_verts;
    double *element_area;
};

/* include/Keff.h:1-9 */
struct Keff_struct {
    double  epsilonA;                /* matrix (region A) area fraction */
    int     no_hvalues;
    double  logr;                    /* step of log10|h| between rows */
    double  logh_first;              /* log10|h| of row 0 */
    double* logh_values;             /* [no_hvalues] */
    double* Keff_values;             /* [4*no_hvalues]: K11, K12, K21, K22 blocks */
};

/* include/soil.h:1-8 */
struct soil_struct {
    double Ksat;
    double thetar;
    double thetas;
    double alpha;
    double n;
    double m;                        /* 1 - 1/n */
};

/* include/user_data.h:1-12 */
struct user_data_struct {
    struct macro_mesh_struct macro_mesh;
    struct micro_mesh_struct micro_mesh;
    struct Keff_struct       effective_conductivity;
    struct soil_struct       soilA;
    struct soil_struct       soilB;
    double boundary_flux;
    int    N;
    int    model_type;
    unsigned int NUM_THREADS;
};

/* include/exprem_options.h:1-14 */
struct exprem_options_struct {
    double AbsTol;
    double RelTol;
    int    MinKrylov;
    int    MaxKrylov;
    double MinStep;
    double MaxStep;
    int    MaxNumSteps;
    double SafetyFac;
    double MinStepDecFac;
    double MaxStepIncFac;
    double InitStep;
};

/* include/exprem_stats.h:1-9 */
struct exprem_stats_struct {
    double  CurrentStep;
    double  CurrentTime;
    int     CurrentKryDim;
    double* CurrentSoln;
    int     NumSteps;
    int     FuncEvals;
};

#endif /* TSR_USE_REFERENCE_HEADERS */

#ifndef __cplusplus
typedef struct macro_mesh_struct     macro_mesh_struct;
typedef struct micro_mesh_struct     micro_mesh_struct;
typedef struct Keff_struct           Keff_struct;
typedef struct soil_struct           soil_struct;
typedef struct user_data_struct      user_data_struct;
typedef struct exprem_options_struct exprem_options_struct;
typedef struct exprem_stats_struct   exprem_stats_struct;
#endif

/* Reference RHS plugin type (include/main.h:21-22,29): output first, no t argument. */
typedef int (*tsr_gfunc_t)(double* g, double* u, struct user_data_struct user_data);

/* ------------------------------------------------------------------------------------
 * Batch view: the problem definition both the HIP path and the CPU oracle consume.
 *
 * The reference carries one problem per user_data_struct (one shared micro mesh, soil B
 * in every micro element, cells hosted by macro elements).  The batch view generalises
 * it without changing those structs: several independent scenarios, several micro
 * templates, cells hosted by macro nodes or elements, per-element soil maps (fixed per
 * template, or from per-cell circular inclusions) and a Dirichlet macro BC (type 5).
 * ------------------------------------------------------------------------------------ */
enum { TSR_HOST_ELEMENT = 0, TSR_HOST_NODE = 1 };
enum { TSR_SOIL_A = 0, TSR_SOIL_B = 1 };
enum { TSR_BC_FLUX = 1, TSR_BC_NOFLUX = 2, TSR_BC_DHDX0 = 3, TSR_BC_DHDY0 = 4, TSR_BC_DIRICHLET = 5 };

struct tsr_scenario {
    struct user_data_struct ud;      /* macro mesh, micro template 0, Keff table, soils, BC flux, N, model_type */
    int       host_kind;             /* TSR_HOST_ELEMENT (reference) or TSR_HOST_NODE */
    int       n_cells;               /* = macro no_elements or no_nodes */
    int       first_cell;            /* index of this scenario's first cell in the global cell arrays */
    int       pad_;
    long long u_offset;              /* first unknown of this scenario in the batched vector */
    double    dirichlet_head;        /* head imposed on nodes of type-5 macro edges */
};

struct tsr_batch_view {
    int n_scen;
    int n_tmpl;
    int n_cells;
    int max_discs;                   /* 0: soil map from tmpl_soil; >0: from cell_discs */
    const struct tsr_scenario*       scen;
    const struct micro_mesh_struct*  tmpl;       /* triangle templates, physical cell-local coordinates */
    const unsigned char* const*      tmpl_soil;  /* [n_tmpl]: per-element soil id, or NULL = all soil B */
    const int*       cell_tmpl;      /* [n_cells] */
    const long long* cell_offset;    /* [n_cells] first unknown of the cell */
    const double*    cell_discs;     /* [n_cells*max_discs*3]: (cx, cy, r) cell-local; r <= 0 = unused */
    double    Ss;                    /* specific storage added to every soil's C(h) [1/m] */
    long long N;                     /* total unknowns */
};

#ifndef __cplusplus
typedef struct tsr_scenario   tsr_scenario;
typedef struct tsr_batch_view tsr_batch_view;
#endif

#ifdef __cplusplus
}
#endif
#endif /* TSR_TYPES_H */
