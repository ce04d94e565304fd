/*
 * ggx_oracle.h -- C interface of the CPU oracle for the per-grain level-set timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a from-scratch FP64 C++/OpenMP restatement of
 * GraGLeS2D's timestep (src/grainhdl.cpp:1049-1141 and the LSbox methods it calls in
 * src/box.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker / the timed CPU baseline.  The product (libggx) never
 * links, loads or calls it.
 */
#ifndef GGX_ORACLE_H
#define GGX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Simulation constants, as grainhdl::setSimulationParameter derives them
 * (src/grainhdl.cpp:56-136) for MicrostructureGenMode 0 (read from file). */
typedef struct ggxo_params {
    int    real_domain_size;      /* R = realDomainSize (grainhdl.cpp:77-78) */
    int    grid_blowup;           /* DomainBorderSize (grainhdl.cpp:134) */
    int    convolution_mode;      /* 0 Laplace, 1 Richardson, 2 Gaussian (Settings.h:41-43) */
    double dt;                    /* timestep after find_correctTimestepSize (grainhdl.cpp:250-274) */
    int    analysis_timestep;     /* Settings::AnalysisTimestep, used by the meanDa reset */
    int    start_time;            /* Settings::StartTime */
    int    isotropic;             /* Settings::IsIsotropicNetwork */
    int    nthreads;              /* OpenMP threads; <= 0 means the OpenMP default */
} ggxo_params;

/* Per-grain summary after a step (fields of LSbox, src/box.h:71-96). */
typedef struct ggxo_grain_info {
    int32_t present;              /* grains[id] != NULL (not yet deleted) */
    int32_t exists;               /* LSbox::m_exists */
    int32_t motion_regular;       /* LSbox::m_isMotionRegular */
    int32_t intersects_boundary;  /* LSbox::m_intersectsBoundaryGrain */
    int32_t xmin, xmax, ymin, ymax; /* bounds of m_outputDistance (the redistanced field) */
    int32_t id_xmin, id_xmax, id_ymin, id_ymax; /* bounds of m_IDLocal */
    int32_t contour_len;          /* m_grainBoundary.getBoundarySegmentCount() */
    int32_t n_neighbours;         /* direct neighbours (buildDirectNeighbours) */
    int32_t n_compare;            /* m_comparisonList.size() */
    int32_t error;                /* non-zero: the reference would have aborted here */
    double  volume, perimeter, energy, mean_da;
} ggxo_grain_info;

int  ggxo_create(const ggxo_params* p, int max_id, void** ctx);
void ggxo_destroy(void* ctx);
/* Add grain `id` (1..max_id) from a closed polygon in unit coordinates (first vertex repeated
 * last) and an orientation quaternion; mirrors the LSbox file constructor (box.cpp:108-158). */
int  ggxo_add_grain(void* ctx, int id, int nverts, const double* xy, const double* q);
/* distanceInitialisation + find_neighbors (grainhdl.cpp:728-753, 1258-1283). */
int  ggxo_init(void* ctx);
/* One pass of the run_sim loop body with loop index `loop` (grainhdl.cpp:1049-1141). */
int  ggxo_step(void* ctx, int loop);
int  ggxo_get_info(void* ctx, int id, ggxo_grain_info* out);
/* Copy the redistanced field (n*n, row-major, row = y) of grain id. */
int  ggxo_grain_phi(void* ctx, int id, double* out);
/* Copy m_IDLocal (its full flat vector). */
int  ggxo_grain_idlocal(void* ctx, int id, uint32_t* out);
/* Copy the contour (x,y pairs, grid-index coordinates). */
int  ggxo_grain_contour(void* ctx, int id, double* xy);
/* Copy neighbour IDs and boundary lengths in first-encounter order. */
int  ggxo_grain_neighbours(void* ctx, int id, int32_t* ids, double* lengths);
/* Copy the comparison list (grain IDs, list order). */
int  ggxo_grain_compare_list(void* ctx, int id, int32_t* ids);
/* Number of non-deleted grains (grainhdl::get_biggestGrainVol currentNrGrains). */
int  ggxo_current_grains(void* ctx);

/* Stand-alone convolution of one n x n box (convolutionGeneratorFFTW, box.cpp:622-695). */
int  ggxo_convolve(int n, int mode, double dt, int R, const double* in, double* out);
/* Stand-alone complex DFT of length n (sign -1 forward, +1 backward, unnormalised). */
int  ggxo_dft(int n, int sign, const double* in_ri, double* out_ri);

#ifdef __cplusplus
}
#endif
#endif
