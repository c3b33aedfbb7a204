/*
 * emu_oracle.h -- FP64 CPU restatement of Emu's flavor-evolution step.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP
 * implementation in emu_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never does.
 *
 * The oracle keeps the reference's data structures on purpose:
 *   - particles are an array of structs {pos[3], rdata[NATTR], id, cpu}
 *     (Source/FlavoredNeutrinoContainer.H:12-26, 64-65), 192 B at NF=2 and
 *     272 B at NF=3;
 *   - the grid is one AMReX-style fab with ghost cells, component-major,
 *     i fastest (Source/main.cpp:67-78), with ParticleToMesh / SumBoundary /
 *     FillBoundary semantics restated in emu_oracle.c.
 *
 * Every function exists twice, suffixed _nf2 and _nf3 (NUM_FLAVORS).
 */
#ifndef EMU_ORACLE_H_
#define EMU_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Run parameters, in the reference's internal units (Source/Parameters.H:11-131):
 * masses in grams, angles in radians, lengths in cm, rho in g/ccm. */
typedef struct or_params {
    int ncell[3];
    int shape_order;            /* SHAPE_FACTOR_ORDER (Make.Emu:2) */
    double plo[3], phi[3];      /* RealBox (main.cpp:58-59) */
    double rho, Ye, T;          /* uniform background (main.cpp:81-84) */
    double mass[3];             /* grams (Parameters.H:81-82,92) */
    double theta12, theta13, theta23, alpha1, alpha2, deltaCP; /* radians */
    double maxError;
    double cfl_factor, flavor_cfl_factor, max_adaptive_speedup;
    int nthreads;               /* 0/1: serial deterministic; >1: OpenMP */
} or_params;

/* Error report for Renormalize / compute_dt (the reference aborts instead). */
enum {
    OR_OK = 0,
    OR_ERR_TRACE = 1,   /* generate_code.py:404-409 */
    OR_ERR_NEGDIAG = 2, /* generate_code.py:418-423 */
    OR_ERR_LENGTH = 3,  /* generate_code.py:433-438 */
    OR_ERR_DT = 4       /* Evolve.cpp:85 */
};

#define OR_DECLARE(NFX)                                                                  \
    int or_nattribs_##NFX(void);                                                         \
    int or_ncomp_##NFX(void);                                                            \
    int64_t or_grid_ndoubles_##NFX(const or_params* p);                                  \
    void or_grid_dims_##NFX(const or_params* p, int* ng, int* nfab);                     \
    void or_mass_matrix_##NFX(const or_params* p, double* M2);                           \
    double or_su_magnitude_##NFX(const double* h);                                       \
    void or_commutator_rhs_##NFX(const double* H, const double* f, double* dfdt);        \
    void or_grid_init_##NFX(const or_params* p, double* grid);                           \
    void or_deposit_to_mesh_##NFX(const or_params* p, const void* parts, int64_t np,     \
                                  double* grid);                                         \
    void or_fill_boundary_##NFX(const or_params* p, double* grid);                       \
    void or_interpolate_rhs_from_mesh_##NFX(const or_params* p, void* rhs, int64_t np,   \
                                            const double* grid);                         \
    int or_compute_dt_##NFX(const or_params* p, const double* grid, double Vvac_max,     \
                            double* dt);                                                 \
    double or_vvac_max_##NFX(const or_params* p, const void* parts, int64_t np);         \
    int or_renormalize_##NFX(const or_params* p, void* parts, int64_t np,                \
                             int64_t* bad_index);                                        \
    void or_sync_location_##NFX(void* parts, int64_t np, int type);                      \
    void or_redistribute_##NFX(const or_params* p, void* parts, int64_t np);             \
    void or_apply_rhs_##NFX(void* parts, int64_t np, double a, const void* dFdt);        \
    int or_rk4_step_##NFX(const or_params* p, void* state, int64_t np, double* grid,     \
                          double dt, double Vvac_max, double* dt_next,                   \
                          int64_t* bad_index);                                           \
    int or_rk_step_##NFX(const or_params* p, void* state, int64_t np, double* grid,       \
                         double dt, double Vvac_max, int nstages, const double* A,       \
                         const double* B, double* dt_next, int64_t* bad_index);          \
    int or_evolve_##NFX(const or_params* p, void* state, int64_t np, double* grid,       \
                        int nsteps, double end_time, double* time, double* dt,           \
                        double* dt_log);

OR_DECLARE(nf2)
OR_DECLARE(nf3)

/* NF-independent pieces (ParticleInterpolator.H, Constants.H). */
void or_interp(double delta, int order, int* first, double* w);
void or_constants(double* out); /* c, hbar, GF, Mp, eV */
void or_roundoff_hi(const or_params* p, double* rhi);

#ifdef __cplusplus
}
#endif

#endif
