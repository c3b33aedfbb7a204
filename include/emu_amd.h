/*
 * emu_amd.h -- C ABI of the MI355X-native flavor-evolution step of Emu.
 *
 * Plain C types only: pointers, sizes, doubles.  Every pointer argument that
 * names particle or grid data is a DEVICE pointer (hipMalloc'd or a torch
 * tensor's data_ptr) unless stated otherwise.  Calls are stream-ordered on the
 * context's stream and asynchronous unless they return a value computed on
 * the device (documented per call).  One host thread per context.
 *
 * Two layers:
 *   1. Per-function entry points with the reference's semantics on the
 *      reference's data (SoA particle attributes in PIdx order, the MultiFab
 *      `state` as one ghost-cell fab in GIdx order).  Each replaces the
 *      reference function cited beside it.
 *   2. A fused, library-owned particle state (emu_state) that advances whole
 *      RK4 steps with the particles sorted by cell, in FP64 ("plain") or with
 *      the bulk in FP32 ("fp32-bulk": FP64 kept for positions, the mesh
 *      deposit accumulation, Renormalize and compute_dt).
 *
 * Return codes: 0 = EMU_OK, otherwise one of emu_status.
 */
#ifndef EMU_AMD_H_
#define EMU_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMU_MAX_ATTRIBS 30 /* PIdx::nattribs at NUM_FLAVORS = 3 */

typedef enum emu_status {
    EMU_OK = 0,
    EMU_ERR_TRACE = 1,       /* Renormalize: |tr f - 1| > 100 maxError (generate_code.py:404-409) */
    EMU_ERR_NEGDIAG = 2,     /* Renormalize: f_ii < -100 maxError (generate_code.py:418-423) */
    EMU_ERR_LENGTH = 3,      /* Renormalize: ||f|_SU - L| > 100 maxError (generate_code.py:433-438) */
    EMU_ERR_DT = 4,          /* compute_dt: no positive CFL factor (Evolve.cpp:85) */
    EMU_ERR_ARG = 10,        /* invalid argument */
    EMU_ERR_HIP = 11,        /* HIP runtime error (see emu_last_error) */
    EMU_ERR_UNSUPPORTED = 12,/* configuration outside what the kernels support */
    EMU_ERR_NOMEM = 13,      /* device allocation failed */
    EMU_ERR_MOVE = 14        /* a particle moved more than one cell in one step */
} emu_status;

enum { EMU_PLAIN_FP64 = 0, EMU_FP32_BULK = 1 };
/* Slot layout of a fused state, or-ed into emu_state_create's precision.  The
 * state keeps its particles sorted by cell; ALIGNED starts every cell's run of
 * slots on a 128-byte boundary of the per-particle arrays (32 slots in
 * FP32-bulk, 16 in FP64), PACKED keeps the runs back to back.  Neither (0)
 * aligns when the cells hold on average at least two boundaries' worth of
 * particles, which bounds the padding at half the particle count. */
enum { EMU_STATE_SLOTS_PACKED = 0x100, EMU_STATE_SLOTS_ALIGNED = 0x200 };

/* Run parameters in the reference's internal units (Source/Parameters.H:11-131
 * after its unit conversions): masses in grams, angles in radians. */
typedef struct emu_params {
    int nf;            /* NUM_FLAVORS, 2 or 3 (Make.Emu:1) */
    int shape_order;   /* SHAPE_FACTOR_ORDER, 0..2 (Make.Emu:2) */
    int ncell[3];
    double plo[3], phi[3];
    double rho, Ye, T; /* uniform background (main.cpp:81-84) */
    double mass[3];
    double theta12, theta13, theta23, alpha1, alpha2, deltaCP;
    double maxError;
    double cfl_factor, flavor_cfl_factor, max_adaptive_speedup;
} emu_params;

/* SoA view of a FlavoredNeutrinoContainer: pos = p.pos(d), rdata[k] =
 * p.rdata(k) in PIdx order (FlavoredNeutrinoContainer.H:12-26). */
typedef struct emu_particles {
    int64_t np;
    double* pos[3];
    double* rdata[EMU_MAX_ATTRIBS];
} emu_particles;

/* The MultiFab `state` (main.cpp:78) as one periodic fab: Array4 layout,
 * i fastest, then j, k, component; lo/hi bound the fab box including ghost
 * cells, in the domain's index space (valid cells are 0..ncell-1). */
typedef struct emu_grid {
    double* data;
    int lo[3], hi[3];
    int ncomp; /* GIdx::ncomp = 3 + 8 NF^2 */
} emu_grid;

typedef struct emu_ctx emu_ctx;
typedef struct emu_state emu_state;

/* ---- context -------------------------------------------------------------
 * Geometry, mass matrix and scratch for one run (evolve_flavor's setup,
 * main.cpp:37-85 and Evolve.H_M2_fill, generate_code.py:176-227). */
int emu_create(const emu_params* params, emu_ctx** out);
int emu_destroy(emu_ctx* ctx);
int emu_set_stream(emu_ctx* ctx, void* hip_stream); /* NULL = default stream */
int emu_sync(emu_ctx* ctx);                         /* hipStreamSynchronize */
const char* emu_last_error(const emu_ctx* ctx);
/* Ghost width per dim of the reference's MultiFab: 1 + (1+order)/2, order 0
 * in 1-cell dims (main.cpp:67-71). */
int emu_grid_ngrow(const emu_ctx* ctx, int ngrow[3]);

/* ---- per-function entry points (FP64, reference semantics) -------------- */

/* deposit_to_mesh (Evolve.cpp:91-126 via amrex::ParticleToMesh): zero the
 * deposit components over valid+ghost cells, deposit N f and N f p/pt with the
 * shape factors, fold ghost deposits onto their periodic images (SumBoundary;
 * ghosts keep their partial sums). */
int emu_deposit_to_mesh(emu_ctx* ctx, const emu_particles* parts, emu_grid* state);

/* emu_deposit_to_mesh without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  Its scratch is a fixed array in
 * the library's device memory (47 MB per GPU, held until the process ends),
 * for at most 4,194,304 particles and 1,048,576 fab cells, ghosts included;
 * larger calls return EMU_ERR_UNSUPPORTED.  All calls on one GPU share that
 * array: calls that overlap on one GPU (other streams, other contexts) race. */
int emu_deposit_to_mesh_static(emu_ctx* ctx, const emu_particles* parts, emu_grid* state);

/* state.FillBoundary(geom.periodicity()) (main.cpp:85,137). */
int emu_fill_boundary(emu_ctx* ctx, emu_grid* state);

/* interpolate_rhs_from_mesh (Evolve.cpp:128-176): 'rhs' holds a copy of the
 * stage particles (main.cpp:148) and is overwritten with d/dt of every
 * attribute.  The grid's ghost cells must be filled. */
int emu_interpolate_rhs_from_mesh(emu_ctx* ctx, emu_particles* rhs, const emu_grid* state);

/* compute_dt (Evolve.cpp:22-89) over the whole fab box, ghosts included.
 * Synchronous: writes *dt on the host. */
int emu_compute_dt(emu_ctx* ctx, const emu_grid* state, double Vvac_max, double* dt);

/* FlavoredNeutrinoContainer::Vvac_max (FlavoredNeutrinoContainerInit.cpp:
 * 657-660).  Synchronous. */
int emu_vvac_max(emu_ctx* ctx, const emu_particles* parts, double* Vvac_max);

/* FlavoredNeutrinoContainer::Renormalize (FlavoredNeutrinoContainer.cpp:
 * 89-113).  Where the reference aborts, returns EMU_ERR_TRACE/NEGDIAG/LENGTH
 * and the lowest offending particle index in *bad_index (-1 if none); that
 * particle is left as it was when the condition was met.  Synchronous. */
int emu_renormalize(emu_ctx* ctx, emu_particles* parts, int64_t* bad_index);

/* FlavoredNeutrinoContainer::SyncLocation (FlavoredNeutrinoContainer.cpp:
 * 8-41): type 0 = CoordinateToPosition, 1 = PositionToCoordinate. */
int emu_sync_location(emu_ctx* ctx, emu_particles* parts, int type);

/* FlavoredNeutrinoContainer::RedistributeLocal (FlavoredNeutrinoContainer.H:
 * 83-90) on one periodic box: shift positions outside [plo, rhi] by one period
 * and clamp into it (AMReX's roundoff domain).  Particle order is kept. */
int emu_redistribute_local(emu_ctx* ctx, emu_particles* parts);

/* ApplyFlavoredNeutrinoRHS::operator() (FlavoredNeutrinoContainer.H:33-47):
 * p.rdata[k] = a * dFdt.rdata[k] + p.rdata[k] for every attribute. */
int emu_apply_rhs(emu_ctx* ctx, emu_particles* parts, double a, const emu_particles* dFdt);

/* copyParticles(src, true) (main.cpp:115,148): attribute-wise device copy. */
int emu_copy_particles(emu_ctx* ctx, emu_particles* dst, const emu_particles* src);

/* One stage combination of the AMReX Runge-Kutta integrator the reference
 * drives (main.cpp:129,202-213): dst = src (copyParticles), then for j = 0 ..
 * n-1 dst = a[j] F[j] + dst (ApplyFlavoredNeutrinoRHS,
 * FlavoredNeutrinoContainer.H:33-47), in one pass and bit for bit that
 * sequence of calls.  F is a host array of n (<= 4) particle views; dst must
 * not alias src or any F[j]. */
int emu_rk_combine(emu_ctx* ctx, emu_particles* dst, const emu_particles* src, int n, const double* a,
                   const emu_particles* F);

/* ---- fused state ----------------------------------------------------------
 * The integrator loop of main.cpp:129-213 (AMReX TimeIntegrator RK4 + the
 * post-step hook) on a library-owned SoA state sorted by cell.  precision is
 * EMU_PLAIN_FP64 or EMU_FP32_BULK, optionally or-ed with one slot layout flag. */
int emu_state_create(emu_ctx* ctx, int64_t np, int precision, emu_state** out);
int emu_state_destroy(emu_state* st);
/* Load particles (device SoA view, any order); computes Vvac_max, deposits and
 * sets the starting dt exactly as main.cpp:115-206.  Synchronous. */
int emu_state_load(emu_state* st, const emu_particles* parts, const int64_t* ids);
/* Replace the Vvac_max that emu_state_load computed (the vacuum term added to
 * compute_dt's flavor speeds, Evolve.cpp:60-61) and recompute the starting dt
 * with it.  The reference sets FlavoredNeutrinoContainer::Vvac_max
 * (FlavoredNeutrinoContainer.H:71) only in InitParticles
 * (FlavoredNeutrinoContainerInit.cpp:657-660), so a run restarted from a
 * checkpoint (main.cpp:103-105) keeps its zero initial value: a restart driver
 * passes 0.  Only between emu_state_load and the first emu_state_evolve
 * (EMU_ERR_ARG afterwards).  Synchronous. */
int emu_state_set_vvac_max(emu_state* st, double vvac_max);
/* Write the state back (sorted order) with each particle's id.  Synchronous. */
int emu_state_store(emu_state* st, emu_particles* parts, int64_t* ids);
/* Advance nsteps RK4 steps (asynchronous; dt stays on the device between
 * steps).  A fixed_dt > 0 replaces compute_dt's value for every step.  Each
 * step's sort plan and compute_dt run on a second stream that the state owns,
 * forked from and joined back into the context's stream with events, so all of
 * a call's work is ordered before later work on the context's stream. */
int emu_state_evolve(emu_state* st, int nsteps, double fixed_dt);
/* Time, last dt, steps taken and the deferred status of the steps since load:
 * EMU_ERR_TRACE/NEGDIAG/LENGTH (Renormalize) or EMU_ERR_MOVE (a particle that
 * moved more than one cell, which only |p| > p_t allows), with *bad_index = the
 * lowest offending particle id; EMU_ERR_DT with *bad_index = -1.  Synchronous. */
int emu_state_info(emu_state* st, double* time, double* dt, int64_t* steps, int64_t* bad_index);
/* Per-launch timing: enable = 1 records a HIP event pair around every launch
 * of the following emu_state_evolve calls, on the stream it runs on (and
 * clears earlier records). */
int emu_state_profile(emu_state* st, int enable);
/* Average milliseconds per launch of kernel class k (0 = stage kernels,
 * 1 = grid fills, 2 = cell sort, 3 = dt reduction) since profiling was
 * enabled.  Synchronous. */
int emu_state_kernel_ms(emu_state* st, int k, double* avg_ms, int64_t* launches);

/* Diagnostic for growth rates: sums[0] = sum over cells of |N_01| and
 * sums[1] = sum over cells of |Nbar_01|, where N_01 (cell) = sum_p w N f01 with
 * the particle shape factors (the N01 and N01bar grid components of
 * deposit_to_mesh, Evolve.cpp:91-126).  Synchronous. */
int emu_state_flavor_amplitude(emu_state* st, double sums[2]);

#ifdef __cplusplus
}
#endif

#endif /* EMU_AMD_H_ */
