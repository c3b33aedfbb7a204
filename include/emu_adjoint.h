/*
 * emu_adjoint.h -- the flavour right-hand side and its vector-Jacobian product
 * on particles kept in a fixed order, for reverse-mode gradients of the RK4
 * evolution (no counterpart in the reference, which has no adjoint).
 *
 * Arrays are device SoA [component][np] in double precision:
 *   pos [3][np], pup [4][np] (pupx, pupy, pupz, pupt), Nn [2][np] (N, Nbar),
 *   Y / k / lam [2 NF^2][np] (nu components then nubar, PIdx order of f),
 *   Ll [2][np] (L, Lbar).
 * 'work' is caller-owned device scratch of emu_adj_work_doubles() doubles.
 */
#ifndef EMU_ADJOINT_H_
#define EMU_ADJOINT_H_

#include "emu_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

int64_t emu_adj_work_doubles(const emu_ctx* ctx, int64_t np);

/* k = d(f, fbar)/dt of interpolate_rhs_from_mesh (Evolve.cpp:128-176) after
 * deposit_to_mesh of Y (Evolve.cpp:91-126) at positions pos. */
int emu_adj_rhs(emu_ctx* ctx, int64_t np, const double* pos, const double* pup, const double* Nn, const double* Y,
                double* k, double* work);

/* lam_Y = (dk/dY)^T lam_k (overwritten); lam_Nbar += (dk/dNbar)^T lam_k when
 * lam_Nbar is not NULL. */
int emu_adj_rhs_vjp(emu_ctx* ctx, int64_t np, const double* pos, const double* pup, const double* Nn,
                    const double* Y, const double* lam_k, double* lam_Y, double* lam_Nbar, double* work);

/* J = sum over cells of |N_01|^2, N_01 = sum_p w N f01 (the nu part of the
 * N01 grid components); dJdY (optional) receives dJ/dY.  'work': 2 ncell + 1
 * doubles at least (either work array of this header).  Synchronous. */
int emu_adj_objective(emu_ctx* ctx, int64_t np, const double* pos, const double* Nn, const double* Y, double* J,
                      double* dJdY, double* work);

/* Positions after one RK4 step of length dt and the periodic shift (the same
 * arithmetic as the fused step). */
int emu_adj_advance_positions(emu_ctx* ctx, int64_t np, double* pos, const double* pup, double dt);

/* Renormalize (FlavoredNeutrinoContainer.cpp:89-113) on the SoA state.
 * Synchronous; returns the first error as emu_renormalize does. */
int emu_adj_renormalize(emu_ctx* ctx, int64_t np, double* Y, const double* Ll, int64_t* bad_index);

/* out = a x + y, elementwise over n doubles. */
int emu_adj_axpy(emu_ctx* ctx, int64_t n, double a, const double* x, const double* y, double* out);

/* Step-level calls: 'work' holds emu_adj_step_work_doubles(ctx, np) doubles. */
int64_t emu_adj_step_work_doubles(const emu_ctx* ctx, int64_t np);

/* One RK4 step of length dt at the step's positions (main.cpp:129-213 with
 * the RK hook of FlavoredNeutrinoContainer.H:33-47; the stage states and
 * accumulator rounded as the AMReX RK4 saxpys): y_out = y_{n+1} from
 * y_in = y_n, Renormalize on y_out (FlavoredNeutrinoContainer.cpp:89-113)
 * with its status deferred to emu_adj_status, then pos advanced in place.
 * y_out must not alias y_in.  Asynchronous. */
int emu_adj_step(emu_ctx* ctx, int64_t np, double* pos, const double* pup, const double* Nn, const double* Ll,
                 const double* y_in, double* y_out, double dt, double* work);

/* The vector-Jacobian product of that step's flavour map at its starting
 * positions (Renormalize passed through): lam_out = (dy_out/dy_in)^T lam_in,
 * lam_Nbar += (dy_out/dNbar)^T lam_in when lam_Nbar is not NULL.  lam_out must
 * not alias lam_in.  Asynchronous. */
int emu_adj_step_vjp(emu_ctx* ctx, int64_t np, const double* pos, const double* pup, const double* Nn,
                     const double* y_in, const double* lam_in, double* lam_out, double* lam_Nbar, double dt,
                     double* work);

/* Doubles of one step's keep buffer: its stage states Y2, Y3, Y4 and its four
 * stage deposits.  No counterpart in the reference (the forward keeps no
 * stage data; this buffer trades memory for the VJP's recompute). */
int64_t emu_adj_step_keep_doubles(const emu_ctx* ctx, int64_t np);

/* emu_adj_step that leaves its stage states and deposits in keep
 * (emu_adj_step_keep_doubles doubles) for emu_adj_step_vjp_kept.  Same
 * y_out, positions and deferred status as emu_adj_step.  Asynchronous. */
int emu_adj_step_keep(emu_ctx* ctx, int64_t np, double* pos, const double* pup, const double* Nn, const double* Ll,
                      const double* y_in, double* y_out, double dt, double* work, double* keep);

/* emu_adj_step_vjp of a step run by emu_adj_step_keep (pos, y_in: that step's
 * start): the stage states and deposits come from keep instead of being
 * recomputed.  Asynchronous. */
int emu_adj_step_vjp_kept(emu_ctx* ctx, int64_t np, const double* pos, const double* pup, const double* Nn,
                          const double* y_in, const double* keep, const double* lam_in, double* lam_out,
                          double* lam_Nbar, double dt, double* work);

/* The FP32-bulk form of the step calls (no counterpart in the reference: the
 * adjoint of the FP32-bulk forward).  The flavour arrays, the particle
 * constants pup, Nn, Ll and the adjoints are float; positions and lam_Nbar stay
 * double.  The gathers, commutators and stage algebra run in FP32, the deposits
 * sum in FP64 and Renormalize runs in FP64, as in the fused FP32-bulk step.
 * Nn holds the weights times wscale, a power of two chosen so that FP32 holds
 * them and the adjoints of |N_01|^2 (which grow as N^2); the potential's
 * coefficient and the matter term take the factor back out, so the evolution
 * is that of the unscaled weights, and the objective and its adjoints are
 * wscale^2 times the unscaled ones.  'work' holds
 * emu_adj_step_work_doubles_f32(ctx, np) doubles, 'keep'
 * emu_adj_step_keep_doubles_f32(ctx, np). */
int64_t emu_adj_step_work_doubles_f32(const emu_ctx* ctx, int64_t np);
int emu_adj_step_f32(emu_ctx* ctx, int64_t np, double* pos, const float* pup, const float* Nn, const float* Ll,
                     const float* y_in, float* y_out, double dt, double wscale, double* work);
int emu_adj_step_vjp_f32(emu_ctx* ctx, int64_t np, const double* pos, const float* pup, const float* Nn,
                         const float* y_in, const float* lam_in, float* lam_out, double* lam_Nbar, double dt,
                         double wscale, double* work);
int64_t emu_adj_step_keep_doubles_f32(const emu_ctx* ctx, int64_t np);
int emu_adj_step_keep_f32(emu_ctx* ctx, int64_t np, double* pos, const float* pup, const float* Nn, const float* Ll,
                          const float* y_in, float* y_out, double dt, double wscale, double* work, double* keep);
int emu_adj_step_vjp_kept_f32(emu_ctx* ctx, int64_t np, const double* pos, const float* pup, const float* Nn,
                              const float* y_in, const double* keep, const float* lam_in, float* lam_out,
                              double* lam_Nbar, double dt, double wscale, double* work);
/* emu_adj_objective on float flavour arrays (J summed in FP64). */
int emu_adj_objective_f32(emu_ctx* ctx, int64_t np, const double* pos, const float* Nn, const float* Y, double* J,
                          float* dJdY, double* work);

/* The particles listed by cell (the cell of their stencil's centre) at
 * positions pos: order[i] is the i-th particle in that order (within a cell,
 * in no set order).  The step calls run fastest on arrays kept in this
 * order: their particle streams are then read cell by cell.  'work' is either
 * step-call work array.  Asynchronous. */
int emu_adj_cell_order(emu_ctx* ctx, int64_t np, const double* pos, int* order, double* work);

/* Rows of np doubles (floats: _f32) through a permutation: mode 0 gathers
 * out[r][i] = in[r][idx[i]], mode 1 scatters out[r][idx[i]] = in[r][i], mode 2
 * adds the scatter into out.  out must not alias in.  Asynchronous. */
int emu_adj_permute(emu_ctx* ctx, int rows, int64_t np, const double* in, const int* idx, double* out, int mode);
int emu_adj_permute_f32(emu_ctx* ctx, int rows, int64_t np, const float* in, const int* idx, float* out, int mode);

/* The first Renormalize error of the emu_adj_step calls since the previous
 * emu_adj_status (then cleared), with the lowest particle index in
 * *bad_index.  Synchronous. */
int emu_adj_status(emu_ctx* ctx, int64_t* bad_index);

#ifdef __cplusplus
}
#endif

#endif
