/*
 * eph_adjoint.h -- gradients of a device-resident fix eph trajectory.
 *
 * No counterpart in USER-EPH: the reference fits beta(rho) by hand.  The loss
 * of an nsteps trajectory of the harness (eph_md.h) with the fix,
 *   L = Ee(after nsteps) + sum_c (T_e,c(after nsteps) - T_target,c)^2,
 * is differentiated in reverse mode with hand-written vector-Jacobian products
 * of every part of the step (Verlet kicks, pair lj/cut, the density and w
 * sweeps, the force sweep of force_prl, alpha(rho), the deposit, the FDM
 * stencil) with respect to the beta(rho) knot values of every element and the
 * initial velocities.  The noise is the Philox stream (seed, tag, step), so a
 * trajectory is a deterministic function of (beta, v0).  The forward pass
 * keeps a checkpoint every `every` steps; the backward pass recomputes each
 * segment from its checkpoint, records it, and runs the adjoint steps in
 * reverse.  Grids with temperature-dependent cells (eph_fdm.h:280-287,
 * 376-382) are differentiated through the energy-space update (the slopes of
 * E_e(T) and its inverse) and through kappa(T), evaluated once per solve at
 * its starting T_e; the sub-step count, a step function of the stability
 * extrema, is held as the forward pass found it.
 */
#ifndef EPH_ADJOINT_H
#define EPH_ADJOINT_H

#include "eph_md.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct eph_adjoint eph_adjoint;

/* md must be attached to ctx (eph_md_attach); the current state is the start of
 * every trajectory.  T_target: [ncells] (NULL: the loss has no grid term). */
int eph_adjoint_create(eph_adjoint** out, eph_md* md, eph_context* ctx, int nsteps, int every,
                       const double* T_target);
void eph_adjoint_destroy(eph_adjoint* a);
const char* eph_adjoint_last_error(const eph_adjoint* a);

/* sizes: elements, beta knots per element, local atoms, T_e cells */
int eph_adjoint_info(const eph_adjoint* a, int* n_elem, int* n_beta, int* nlocal, int* ncells);
/* beta knot values of the .beta file, [n_elem][n_beta] */
int eph_adjoint_get_beta(const eph_adjoint* a, double* beta);

/* the start's local velocities, [nlocal][3] */
int eph_adjoint_get_v0(const eph_adjoint* a, double* v0);

/* Run the trajectory from the start with beta knots [n_elem][n_beta] (NULL:
 * the file's) and initial local velocities [nlocal][3] (NULL: the start's).
 * *loss = L; parts (may be NULL) = (Ee, grid term). */
int eph_adjoint_forward(eph_adjoint* a, const double* beta, const double* v0, double* loss,
                        double* parts);

/* dL/dbeta [n_elem][n_beta] and dL/dv0 [nlocal][3] of the last forward call */
int eph_adjoint_backward(eph_adjoint* a, double* g_beta, double* g_v0);

/* Host: Jacobian of the table eph_spline_build makes, jac[k][c][m] =
 * d(coefficient c of interval k)/d y_m, n x 4 x n (no device needed). */
int eph_spline_jacobian(double dx, const double* y, int n, double* jac);

#ifdef __cplusplus
}
#endif

#endif /* EPH_ADJOINT_H */
