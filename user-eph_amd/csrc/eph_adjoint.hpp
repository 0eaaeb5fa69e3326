// eph_adjoint.hpp -- device views of the reverse-mode (adjoint) step.
#pragma once

#include "eph_device.hpp"

namespace eph {

// What one recorded forward step left behind (all device pointers).
struct AdjStepView {
  const ForceRec* rec = nullptr;  // x', (w, rho), (xi, alpha) of all atoms at post_force
  const PosVel* pv = nullptr;     // x', vh of all atoms at post_force
  const double4* vn = nullptr;    // v' after final_integrate [nlocal]
  const double* fE = nullptr;     // f_EPH, f_RNG [3][fstride]
  const double* fR = nullptr;
  int fstride = 0;
  const int* act = nullptr;       // active list (ELL, stride of AtomsView) and counts
  const int* nact = nullptr;
  const int* lj = nullptr;        // pair list (ELL) and counts
  const int* ljn = nullptr;
  const int* owner = nullptr;     // ghost -> owner
  int nghost = 0;
};

// Adjoint accumulators (zeroed by the host where a step needs it).
struct AdjBuffers {
  double4* xb = nullptr;   // dL/dx' [nall]
  double4* vhb = nullptr;  // dL/dvh [nall]
  double4* wb = nullptr;   // dL/dw [nall]
  double4* vnb = nullptr;  // dL/dv' [nlocal]
  double4* fEb = nullptr;  // dL/df_EPH from the deposit [nlocal]
  double4* fRb = nullptr;  // dL/df_RNG from the deposit [nlocal]
  double4* ftb = nullptr;  // dL/df_total [nlocal]
  double* ab = nullptr;    // dL/dalpha [nall]
  double* rb = nullptr;    // dL/drho [nall]
  double* Tb = nullptr;    // dL/dT_e at post_force [G]
  double* Gb = nullptr;    // dL/d(alpha table coefficients) [n_elem][n_beta][4]
};

struct AdjStepFlags {
  int fric = 0, rnd = 0, addE = 0, addR = 0, noint = 0, model4 = 0;
  double lj1 = 0, lj2 = 0, ljrc2 = 0;
};

// kb (dL/d kappa of the step's solve) is gathered only for grids with
// temperature-dependent cells; nullptr otherwise
hipError_t launch_adj_fdm(const GridView& g, const double* Tin, const double* Tb_out, double* Tb_in,
                          double* dTb, double* kb, double inner_dt, hipStream_t s);
// Tb += kb dkappa/dT at the solve's starting T_e (T0), temperature-dependent cells
hipError_t launch_adj_kappa_T(const GridView& g, const double* T0, const double* kb, double* Tb,
                              hipStream_t s);
hipError_t launch_adj_step(const AdjStepView& sv, const AtomsView& a, const TablesView& t,
                           const GridView& g, const StepParams& p, const AdjStepFlags& f,
                           const double* dTb, const AdjBuffers& b, hipStream_t s);

}  // namespace eph
