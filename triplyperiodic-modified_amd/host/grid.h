// grid.h -- host mirror of the reference's grid class (grid.h:14-146).
//
// The same constructor, public fields and substep methods, so that the
// reference's driver loop (simulation.cpp:57-93) runs unchanged on top of
// the device library: every method forwards to one C ABI entry point, the
// public tensors are handles to device-resident fields, and the scalars are
// refreshed from the device after each method that changes them.
#pragma once

#include <fstream>

#include "fields.h"

class grid {
  params *param_;
  proc *pc_;
  gridsize *size_;
  communicator *com_;
  tpx_ctx *ctx_ = nullptr;
  poisson PS_;
  bool Is_touch_ = false;
  // the statistics streams (grid.h:22-49); GasMaxDiffCFL is never opened
  // there, so its lines go nowhere, as here
  std::ofstream stat_Tg, stat_Tp, stat_HT, stat_TKE2, stat_TKE_U, stat_TKE_V, stat_TKE_W,
      stat_Passive_Scalar_mean, stat_TKEV_V1, stat_TKEV_V2, stat_TKEV_V3, stat_P0, stat_CMax, stat_CMin,
      stat_CMean, stat_RhoMax, stat_RhoMin, stat_RhoMean, stat_ParticleMaxCFL, stat_GasMaxCFL,
      stat_GasMaxDiffCFL, stat_NumIteration, stat_BalanceIndex, stat_trajectory;
  // fields that only Store writes when they are not advanced on the device:
  // the passive scalar's and vector's initial values (grid.cpp:256-458) and Q
  std::vector<double> ps_host_, rv_host_[3], q_host_;
  void pull_scalars();
  void initial_vector();
  void write_ghosted(const std::vector<double> &f, const std::string &name);
  void ok(int rc, const char *what) { tpx_check(ctx_, rc, what); }
  [[noreturn]] void unsupported(const char *what);

 public:
  particle part;
  tensor1 RU, RU_int, RU_new, RU_np1, RHS_RU;
  tensor1 RV, RV_int, RV_new, RV_np1;
  tensor0 P, RHS_Pois, C, Rho, Rho_int, Rho_new, Rho_np1, divergence, RHS_Part_Temp;
  tensor0 Q, RHS_Pois_Q;
  tensor0 Passive_Scalar, Passive_Scalar_int, Passive_Scalar_new, Passive_Scalar_np1;
  double P0 = 0, P0_int = 0, P0_new = 0, P0_np1 = 0, dP0_dt = 0;
  double T_cur = 0;
  int num_timestep = 0;
  int RK4_count = 0;

  grid(gridsize *size, params *param, proc *pc, communicator *com);
  ~grid();
  grid(const grid &) = delete;
  grid &operator=(const grid &) = delete;
  tpx_ctx *context() { return ctx_; }

  void Initialize();          // grid.cpp:117-486 (Initial = 0..5)
  void Store();               // grid.cpp:489-593
  void TimeAdvance();         // grid.cpp:595-606
  void Update_Rho();          // grid.cpp:608-614
  void Update_RU_WOP();       // grid.cpp:776-814
  void Update_P0();           // grid.cpp:816-830
  void Compute_Div_U_new();   // grid.cpp:832-839
  void Compute_RHS_Pois();    // grid.cpp:841-849
  void Solve_Poisson();       // grid.cpp:851-855
  void Update_RU_WP();        // grid.cpp:857-863
  void Update_Particle();     // grid.cpp:865-888
  void Update_Passive_Scalar();  // grid.cpp:711-745
  void TimeAdvance_RK4();     // grid.cpp:890-900
  void Update_RV_WOQ();       // grid.cpp:646-692
  void Compute_RHS_Pois_Q();  // grid.cpp:694-701
  void Solve_Poisson_Q();     // grid.cpp:703-707
  void Update_RV_WQ();        // grid.cpp:709-715
  void Statistics();          // grid.cpp:902-982 (with part.trajectory)
  bool Touch();               // grid.cpp:94-108
  int num_iteration() const { return PS_.num_iteration(); }

  // elongated-box copy: outside this library
  void CopyBox() { unsupported("CopyBox"); }
};
