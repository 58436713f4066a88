// params.h -- host mirror of the reference's input-file parameters
// (params.h:8-158, params.cpp:6-86).  Same class name, constructor and
// accessors, so a driver written against the reference compiles unchanged;
// to_abi() hands the step's subset to the device library.
#pragma once

#include <cmath>
#include <string>

#include "../../include/tpx.h"

const double PI(3.141592653589793);        // params.h:5
const double TWO_PI(2 * 3.141592653589793);

class params {
  int Nx_tot_, Ny_tot_, Nz_tot_;
  int Np_track_;
  int bs_;
  double T_final_, dt_, threshold_;
  bool Statistics_, Stat_print_;
  int Data_freq_fast_, Data_freq_slow_;
  std::string Data_dir_, Stat_dir_;
  int Initial_, Initial_C_, Initial_V_;
  int PreCond_, Solver_;
  int cooling_;
  int Iteration_, Iteration1_, Iteration2_;
  double epsilon_;
  int TWC_, ParticleGravity_;
  double Lx_, Ly_, Lz_;
  int elongated_box_;
  int solve_for_scalar_, solve_for_vector_;
  double gx_, gy_, gz_, gx1_, gy1_, gz1_, gx2_, gy2_, gz2_;
  double A_, A1_, A2_;
  double Cp_, Cv_, R_;
  double Mu0_, k_;
  double Rho0_, U0_, V0_, W0_;
  double Rho_forV_, V0_1_, V0_2_, V0_3_;
  double T0_;
  double np0_, Dp_, Rhop_, Cvp_, epsilonp_;
  double I0_, I01_, I02_;
  double D_M_, eta_, s1t_, s2t_;
  double Nu_, Tp_, mp_, P0_, W_;
  bool ok_ = false;

 public:
  // Reads the positional input file: '#' lines and empty lines are skipped,
  // every other line joins one whitespace-separated stream (params.cpp:6-59).
  explicit params(const char *address);
  // After the threshold time the *2 coefficients apply (params.cpp:75-86).
  void update(double time);
  bool ok() const { return ok_; }   // the file opened and every group was read

  double Lx() const { return Lx_; }
  double Ly() const { return Ly_; }
  double Lz() const { return Lz_; }
  double elongated_box() const { return elongated_box_; }
  double solve_for_scalar() const { return solve_for_scalar_; }
  double solve_for_vector() const { return solve_for_vector_; }
  double Tp() const { return Tp_; }
  double R() const { return R_; }
  double P0() const { return P0_; }
  double mp() const { return mp_; }
  double Nu() const { return Nu_; }
  double I0() const { return I0_; }
  double I01() const { return I01_; }
  double I02() const { return I02_; }
  double D_M() const { return D_M_; }
  double S1_type() const { return s1t_; }
  double eta0() const { return eta_; }
  double S2_type() const { return s2t_; }
  double epsilonp() const { return epsilonp_; }
  double Cvp() const { return Cvp_; }
  double Rhop() const { return Rhop_; }
  double Dp() const { return Dp_; }
  double np0() const { return np0_; }
  double k() const { return k_; }
  double Cv() const { return Cv_; }
  double Cp() const { return Cp_; }
  double Mu0() const { return Mu0_; }
  double T0() const { return T0_; }
  double U0() const { return U0_; }
  double V0() const { return V0_; }
  double W0() const { return W0_; }
  double V0_1() const { return V0_1_; }
  double V0_2() const { return V0_2_; }
  double V0_3() const { return V0_3_; }
  double Rho0() const { return Rho0_; }
  double Rho_forV() const { return Rho_forV_; }
  double A() const { return A_; }
  double A1() const { return A1_; }
  double A2() const { return A2_; }
  double gx1() const { return gx1_; }
  double gx2() const { return gx2_; }
  double gy1() const { return gy1_; }
  double gy2() const { return gy2_; }
  double gz1() const { return gz1_; }
  double gz2() const { return gz2_; }
  double gx() const { return gx_; }
  double gy() const { return gy_; }
  double gz() const { return gz_; }
  double Vol() const { return Lx_ * Ly_ * Lz_; }
  int Np_track() const { return Np_track_; }
  int N0() const { return (int)(np0_ * Vol()); }          // params.h:120
  double Re() const { return Rho0_ * U0_ * W_ / Mu0_; }
  double Fr() const { return U0_ / std::sqrt(gz_ * W_); }
  double St() const { return N0() * mp_ / (Lx_ * Ly_ * Lz_ * Rho0_); }  // mass loading (params.h:123)
  double Pr() const { return Mu0_ * Cp_ / k_; }
  double Sc() const { return Rho0_ / (Mu0_ * D_M_); }
  double gamma() const { return Cp_ / Cv_; }
  double GAMMA() const { return Cvp_ / Cv_; }
  int Nx_tot() const { return Nx_tot_; }
  int Ny_tot() const { return Ny_tot_; }
  int Nz_tot() const { return Nz_tot_; }
  int bs() const { return bs_; }
  double T_final() const { return T_final_; }
  double dt() const { return dt_; }
  double threshold() const { return threshold_; }
  int data_freq_fast() const { return Data_freq_fast_; }
  int data_freq_slow() const { return Data_freq_slow_; }
  std::string data_dir() const { return Data_dir_; }
  std::string stat_dir() const { return Stat_dir_; }
  bool Statistics() const { return Statistics_; }
  bool Stat_print() const { return Stat_print_; }
  int Initial() const { return Initial_; }
  int Initial_C() const { return Initial_C_; }
  int Initial_V() const { return Initial_V_; }
  int PreCond() const { return PreCond_; }
  int Solver() const { return Solver_; }
  int cooling() const { return cooling_; }
  int Iteration() const { return Iteration_; }
  int Iteration1() const { return Iteration1_; }
  int Iteration2() const { return Iteration2_; }
  double epsilon() const { return epsilon_; }
  int TWC() const { return TWC_; }
  int ParticleGravity() const { return ParticleGravity_; }

  // the parameters the device step reads
  tpx_params to_abi() const;
};

// A single-process stand-in for the reference's MPI process grid (proc.h):
// this library runs one context per GPU with the whole box on it.
class proc {
 public:
  bool IsRoot() const { return true; }
  int RANK() const { return 0; }
  int TOT() const { return 1; }
  int ROOT() const { return 0; }
  int NX() const { return 1; }
  int NY() const { return 1; }
  int NZ() const { return 1; }
};

// Local grid extents (gridsize.h:5-56) for the single-process case: the
// local box is the whole box.
class gridsize {
  int Nx_tot_, Ny_tot_, Nz_tot_, bs_;
  double Lx_, Ly_, Lz_, dx_, dy_, dz_;

 public:
  gridsize(params *p, proc *pc);
  int Nx() const { return Nx_tot_; }
  int Ny() const { return Ny_tot_; }
  int Nz() const { return Nz_tot_; }
  int Nx_tot() const { return Nx_tot_; }
  int Ny_tot() const { return Ny_tot_; }
  int Nz_tot() const { return Nz_tot_; }
  int size_tot() const { return Nx_tot_ * Ny_tot_ * Nz_tot_; }
  int size() const { return Nx_tot_ * Ny_tot_ * Nz_tot_; }
  int bs() const { return bs_; }
  double Lx() const { return Lx_; }
  double Ly() const { return Ly_; }
  double Lz() const { return Lz_; }
  double dx() const { return dx_; }
  double dy() const { return dy_; }
  double dz() const { return dz_; }
  double Vcell() const { return dx_ * dy_ * dz_; }
  int il() const { return 0; }
  int ih() const { return Nx_tot_ - 1; }
  int jl() const { return 0; }
  int jh() const { return Ny_tot_ - 1; }
  int kl() const { return 0; }
  int kh() const { return Nz_tot_ - 1; }
  double xl() const { return 0.0; }
  double xh() const { return Lx_; }
  double yl() const { return 0.0; }
  double yh() const { return Ly_; }
  double zl() const { return 0.0; }
  double zh() const { return Lz_; }
};
