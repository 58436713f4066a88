// params.cpp -- input-file parser and derived coefficients (params.cpp:6-86).
#include "params.h"

#include <fstream>
#include <sstream>

params::params(const char *address) {
  std::ifstream in(address);
  std::ostringstream kept;
  std::string line;
  bool opened = (bool)in;
  while (opened && std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    kept << line << '\n';
  }
  std::istringstream s(kept.str());
  s >> Nx_tot_ >> Ny_tot_ >> Nz_tot_;
  s >> Lx_ >> Ly_ >> Lz_;
  s >> elongated_box_;
  s >> solve_for_scalar_;
  s >> solve_for_vector_;
  s >> bs_;
  s >> Np_track_;
  s >> T_final_ >> dt_ >> threshold_;
  s >> Statistics_ >> Stat_print_;
  s >> Data_freq_fast_ >> Data_freq_slow_;
  s >> Data_dir_ >> Stat_dir_;
  s >> Initial_;
  s >> Initial_C_;
  s >> Initial_V_;
  s >> PreCond_ >> Solver_;
  s >> cooling_;
  s >> Iteration1_ >> Iteration2_ >> epsilon_;
  s >> TWC_ >> ParticleGravity_;
  s >> A1_ >> A2_;
  s >> Rho0_;
  s >> Rho_forV_;
  s >> U0_ >> V0_ >> W0_;
  s >> V0_1_ >> V0_2_ >> V0_3_;
  s >> T0_;
  s >> Mu0_ >> k_;
  s >> gx1_ >> gy1_ >> gz1_ >> gx2_ >> gy2_ >> gz2_;
  s >> Cp_ >> Cv_;
  s >> np0_;
  s >> Dp_ >> Rhop_;
  s >> Cvp_;
  s >> epsilonp_;
  s >> Nu_;
  s >> I01_ >> I02_;
  s >> D_M_;
  s >> eta_;
  s >> s1t_;
  s >> s2t_;
  ok_ = opened && !s.fail();
  // derived quantities (params.cpp:61-72)
  R_ = Cp_ - Cv_;
  P0_ = Rho0_ * T0_ * R_;
  Tp_ = Rhop_ * Dp_ * Dp_ / (18 * Mu0_);
  mp_ = Rhop_ * PI * Dp_ * Dp_ * Dp_ / 6.;
  W_ = Ly_;
  Iteration_ = Iteration1_;
  A_ = A1_;
  gx_ = gx1_;
  gy_ = gy1_;
  gz_ = gz1_;
  I0_ = I01_;
}

void params::update(double time) {
  if (time > threshold_) {
    Iteration_ = Iteration2_;
    A_ = A2_;
    gx_ = gx2_;
    gy_ = gy2_;
    gz_ = gz2_;
    I0_ = I02_;
  }
}

tpx_params params::to_abi() const {
  tpx_params p{};
  p.nx = Nx_tot_; p.ny = Ny_tot_; p.nz = Nz_tot_;
  p.lx = Lx_; p.ly = Ly_; p.lz = Lz_;
  p.bs = bs_;
  p.dt = dt_; p.threshold = threshold_;
  p.iteration1 = Iteration1_; p.iteration2 = Iteration2_;
  p.epsilon = epsilon_;
  p.twc = TWC_; p.particle_gravity = ParticleGravity_; p.cooling = cooling_;
  p.a1 = A1_; p.a2 = A2_; p.rho0 = Rho0_;
  p.u0 = U0_; p.v0 = V0_; p.w0 = W0_;
  p.t0 = T0_; p.mu0 = Mu0_; p.k = k_;
  p.gx1 = gx1_; p.gy1 = gy1_; p.gz1 = gz1_; p.gx2 = gx2_; p.gy2 = gy2_; p.gz2 = gz2_;
  p.cp = Cp_; p.cv = Cv_; p.np0 = np0_; p.dp = Dp_; p.rhop = Rhop_; p.cvp = Cvp_;
  p.epsilonp = epsilonp_; p.nu = Nu_; p.i01 = I01_; p.i02 = I02_;
  p.solve_for_scalar = solve_for_scalar_ ? 1 : 0; p.s1_type = s1t_ != 0 ? 1 : 0; p.d_m = D_M_;
  p.solve_for_vector = solve_for_vector_ ? 1 : 0; p.s2_type = s2t_ != 0 ? 1 : 0;
  p.eta = eta_; p.rho_forv = Rho_forV_;
  return p;
}

gridsize::gridsize(params *p, proc *) {
  Nx_tot_ = p->Nx_tot(); Ny_tot_ = p->Ny_tot(); Nz_tot_ = p->Nz_tot();
  bs_ = p->bs();
  Lx_ = p->Lx(); Ly_ = p->Ly(); Lz_ = p->Lz();
  dx_ = Lx_ / Nx_tot_; dy_ = Ly_ / Ny_tot_; dz_ = Lz_ / Nz_tot_;  // gridsize.cpp:19-25
}
