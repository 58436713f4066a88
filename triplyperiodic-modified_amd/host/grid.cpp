// grid.cpp -- host mirror of grid's substep methods, forwarding to the C ABI.
#include "grid.h"

#include <cmath>
#include <iostream>
#include <sstream>

grid::grid(gridsize *size, params *param, proc *pc, communicator *com)
    : param_(param), pc_(pc), size_(size), com_(com) {
  tpx_params p = param_->to_abi();
  int64_t cap = param_->N0() > 0 ? param_->N0() : 1;
  tpx_check(nullptr, tpx_create(&p, cap, &ctx_), "grid (tpx_create)");
  part.bind(&ctx_, param_);
  RU.bind(&ctx_, size_, TPX_RU);
  RU_int.bind(&ctx_, size_, TPX_RU_INT);
  RU_new.bind(&ctx_, size_, TPX_RU_NEW);
  RU_np1.bind(&ctx_, size_, TPX_RU_NP1);
  RHS_RU.bind(&ctx_, size_, TPX_RHS_RU);
  P.bind(&ctx_, size_, TPX_P, 0);
  RHS_Pois.bind(&ctx_, size_, TPX_RHS_POIS, 0);
  C.bind(&ctx_, size_, TPX_C, 0);
  Rho.bind(&ctx_, size_, TPX_RHO, 0);
  Rho_int.bind(&ctx_, size_, TPX_RHO_INT, 0);
  Rho_new.bind(&ctx_, size_, TPX_RHO_NEW, 0);
  Rho_np1.bind(&ctx_, size_, TPX_RHO_NP1, 0);
  divergence.bind(&ctx_, size_, TPX_DIVERGENCE, 0);
  RHS_Part_Temp.bind(&ctx_, size_, TPX_RHS_PART_TEMP, 0);
  Passive_Scalar.bind(&ctx_, size_, TPX_PASSIVE_SCALAR, 0);
  Passive_Scalar_int.bind(&ctx_, size_, TPX_PASSIVE_SCALAR_INT, 0);
  Passive_Scalar_new.bind(&ctx_, size_, TPX_PASSIVE_SCALAR_NEW, 0);
  Passive_Scalar_np1.bind(&ctx_, size_, TPX_PASSIVE_SCALAR_NP1, 0);
  if (param_->Statistics()) {
    ok(tpx_statistics_enable(ctx_, 1), "tpx_statistics_enable");
    if (pc_->IsRoot()) {  // grid.cpp:32-57 and open_stat_file (grid.cpp:1147-1154)
      auto open = [&](const char *name, std::ofstream &f) {
        f.open(param_->stat_dir() + name + ".dat", std::ios::out | std::ios::app);
      };
      open("Tg", stat_Tg);
      open("Tp", stat_Tp);
      open("HT", stat_HT);
      open("RU", stat_TKE2);
      open("TKE_U", stat_TKE_U);
      open("TKE_V", stat_TKE_V);
      open("TKE_W", stat_TKE_W);
      open("Passive_Scalar_mean", stat_Passive_Scalar_mean);
      open("TKE_V1", stat_TKEV_V1);
      open("TKE_V2", stat_TKEV_V2);
      open("TKE_V3", stat_TKEV_V3);
      open("P0", stat_P0);
      open("C_Max", stat_CMax);
      open("C_Min", stat_CMin);
      open("C_Mean", stat_CMean);
      open("Rho_Max", stat_RhoMax);
      open("Rho_Min", stat_RhoMin);
      open("Rho_Mean", stat_RhoMean);
      open("Max_CFL_Vp", stat_ParticleMaxCFL);
      open("Max_CFL_U", stat_GasMaxCFL);
      open("Num_Iteration", stat_NumIteration);
      open("Balance_Index", stat_BalanceIndex);
    }
  }
}

grid::~grid() {
  if (ctx_) tpx_destroy(ctx_);
}

void grid::unsupported(const char *what) {
  throw std::runtime_error(std::string(what) + " is outside this build (the RK4 step of the particle-laden gas only)");
}

void grid::pull_scalars() {
  double s[TPX_S_COUNT];
  ok(tpx_get_scalars(ctx_, s), "tpx_get_scalars");
  P0 = s[TPX_S_P0];
  P0_int = s[TPX_S_P0_INT];
  P0_new = s[TPX_S_P0_NEW];
  P0_np1 = s[TPX_S_P0_NP1];
  dP0_dt = s[TPX_S_DP0_DT];
  T_cur = s[TPX_S_T_CUR];
  num_timestep = (int)s[TPX_S_NUM_TIMESTEP];
}

void grid::Initialize() {
  int init = param_->Initial();
  const gridsize &g = *size_;
  if (init == 1) {
    // restart files in the working directory (grid.cpp:221-238)
    com_->read(RU, "Restart_RU.bin");
    com_->read(Rho, "Restart_Rho.bin");
    com_->read(P, "Restart_P.bin");
    part.Load_All();
    std::ifstream numbers("Restart_numbers.dat");
    if (!numbers) throw std::runtime_error("cannot read Restart_numbers.dat");
    double p0 = 0, t = 0;
    int n = 0;
    numbers >> p0 >> t >> n;
    ok(tpx_set_restart_numbers(ctx_, p0, t, n), "restart numbers");
  } else if (init >= 2 && init <= 5) {
    // uniform density, zero pressure and one of the analytic velocity
    // fields of grid.cpp:122-212, particles at rest from load_random
    std::vector<double> ru[3];
    for (auto &c : ru) c.assign(RU.x.ghosted_size(), 0.0);
    int b = g.bs();
    long gx = g.Nx() + 2 * b, gy = g.Ny() + 2 * b;
    double kx = TWO_PI / g.Lx(), ky = TWO_PI / g.Ly(), kz = TWO_PI / g.Lz();
    for (int k = 0; k < g.Nz(); ++k)
      for (int j = 0; j < g.Ny(); ++j)
        for (int i = 0; i < g.Nx(); ++i) {
          long I = ((long)(k + b) * gy + (j + b)) * gx + (i + b);
          double X = g.dx() / 2. + i * g.dx(), Y = g.dy() / 2. + j * g.dy(), Z = g.dz() / 2. + k * g.dz();
          if (init == 5) {
            ru[0][I] = ky * std::sin(kx * (X - g.dx() / 2.)) * std::cos(ky * Y);
            ru[1][I] = -kx * std::cos(kx * X) * std::sin(ky * (Y - g.dy() / 2.));
          } else if (init == 4) {
            ru[0][I] = -kz * std::cos(kz * Z) * std::sin(kx * (X - g.dx() / 2.));
            ru[2][I] = kx * std::sin(kz * (Z - g.dz() / 2.)) * std::cos(kx * X);
          } else if (init == 3) {
            ru[1][I] = kz * std::sin(ky * (Y - g.dy() / 2.)) * std::cos(kz * Z);
            ru[2][I] = -ky * std::cos(ky * Y) * std::sin(kz * (Z - g.dz() / 2.));
          }
        }
    RU.x.upload(ru[0].data());
    RU.y.upload(ru[1].data());
    RU.z.upload(ru[2].data());
    std::vector<double> rho(Rho.ghosted_size(), param_->Rho0()), zero(P.ghosted_size(), 0.0);
    Rho.upload(rho.data());
    P.upload(zero.data());
    part.load_random(g.Lx(), g.Ly(), g.Lz());
    ok(tpx_set_restart_numbers(ctx_, param_->P0(), 0.0, 0), "restart numbers");
  } else {
    unsupported("Initial = 0 (initial turbulence from U/V/W files)");
  }
  // the passive scalar's initial field (grid.cpp:385-457), read or set when
  // it is solved: Initial_C 0 and 1 read Passive_Scalar.bin and
  // Restart_Passive_Scalar.bin, 3-5 are products of two sines
  int ic = param_->Initial_C();
  if (param_->solve_for_scalar()) {
    if (ic == 0) com_->read(Passive_Scalar, "Passive_Scalar.bin");
    else if (ic == 1) com_->read(Passive_Scalar, "Restart_Passive_Scalar.bin");
    else if (ic >= 3 && ic <= 5) {
      std::vector<double> c(Passive_Scalar.ghosted_size(), 0.0);
      int b = g.bs();
      long gx = g.Nx() + 2 * b, gy = g.Ny() + 2 * b;
      double kx = TWO_PI / g.Lx(), ky = TWO_PI / g.Ly(), kz = TWO_PI / g.Lz();
      for (int k = 0; k < g.Nz(); ++k)
        for (int j = 0; j < g.Ny(); ++j)
          for (int i = 0; i < g.Nx(); ++i) {
            double X = g.dx() / 2. + i * g.dx(), Y = g.dy() / 2. + j * g.dy(), Z = g.dz() / 2. + k * g.dz();
            double v = ic == 5 ? std::sin(kx * X) * std::sin(ky * Y)
                     : ic == 4 ? std::sin(kz * Z) * std::sin(kx * X) : std::sin(ky * Y) * std::sin(kz * Z);
            c[((long)(k + b) * gy + (j + b)) * gx + (i + b)] = v;
          }
      Passive_Scalar.upload(c.data());
    }
  }
  ok(tpx_initialize(ctx_), "Initialize");
  pull_scalars();
}

void grid::Update_Rho() {
  // the driver calls PARAM.update(T_cur) before the substeps; the device
  // keeps its own copy of the switched coefficients and applies the same rule
  if (RK4_count == 0) ok(tpx_params_update(ctx_), "params update");
  ok(tpx_update_rho(ctx_, RK4_count), "Update_Rho");
}

void grid::Update_P0() {
  ok(tpx_update_p0(ctx_, RK4_count), "Update_P0");
}

void grid::Update_Particle() { ok(tpx_update_particle(ctx_, RK4_count), "Update_Particle"); }
void grid::Update_Passive_Scalar() { ok(tpx_update_passive_scalar(ctx_, RK4_count), "Update_Passive_Scalar"); }
void grid::Update_RU_WOP() { ok(tpx_update_ru_wop(ctx_, RK4_count), "Update_RU_WOP"); }
void grid::Compute_Div_U_new() { ok(tpx_compute_div_u_new(ctx_, RK4_count), "Compute_Div_U_new"); }
void grid::Compute_RHS_Pois() { ok(tpx_compute_rhs_pois(ctx_, RK4_count), "Compute_RHS_Pois"); }

void grid::Solve_Poisson() {
  tensor1 C_face;  // the device forms C = 1/Rho_face itself (grid.cpp:853)
  PS_.Solve(C_face, P, RHS_Pois);
}

void grid::Update_RU_WP() { ok(tpx_update_ru_wp(ctx_, RK4_count), "Update_RU_WP"); }

void grid::TimeAdvance_RK4() {
  ok(tpx_time_advance_rk4(ctx_), "TimeAdvance_RK4");
}

void grid::TimeAdvance() {
  ok(tpx_time_advance(ctx_), "TimeAdvance");
  pull_scalars();
}

// grid::Statistics (grid.cpp:902-982): the reductions run on the device
// (tpx_statistics); the particle trajectory file (part.trajectory) is not
// written by this build.
void grid::Statistics() {
  if (!param_->Statistics()) return;
  double st[TPX_ST_COUNT];
  ok(tpx_statistics(ctx_, st), "tpx_statistics");
  pull_scalars();
  if (!pc_->IsRoot()) return;
  auto line = [&](std::ofstream &f, double v) { f << T_cur << " " << v << std::endl; };
  line(stat_Tg, st[TPX_ST_TG_MEAN]);
  line(stat_Tp, st[TPX_ST_TP_MEAN]);
  line(stat_HT, st[TPX_ST_HT_MEAN]);
  line(stat_TKE2, st[TPX_ST_TKE2]);
  line(stat_TKE_U, st[TPX_ST_TKE_U]);
  line(stat_TKE_V, st[TPX_ST_TKE_V]);
  line(stat_TKE_W, st[TPX_ST_TKE_W]);
  line(stat_Passive_Scalar_mean, st[TPX_ST_PASSIVE_SCALAR_MEAN]);
  line(stat_TKEV_V1, st[TPX_ST_TKEV_V1]);
  line(stat_TKEV_V2, st[TPX_ST_TKEV_V2]);
  line(stat_TKEV_V3, st[TPX_ST_TKEV_V3]);
  line(stat_P0, st[TPX_ST_P0]);
  line(stat_CMax, st[TPX_ST_C_MAX]);
  line(stat_CMin, st[TPX_ST_C_MIN]);
  line(stat_CMean, st[TPX_ST_C_MEAN]);
  line(stat_RhoMax, st[TPX_ST_RHO_MAX]);
  line(stat_RhoMin, st[TPX_ST_RHO_MIN]);
  line(stat_RhoMean, st[TPX_ST_RHO_MEAN]);
  line(stat_ParticleMaxCFL, st[TPX_ST_PARTICLE_MAX_CFL]);
  line(stat_GasMaxCFL, st[TPX_ST_GAS_MAX_CFL]);
  line(stat_GasMaxDiffCFL, st[TPX_ST_GAS_MAX_DIFF_CFL]);
  line(stat_NumIteration, st[TPX_ST_NUM_ITERATION]);
  line(stat_BalanceIndex, st[TPX_ST_BALANCE_INDEX]);
  if (!param_->Stat_print()) return;
  const double tke = st[TPX_ST_TKE_U] + st[TPX_ST_TKE_V] + st[TPX_ST_TKE_W];
  std::cout << std::endl << "::::::::::TIME=" << T_cur << "::::::::::STEP=" << num_timestep << "::::::::::" << std::endl;
  std::cout << "*** Rho_min=" << st[TPX_ST_RHO_MIN] << "  ,  Rho_max=" << st[TPX_ST_RHO_MAX]
            << "  ,  Rho_mean=" << st[TPX_ST_RHO_MEAN] << std::endl;
  std::cout << "*** C_min=" << st[TPX_ST_C_MIN] << "  ,  C_max=" << st[TPX_ST_C_MAX] << "  ,  C_mean=" << st[TPX_ST_C_MEAN]
            << std::endl;
  std::cout << "*** Particle Maximum CFL=" << st[TPX_ST_PARTICLE_MAX_CFL] << "  ,  Gas Maximum CFL="
            << st[TPX_ST_GAS_MAX_CFL] << "  ,  Gas Maximum diffusive CFL=" << st[TPX_ST_GAS_MAX_DIFF_CFL] << std::endl;
  std::cout << "*** P0=" << st[TPX_ST_P0] << "   ,   Number of Poisson solve iterations=" << st[TPX_ST_NUM_ITERATION]
            << std::endl;
  std::cout << "*** Twice TKE_U =" << st[TPX_ST_TKE_U] << "  ,  TKE_V =" << st[TPX_ST_TKE_V] << "  ,  TKE_W ="
            << st[TPX_ST_TKE_W] << "  , Twice TKE =" << tke << "  , Twice TKE2=" << st[TPX_ST_TKE2] << std::endl;
  std::cout << "*** Passive_Scalar_mean =" << st[TPX_ST_PASSIVE_SCALAR_MEAN] << std::endl;
  std::cout << "*** Twice TKEV_V1 =" << st[TPX_ST_TKEV_V1] << "  ,  TKEV_V2 =" << st[TPX_ST_TKEV_V2] << "  ,  TKEV_V3 ="
            << st[TPX_ST_TKEV_V3] << "  , Twice TKEV =" << 0.0 << "  , Twice VV=" << st[TPX_ST_VV] << std::endl;
  std::cout << "*** Particle u_max=" << st[TPX_ST_VP_MAX] << "  ,  Gas interpolated u_max=" << st[TPX_ST_UG_MAX]
            << "  ,  Gas u_max=" << st[TPX_ST_U_MAX] << std::endl;
  std::cout << "Mean energy transferred from particle to gas=" << -st[TPX_ST_HT_MEAN]
            << "  ,  Balance_Index=" << st[TPX_ST_BALANCE_INDEX] << std::endl;
  std::cout << "*** Particle Tmax =" << st[TPX_ST_TP_MAX] << std::endl;
}

bool grid::Touch() {
  bool ans = false;
  std::ifstream f("touch.check");
  if (f) f >> ans;
  if (ans) std::cout << "Someone touched me! Simulation stopped." << std::endl;
  Is_touch_ = ans;
  Store();
  return ans;
}

void grid::Store() {
  auto name = [&](const char *stem, int n, const char *ext) {
    std::ostringstream s;
    s << param_->data_dir() << stem << "_" << n << ext;
    return s.str();
  };
  if (num_timestep % param_->data_freq_fast() == 0 || Is_touch_) {
    com_->write(RU, name("RU", num_timestep, ".bin").c_str());
    // the passive scalar as C_n.bin (grid.cpp:538-544; RV_n.bin: no passive vector here)
    if (param_->solve_for_scalar()) com_->write(Passive_Scalar, name("C", num_timestep, ".bin").c_str());
    std::ofstream numbers(name("numbers", num_timestep, ".dat"));
    numbers.precision(16);
    numbers << P0 << std::endl << T_cur << std::endl << num_timestep << std::endl;
  }
  if (num_timestep % param_->data_freq_slow() == 0 || Is_touch_) {
    part.Store_All(num_timestep);
    com_->write(P, name("P", num_timestep, ".bin").c_str());
    com_->write(Rho, name("Rho", num_timestep, ".bin").c_str());
  }
}
