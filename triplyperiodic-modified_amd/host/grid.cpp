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
  RV.bind(&ctx_, size_, TPX_RV);
  RV_int.bind(&ctx_, size_, TPX_RV_INT);
  RV_new.bind(&ctx_, size_, TPX_RV_NEW);
  RV_np1.bind(&ctx_, size_, TPX_RV_NP1);
  Q.bind(&ctx_, size_, TPX_Q, 0);
  RHS_Pois_Q.bind(&ctx_, size_, TPX_RHS_POIS_Q, 0);
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
      // particle.cpp:34-40: the trajectory stream lives in the stat directory
      if (param_->Np_track() > 0)
        stat_trajectory.open(param_->stat_dir() + "Part_trajectory.dat", std::ios::out | std::ios::app);
    }
  }
}

grid::~grid() {
  if (ctx_) tpx_destroy(ctx_);
}

void grid::unsupported(const char *what) {
  throw std::runtime_error(std::string(what) + " is outside this library (the RK4 step of the particle-laden gas only)");
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
  } else if (init == 0) {
    // velocity from U.bin, V.bin, W.bin times Rho0, uniform density, zero
    // pressure, particles at rest from load_random (grid.cpp:240-253)
    tensor0 *comp[3] = {&RU.x, &RU.y, &RU.z};
    const char *file[3] = {"U.bin", "V.bin", "W.bin"};
    for (int c = 0; c < 3; ++c) {
      std::vector<double> f = com_->read_host(file[c]);
      for (auto &v : f) v *= param_->Rho0();  // RU *= Rho0 (tensor1 operator*=, all elements)
      comp[c]->upload(f.data());
    }
    std::vector<double> rho(Rho.ghosted_size(), param_->Rho0()), zero(P.ghosted_size(), 0.0);
    Rho.upload(rho.data());
    P.upload(zero.data());
    part.load_random(g.Lx(), g.Ly(), g.Lz());
    ok(tpx_set_restart_numbers(ctx_, param_->P0(), 0.0, 0), "restart numbers");
  } else {
    unsupported("this value of Initial");
  }
  initial_vector();
  // the passive scalar's initial field (grid.cpp:385-457), read or set when
  // it is solved: Initial_C 0 and 1 read Passive_Scalar.bin and
  // Restart_Passive_Scalar.bin, 3-5 are products of two sines
  // the passive scalar's initial field (grid.cpp:385-457), set whether or not
  // it is solved (Store writes it as C_<n>.bin either way): Initial_C 0 and 1
  // read Passive_Scalar.bin and Restart_Passive_Scalar.bin, 3-5 are products
  // of two sines; 2 leaves it unset in the reference (zero here).  Unsolved,
  // it stays on the host.
  int ic = param_->Initial_C();
  std::vector<double> c0(Passive_Scalar.ghosted_size(), 0.0);
  if (ic == 0) c0 = com_->read_host("Passive_Scalar.bin");
  else if (ic == 1) c0 = com_->read_host("Restart_Passive_Scalar.bin");
  else if (ic >= 3 && ic <= 5) {
    int b = g.bs();
    long gx = g.Nx() + 2 * b, gy = g.Ny() + 2 * b;
    double kx = TWO_PI / g.Lx(), ky = TWO_PI / g.Ly(), kz = TWO_PI / g.Lz();
    for (int k = 0; k < g.Nz(); ++k)
      for (int j = 0; j < g.Ny(); ++j)
        for (int i = 0; i < g.Nx(); ++i) {
          double X = g.dx() / 2. + i * g.dx(), Y = g.dy() / 2. + j * g.dy(), Z = g.dz() / 2. + k * g.dz();
          double v = ic == 5 ? std::sin(kx * X) * std::sin(ky * Y)
                   : ic == 4 ? std::sin(kz * Z) * std::sin(kx * X) : std::sin(ky * Y) * std::sin(kz * Z);
          c0[((long)(k + b) * gy + (j + b)) * gx + (i + b)] = v;
        }
  }
  if (param_->solve_for_scalar()) Passive_Scalar.upload(c0.data());
  else ps_host_ = c0;
  ok(tpx_initialize(ctx_), "Initialize");
  pull_scalars();
}

// the passive vector's initial field and Q (grid.cpp:256-383): Initial_V 0
// reads V1/V2/V3.bin times Rho_forV, 1 reads Restart_RV.bin and
// Restart_Q.bin, 2 is zero, 3-5 are the reference's analytic fields; Q is
// zero except when read.  Unsolved, both stay on the host for Store.
void grid::initial_vector() {
  const gridsize &g = *size_;
  const size_t m = com_->ghosted_count();
  std::vector<double> rv[3] = {std::vector<double>(m, 0.0), std::vector<double>(m, 0.0), std::vector<double>(m, 0.0)};
  std::vector<double> q(m, 0.0);
  int iv = param_->Initial_V();
  if (iv == 0) {
    const char *file[3] = {"V1.bin", "V2.bin", "V3.bin"};
    for (int c = 0; c < 3; ++c) {
      rv[c] = com_->read_host(file[c]);
      for (auto &v : rv[c]) v *= param_->Rho_forV();
    }
  } else if (iv == 1) {
    for (int c = 0; c < 3; ++c) rv[c] = com_->read_host("Restart_RV.bin", c);
    q = com_->read_host("Restart_Q.bin");
  } else if (iv >= 3 && iv <= 5) {
    int b = g.bs();
    long gx = g.Nx() + 2 * b, gy = g.Ny() + 2 * b;
    double kx = TWO_PI / g.Lx(), ky = TWO_PI / g.Ly(), kz = TWO_PI / g.Lz();
    for (int k = 0; k < g.Nz(); ++k)
      for (int j = 0; j < g.Ny(); ++j)
        for (int i = 0; i < g.Nx(); ++i) {
          long I = ((long)(k + b) * gy + (j + b)) * gx + (i + b);
          double X = g.dx() / 2. + i * g.dx(), Y = g.dy() / 2. + j * g.dy(), Z = g.dz() / 2. + k * g.dz();
          if (iv == 5) {
            rv[0][I] = ky * std::sin(kx * (X - g.dx() / 2.)) * std::cos(ky * Y);
            rv[1][I] = -kx * std::cos(kx * X) * std::sin(ky * (Y - g.dy() / 2.));
          } else if (iv == 4) {
            rv[0][I] = -kz * std::cos(kz * Z) * std::sin(kx * (X - g.dx() / 2.));
            rv[2][I] = kx * std::sin(kz * (Z - g.dz() / 2.)) * std::cos(kx * X);
          } else {
            rv[1][I] = kz * std::sin(ky * (Y - g.dy() / 2.)) * std::cos(kz * Z);
            rv[2][I] = -ky * std::cos(ky * Y) * std::sin(kz * (Z - g.dz() / 2.));
          }
        }
  }
  if (param_->solve_for_vector()) {
    RV.x.upload(rv[0].data());
    RV.y.upload(rv[1].data());
    RV.z.upload(rv[2].data());
    Q.upload(q.data());
  } else {
    for (int c = 0; c < 3; ++c) rv_host_[c] = rv[c];
    q_host_ = q;
  }
}

void grid::write_ghosted(const std::vector<double> &f, const std::string &name) {
  com_->write_host(f, name.c_str(), false);
}

void grid::Update_RV_WOQ() { ok(tpx_update_rv_woq(ctx_, RK4_count), "Update_RV_WOQ"); }
void grid::Compute_RHS_Pois_Q() { ok(tpx_compute_rhs_pois_q(ctx_, RK4_count), "Compute_RHS_Pois_Q"); }
void grid::Solve_Poisson_Q() {
  // PS_.Solve(1/Rho_forV, Q, RHS_Pois_Q) (grid.cpp:703-707), on the device
  int it = 0;
  ok(tpx_solve_poisson_q(ctx_, RK4_count, &it), "Solve_Poisson_Q");
}
void grid::Update_RV_WQ() { ok(tpx_update_rv_wq(ctx_, RK4_count), "Update_RV_WQ"); }

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
// (tpx_statistics), the tracked particles' positions for the trajectory file
// (part.trajectory, particle.cpp:717-745) come from tpx_track_particles.
void grid::Statistics() {
  if (!param_->Statistics()) return;
  double st[TPX_ST_COUNT];
  ok(tpx_statistics(ctx_, st), "tpx_statistics");
  pull_scalars();
  if (!param_->solve_for_scalar()) {
    // Passive_Scalar.mean_squares() of the unsolved field (grid.cpp:913)
    const gridsize &g = *size_;
    const int b = g.bs();
    const long gx = g.Nx() + 2 * b, gy = g.Ny() + 2 * b;
    double sq = 0.0;
    for (int i = 0; i < g.Nx(); ++i)
      for (int j = 0; j < g.Ny(); ++j)
        for (int k = 0; k < g.Nz(); ++k) {
          const double v = ps_host_[((long)(k + b) * gy + (j + b)) * gx + (i + b)];
          sq += v * v;
        }
    st[TPX_ST_PASSIVE_SCALAR_MEAN] = sq / (double)g.size_tot();
  }
  if (!pc_->IsRoot()) return;
  if (param_->Np_track() > 0) {
    // part.trajectory(T_cur) (grid.cpp:939, particle.cpp:717-745): flags 0, K,
    // 2K, ... below N0 with K = ceil(N0 / Np_track)
    const int n0 = param_->N0(), K = (n0 + param_->Np_track() - 1) / param_->Np_track();
    const int count = K > 0 ? (n0 + K - 1) / K : 0;
    std::vector<double> xyz(3 * (size_t)count);
    if (count > 0) ok(tpx_track_particles(ctx_, K, count, xyz.data()), "part.trajectory");
    stat_trajectory << T_cur << " ";
    for (double v : xyz) stat_trajectory << v << " ";
    stat_trajectory << std::endl;
  }
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
  const gridsize &g = *size_;
  const size_t m = com_->ghosted_count();
  const int b = g.bs();
  const long gx = g.Nx() + 2 * b, gy = g.Ny() + 2 * b;
  auto interior = [&](auto &&fn) {
    for (int k = 0; k < g.Nz(); ++k)
      for (int j = 0; j < g.Ny(); ++j)
        for (int i = 0; i < g.Nx(); ++i) fn(((long)(k + b) * gy + (j + b)) * gx + (i + b), i, j, k);
  };
  if (num_timestep == 1) {
    // grid.cpp:491-520: the cell-centre coordinates XYZ.bin, and the sources
    // S1.bin (the scalar's C_Source minus its mean when it is solved with
    // S1_type 0) and S2.bin (V_Source minus its means for the vector with
    // S2_type 0); zero where the reference never computes them
    std::vector<double> xyz[3] = {std::vector<double>(m, 0.0), std::vector<double>(m, 0.0), std::vector<double>(m, 0.0)};
    std::vector<double> s1(m, 0.0), s2y(m, 0.0), zero(m, 0.0);
    interior([&](long I, int i, int j, int k) {
      xyz[0][I] = g.dx() / 2. + i * g.dx();
      xyz[1][I] = g.dy() / 2. + j * g.dy();
      xyz[2][I] = g.dz() / 2. + k * g.dz();
    });
    const std::string xyzf = param_->data_dir() + "XYZ.bin";
    com_->write_host(xyz[0], xyzf.c_str(), false);
    com_->write_host(xyz[1], xyzf.c_str(), true);
    com_->write_host(xyz[2], xyzf.c_str(), true);
    auto cos_minus_mean = [&](std::vector<double> &f, bool centre_first) {
      double sum = 0.0;
      interior([&](long I, int i, int, int) {
        const double X = centre_first ? g.dx() / 2. + i * g.dx() : i * g.dx() + g.dx() / 2.;
        f[I] = std::cos(X);
      });
      for (int i = 0; i < g.Nx(); ++i)  // tensor0::sum's loop order (i, j, k)
        for (int j = 0; j < g.Ny(); ++j)
          for (int k = 0; k < g.Nz(); ++k) sum += f[((long)(k + b) * gy + (j + b)) * gx + (i + b)];
      const double mean = sum / (double)g.size_tot();
      interior([&](long I, int, int, int) { f[I] -= mean; });
    };
    if (param_->solve_for_scalar() && param_->S1_type() == 0) cos_minus_mean(s1, true);
    if (param_->solve_for_vector() && param_->S2_type() == 0) cos_minus_mean(s2y, false);
    write_ghosted(s1, param_->data_dir() + "S1.bin");
    const std::string s2f = param_->data_dir() + "S2.bin";
    com_->write_host(zero, s2f.c_str(), false);
    com_->write_host(s2y, s2f.c_str(), true);
    com_->write_host(zero, s2f.c_str(), true);
  }
  if (num_timestep % param_->data_freq_fast() == 0 || Is_touch_) {
    // grid.cpp:523-557: RU, RV and the passive scalar C every fast dump
    com_->write(RU, name("RU", num_timestep, ".bin").c_str());
    const std::string rvf = name("RV", num_timestep, ".bin");
    if (param_->solve_for_vector()) com_->write(RV, rvf.c_str());
    else {
      com_->write_host(rv_host_[0], rvf.c_str(), false);
      com_->write_host(rv_host_[1], rvf.c_str(), true);
      com_->write_host(rv_host_[2], rvf.c_str(), true);
    }
    if (param_->solve_for_scalar()) com_->write(Passive_Scalar, name("C", num_timestep, ".bin").c_str());
    else write_ghosted(ps_host_, name("C", num_timestep, ".bin"));
    std::ofstream numbers(name("numbers", num_timestep, ".dat"));
    numbers.precision(16);
    numbers << P0 << std::endl << T_cur << std::endl << num_timestep << std::endl;
  }
  if (num_timestep % param_->data_freq_slow() == 0 || Is_touch_) {
    // grid.cpp:560-592: particles, P, Q and Rho every slow dump
    part.Store_All(num_timestep);
    com_->write(P, name("P", num_timestep, ".bin").c_str());
    if (param_->solve_for_vector()) com_->write(Q, name("Q", num_timestep, ".bin").c_str());
    else write_ghosted(q_host_, name("Q", num_timestep, ".bin"));
    com_->write(Rho, name("Rho", num_timestep, ".bin").c_str());
  }
}
