// simulation.cpp -- the driver loop of the reference (simulation.cpp:16-120)
// on the device library, with Update_Particle enabled and one process per
// GPU (no MPI: the box lives on one device).
//
//   tpx_sim <input file>            run to T_final (or until touch.check says 1)
//   tpx_sim --print-params <file>   parse the file and print the values read
#include <chrono>
#include <cstring>
#include <iomanip>
#include <iostream>

#include "grid.h"

using namespace std::chrono;

static int print_params(const char *file) {
  params P(file);
  if (!P.ok()) {
    std::cerr << "could not read every parameter group from " << file << std::endl;
    return 1;
  }
  tpx_params a = P.to_abi();
  std::cout << std::setprecision(17);
  std::cout << "nx " << a.nx << "\nny " << a.ny << "\nnz " << a.nz << "\nlx " << a.lx << "\nly " << a.ly
            << "\nlz " << a.lz << "\nbs " << a.bs << "\ndt " << a.dt << "\nthreshold " << a.threshold
            << "\niteration1 " << a.iteration1 << "\niteration2 " << a.iteration2 << "\nepsilon " << a.epsilon
            << "\ntwc " << a.twc << "\nparticle_gravity " << a.particle_gravity << "\ncooling " << a.cooling
            << "\na1 " << a.a1 << "\na2 " << a.a2 << "\nrho0 " << a.rho0 << "\nu0 " << a.u0 << "\nv0 " << a.v0
            << "\nw0 " << a.w0 << "\nt0 " << a.t0 << "\nmu0 " << a.mu0 << "\nk " << a.k << "\ngx1 " << a.gx1
            << "\ngy1 " << a.gy1 << "\ngz1 " << a.gz1 << "\ngx2 " << a.gx2 << "\ngy2 " << a.gy2 << "\ngz2 "
            << a.gz2 << "\ncp " << a.cp << "\ncv " << a.cv << "\nnp0 " << a.np0 << "\ndp " << a.dp << "\nrhop "
            << a.rhop << "\ncvp " << a.cvp << "\nepsilonp " << a.epsilonp << "\nnu " << a.nu << "\ni01 "
            << a.i01 << "\ni02 " << a.i02 << "\nsolve_for_scalar " << a.solve_for_scalar << "\ns1_type "
            << a.s1_type << "\nd_m " << a.d_m << "\nsolve_for_vector " << a.solve_for_vector
            << "\ns2_type " << a.s2_type << "\neta " << a.eta << "\nrho_forv " << a.rho_forv << "\nT_final " << P.T_final() << "\nN0 " << P.N0() << "\nTp "
            << P.Tp() << "\nmp " << P.mp() << "\nP0 " << P.P0() << "\nInitial " << P.Initial() << std::endl;
  return 0;
}

int main(int argc, char *argv[]) {
  if (argc == 3 && std::strcmp(argv[1], "--print-params") == 0) return print_params(argv[2]);
  if (argc != 2) {
    std::cout << "Please supply a parameter file!" << std::endl;
    return 1;
  }
  try {
    params PARAM(argv[1]);
    if (!PARAM.ok()) {
      std::cout << "could not read the parameter file " << argv[1] << std::endl;
      return 1;
    }
    double Rho0_ = PARAM.Rho0();
    double RhoU_[3] = {PARAM.U0() * Rho0_, PARAM.V0() * Rho0_, PARAM.W0() * Rho0_};
    proc PROC;
    gridsize GSIZE(&PARAM, &PROC);
    communicator COMM(&GSIZE, &PARAM, &PROC);
    grid GRID(&GSIZE, &PARAM, &PROC, &COMM);
    GRID.Initialize();
    auto start = high_resolution_clock::now();
    int steps = 0;
    do {
      PARAM.update(GRID.T_cur);
      for (GRID.RK4_count = 0; GRID.RK4_count < 4; GRID.RK4_count++) {
        GRID.Update_Rho();
        GRID.Update_P0();
        GRID.Update_Particle();
        if (PARAM.solve_for_scalar()) GRID.Update_Passive_Scalar();
        if (PARAM.solve_for_vector()) {
          GRID.Update_RV_WOQ();
          GRID.Compute_RHS_Pois_Q();
          GRID.Solve_Poisson_Q();
          GRID.Update_RV_WQ();
        }
        GRID.Update_RU_WOP();
        GRID.Compute_Div_U_new();
        GRID.Compute_RHS_Pois();
        GRID.Solve_Poisson();
        GRID.Update_RU_WP();
        GRID.TimeAdvance_RK4();
      }
      GRID.RU_np1.make_mean_U0(RhoU_);
      if (PARAM.elongated_box() == 1) {
        GRID.RU_np1.y.kill_strong_modes();
        GRID.RU_np1.z.kill_strong_modes();
      }
      if (PARAM.elongated_box() == 2) {
        if (GRID.num_timestep % 100 == 0) GRID.CopyBox();
      }
      GRID.TimeAdvance();
      GRID.Statistics();
      ++steps;
    } while ((GRID.T_cur < PARAM.T_final()) && (!GRID.Touch()));
    tpx_synchronize(GRID.context());
    double seconds = duration_cast<duration<double>>(high_resolution_clock::now() - start).count();
    std::cout << "run time: " << seconds << " average time per RK4 step : " << seconds / steps << std::endl;
  } catch (const std::exception &e) {
    std::cout << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
