// src/simulation_data.h:25-100 / src/simulation_data.cpp:20-138: the
// immutable constants of a body, same class and accessor names.
#ifndef MFREE_SIMULATION_DATA_H_
#define MFREE_SIMULATION_DATA_H_

#include <math.h>

class physical_constants {
  double m_nu = 0., m_E = 0., m_rho0 = 0.;

 public:
  physical_constants() {}
  physical_constants(double nu, double E, double rho0) : m_nu(nu), m_E(E), m_rho0(rho0) {}
  double nu() const { return m_nu; }
  double E() const { return m_E; }
  double G() const { return m_E / (2. * (1. + m_nu)); }
  double K() const { return 2.0 * G() * (1 + m_nu) / (3 * (1 - 2 * m_nu)); }
  double rho0() const { return m_rho0; }
  double c0() const { return sqrt(K() / m_rho0); }
};

class constants_monaghan {
  double m_wdeltap = 0., m_corr_exp = 0., m_eps = 0.;

 public:
  constants_monaghan() {}
  constants_monaghan(double wdeltap, double corr_exp, double eps)
      : m_wdeltap(wdeltap), m_corr_exp(corr_exp), m_eps(eps) {}
  double mghn_wdeltap() const { return m_wdeltap; }
  double mghn_corr_exp() const { return m_corr_exp; }
  double mghn_eps() const { return m_eps; }
};

class constants_artificial_viscosity {
  double m_alpha = 0., m_beta = 0., m_eta = 0.;

 public:
  constants_artificial_viscosity() {}
  constants_artificial_viscosity(double alpha, double beta, double eta) : m_alpha(alpha), m_beta(beta), m_eta(eta) {}
  double artvisc_alpha() const { return m_alpha; }
  double artvisc_beta() const { return m_beta; }
  double artvisc_eta() const { return m_eta; }
};

class correction_constants {
  double m_xsph_eps = 0.;
  bool m_do_smooth = false;
  constants_monaghan m_mghn;
  constants_artificial_viscosity m_av;

 public:
  correction_constants() {}
  correction_constants(constants_monaghan m, constants_artificial_viscosity av, double xsph_eps)
      : m_xsph_eps(xsph_eps), m_mghn(m), m_av(av) {}
  correction_constants(constants_monaghan m, constants_artificial_viscosity av, double xsph_eps, bool smooth)
      : m_xsph_eps(xsph_eps), m_do_smooth(smooth), m_mghn(m), m_av(av) {}
  double xsph_eps() const { return m_xsph_eps; }
  bool do_smooth() const { return m_do_smooth; }
  constants_monaghan get_monaghan_const() const { return m_mghn; }
  constants_artificial_viscosity get_art_visc_const() const { return m_av; }
};

class simulation_data {
  physical_constants m_pc;
  correction_constants m_cc;

 public:
  simulation_data() {}
  simulation_data(physical_constants pc, correction_constants cc) : m_pc(pc), m_cc(cc) {}
  physical_constants get_physical_constants() const { return m_pc; }
  correction_constants get_correction_constants() const { return m_cc; }
};

#endif
