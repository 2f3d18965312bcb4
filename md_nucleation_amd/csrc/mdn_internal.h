// mdn_internal.h -- host-side types shared by the front end and the engine.
#pragma once
#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "mdn.h"

namespace mdn {

struct AtomType {
  std::string name;
  bool is_h = false;
  double mass = 0, eps = 0, sigma = 0;
};

// One molecule species (ion, gas or vapor) with its local topology.
struct Species {
  int natoms = 0;
  std::vector<double> mass, charge, tmpl_x;
  std::vector<int> type;
  std::vector<int> bond, angle, dih, rigid;  // 3, 4, 5, 3 ints per term
  double molmass = 0;
};

// Parsed system definition: what readCondFile, readIonFile, readVaporFile,
// setPhysicalProp and setCrossPotentials leave in Variables/MD/Physical.
struct SysDef {
  std::vector<AtomType> atypes;
  std::vector<double> pair_coeff, t_eps, t_sig, t_mass;
  std::vector<double> btype, ctype, dcoeff;
  std::vector<int> dmulti;
  Species sp[3];
  int n_gas = 1, n_vap = 1, n_mol = 0, n_atoms = 0;
  int ion_type_base = 0;
  double dt = 0.5, T = 300, pressure = 1e5, d_size = 0, cutoff = 20, margin = 10, ML2 = 900;
  double Mgas = 28.0, m_gas_kg = 0, m_vap_kg = 0;
  int gastype = 2, gas_spin = 1;
  bool has_gas_line = false, n_gas_default = false;
  uint32_t flags = 0;
  double efield[3] = {0, 0, 0};
  int shake = 0, shake_max_iter = 1000;
  int64_t total_steps = 0, relax_steps = 0;
  int observe = 10000000;
  std::string output_name, nvt_ion, nvt_gas;
  double cluster_radius = 0, cluster_density = 0;
  int nvt_ion_mode = 0;  // 0 off, 1 velocity scaling, 2 Nose-Hoover
  double Tnh_ion = 0;

  void finalize_views();
  int mol_offset(int i) const;
};

void load_sysdef(SysDef &d, const std::string &cond);
void init_state(const SysDef &d, uint64_t seed, mdn_state *st);

int set_error(int code, const std::string &msg);

// the vi term over S snapshots with per-atom ion parameters (device
// pointers; mdn_adjoint.hip)
struct ViBatch {
  const double *x, *w;  // [S][n][3], [S][n]; atoms: n_ion ion atoms, then n_vap vapor atoms
  int S, n_ion, n_vap;
  const double *sig_i, *eps_i, *q_i;  // [n_ion]
  const double *sig_v, *eps_v, *q_v;  // [n_vap]
  const uint8_t *vin;                 // [S][n_vap] vapor atom in the domain, or null (all)
  double K;                           // qqrd2e
};
// kernels queued on `stream` (a hipStream_t): F [S][n][3] written (not
// accumulated), Epart [S][n_ion] scratch, E [S] (may be null)
int vi_forward_dev(const ViBatch &b, double *F, double *Epart, double *E, void *stream);
// HU [S][n][3] written: the derivative of the forces along U
int vi_jvp_dev(const ViBatch &b, const double *U, double *HU, void *stream);
// g_sigma, g_eps, g_q [n_ion] written from dL/dE (gE, may be null) and dL/dF
// (gF [S][n][3], may be null); part [S][n_ion][3] scratch
int vi_adjoint_dev(const ViBatch &b, const double *gE, const double *gF, double *part, double *g_sigma,
                   double *g_eps, double *g_q, void *stream);

// engine facts the host run driver reads (committed context)
int engine_replicas(const mdn_ctx *c);  // -1 when not committed
const SysDef &engine_sysdef(const mdn_ctx *c, int r);
int engine_log_cap(const mdn_ctx *c);

}  // namespace mdn
