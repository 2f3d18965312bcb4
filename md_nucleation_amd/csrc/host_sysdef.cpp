// host_sysdef.cpp -- host front end of libmdn: condition / ion / vapor file
// parsing, parameter tables and initial placement.  This is the part of the
// reference that stays host C++ (initialization); the per-step path runs on
// the GPU (mdn_engine.hip).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "mdn.h"
#include "mdn_internal.h"

namespace mdn {

static const double kNw = 6.02e23;   // constants.hpp:28
static const double kKb = 1.38e-23;  // constants.hpp:29
static const double kMN2 = 28.0, kMHe = 4.0026, kMAr = 39.948;  // constants.hpp:34

static std::vector<std::string> split_ws(const std::string &s) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : s) {
    if (ch == ' ' || ch == '\t' || ch == '\r' || ch == '\n') {
      if (!cur.empty()) out.push_back(cur), cur.clear();
    } else {
      cur.push_back(ch);
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

static std::vector<std::string> split_tab(const std::string &s) {
  if (s.find('\t') == std::string::npos) return split_ws(s);
  std::vector<std::string> out;
  std::string cur;
  for (char ch : s) {
    if (ch == '\t') {
      out.push_back(cur);
      cur.clear();
    } else if (ch != '\r') {
      cur.push_back(ch);
    }
  }
  out.push_back(cur);
  // trailing empty fields are not data
  while (!out.empty() && out.back().empty()) out.pop_back();
  return out;
}

static std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  return s.substr(a, b - a + 1);
}

static std::string dirname_of(const std::string &p) {
  size_t k = p.find_last_of('/');
  return k == std::string::npos ? std::string(".") : p.substr(0, k);
}

static std::string resolve(const std::string &base, const std::string &f) {
  if (f.empty() || f[0] == '/') return f;
  std::ifstream t(f);
  if (t.good()) return f;  // relative to the working directory, as the reference
  return base + "/" + f;
}

// Element name of a GAFF atom type, as readIonFile.cpp:48-67 and
// readVaporFile.cpp:32-51 assign it (vapor names carry "(v)", except P);
// types outside the table keep their own name.  SHAKE H treats an atom as
// hydrogen by this name ("H" in the ion, "H(v)" in the vapor).
static std::string element_name(const std::string &t, bool vapor) {
  struct Group { const char *el; const char *types; };
  static const Group groups[] = {
      {"C", "#c #cs #c1 #c2 #c3 #ca #cp #cq #cc #cd #ce #cf #cg #ch #cx #cy #cu #cv #cz"},
      {"H", "#h1 #h2 #h3 #h4 #h5 #ha #hc #hn #ho #hp #hs #hw #hx"},
      {"F", "f"}, {"Cl", "cl"}, {"Br", "br"}, {"I", "i"},
      {"N", "#n #n1 #n2 #n3 #n4 #na #nb #nc #nd #ne #nf #nh #no #ns #nt #nx #ny #nz #n+ #nu #nv #n7 #n8 #n9"},
      {"O", "#o #oh #os #ow"},
      {"P", "#p2 #p3 #p4 #p5 #pb #pc #pd #pe #pf #px #py"},
      {"S", "#s #s2 #s4 #s6 #sh #ss #sx #sy"},
  };
  for (const Group &g : groups) {
    std::istringstream ss(g.types);
    std::string w;
    while (ss >> w)
      if (w == t) return std::string(g.el) + (vapor && std::string(g.el) != "P" ? "(v)" : "");
  }
  return t;
}

static double num(const std::vector<std::string> &r, size_t k) {
  if (k >= r.size()) throw std::runtime_error("missing column " + std::to_string(k));
  return std::stod(r[k]);
}

// Variables::readIonFile (readIonFile.cpp:4-180) / readVaporFile
// (readVaporFile.cpp:4-131): sections, type offsets, SHAKE-H split.
static void read_molecule_file(SysDef &d, const std::string &path, int sp, bool shakeH) {
  std::ifstream in(path);
  if (!in.good()) throw std::runtime_error("cannot open " + path);
  int atOff = (int)d.atypes.size() - 1, btOff = (int)d.btype.size() / 2 - 1;
  int ctOff = (int)d.ctype.size() / 2 - 1, dtOff = (int)d.dmulti.size() - 1;
  Species &S = d.sp[sp];
  S = Species();
  std::vector<std::array<int, 3>> bonds_all;
  std::string line;
  int sec = 0;
  while (std::getline(in, line)) {
    std::string t = trim(line);
    if (t.empty()) continue;
    if (t == "atom type") { sec = 1; continue; }
    if (t == "bond type") { sec = 2; continue; }
    if (t == "angle type") { sec = 3; continue; }
    if (t == "dihedral type") { sec = 4; continue; }
    if (t == "atoms") { sec = 5; continue; }
    if (t == "bonds") { sec = 6; continue; }
    if (t == "angles") { sec = 7; continue; }
    if (t == "dihedrals") { sec = 8; continue; }
    std::vector<std::string> r = split_tab(line);
    if (r.empty()) continue;
    if (sec == 1) {
      AtomType at;
      at.name = element_name(r.size() > 1 ? r[1] : "", sp == 2);
      at.is_h = at.name == (sp == 2 ? "H(v)" : "H");
      at.mass = num(r, 2);
      at.eps = num(r, 3);
      at.sigma = num(r, 4);
      d.atypes.push_back(at);
    } else if (sec == 2) {
      d.btype.push_back(num(r, 2));
      d.btype.push_back(num(r, 3));
    } else if (sec == 3) {
      d.ctype.push_back(num(r, 2));
      d.ctype.push_back(num(r, 3) / 180.0 * M_PI);
    } else if (sec == 4) {
      d.dmulti.push_back((int)num(r, 2));
      double c[15];
      for (int k = 0; k < 3; k++) {
        c[5 * k + 0] = num(r, 3 + 3 * k);
        c[5 * k + 1] = num(r, 4 + 3 * k);
        c[5 * k + 2] = num(r, 5 + 3 * k);
        c[5 * k + 3] = cos(num(r, 5 + 3 * k) / 180.0 * M_PI);
        c[5 * k + 4] = sin(num(r, 5 + 3 * k) / 180.0 * M_PI);
      }
      d.dcoeff.insert(d.dcoeff.end(), c, c + 15);
    } else if (sec == 5) {
      int ty = (int)num(r, 1) + atOff;
      if (ty < 0 || ty >= (int)d.atypes.size()) throw std::runtime_error("atom type out of range in " + path);
      S.type.push_back(ty);
      S.mass.push_back(d.atypes[ty].mass);
      S.charge.push_back(num(r, 2));
      S.tmpl_x.push_back(num(r, 3));
      S.tmpl_x.push_back(num(r, 4));
      S.tmpl_x.push_back(num(r, 5));
    } else if (sec == 6) {
      int i = (int)num(r, 0) - 1, j = (int)num(r, 1) - 1, ty = (int)num(r, 2) + btOff;
      if (i < 0 || j < 0 || i >= (int)S.type.size() || j >= (int)S.type.size())
        throw std::runtime_error("bond atom out of range in " + path);
      bool rigid = shakeH && (d.atypes[S.type[i]].is_h || d.atypes[S.type[j]].is_h);
      std::vector<int> &dst = rigid ? S.rigid : S.bond;
      dst.push_back(i); dst.push_back(j); dst.push_back(ty);
    } else if (sec == 7) {
      for (int k = 0; k < 3; k++) S.angle.push_back((int)num(r, k) - 1);
      S.angle.push_back((int)num(r, 3) + ctOff);
    } else if (sec == 8) {
      for (int k = 0; k < 4; k++) S.dih.push_back((int)num(r, k) - 1);
      S.dih.push_back((int)num(r, 4) + dtOff);
    }
  }
  S.natoms = (int)S.type.size();
  double tot = 0;
  for (double m : S.mass) tot += m;
  S.molmass = tot;  // ion: vars->mass[0] (readIonFile.cpp:175-177); vapor: pp->Mvapor
}

// MD::readCondFile (setCondition.cpp:5-254) + setPhysicalProp
// (PhysicalProp.cpp:4-39) + setCrossPotentials + setPotential's selection.
void load_sysdef(SysDef &d, const std::string &cond) {
  std::ifstream in(cond);
  if (!in.good()) throw std::runtime_error("cannot open condition file " + cond);
  std::string base = dirname_of(cond);
  std::string ion_file, vap_file;
  d = SysDef();
  d.dt = 0.5; d.cutoff = 20.0; d.margin = 10.0; d.observe = 10000000;  // defaultVariables.cpp:5-22
  d.T = 300; d.pressure = 1e5; d.gastype = 2; d.n_gas = 1; d.n_vap = 1;
  bool shakeH = false;
  uint32_t flags = MDN_VAPOR_INTRA | MDN_GAS_INTRA;  // flags.cpp:45-47
  std::string line;
  while (std::getline(in, line)) {
    std::vector<std::string> r = split_ws(line);
    if (r.empty()) continue;
    const std::string &k = r[0];
    auto arg = [&](size_t i) -> const std::string & {
      if (i >= r.size()) throw std::runtime_error("missing argument in line: " + line);
      return r[i];
    };
    if (k == "Input") ion_file = resolve(base, arg(1));
    else if (k == "Vapor") { vap_file = resolve(base, arg(1)); d.n_vap = std::stoi(arg(2)); }
    else if (k == "VaporStickPositions") { /* parsed, unused by the reference */ }
    else if (k == "Gas") {
      AtomType at;
      const std::string &g = arg(1);
      Species &G = d.sp[1];
      G = Species();
      if (g == "He") {
        d.gastype = 1; at.mass = 4.027; at.name = "He(g)"; at.eps = 0.0203; at.sigma = 2.556;
        G.tmpl_x = {0, 0, 0}; G.charge = {0}; G.type = {0}; G.mass = {at.mass};
        d.Mgas = kMHe;
      } else if (g == "N2") {
        d.gastype = 2; at.mass = 14.01; at.name = "N(g)"; at.eps = 0.1098; at.sigma = 3.27351824993;
        G.tmpl_x = {-0.549, 0, 0, 0.549, 0, 0}; G.charge = {0, 0}; G.type = {0, 0};
        G.mass = {at.mass, at.mass};
        d.btype.push_back(1221.7); d.btype.push_back(1.098);
        G.bond = {0, 1, 0};
        d.Mgas = kMN2;
      } else if (g == "N2monoatomic") {
        d.gastype = 3; at.mass = 28.02; at.name = "N2(g)"; at.eps = 0.14397; at.sigma = 3.798;
        G.tmpl_x = {0, 0, 0}; G.charge = {0}; G.type = {0}; G.mass = {at.mass};
        d.Mgas = kMN2;
      } else if (g == "Ar") {
        // the reference's Ar copies N2-mono parameters and no mass
        // (setCondition.cpp:102-113); argon proper is used here
        d.gastype = 4; at.mass = kMAr; at.name = "Ar(g)"; at.eps = 0.238; at.sigma = 3.4;
        G.tmpl_x = {0, 0, 0}; G.charge = {0}; G.type = {0}; G.mass = {at.mass};
        d.Mgas = kMAr;
      } else {
        throw std::runtime_error("unknown gas " + g);
      }
      d.atypes.push_back(at);
      G.natoms = (int)G.type.size();
      d.n_gas = std::stoi(arg(2));
      d.has_gas_line = true;
    }
    else if (k == "Temperature") d.T = std::stod(arg(1));
    else if (k == "Pressure") d.pressure = std::stod(arg(1));
    else if (k == "dt") d.dt = std::stod(arg(1));
    else if (k == "TotalSteps") d.total_steps = (int64_t)std::stod(arg(1));
    else if (k == "RelaxSteps") d.relax_steps = (int64_t)std::stod(arg(1));
    else if (k == "CutOff") d.cutoff = std::stod(arg(1));
    else if (k == "Margin") d.margin = std::stod(arg(1));
    else if (k == "Output") { d.output_name = arg(1); d.observe = std::stoi(arg(2)); }
    else if (k == "NVTion") d.nvt_ion = arg(1);
    else if (k == "NVTgas") d.nvt_gas = arg(1);
    else if (k == "Interactions") continue;
    else if (k == "Gyration") { /* output only */ }
    else if (k == "SHAKE") { if (arg(1) == "H") shakeH = true; }
    else if (k == "VaporCluster") { d.cluster_radius = std::stod(arg(1)); d.cluster_density = std::stod(arg(2)); }
    else if (k == "Seed") { /* accepted for scripts that carry it */ }
    else if (k == "gg") { if (arg(1) == "LJ") flags |= MDN_GG; else if (arg(1) == "OFF") flags &= ~MDN_GG; }
    else if (k == "gi" || k == "ig") {
      if (arg(1) == "LJ") flags |= MDN_GI;
      else if (arg(1) == "OFF") flags &= ~MDN_GI;
    }
    else if (k == "ion") {
      const std::string &v = arg(1);
      if (v == "AMBER") flags |= MDN_ION_AMBER;
      else if (v == "Stilinger-Weber") flags |= MDN_ION_SW;
      else if (v == "Tersoff") flags |= MDN_ION_TERSOFF;
      else if (v == "Born-Mayer-Huggins-NaCl") flags |= MDN_ION_BMH;
      else throw std::runtime_error("unknown ion potential " + v);
    }
    else if (k == "vi" || k == "iv") { if (arg(1) == "LJcoul") flags |= MDN_VI; else if (arg(1) == "OFF") flags &= ~MDN_VI; }
    else if (k == "vv") { if (arg(1) == "LJcoul") flags |= MDN_VV; else if (arg(1) == "OFF") flags &= ~MDN_VV; }
    else if (k == "gv" || k == "vg") { if (arg(1) == "LJ") flags |= MDN_GV; else if (arg(1) == "OFF") flags &= ~MDN_GV; }
    else if (k == "Efield") {
      flags |= MDN_EFIELD;
      d.efield[0] = std::stod(arg(1)); d.efield[1] = std::stod(arg(2)); d.efield[2] = std::stod(arg(3));
    }
  }
  if (ion_file.empty()) throw std::runtime_error("condition file has no Input line");
  d.ion_type_base = (int)d.atypes.size();
  read_molecule_file(d, ion_file, 0, shakeH);
  if (!vap_file.empty()) {
    read_molecule_file(d, vap_file, 2, shakeH);
  } else {
    d.sp[2] = Species();
    d.n_vap = 0;
  }
  if (!d.has_gas_line) d.sp[1] = Species();  // default N2 gastype with no gas atoms
  d.sp[1].molmass = d.Mgas;  // initializationGas.cpp:61
  // PhysicalProp.cpp:336-338
  d.m_gas_kg = d.Mgas / kNw / 1000.0;
  d.m_vap_kg = d.sp[2].molmass / kNw / 1000.0;
  d.gas_spin = d.gastype == 2 ? 1 : 0;
  d.flags = flags;
  d.shake = shakeH ? 1 : 0;
  d.shake_max_iter = 1000;
  // NVTion (setCondition.cpp:158-172): OFF, "scale" (velocity scaling), or
  // a Nose-Hoover temperature
  if (d.nvt_ion.empty() || d.nvt_ion == "OFF") d.nvt_ion_mode = 0;
  else if (d.nvt_ion == "scale") d.nvt_ion_mode = 1;
  else {
    char *end = nullptr;
    d.Tnh_ion = strtod(d.nvt_ion.c_str(), &end);
    if (end == d.nvt_ion.c_str()) throw std::runtime_error("NVTion: expected OFF, scale or a temperature");
    d.nvt_ion_mode = 2;
  }
  // setCondition.cpp:247-250
  d.d_size = pow(d.n_gas * kKb * d.T / d.pressure, 1 / 3.0) * 1e10;
  d.ML2 = (d.cutoff + d.margin) * (d.cutoff + d.margin);
  // setCrossPotential.cpp:4-23
  int n = (int)d.atypes.size();
  d.pair_coeff.assign((size_t)n * n * 2, 0.0);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double epu = sqrt(d.atypes[i].eps * d.atypes[j].eps);
      double sig = (d.atypes[i].sigma + d.atypes[j].sigma) * 0.5;
      d.pair_coeff[(i * n + j) * 2 + 0] = 48 * epu * pow(sig, 12.0);
      d.pair_coeff[(i * n + j) * 2 + 1] = 24 * epu * pow(sig, 6.0);
    }
  d.finalize_views();
}

void SysDef::finalize_views() {
  int n = (int)atypes.size();
  t_eps.resize(n); t_sig.resize(n); t_mass.resize(n);
  for (int i = 0; i < n; i++) { t_eps[i] = atypes[i].eps; t_sig[i] = atypes[i].sigma; t_mass[i] = atypes[i].mass; }
  n_mol = 1 + n_gas + n_vap;
  n_atoms = sp[0].natoms + n_gas * sp[1].natoms + n_vap * sp[2].natoms;
}

int SysDef::mol_offset(int i) const {
  if (i == 0) return 0;
  if (i <= n_gas) return sp[0].natoms + (i - 1) * sp[1].natoms;
  return sp[0].natoms + n_gas * sp[1].natoms + (i - 1 - n_gas) * sp[2].natoms;
}

// variables.hpp:94-98
static void rotation(double &X, double &Y, double &Z, double A, double B, double C, double x, double y, double z) {
  X = cos(C) * (x * cos(B) + y * sin(A) * sin(B) - z * cos(A) * sin(B)) + sin(C) * (y * cos(A) + z * sin(A));
  Y = -sin(C) * (x * cos(B) + y * sin(A) * sin(B) - z * cos(A) * sin(B)) + cos(C) * (y * cos(A) + z * sin(A));
  Z = x * sin(B) - y * sin(A) * cos(B) + z * cos(A) * cos(B);
}

// periodic minimum image (periodic.cpp:134-142)
static void adjust_periodic(double &dx, double &dy, double &dz, double L) {
  double h = 0.5 * L;
  if (dx < -h) dx += L; if (dx > h) dx -= L;
  if (dy < -h) dy += L; if (dy > h) dy -= L;
  if (dz < -h) dz += L; if (dz > h) dz -= L;
}

// Uniform grid over the periodic box for minimum-distance rejection; it
// replaces the O(N^2) scan of initializationGas.cpp:36-89.
struct CellGrid {
  double L, cell;
  int n;
  std::vector<std::vector<std::array<double, 3>>> bins;
  CellGrid(double L_, double cutoff) : L(L_) {
    n = std::max(1, std::min(256, (int)floor(L / std::max(cutoff, 1e-3))));
    cell = L / n;
    bins.resize((size_t)n * n * n);
  }
  int idx(double v) const {
    int k = (int)floor((v + 0.5 * L) / cell);
    k %= n; if (k < 0) k += n;
    return k;
  }
  void add(const double *p) { bins[((size_t)idx(p[0]) * n + idx(p[1])) * n + idx(p[2])].push_back({p[0], p[1], p[2]}); }
  bool clear_of(const double *p, double dmin) const {
    int ix = idx(p[0]), iy = idx(p[1]), iz = idx(p[2]);
    int span = n >= 3 ? 1 : 0;
    for (int a = -span; a <= span; a++)
      for (int b = -span; b <= span; b++)
        for (int c = -span; c <= span; c++) {
          int jx = (ix + a + n) % n, jy = (iy + b + n) % n, jz = (iz + c + n) % n;
          for (auto &q : bins[((size_t)jx * n + jy) * n + jz]) {
            double dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
            adjust_periodic(dx, dy, dz, L);
            if (dx * dx + dy * dy + dz * dz <= dmin * dmin) return false;
          }
          if (n < 3) return true;
        }
    return true;
  }
};

void init_state(const SysDef &d, uint64_t seed, mdn_state *st) {
  std::mt19937_64 rng(seed * 0x9E3779B97F4A7C15ull + 12345);
  std::uniform_real_distribution<double> U01(0.0, 1.0);
  std::normal_distribution<double> N01(0.0, 1.0);
  int nm = d.n_mol, na = d.n_atoms;
  memset(st->mol_x, 0, sizeof(double) * 3 * nm);
  memset(st->mol_v, 0, sizeof(double) * 3 * nm);
  memset(st->region, 0, nm);
  memset(st->atom_x, 0, sizeof(double) * 3 * na);
  memset(st->atom_v, 0, sizeof(double) * 3 * na);
  memset(st->atom_f, 0, sizeof(double) * 3 * na);
  for (int a = 0; a < na; a++) st->atom_w[a] = 0;
  st->loopPair = 0; st->loop_update = 0; st->rebuild_count = 0;
  st->n_gi = st->n_gv = st->n_gg = 0;
  for (int k = 0; k < MDN_U_N; k++) st->U[k] = 0;
  const Species &I = d.sp[0];
  // Variables::ionInitial (variables.cpp:12-30): random rotation about the
  // origin and per-atom Maxwell velocities; ion weights are 1
  // (readIonFile.cpp:124).
  double A = U01(rng) * 2 * M_PI, B = U01(rng) * 2 * M_PI, C = U01(rng) * 2 * M_PI;
  for (int i = 0; i < I.natoms; i++) {
    double x = I.tmpl_x[3 * i], y = I.tmpl_x[3 * i + 1], z = I.tmpl_x[3 * i + 2];
    rotation(st->atom_x[3 * i], st->atom_x[3 * i + 1], st->atom_x[3 * i + 2], A, B, C, x, y, z);
    double matom = I.mass[i] * 1e-3 / kNw;
    double sg = sqrt(kKb * d.T / matom);
    for (int k = 0; k < 3; k++) st->atom_v[3 * i + k] = sg * N01(rng) * 1e-5;
    st->atom_w[i] = 1;
  }
  st->region[0] = 1;
  const double dis = 10.0;  // initializationGas.cpp:25
  CellGrid grid(d.d_size, dis);
  for (int i = 0; i < I.natoms; i++) grid.add(&st->atom_x[3 * i]);
  double origin[3] = {0, 0, 0};
  grid.add(origin);
  // vapor cluster pre-packed around the ion (extension keyword VaporCluster):
  // lattice sites closest to the ion at the given density, each molecule
  // placed in domain with a rotation that keeps atoms >= 1.8 A apart.
  int first_vap = 1 + d.n_gas;
  int placed_vap = 0;
  if (d.cluster_radius > 0 && d.n_vap > 0) {
    const Species &V = d.sp[2];
    double vol = V.molmass / d.cluster_density * 1.66054;  // A^3 per molecule
    double a = cbrt(vol);
    int m = (int)ceil(d.cluster_radius / a) + 1;
    std::vector<std::array<double, 4>> sites;
    for (int ix = -m; ix <= m; ix++)
      for (int iy = -m; iy <= m; iy++)
        for (int iz = -m; iz <= m; iz++) {
          double p[3] = {ix * a, iy * a, iz * a};
          double r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
          if (r2 > d.cluster_radius * d.cluster_radius) continue;
          double dmin2 = 1e30;
          for (int i = 0; i < I.natoms; i++) {
            double dx = p[0] - st->atom_x[3 * i], dy = p[1] - st->atom_x[3 * i + 1], dz = p[2] - st->atom_x[3 * i + 2];
            dmin2 = std::min(dmin2, dx * dx + dy * dy + dz * dz);
          }
          if (dmin2 < 9.0) continue;
          sites.push_back({p[0], p[1], p[2], r2});
        }
    std::sort(sites.begin(), sites.end(), [](const std::array<double, 4> &u, const std::array<double, 4> &v) { return u[3] < v[3]; });
    int ncl = std::min<int>(d.n_vap, (int)sites.size());
    std::vector<double> placed;  // atoms already placed (cluster)
    for (int i = 0; i < I.natoms; i++) placed.insert(placed.end(), &st->atom_x[3 * i], &st->atom_x[3 * i] + 3);
    double mk = d.m_vap_kg, sgm = sqrt(kKb * d.T / mk);
    const int nv = V.natoms;
    std::vector<double> bx(3 * (size_t)nv), tx(3 * (size_t)nv);
    for (int s = 0; s < ncl; s++) {
      int mol = first_vap + s;
      int o = d.mol_offset(mol);
      double best = -1;
      for (int tr = 0; tr < 200; tr++) {
        double a1 = U01(rng) * 2 * M_PI, b1 = U01(rng) * 2 * M_PI, c1 = U01(rng) * 2 * M_PI;
        for (int k = 0; k < nv; k++) {
          rotation(tx[3 * k], tx[3 * k + 1], tx[3 * k + 2], a1, b1, c1, V.tmpl_x[3 * k], V.tmpl_x[3 * k + 1], V.tmpl_x[3 * k + 2]);
          for (int q = 0; q < 3; q++) tx[3 * k + q] += sites[s][q];
        }
        double dmin2 = 1e30;
        size_t np = placed.size() / 3;
        for (size_t p = 0; p < np; p++) {
          double ddx = placed[3 * p] - sites[s][0], ddy = placed[3 * p + 1] - sites[s][1], ddz = placed[3 * p + 2] - sites[s][2];
          if (ddx * ddx + ddy * ddy + ddz * ddz > 100.0) continue;
          for (int k = 0; k < nv; k++) {
            double ex = placed[3 * p] - tx[3 * k], ey = placed[3 * p + 1] - tx[3 * k + 1], ez = placed[3 * p + 2] - tx[3 * k + 2];
            dmin2 = std::min(dmin2, ex * ex + ey * ey + ez * ez);
          }
        }
        if (dmin2 > best) { best = dmin2; bx = tx; }
        if (dmin2 >= 1.8 * 1.8) break;
      }
      double vc[3] = {sgm * N01(rng) * 1e-5, sgm * N01(rng) * 1e-5, sgm * N01(rng) * 1e-5};
      for (int k = 0; k < nv; k++) {
        for (int q = 0; q < 3; q++) { st->atom_x[3 * (o + k) + q] = bx[3 * k + q]; st->atom_v[3 * (o + k) + q] = vc[q]; }
        placed.insert(placed.end(), &bx[3 * k], &bx[3 * k] + 3);
      }
      for (int q = 0; q < 3; q++) { st->mol_x[3 * mol + q] = sites[s][q]; st->mol_v[3 * mol + q] = vc[q]; }
      st->region[mol] = 1;
      grid.add(sites[s].data());
      placed_vap++;
    }
  }
  // MD::initialization_gas (initializationGas.cpp:24-90): uniform COMs with
  // minimum distance `dis` (periodic), Maxwell COM velocities.  The
  // reference retries without end; here a molecule that finds no free place
  // in 10^5 tries is an error (the box is too crowded for the 10 A rule).
  auto place = [&](int mol, double m_kg) {
    double sg = sqrt(kKb * d.T / m_kg);
    for (int tries = 0;; tries++) {
      if (tries == 100000)
        throw std::runtime_error("molecule " + std::to_string(mol) + " found no place at least " +
                                 std::to_string(dis) + " A from the others in 100000 tries (box " +
                                 std::to_string(d.d_size) + " A too small for the molecule count)");
      double X[3];
      for (int k = 0; k < 3; k++) X[k] = (U01(rng) - 0.5) * d.d_size;
      if (grid.clear_of(X, dis)) {
        for (int k = 0; k < 3; k++) st->mol_x[3 * mol + k] = X[k];
        for (int k = 0; k < 3; k++) st->mol_v[3 * mol + k] = sg * N01(rng) * 1e-5;
        grid.add(X);
        return;
      }
    }
  };
  for (int g = 1; g <= d.n_gas; g++) place(g, d.m_gas_kg);
  // MD::initialization_vapor (initializationVapor.cpp:14-77)
  for (int v = first_vap + placed_vap; v < nm; v++) place(v, d.m_vap_kg);
}

}  // namespace mdn

using namespace mdn;

extern "C" {

int mdn_sysdef_load(const char *cond_path, mdn_sysdef **out) {
  if (!cond_path || !out) return set_error(MDN_E_ARG, "mdn_sysdef_load: null argument");
  SysDef *d = new SysDef();
  try {
    load_sysdef(*d, cond_path);
  } catch (const std::exception &e) {
    delete d;
    return set_error(MDN_E_IO, std::string("mdn_sysdef_load: ") + e.what());
  }
  *out = reinterpret_cast<mdn_sysdef *>(d);
  return MDN_OK;
}

int mdn_sysdef_view(const mdn_sysdef *h, mdn_system *o) {
  if (!h || !o) return set_error(MDN_E_ARG, "mdn_sysdef_view: null argument");
  const SysDef &d = *reinterpret_cast<const SysDef *>(h);
  memset(o, 0, sizeof(*o));
  o->n_types = (int)d.atypes.size();
  o->pair_coeff = d.pair_coeff.data();
  o->type_eps = d.t_eps.data(); o->type_sigma = d.t_sig.data(); o->type_mass = d.t_mass.data();
  o->n_btypes = (int)d.btype.size() / 2; o->btype = d.btype.data();
  o->n_ctypes = (int)d.ctype.size() / 2; o->ctype = d.ctype.data();
  o->n_dtypes = (int)d.dmulti.size(); o->dmulti = d.dmulti.data(); o->dcoeff = d.dcoeff.data();
  for (int s = 0; s < 3; s++) {
    const Species &S = d.sp[s];
    o->natoms[s] = S.natoms;
    o->mass[s] = S.mass.data(); o->charge[s] = S.charge.data(); o->type[s] = S.type.data();
    o->tmpl_x[s] = S.tmpl_x.data();
    o->nbond[s] = (int)S.bond.size() / 3; o->bond[s] = S.bond.data();
    o->nangle[s] = (int)S.angle.size() / 4; o->angle[s] = S.angle.data();
    o->ndih[s] = (int)S.dih.size() / 5; o->dih[s] = S.dih.data();
    o->nrigid[s] = (int)S.rigid.size() / 3; o->rigid[s] = S.rigid.data();
    o->molmass[s] = S.molmass;
  }
  o->n_gas = d.n_gas; o->n_vap = d.n_vap; o->ion_type_base = d.ion_type_base;
  o->dt = d.dt; o->T = d.T; o->pressure = d.pressure; o->d_size = d.d_size; o->cutoff = d.cutoff;
  o->margin = d.margin; o->ML2 = d.ML2; o->m_gas_kg = d.m_gas_kg; o->m_vap_kg = d.m_vap_kg;
  o->gastype = d.gastype; o->gas_spin = d.gas_spin; o->flags = d.flags;
  for (int k = 0; k < 3; k++) o->efield[k] = d.efield[k];
  o->shake = d.shake; o->shake_max_iter = d.shake_max_iter;
  o->total_steps = d.total_steps; o->relax_steps = d.relax_steps; o->observe = d.observe;
  o->cluster_radius = d.cluster_radius;
  o->nvt_ion = d.nvt_ion_mode; o->Tnh_ion = d.Tnh_ion;
  return MDN_OK;
}

int mdn_sysdef_sizes(const mdn_sysdef *h, int32_t *n_mol, int32_t *n_atoms) {
  if (!h) return set_error(MDN_E_ARG, "mdn_sysdef_sizes: null sysdef");
  const SysDef &d = *reinterpret_cast<const SysDef *>(h);
  if (n_mol) *n_mol = d.n_mol;
  if (n_atoms) *n_atoms = d.n_atoms;
  return MDN_OK;
}

int mdn_sysdef_set_shake(mdn_sysdef *h, int32_t shake, int32_t max_iter) {
  if (!h || max_iter <= 0) return set_error(MDN_E_ARG, "mdn_sysdef_set_shake: bad argument");
  if (shake < 0 || shake > 2) return set_error(MDN_E_ARG, "mdn_sysdef_set_shake: mode must be 0, 1 or 2");
  SysDef &d = *reinterpret_cast<SysDef *>(h);
  d.shake = shake;
  d.shake_max_iter = max_iter;
  return MDN_OK;
}

int mdn_sysdef_set_flags(mdn_sysdef *h, uint32_t flags) {
  if (!h) return set_error(MDN_E_ARG, "mdn_sysdef_set_flags: null sysdef");
  reinterpret_cast<SysDef *>(h)->flags = flags;
  return MDN_OK;
}

int mdn_sysdef_init_state(const mdn_sysdef *h, uint64_t seed, mdn_state *st) {
  if (!h || !st || !st->mol_x || !st->mol_v || !st->region || !st->atom_x || !st->atom_v || !st->atom_f || !st->atom_w)
    return set_error(MDN_E_ARG, "mdn_sysdef_init_state: null argument");
  try {
    init_state(*reinterpret_cast<const SysDef *>(h), seed, st);
  } catch (const std::exception &e) {
    return set_error(MDN_E_ARG, std::string("mdn_sysdef_init_state: ") + e.what());
  }
  return MDN_OK;
}

void mdn_sysdef_free(mdn_sysdef *h) { delete reinterpret_cast<SysDef *>(h); }

}  // extern "C"
