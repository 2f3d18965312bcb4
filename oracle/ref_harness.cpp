// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A driver linked against the reference's own sources (compiled where they
// lie under the reference root by oracle/Makefile, output in oracle/_ref/).
// It constructs the reference MD object from a condition file (the
// reference's parser and initialization), overwrites its dynamic state with a
// state given in a text file, runs one operation of the reference and writes
// the results.  Used only to produce golden vectors (tests/golden/) that pin
// the oracle, and as the "reference" CPU baseline.
//
// Text format (both directions): one array per line,
//   <name> <count> <v0> <v1> ...      (doubles printed with %.17g)
//
// Operations:
//   sys                 dump parsed tables/topology (parser golden)
//   init                dump the state the constructor leaves (after its make_pair)
//   forces              makePair of every InterInter + all computes, eflag=1
//   steps N             N calls of MD::verlet (caller picks loop_update so that
//                       no rebuild happens: rebuilds draw std::random_device)
//   make_pair           one MD::make_pair (lists / classification / COMs)
//   shake MOL           MD::shake on molecule MOL with oldPosition = state
//                       atom_x and current positions = array x_new
//   tersoff | sw        neighbour list by the potential's own make_pair, then
//                       its compute() alone (forces are x23.06 in the reference)
//   born BASE           PotentialBorn::compute alone on the ion, with the ion's
//                       atom types shifted down by BASE first (its coefficient
//                       table is indexed by ion-local type 0 = Na, 1 = Cl)
//   rotation            Variables::ROTATION on rows of array rot_in
//   time N              N verlet steps as the reference runs them (with its own
//                       random rebuilds); prints wall time and in-domain
//                       atom-steps
//   run                 MD::make_pair (the constructor's, md.cpp:34) then
//                       MD::run (run.cpp): relaxation, re-centre, setNVE and
//                       the main loop; its output files land in the working
//                       directory, the final state on stdout
#include <dirent.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/time.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <vector>

// the reference's SW neighbour builder is private; the driver calls it
#define private public
#include "md.hpp"
#include "potential/BMHFT/potentialBMHFT.hpp"
#undef private

typedef std::map<std::string, std::vector<double>> Arrays;

static Arrays read_arrays(const char *path) {
  Arrays a;
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string name;
    long n;
    if (!(ss >> name >> n)) continue;
    std::vector<double> v((size_t)n);
    for (long i = 0; i < n; i++) {
      std::string tok;
      ss >> tok;
      v[(size_t)i] = strtod(tok.c_str(), nullptr);
    }
    a[name] = v;
  }
  return a;
}

static void put(FILE *f, const char *name, const std::vector<double> &v) {
  fprintf(f, "%s %zu", name, v.size());
  for (double x : v) fprintf(f, " %.17g", x);
  fprintf(f, "\n");
}

static const std::vector<double> &need(Arrays &a, const char *k) {
  if (!a.count(k)) {
    fprintf(stderr, "ref_harness: missing array %s\n", k);
    exit(2);
  }
  return a[k];
}

static void load_state(MD *md, Arrays &a) {
  Variables *v = md->vars;
  int nmol = (int)v->mass.size();
  const std::vector<double> &mx = need(a, "mol_x"), &mv = need(a, "mol_v"), &rg = need(a, "region");
  const std::vector<double> &ax = need(a, "atom_x"), &av = need(a, "atom_v"), &af = need(a, "atom_f"),
                            &aw = need(a, "atom_w");
  if ((int)mx.size() != 3 * nmol) {
    fprintf(stderr, "ref_harness: state has %zu molecules, reference has %d\n", mx.size() / 3, nmol);
    exit(3);
  }
  size_t o = 0;
  for (int i = 0; i < nmol; i++) {
    for (int k = 0; k < 3; k++) v->position[i][k] = mx[3 * i + k], v->velocity[i][k] = mv[3 * i + k];
    v->region[i] = rg[i] != 0;
    int na = (int)v->massAtom[i].size();
    for (int ii = 0; ii < na; ii++, o++) {
      if (3 * o + 2 >= ax.size()) {
        fprintf(stderr, "ref_harness: atom count mismatch\n");
        exit(3);
      }
      for (int k = 0; k < 3; k++) {
        v->positionAtom[i][ii][k] = ax[3 * o + k];
        v->velocityAtom[i][ii][k] = av[3 * o + k];
        v->forceAtom[i][ii][k] = af[3 * o + k];
      }
      v->weight[i][ii] = aw[o];
    }
  }
  if (3 * o != ax.size()) {
    fprintf(stderr, "ref_harness: atom count mismatch (%zu vs %zu)\n", 3 * o, ax.size());
    exit(3);
  }
  v->gas_in.clear(); v->gas_out.clear(); v->vapor_in.clear(); v->vapor_out.clear();
  for (int i : v->group[1]) (v->region[i] ? v->gas_in : v->gas_out).push_back(i);
  for (int i : v->group[2]) (v->region[i] ? v->vapor_in : v->vapor_out).push_back(i);
  md->loopPair = (int)need(a, "loopPair")[0];
  md->loop_update = (int)need(a, "loop_update")[0];
  md->margin_length = need(a, "margin_length")[0];
  // Variables never initialises zeta_ion; the state sets it (default 0)
  v->zeta_ion = a.count("zeta_ion") ? a["zeta_ion"][0] : 0.0;
  v->Uzero();
}

static void dump_state(FILE *f, MD *md) {
  Variables *v = md->vars;
  int nmol = (int)v->mass.size();
  std::vector<double> mx, mv, rg, ax, av, af, aw;
  for (int i = 0; i < nmol; i++) {
    for (int k = 0; k < 3; k++) mx.push_back(v->position[i][k]), mv.push_back(v->velocity[i][k]);
    rg.push_back(v->region[i] ? 1 : 0);
    for (size_t ii = 0; ii < v->massAtom[i].size(); ii++) {
      for (int k = 0; k < 3; k++) {
        ax.push_back(v->positionAtom[i][ii][k]);
        av.push_back(v->velocityAtom[i][ii][k]);
        af.push_back(v->forceAtom[i][ii][k]);
      }
      aw.push_back(v->weight[i][ii]);
    }
  }
  put(f, "mol_x", mx); put(f, "mol_v", mv); put(f, "region", rg);
  put(f, "atom_x", ax); put(f, "atom_v", av); put(f, "atom_f", af); put(f, "atom_w", aw);
  put(f, "loopPair", {(double)md->loopPair});
  put(f, "loop_update", {(double)md->loop_update});
  const Potentials &U = v->Utotal;
  put(f, "U", {U.Uion, U.Ugas, U.Uvap, U.Ugi, U.Ugg, U.Uvg, U.Uvi, U.Uvv});
  put(f, "zeta_ion", {v->zeta_ion});
  for (auto *p : md->InterInter) {
    std::vector<double> l;
    for (auto &q : p->pairList) l.push_back(q.i), l.push_back(q.j);
    std::string nm = "list_" + std::string(dynamic_cast<PotentialGasIon *>(p) ? "gi" :
                                           dynamic_cast<PotentialVaporGas *>(p) ? "gv" :
                                           dynamic_cast<PotentialVaporVapor *>(p) ? "vv" :
                                           dynamic_cast<PotentialGasGas *>(p) ? "gg" : "other");
    put(f, nm.c_str(), l);
  }
  std::vector<double> gi(v->gas_in.begin(), v->gas_in.end()), vi(v->vapor_in.begin(), v->vapor_in.end());
  put(f, "gas_in", gi);
  put(f, "vapor_in", vi);
}

static void dump_sys(FILE *f, MD *md) {
  Variables *v = md->vars;
  int nt = (int)v->atypes.size();
  std::vector<double> pc, te, ts, tm;
  for (int i = 0; i < nt; i++) {
    te.push_back(v->atypes[i].coeff1); ts.push_back(v->atypes[i].coeff2); tm.push_back(v->atypes[i].mass);
    for (int j = 0; j < nt; j++) pc.push_back(v->pair_coeff[i][j][0]), pc.push_back(v->pair_coeff[i][j][1]);
  }
  put(f, "pair_coeff", pc); put(f, "type_eps", te); put(f, "type_sigma", ts); put(f, "type_mass", tm);
  std::vector<double> bt, ct, dc, dm;
  for (auto &b : v->btypes) bt.push_back(b.coeff[0]), bt.push_back(b.coeff[1]);
  for (auto &c : v->ctypes) ct.push_back(c.coeff[0]), ct.push_back(c.coeff[1]);
  for (auto &d : v->dtypes) {
    dm.push_back(d.multi);
    for (int k = 0; k < 15; k++) dc.push_back(d.coeff[k]);
  }
  put(f, "btype", bt); put(f, "ctype", ct); put(f, "dcoeff", dc); put(f, "dmulti", dm);
  int ng = (int)v->group[1].size(), nv = (int)v->group[2].size();
  int mols[3] = {0, ng ? v->group[1][0] : -1, nv ? v->group[2][0] : -1};
  const char *sn[3] = {"ion", "gas", "vap"};
  for (int s = 0; s < 3; s++) {
    int i = mols[s];
    std::vector<double> m, q, t, b, a, dd, r;
    if (i >= 0) {
      for (size_t k = 0; k < v->massAtom[i].size(); k++) m.push_back(v->massAtom[i][k]);
      for (size_t k = 0; k < v->chargeAtom[i].size(); k++) q.push_back(v->chargeAtom[i][k]);
      for (size_t k = 0; k < v->typeAtom[i].size(); k++) t.push_back(v->typeAtom[i][k]);
      for (auto &x : v->bonds[i]) b.insert(b.end(), x.begin(), x.end());
      for (auto &x : v->angles[i]) a.insert(a.end(), x.begin(), x.end());
      for (auto &x : v->dihedrals[i]) dd.insert(dd.end(), x.begin(), x.end());
      for (auto &x : v->rigid_pairs[i]) r.insert(r.end(), x.begin(), x.end());
    }
    std::string p(sn[s]);
    put(f, (p + "_mass").c_str(), m); put(f, (p + "_charge").c_str(), q); put(f, (p + "_type").c_str(), t);
    put(f, (p + "_bond").c_str(), b); put(f, (p + "_angle").c_str(), a); put(f, (p + "_dih").c_str(), dd);
    put(f, (p + "_rigid").c_str(), r);
    put(f, (p + "_molmass").c_str(), {i >= 0 ? v->mass[i] : 0.0});
  }
  std::vector<double> tg, tv;
  for (auto &x : v->positionGas) tg.insert(tg.end(), x.begin(), x.end());
  for (auto &x : v->positionVapor) tv.insert(tv.end(), x.begin(), x.end());
  put(f, "gas_tmpl", tg); put(f, "vap_tmpl", tv);
  put(f, "scalars", {md->dt, md->T, md->p, md->d_size, md->CUTOFF, md->MARGIN, md->ML2, md->pp->mgas,
                     md->pp->mvapor, (double)md->gastype, (double)ng, (double)nv});
  put(f, "flags", {(double)md->flags->force_lj, (double)md->flags->inter_vi, (double)md->flags->inter_gg,
                   (double)md->flags->inter_vg, (double)md->flags->inter_vv, (double)md->flags->intra_AMBER,
                   (double)md->flags->force_born, (double)md->flags->force_ters, (double)md->flags->force_sw,
                   (double)md->flags->efield, (double)md->flags->vapor_intra, (double)md->flags->gas_intra});
  // NVTion (setCondition.cpp:158-172): 0 off, 1 velocity scaling, 2 Nose-Hoover
  put(f, "nvt", {(double)(md->flags->nose_hoover_ion ? 2 : md->flags->velocity_scaling ? 1 : 0), md->pp->Tnh_ion});
}

static double now() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: ref_harness <cond> <state|-> <op> [arg] > out\n");
    return 1;
  }
  const char *cond = argv[1], *statef = argv[2], *op = argv[3];
  long arg = argc > 4 ? atol(argv[4]) : 0;
  // results go to the original stdout; the reference's own reports are muted
  FILE *out = fdopen(dup(1), "w");
  if (!freopen("/dev/null", "w", stdout)) return 4;
  MD *md = new MD((char *)cond, 0);
  md->vars->zeta_ion = 0;
  md->margin_length = md->MARGIN;
  std::string o(op);
  if (o == "sys") {
    dump_sys(out, md);
    put(out, "efield", {md->Ecoeff[0], md->Ecoeff[1], md->Ecoeff[2]});
    fflush(out);
    return 0;
  }
  if (o == "init") {
    dump_state(out, md);
    fflush(out);
    return 0;
  }
  if (o == "time") {
    // reference stepping loop as MD::run does it (relaxation skipped)
    md->setNVE();
    md->loopPair = 0;
    md->make_pair();
    Variables *v = md->vars;
    long long atom_steps = 0;
    auto n_in = [&]() {
      long long n = 0;
      for (size_t i = 0; i < v->mass.size(); i++)
        if (v->region[i]) n += (long long)v->massAtom[i].size();
      return n;
    };
    long long cur = n_in();
    double t0 = now();
    for (long s = 0; s < arg; s++) {
      md->verlet();
      if (md->loopPair == 0) cur = n_in();
      atom_steps += cur;
    }
    double t1 = now();
    fprintf(out, "seconds 1 %.9g\natom_steps 1 %lld\n", t1 - t0, atom_steps);
    fflush(out);
    return 0;
  }
  if (o == "rotation") {
    Arrays a = read_arrays(statef);
    const std::vector<double> &r = need(a, "rot_in");
    std::vector<double> res;
    for (size_t k = 0; k + 5 < r.size(); k += 6) {
      double X, Y, Z;
      md->vars->ROTATION(X, Y, Z, r[k], r[k + 1], r[k + 2], r[k + 3], r[k + 4], r[k + 5]);
      res.push_back(X), res.push_back(Y), res.push_back(Z);
    }
    put(out, "rot_out", res);
    fflush(out);
    return 0;
  }
  Arrays a = read_arrays(statef);
  load_state(md, a);
  if (o == "forces") {
    for (auto *p : md->InterInter) p->makePair(md->vars);
    md->flags->eflag = 1;
    md->vars->Uzero();
    for (auto *p : md->InterInter) p->compute(md->vars, md->flags);
    for (auto *p : md->IntraInter) p->compute(md->vars, md->flags);
    md->flags->eflag = 0;
    dump_state(out, md);
  } else if (o == "steps") {
    for (auto *p : md->InterInter) p->makePair(md->vars);
    long every = a.count("eflag_every") ? (long)a["eflag_every"][0] : 0;
    for (long s = 0; s < arg; s++) {
      if (every > 0 && (s + 1) % every == 0) {
        md->flags->eflag = 1;
        md->vars->Uzero();
      }
      md->verlet();
    }
    dump_state(out, md);
  } else if (o == "make_pair") {
    md->make_pair();
    dump_state(out, md);
  } else if (o == "run") {
    md->itime = 0;
    md->make_pair();
    md->margin_length = md->MARGIN;
    md->run(argv);
    dump_state(out, md);
  } else if (o == "shake") {
    int mol = (int)arg;
    Variables *v = md->vars;
    v->oldPosition.clear();
    for (auto &p : v->positionAtom[mol]) v->oldPosition.push_back(p);
    const std::vector<double> &xn = need(a, "x_new");
    for (size_t ii = 0; ii < v->positionAtom[mol].size(); ii++)
      for (int k = 0; k < 3; k++) v->positionAtom[mol][ii][k] = xn[3 * ii + k];
    md->shake(mol);
    std::vector<double> res;
    for (auto &p : v->positionAtom[mol]) res.insert(res.end(), p.begin(), p.end());
    put(out, "x_shake", res);
  } else if (o == "gg") {
    // PotentialGasGas::compute alone on the pair list given in gg_list
    // (molecule ids); every force cleared first.  Its own makePair pairs
    // loop counters (GasGas.cpp:50), so the list comes from the caller;
    // the list the reference's makePair builds is dumped as list_gg_ref.
    Variables *v = md->vars;
    PotentialGasGas g(md->ML2);
    g.makePair(v);
    std::vector<double> own;
    for (auto &q : g.pairList) own.push_back(q.i), own.push_back(q.j);
    g.pairList.clear();
    const std::vector<double> &l = need(a, "gg_list");
    for (size_t k = 0; k + 1 < l.size(); k += 2) {
      Pair p;
      p.i = (int)l[k];
      p.j = (int)l[k + 1];
      g.pairList.push_back(p);
    }
    for (auto &mol : v->forceAtom)
      for (auto &f : mol) f[0] = f[1] = f[2] = 0;
    v->Uzero();
    md->flags->eflag = 1;
    g.compute(v, md->flags);
    md->flags->eflag = 0;
    dump_state(out, md);
    put(out, "list_gg_ref", own);
  } else if (o == "born") {
    Variables *v = md->vars;
    for (auto &f : v->forceAtom[0]) f[0] = f[1] = f[2] = 0;
    for (auto &t : v->typeAtom[0]) t -= (int)arg;
    v->Uzero();
    md->flags->eflag = 1;
    PotentialBorn b;
    b.compute(v, md->flags);
    std::vector<double> fo;
    for (auto &f : v->forceAtom[0]) fo.insert(fo.end(), f.begin(), f.end());
    put(out, "ion_f", fo);
    put(out, "U_ion", {v->Utotal.Uion});
  } else if (o == "tersoff" || o == "sw") {
    Variables *v = md->vars;
    for (auto &f : v->forceAtom[0]) f[0] = f[1] = f[2] = 0;
    double U = 0;
    if (o == "tersoff") {
      PotentialTersoff t;
      t.make_pair(v);
      t.compute(v, md->flags);
      U = t.U;
    } else {
      PotentialSW s;
      s.make_pair(v);
      s.compute(v, md->flags);
      U = s.U;
    }
    std::vector<double> fo;
    for (auto &f : v->forceAtom[0]) fo.insert(fo.end(), f.begin(), f.end());
    put(out, "ion_f", fo);
    put(out, "U_raw", {U});
  } else {
    fprintf(stderr, "ref_harness: unknown op %s\n", op);
    return 1;
  }
  fflush(out);
  return 0;
}
