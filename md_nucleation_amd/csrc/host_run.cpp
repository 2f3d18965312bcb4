// host_run.cpp -- MD::run (run.cpp:18-75) over a libmdn context: the
// relaxation and main loops of every replica stepped together through
// mdn_run_batched, with the reference's observables and output files
// (output.cpp, observer.cpp) produced from device state at the reference's
// cadence.  Uses only the public entry points plus the parsed system of each
// replica (type names, cadences, file names).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "mdn.h"
#include "mdn_internal.h"

namespace mdn {
namespace {

constexpr long kLogger = 10000;  // run.cpp:20

struct FileCloser {
  void operator()(FILE *f) const { if (f) fclose(f); }
};
using File = std::unique_ptr<FILE, FileCloser>;

// one replica's files and host buffers
struct Out {
  const SysDef *d = nullptr;
  int calc = 0;
  File k, u, ion, vin, vout, dump;
  std::vector<double> mol_x, mol_v, atom_x, atom_v, atom_f, atom_w;
  std::vector<uint8_t> region;
  mdn_state st{};
};

File open_file(const std::string &path, int *rc) {
  FILE *f = fopen(path.c_str(), "w");
  if (!f) *rc = set_error(MDN_E_IO, "mdn_run: cannot write " + path);
  return File(f);
}

std::string join(const std::string &dir, const std::string &name) {
  if (dir.empty()) return name;
  return dir.back() == '/' ? dir + name : dir + "/" + name;
}

// MD::display's file part (output.cpp:67-108): K_<n>.dat and U_<n>.dat; the
// console block when asked
void display(Out &o, const double *obs, double time, bool verbose) {
  const double *U = o.st.U;
  const double K = obs[0] + obs[1] + obs[2];
  fprintf(o.k.get(), "%e %e %e %e %e\n", time, obs[0], obs[1], obs[2], K);
  fprintf(o.u.get(), "%e %e %e %e %e %e %e %e %e\n", time, U[MDN_U_ION], U[MDN_U_GAS], U[MDN_U_VAP],
          U[MDN_U_GI], U[MDN_U_GG], U[MDN_U_VG], U[MDN_U_VI], U[MDN_U_VV]);
  if (verbose) {
    double Us = 0;
    for (int k = 0; k < MDN_U_N; k++) Us += U[k];
    printf("[%d] ---------------- TIME = %g ps ----------------\n", o.calc, time / 1000.0);
    printf("[%d]   Kion = %1.2e  Kgas = %1.2e  Kvap = %1.2e\n[%d]   Tion = %1.2f  Tgas = %1.2f  Tvap = %1.2f\n",
           o.calc, obs[0], obs[1], obs[2], o.calc, obs[3], obs[4], obs[5]);
    printf("[%d]   K = %1.2e  U = %1.2e\n", o.calc, K, Us);
    fflush(stdout);
  }
}

// MD::export_dump_close (output.cpp:132-170): the in-domain atoms, positions
// relative to the ion COM (dump_fix = 1, flags.cpp)
void dump(Out &o, double time) {
  if (!o.dump) return;
  const SysDef &d = *o.d;
  FILE *f = o.dump.get();
  int count = (int)time;
  int n = 0;
  for (int i = 0; i < d.n_mol; i++)
    if (o.region[i]) n += d.sp[i == 0 ? 0 : (i <= d.n_gas ? 1 : 2)].natoms;
  const double h = d.d_size * 0.5;
  fprintf(f,
          "ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n%e %e\n%e %e\n%e %e\nITEM: "
          "ATOMS id type x y z vx vy vz\n",
          count, n, -h, h, -h, h, -h, h);
  const double X = o.mol_x[0], Y = o.mol_x[1], Z = o.mol_x[2];
  int id = 0;
  for (int i = 0; i < d.n_mol; i++) {
    if (!o.region[i]) continue;
    const int sp = i == 0 ? 0 : (i <= d.n_gas ? 1 : 2);
    const int a0 = d.mol_offset(i);
    const Species &S = d.sp[sp];
    for (int ii = 0; ii < S.natoms; ii++, id++) {
      const double *x = &o.atom_x[3 * (size_t)(a0 + ii)], *v = &o.atom_v[3 * (size_t)(a0 + ii)];
      fprintf(f, "%d %s %f %f %f %f %f %f\n", id, d.atypes[S.type[ii]].name.c_str(), x[0] - X, x[1] - Y, x[2] - Z,
              v[0], v[1], v[2]);
    }
  }
}

// MD::output (output.cpp:33-41): the ion COM every kLogger steps
void output_ion(Out &o, double time) {
  fprintf(o.ion.get(), "%e %e %e %e %e %e %e\n", time / 1e6, o.mol_x[0], o.mol_x[1], o.mol_x[2], o.mol_v[0],
          o.mol_v[1], o.mol_v[2]);
}

}  // namespace
}  // namespace mdn

using namespace mdn;

extern "C" int mdn_run(mdn_ctx *c, const char *outdir, int32_t calc0, int32_t precision, uint32_t opts) {
  const int n = engine_replicas(c);
  if (n < 0) return set_error(MDN_E_STATE, "mdn_run: call mdn_commit (and mdn_make_pair) first");
  if (precision != MDN_FP64 && precision != MDN_MIXED) return set_error(MDN_E_ARG, "mdn_run: bad precision");
  const SysDef &d0 = engine_sysdef(c, 0);
  int n_vap_max = 1;
  for (int r = 0; r < n; r++) {
    const SysDef &d = engine_sysdef(c, r);
    if (d.relax_steps != d0.relax_steps || d.total_steps != d0.total_steps || d.observe != d0.observe)
      return set_error(MDN_E_ARG, "mdn_run: replicas differ in RelaxSteps, TotalSteps or the Output cadence");
    n_vap_max = std::max(n_vap_max, d.n_vap);
  }
  const long observe = d0.observe;
  if (observe <= 0) return set_error(MDN_E_ARG, "mdn_run: the Output cadence must be positive");
  const bool verbose = (opts & MDN_RUN_VERBOSE) != 0;
  int rc = MDN_OK;
  // a rebuild records at most one event per vapor molecule and happens at
  // most once per step: chunks of cap / n_vap steps never overflow the log
  if (engine_log_cap(c) == 0 && (rc = mdn_set_vapor_log(c, 250 * n_vap_max + 64))) return rc;
  const long chunk_max = std::max(1, engine_log_cap(c) / n_vap_max);

  std::vector<Out> out(n);
  const std::string dir = outdir ? outdir : "";
  for (int r = 0; r < n; r++) {
    Out &o = out[r];
    o.d = &engine_sysdef(c, r);
    o.calc = calc0 + r;
    const SysDef &d = *o.d;
    char buf[256];
    // output_initial (output.cpp:10-31) and the Output line (setCondition.cpp:145-157)
    snprintf(buf, sizeof buf, "ion_%d_%d.dat", (int)d.T, o.calc);
    o.ion = open_file(join(dir, buf), &rc);
    snprintf(buf, sizeof buf, "vapor_collision_%d.dat", o.calc);
    open_file(join(dir, buf), &rc);
    snprintf(buf, sizeof buf, "vapor_in_%d.dat", o.calc);
    o.vin = open_file(join(dir, buf), &rc);
    snprintf(buf, sizeof buf, "vapor_out_%d.dat", o.calc);
    o.vout = open_file(join(dir, buf), &rc);
    snprintf(buf, sizeof buf, "K_%d.dat", o.calc);
    o.k = open_file(join(dir, buf), &rc);
    snprintf(buf, sizeof buf, "U_%d.dat", o.calc);
    o.u = open_file(join(dir, buf), &rc);
    if (!d.output_name.empty()) o.dump = open_file(join(dir, d.output_name + "_" + std::to_string(o.calc) + ".dump"), &rc);
    if (rc) return rc;
    o.mol_x.resize(3 * (size_t)d.n_mol); o.mol_v.resize(3 * (size_t)d.n_mol); o.region.resize(d.n_mol);
    o.atom_x.resize(3 * (size_t)d.n_atoms); o.atom_v.resize(3 * (size_t)d.n_atoms);
    o.atom_f.resize(3 * (size_t)d.n_atoms); o.atom_w.resize(d.n_atoms);
    o.st.mol_x = o.mol_x.data(); o.st.mol_v = o.mol_v.data(); o.st.region = o.region.data();
    o.st.atom_x = o.atom_x.data(); o.st.atom_v = o.atom_v.data(); o.st.atom_f = o.atom_f.data();
    o.st.atom_w = o.atom_w.data();
  }
  std::vector<mdn_vapor_event> ev(std::max(engine_log_cap(c), 1));
  // the vapor events since the last drain, per file in the reference's order
  // (rebuild by rebuild, molecules in id order)
  auto drain = [&]() -> int {
    for (int r = 0; r < n; r++) {
      int32_t m = 0;
      int e = mdn_drain_vapor_log(c, r, ev.data(), (int32_t)ev.size(), &m);
      if (e) return e;
      std::sort(ev.begin(), ev.begin() + m, [](const mdn_vapor_event &a, const mdn_vapor_event &b) {
        return a.step != b.step ? a.step < b.step : a.mol < b.mol;
      });
      const double dt = out[r].d->dt;
      for (int k = 0; k < m; k++) {
        const mdn_vapor_event &x = ev[k];
        fprintf(x.kind == 0 ? out[r].vin.get() : out[r].vout.get(), "%d %e %e %e %e %e %e %e\n", x.mol,
                (double)x.step * dt, x.dx[0], x.dx[1], x.dx[2], x.dv[0], x.dv[1], x.dv[2]);
      }
    }
    return MDN_OK;
  };
  std::vector<double> obs(6 * (size_t)n);
  auto fetch = [&]() -> int {
    for (int r = 0; r < n; r++)
      if (int e = mdn_get_state(c, r, &out[r].st)) return e;
    return MDN_OK;
  };
  // display + export_dump_close of every replica
  auto observe_all = [&](double time) -> int {
    int e = fetch();
    if (!e) e = mdn_observe(c, -1, obs.data());
    if (e) return e;
    for (int r = 0; r < n; r++) {
      display(out[r], &obs[6 * (size_t)r], time, verbose);
      dump(out[r], time);
    }
    return MDN_OK;
  };
  auto next_stop = [&](long it, long end, long a, long b) {
    long s = std::min(end, (it / a + 1) * a);
    if (b > 0) s = std::min(s, (it / b + 1) * b);
    return std::min(s, it + chunk_max);
  };
  // events of the constructor's make_pair (md.cpp:34), if it was logged
  if ((rc = drain())) return rc;

  // ---- thermal relaxation (run.cpp:18-28) ----
  double time = 0;  // vars->time (variables.hpp:11)
  const long relax = (long)d0.relax_steps, total = (long)d0.total_steps;
  if ((rc = mdn_set_step(c, -1, 0))) return rc;
  for (long it = 0; it < relax;) {
    if (it % observe == 0) {
      if ((rc = observe_all(time))) return rc;
      time = it * d0.dt;
    }
    long nx = next_stop(it, relax, observe, 0);
    // no energy steps in the relaxation (run.cpp:30 is commented out)
    if ((rc = mdn_run_batched(c, (int32_t)(nx - it), 0, it, precision))) return rc;
    if ((rc = drain())) return rc;
    it = nx;
  }
  // ---- re-initialisation (run.cpp:32-56) ----
  time = 0;
  if ((rc = mdn_recentre(c, -1))) return rc;
  if ((rc = mdn_set_step(c, -1, 0))) return rc;
  for (int r = 0; r < n; r++)
    if ((rc = mdn_make_pair(c, r))) return rc;
  if ((rc = drain())) return rc;
  if ((rc = mdn_set_thermostat(c, -1, MDN_NVE, 0))) return rc;
  // ---- main simulation (run.cpp:58-74) ----
  for (long it = 0; it < total;) {
    if (it % kLogger == 0) {
      if ((rc = fetch())) return rc;
      for (int r = 0; r < n; r++) output_ion(out[r], time);
      time += kLogger * d0.dt;
    }
    if (it % observe == 0)
      if ((rc = observe_all(time))) return rc;
    long nx = next_stop(it, total, observe, kLogger);
    // energies on the steps with (itime + 1) % OBSERVE == 0 (run.cpp:68-71)
    if ((rc = mdn_run_batched(c, (int32_t)(nx - it), (int32_t)observe, it, precision))) return rc;
    if ((rc = drain())) return rc;
    it = nx;
  }
  for (int r = 0; r < n; r++)
    for (FILE *f : {out[r].k.get(), out[r].u.get(), out[r].ion.get(), out[r].vin.get(), out[r].vout.get(),
                    out[r].dump.get()})
      if (f && fflush(f)) return set_error(MDN_E_IO, "mdn_run: write failed");
  return MDN_OK;
}
