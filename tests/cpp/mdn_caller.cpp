// A C++ caller of libmdn through include/mdn.h only, shaped as the
// reference's own code would bind it:
//   * GpuMD mirrors the MD object (md.hpp) with the patch that hands its
//     state to the device at the end of MD::MD (md.cpp:34 make_pair) and
//     turns MD::verlet (verlet.cpp:20-39) into one mdn_verlet step, driven by
//     a stepping loop with run.cpp:68-71's energy cadence;
//   * batched mode is main.cpp:21-30's replica fan-out (calculation number
//     seed + i) as one context and mdn_run_batched;
//   * errors mode checks the host-side error contract (no device needed).
//
// usage:
//   mdn_caller percall <cond> <seed> <nsteps> <every> <out.bin>
//   mdn_caller batched <cond> <seed> <nrep> <nsteps> <every> <out.bin>
//   mdn_caller errors  <cond>
//
// out.bin: int32 n_rep, n_mol, n_atoms; then per replica int64
// rebuild_count, float64 atom_x[3 n_atoms], atom_v[3 n_atoms],
// atom_f[3 n_atoms], U[MDN_U_N], uint8 region[n_mol].
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mdn.h"

namespace {

void ok(int rc, const char *what) {
  if (rc != MDN_OK) throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + mdn_last_error());
}

// caller-owned host state of one replica (the reference's Variables arrays,
// flattened in molecule order)
struct HostState {
  std::vector<double> mol_x, mol_v, atom_x, atom_v, atom_f, atom_w;
  std::vector<uint8_t> region;
  std::vector<int32_t> gi, gv, gg;
  mdn_state st{};
  HostState(int n_mol, int n_atoms, int cap)
      : mol_x(3 * n_mol), mol_v(3 * n_mol), atom_x(3 * n_atoms), atom_v(3 * n_atoms), atom_f(3 * n_atoms),
        atom_w(n_atoms), region(n_mol), gi(2 * cap), gv(2 * cap), gg(2 * cap) {
    st.mol_x = mol_x.data(); st.mol_v = mol_v.data(); st.region = region.data();
    st.atom_x = atom_x.data(); st.atom_v = atom_v.data(); st.atom_f = atom_f.data(); st.atom_w = atom_w.data();
    st.gi = gi.data(); st.gv = gv.data(); st.gg = gg.data();
    st.cap_gi = st.cap_gv = st.cap_gg = cap;
  }
};

struct SysDef {
  mdn_sysdef *h = nullptr;
  int n_mol = 0, n_atoms = 0;
  explicit SysDef(const char *cond) {
    ok(mdn_sysdef_load(cond, &h), "mdn_sysdef_load");
    ok(mdn_sysdef_sizes(h, &n_mol, &n_atoms), "mdn_sysdef_sizes");
  }
  ~SysDef() { mdn_sysdef_free(h); }
};

// MD with the device patch: the constructor ends by handing its state to the
// GPU and building the pair lists; verlet() is one device step
struct GpuMD {
  SysDef &sd;
  HostState hs;
  mdn_ctx *ctx = nullptr;
  int grep = 0;
  int64_t itime = 0;
  GpuMD(SysDef &d, uint64_t calculation_number) : sd(d), hs(d.n_mol, d.n_atoms, 20000) {
    ok(mdn_sysdef_init_state(sd.h, calculation_number, &hs.st), "mdn_sysdef_init_state");
    ok(mdn_create(0, &ctx), "mdn_create");
    ok(mdn_add_replica(ctx, sd.h, &hs.st, calculation_number, 0, &grep), "mdn_add_replica");
    ok(mdn_commit(ctx), "mdn_commit");
    ok(mdn_make_pair(ctx, grep), "mdn_make_pair");  // md.cpp:34
  }
  ~GpuMD() { mdn_destroy(ctx); }
  // MD::verlet; energies on the steps MD::run asks for them (run.cpp:68-71)
  void verlet(int every) {
    ok(mdn_verlet(ctx, grep, 1, every, itime, MDN_FP64), "mdn_verlet");
    itime++;
  }
  void read_back() { ok(mdn_get_state(ctx, grep, &hs.st), "mdn_get_state"); }
};

void write_header(FILE *f, int n_rep, const SysDef &sd) {
  int32_t h[3] = {n_rep, sd.n_mol, sd.n_atoms};
  fwrite(h, sizeof h, 1, f);
}

void write_state(FILE *f, const HostState &hs) {
  int64_t rb = hs.st.rebuild_count;
  fwrite(&rb, sizeof rb, 1, f);
  fwrite(hs.atom_x.data(), sizeof(double), hs.atom_x.size(), f);
  fwrite(hs.atom_v.data(), sizeof(double), hs.atom_v.size(), f);
  fwrite(hs.atom_f.data(), sizeof(double), hs.atom_f.size(), f);
  fwrite(hs.st.U, sizeof(double), MDN_U_N, f);
  fwrite(hs.region.data(), 1, hs.region.size(), f);
}

int run_percall(const char *cond, uint64_t seed, int nsteps, int every, const char *out) {
  SysDef sd(cond);
  GpuMD md(sd, seed);
  for (int i = 0; i < nsteps; i++) md.verlet(every);
  md.read_back();
  FILE *f = fopen(out, "wb");
  if (!f) throw std::runtime_error(std::string("cannot write ") + out);
  write_header(f, 1, sd);
  write_state(f, md.hs);
  fclose(f);
  std::printf("percall: %d steps, %lld rebuilds\n", nsteps, (long long)md.hs.st.rebuild_count);
  return 0;
}

int run_batched(const char *cond, uint64_t seed, int nrep, int nsteps, int every, const char *out) {
  SysDef sd(cond);
  std::vector<HostState> hs;
  hs.reserve(nrep);
  mdn_ctx *ctx = nullptr;
  ok(mdn_create(0, &ctx), "mdn_create");
  // main.cpp:21-30: replica i is calculation number seed + i
  for (int i = 0; i < nrep; i++) {
    hs.emplace_back(sd.n_mol, sd.n_atoms, 20000);
    ok(mdn_sysdef_init_state(sd.h, seed + i, &hs[i].st), "mdn_sysdef_init_state");
    ok(mdn_add_replica(ctx, sd.h, &hs[i].st, seed + i, (uint32_t)i, nullptr), "mdn_add_replica");
  }
  ok(mdn_commit(ctx), "mdn_commit");
  for (int i = 0; i < nrep; i++) ok(mdn_make_pair(ctx, i), "mdn_make_pair");
  ok(mdn_run_batched(ctx, nsteps, every, 0, MDN_FP64), "mdn_run_batched");
  FILE *f = fopen(out, "wb");
  if (!f) throw std::runtime_error(std::string("cannot write ") + out);
  write_header(f, nrep, sd);
  for (int i = 0; i < nrep; i++) {
    int64_t steps = 0, rebuilds = 0, fails = 0;
    int32_t iters = 0;
    ok(mdn_get_counters(ctx, i, &steps, &rebuilds, &iters, &fails), "mdn_get_counters");
    ok(mdn_get_state(ctx, i, &hs[i].st), "mdn_get_state");
    write_state(f, hs[i]);
  }
  fclose(f);
  mdn_destroy(ctx);
  std::printf("batched: %d replicas, %d steps\n", nrep, nsteps);
  return 0;
}

// the host-side error contract: a negative mdn_status and a message
int run_errors(const char *cond) {
  int fails = 0;
  auto expect = [&](int rc, int want, const char *what) {
    const char *msg = mdn_last_error();
    bool good = rc == want && msg && std::strlen(msg) > 0;
    std::printf("%s: rc %d (%s) \"%s\"\n", what, rc, good ? "as expected" : "UNEXPECTED", msg ? msg : "");
    fails += !good;
  };
  mdn_sysdef *bad = nullptr;
  expect(mdn_sysdef_load("/nonexistent/condition.cond", &bad), MDN_E_IO, "load of a missing file");
  SysDef sd(cond);
  expect(mdn_sysdef_set_shake(sd.h, 7, 100), MDN_E_ARG, "SHAKE mode 7");
  mdn_state empty{};
  expect(mdn_sysdef_init_state(sd.h, 1, &empty), MDN_E_ARG, "init_state without buffers");
  HostState hs(sd.n_mol, sd.n_atoms, 16);
  ok(mdn_sysdef_init_state(sd.h, 1, &hs.st), "mdn_sysdef_init_state");
  int ion = 0;
  for (int m = 0; m < sd.n_mol; m++) ion += hs.region[m];
  std::printf("init_state: %d molecules, %d atoms, %d in the domain\n", sd.n_mol, sd.n_atoms, ion);
  if (fails == 0) std::printf("errors: ok\n");
  return fails ? 1 : 0;
}

}  // namespace

int main(int argc, char **argv) {
  try {
    std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "percall" && argc == 7)
      return run_percall(argv[2], std::strtoull(argv[3], nullptr, 10), std::atoi(argv[4]), std::atoi(argv[5]), argv[6]);
    if (mode == "batched" && argc == 8)
      return run_batched(argv[2], std::strtoull(argv[3], nullptr, 10), std::atoi(argv[4]), std::atoi(argv[5]),
                         std::atoi(argv[6]), argv[7]);
    if (mode == "errors" && argc == 3) return run_errors(argv[2]);
    std::fprintf(stderr, "usage: mdn_caller percall|batched|errors ...\n");
    return 2;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "mdn_caller: %s\n", e.what());
    return 1;
  }
}
