/*
 * mdn.h -- C ABI of libmdn, the MI355X (gfx950) per-step force and
 * integration path of the effective-domain nucleation MD method of
 * tamadate/MD_nucleation.
 *
 * Conventions: every entry point returns 0 on success and a negative
 * mdn_status on failure (mdn_last_error() gives the message of the calling
 * thread's last failure).  Host buffers are owned by the caller and copied;
 * device memory is owned by the library.  One mdn_ctx per host thread; a
 * context is not thread-safe.  Units follow the reference: A, fs, g/mol,
 * kcal/mol, e.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   mdn_sysdef_load        MD::readCondFile (src/initialization/setCondition.cpp:5-254),
 *                          Variables::readIonFile (readIonFile.cpp:4-180),
 *                          Variables::readVaporFile (readVaporFile.cpp:4-131),
 *                          Variables::setCrossPotentials (setCrossPotential.cpp:4-23),
 *                          MD::setPotential (setPotential.cpp:7-111)
 *   mdn_sysdef_init_state  Variables::ionInitial (variables.cpp:12-30),
 *                          MD::initialization_gas (initializationGas.cpp:24-90),
 *                          MD::initialization_vapor (initializationVapor.cpp:14-77)
 *   mdn_make_pair          MD::make_pair (src/pairList/pairlist.cpp:16-49) with
 *                          updateInCenters, update_gas_in, update_vapor_in and
 *                          every Potential::makePair (basePotential.hpp:13)
 *   mdn_compute            Potential::compute (basePotential.hpp:12) of the
 *                          registered potentials, as called at verlet.cpp:31-32
 *   mdn_velocity_calculation  MD::velocity_calculation (verlet.cpp:47-65)
 *   mdn_update_position    MD::update_position (verlet.cpp:73-136) and
 *                          MD::updatePositionSHAKE + MD::shake (shake.cpp:6-92)
 *   mdn_verlet             MD::verlet (verlet.cpp:20-39) x nsteps, with the
 *                          energy cadence of MD::run (run.cpp:68-71)
 *   mdn_run_batched        the OpenMP replica fan-out of main (main.cpp:21-30)
 *                          around MD::run's stepping loop, as one persistent
 *                          launch over all replicas
 */
#ifndef MDN_H
#define MDN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mdn_status {
  MDN_OK = 0,
  MDN_E_ARG = -1,       /* invalid argument */
  MDN_E_IO = -2,        /* input file missing or malformed */
  MDN_E_HIP = -3,       /* HIP runtime error */
  MDN_E_CAPACITY = -4,  /* a list or pool exceeded its capacity */
  MDN_E_SHAKE = -5,     /* reserved (SHAKE failures are counted, see mdn_get_counters) */
  MDN_E_STATE = -6      /* call out of order (e.g. before mdn_commit) */
} mdn_status;

/* interaction bits: the FLAG fields set by the Interactions block
 * (flags.hpp:3-66, setCondition.cpp:186-228) */
enum {
  MDN_GI = 1 << 0,  MDN_VI = 1 << 1,  MDN_GG = 1 << 2,  MDN_GV = 1 << 3,
  MDN_VV = 1 << 4,  MDN_ION_AMBER = 1 << 5, MDN_ION_BMH = 1 << 6,
  MDN_ION_TERSOFF = 1 << 7, MDN_ION_SW = 1 << 8, MDN_EFIELD = 1 << 9,
  MDN_VAPOR_INTRA = 1 << 10, MDN_GAS_INTRA = 1 << 11
};
/* energy buckets, order of struct Potentials (struct.hpp:21-30) */
enum { MDN_U_ION = 0, MDN_U_GAS, MDN_U_VAP, MDN_U_GI, MDN_U_GG, MDN_U_VG,
       MDN_U_VI, MDN_U_VV, MDN_U_N };
/* precision of the pair and bonded arithmetic */
enum { MDN_FP64 = 0, MDN_MIXED = 1 };
/* ion thermostat (flags velocity_scaling / nose_hoover_ion, flags.hpp:5-8) */
enum { MDN_NVE = 0, MDN_VSCALE = 1, MDN_NOSE_HOOVER = 2 };

typedef struct mdn_sysdef mdn_sysdef;
typedef struct mdn_ctx mdn_ctx;

/* Read-only view of a parsed system definition (arrays owned by the sysdef).
 * Species 0 = ion, 1 = carrier gas, 2 = vapor.  Molecule order: ion, the
 * n_gas gas molecules, the n_vap vapor molecules (group[0..2] of the
 * reference).  Atom arrays hold every molecule's atoms in that order. */
typedef struct mdn_system {
  int32_t n_types;
  const double *pair_coeff;  /* [n_types*n_types*2] = 48 e s^12, 24 e s^6 */
  const double *type_eps, *type_sigma, *type_mass; /* [n_types] */
  int32_t n_btypes; const double *btype;   /* [n*2] k, r0 */
  int32_t n_ctypes; const double *ctype;   /* [n*2] k, theta0 (rad) */
  int32_t n_dtypes; const int32_t *dmulti; const double *dcoeff; /* [n*15] */
  int32_t natoms[3];
  const double *mass[3], *charge[3];
  const int32_t *type[3];
  const double *tmpl_x[3];
  int32_t nbond[3];  const int32_t *bond[3];
  int32_t nangle[3]; const int32_t *angle[3];
  int32_t ndih[3];   const int32_t *dih[3];
  int32_t nrigid[3]; const int32_t *rigid[3];
  double molmass[3];
  int32_t n_gas, n_vap;
  int32_t ion_type_base;
  double dt, T, pressure, d_size, cutoff, margin, ML2;
  double m_gas_kg, m_vap_kg;
  int32_t gastype, gas_spin;
  uint32_t flags;
  double efield[3];
  int32_t shake, shake_max_iter;
  int64_t total_steps, relax_steps;
  int32_t observe;
  double cluster_radius;     /* >0: vapor pre-packed around the ion */
  /* ion thermostat of the relaxation phase ("NVTion", setCondition.cpp:
   * 158-172): 0 off, 1 velocity scaling to 300 K (nvt.cpp:5-16), 2
   * Nose-Hoover at Tnh_ion (nvt.cpp:19-35) */
  int32_t nvt_ion;
  double Tnh_ion;
} mdn_system;

/* Caller-owned host state of one replica (sizes from mdn_sysdef_sizes). */
typedef struct mdn_state {
  double *mol_x, *mol_v;      /* [n_mol*3] molecule COM (vars->position/velocity) */
  uint8_t *region;            /* [n_mol] */
  double *atom_x, *atom_v, *atom_f; /* [n_atoms*3] */
  double *atom_w;             /* [n_atoms] blend weights */
  int32_t loopPair, loop_update;
  int64_t rebuild_count;
  int32_t cap_gi, cap_gv, cap_gg;
  int32_t n_gi, n_gv, n_gg;
  int32_t *gi, *gv, *gg;      /* [cap*2] pair lists (molecule, ion atom) etc. */
  double U[MDN_U_N];
  double zeta_ion;            /* Nose-Hoover friction of the ion (vars->zeta_ion) */
} mdn_state;

const char *mdn_last_error(void);

/* ---- host front end (initialization stays host C++) ---- */
int mdn_sysdef_load(const char *cond_path, mdn_sysdef **out);
int mdn_sysdef_view(const mdn_sysdef *d, mdn_system *out);
int mdn_sysdef_sizes(const mdn_sysdef *d, int32_t *n_mol, int32_t *n_atoms);
/* SHAKE mode: 0 off (the reference as shipped never calls shake(),
 * verlet.cpp:24); 1 (default under "SHAKE H") Gauss-Seidel SHAKE over the
 * rigid pairs with the textbook update (shake.cpp:67 multiplies by the
 * inverse reduced mass where SHAKE divides; both reach the same constraint
 * solution) plus the velocity correction v = (x_shaken - x_old)/dt of
 * constrained atoms; 2 shake.cpp's own update, positions only.  max_iter caps
 * the iteration. */
int mdn_sysdef_set_shake(mdn_sysdef *d, int32_t shake, int32_t max_iter);
int mdn_sysdef_set_flags(mdn_sysdef *d, uint32_t flags);
int mdn_sysdef_init_state(const mdn_sysdef *d, uint64_t seed, mdn_state *st);
void mdn_sysdef_free(mdn_sysdef *d);

/* ---- device engine ---- */
int mdn_create(int device, mdn_ctx **out);
void mdn_destroy(mdn_ctx *c);
/* adds a replica (copies sysdef tables and state); seed/replica_id key the
 * counter-based random stream of its rebuilds */
int mdn_add_replica(mdn_ctx *c, const mdn_sysdef *d, const mdn_state *st,
                    uint64_t seed, uint32_t replica_id, int32_t *handle);
/* allocates device memory for all replicas and uploads them */
int mdn_commit(mdn_ctx *c);
/* replaces one replica's state (which starts at a step boundary, with no
 * capacity flags); the other replicas keep theirs */
int mdn_set_state(mdn_ctx *c, int32_t replica, const mdn_state *st);
/* copies the replica's state out; MDN_E_CAPACITY (after copying) when a
 * pair list overflowed its capacity on the device (the lists are then
 * truncated) */
int mdn_get_state(mdn_ctx *c, int32_t replica, mdn_state *st);
/* ion thermostat of MD::verlet (verlet.cpp:22, 27, 36; nvt.cpp:5-35) for one
 * replica or all (replica = -1).  MDN_NVE is MD::setNVE (nvt.cpp:38-43, called
 * by run.cpp:56 after the relaxation); MDN_VSCALE scales the ion to 300 K
 * after the first half-kick; MDN_NOSE_HOOVER at Tnh (K) resets zeta_ion to 0.
 * A replica starts in the mode of its condition file's NVTion line. */
int mdn_set_thermostat(mdn_ctx *c, int32_t replica, int32_t mode, double Tnh);

/* ---- observables and outputs (output.cpp, observer.cpp, run.cpp) ------- */
/* Observer::computeProps (observer.cpp:5-36) for one replica or all (-1):
 * out[6 per replica] = Kion, K_g, K_v (kcal/mol), Tion, T_g, T_v (K).  Gas
 * and vapor use every molecule's COM velocity (updated at rebuilds, as in the
 * reference) and declared molecule mass. */
int mdn_observe(mdn_ctx *c, int32_t replica, double *out);

/* One MD::Ovin / MD::Ovout record (output.cpp:43-64, written at each rebuild
 * by updateVaporIn.cpp:33, 39) */
typedef struct mdn_vapor_event {
  int32_t mol;     /* molecule id (0 = ion, gas 1..n_gas, vapor after) */
  int32_t kind;    /* 0: was in and stays in (vapor_in file), 1: leaves (vapor_out) */
  int64_t step;    /* the step counter (MD::itime) at the rebuild; time = step * dt */
  double dx[3];    /* position[mol] - position[0] */
  double dv[3];    /* velocity[mol] - velocity[0] */
} mdn_vapor_event;
/* records up to cap events per replica between drains (0: off, the
 * default); re-uploads the context when called after mdn_commit.  Events not
 * drained are lost by mdn_set_state and mdn_set_vapor_log. */
int mdn_set_vapor_log(mdn_ctx *c, int32_t cap);
/* copies the replica's events recorded since the last drain (in rebuild
 * order, unordered within one rebuild) and clears them; MDN_E_CAPACITY if
 * out holds fewer than recorded (nothing drained) or if events were dropped
 * (the kept ones are returned) */
int mdn_drain_vapor_log(mdn_ctx *c, int32_t replica, mdn_vapor_event *out, int32_t cap, int32_t *n);
/* sets the step counter (MD::itime, run.cpp:22, 41, 61) that stamps vapor
 * events; every step advances it by one; replica -1: all */
int mdn_set_step(mdn_ctx *c, int32_t replica, int64_t step);
/* the re-centre of run.cpp:42-53 as written (only position[0] becomes 0) */
int mdn_recentre(mdn_ctx *c, int32_t replica);

/* MD::run (run.cpp:18-75) for every replica of the context as calculation
 * number calc0 + r (main.cpp:25-29): RelaxSteps under the NVTion thermostat
 * with display/dump every Output steps, the re-centre, MD::make_pair,
 * MD::setNVE, then TotalSteps with the ion file every 10000 steps and
 * display/dump every Output steps, energies on the steps before.  Writes the
 * reference's files under outdir (output.cpp): K_<n>.dat, U_<n>.dat,
 * ion_<T>_<n>.dat, <Output name>_<n>.dump, vapor_in_<n>.dat,
 * vapor_out_<n>.dat (and an empty vapor_collision_<n>.dat).  Steps through
 * mdn_run_batched in the given precision.  Call it where MD::MD ends: after
 * mdn_commit and mdn_make_pair, with the vapor log enabled first if the
 * constructor's rebuild is to be logged (mdn_run enables it otherwise).
 * Replicas must agree on RelaxSteps, TotalSteps and the Output cadence.
 * opts: MDN_RUN_VERBOSE prints each display() block to stdout. */
enum { MDN_RUN_VERBOSE = 1 };
int mdn_run(mdn_ctx *c, const char *outdir, int32_t calc0, int32_t precision, uint32_t opts);

/* atom positions (reference order, [n_atoms][3]) into the replica's device
 * state, nothing else re-uploaded; and the replica's atom forces back.  No
 * reference counterpart: for callers that move atoms themselves between
 * force evaluations (the trajectory-segment adjoint) */
int mdn_set_positions(mdn_ctx *c, int32_t replica, const double *atom_x);
int mdn_get_forces(mdn_ctx *c, int32_t replica, double *atom_f);
/* derivative of the forces of the selected terms (bits as mdn_compute; gi,
 * vi, gg, gv, vv, AMBER, vapor and gas intra; Efield adds nothing; BMH,
 * Tersoff and SW: MDN_E_ARG) along the atom displacement atom_u ([n_atoms][3],
 * reference order) at the replica's positions, in FP64: atom_hu = d/de
 * F(x + e u) at e = 0, the Hessian of -U applied to u, by forward-mode
 * differentiation of the force kernels (exact to rounding).  Pair lists and
 * blend weights are held as they stand.  No reference counterpart: the
 * second-order terms of the trajectory-segment adjoint. */
int mdn_force_jvp(mdn_ctx *c, int32_t replica, uint32_t bits, const double *atom_u, double *atom_hu);

/* drop-in path: one replica, FP64, one launch per reference call */
int mdn_make_pair(mdn_ctx *c, int32_t replica);
int mdn_zero_forces(mdn_ctx *c, int32_t replica);
int mdn_compute(mdn_ctx *c, int32_t replica, uint32_t bits, int32_t eflag,
                int32_t precision);
int mdn_velocity_calculation(mdn_ctx *c, int32_t replica);
int mdn_update_position(mdn_ctx *c, int32_t replica);
int mdn_verlet(mdn_ctx *c, int32_t replica, int32_t nsteps,
               int32_t eflag_every, int64_t step0, int32_t precision);

/* mdn_verlet with each step's kernel sequence captured once into a hipGraph
 * (per replica, precision and energy flag) and replayed; the same arithmetic
 * as mdn_verlet.  No reference counterpart: launch-overhead relief for a
 * single large replica. */
int mdn_verlet_graph(mdn_ctx *c, int32_t replica, int32_t nsteps,
                     int32_t eflag_every, int64_t step0, int32_t precision);

/* batched path: all replicas, nsteps in one persistent launch.
 * A replica whose in-domain atoms exceed the launch's LDS capacity is
 * flagged (mdn_get_counters returns MDN_E_CAPACITY) and skipped from then
 * on (until mdn_set_state gives it a new state).  When that happens at a
 * launch start its state is that of the last completed step; when it
 * happens at a rebuild inside a step, the stored state is that step in
 * progress: positions, blend weights and pair lists of the step after its
 * rebuild, velocities after the first half-kick, forces cleared.
 * mdn_verlet and mdn_verlet_graph then finish that step first (forces,
 * second half-kick), counting it as the first of their nsteps;
 * mdn_make_pair and mdn_update_position return MDN_E_STATE until it is
 * finished (mdn_compute followed by mdn_velocity_calculation also finishes
 * it). */
int mdn_run_batched(mdn_ctx *c, int32_t nsteps, int32_t eflag_every,
                    int64_t step0, int32_t precision);

/* launch shape of mdn_run_batched: threads per workgroup (64..512; 0 = split:
 * the heaviest replicas in 512-thread workgroups, one per CU, the rest in
 * 128-thread workgroups, the default), LDS
 * capacity in in-domain atoms per replica (0 = sized from the current
 * in-domain counts), grid size (0 = one workgroup per replica), steps per
 * launch (0 = all; between launches the LDS capacity is re-sized).  No
 * reference counterpart: the batched path's tuning knobs. */
int mdn_set_batch_shape(mdn_ctx *c, int32_t threads, int32_t lds_cap_atoms, int32_t grid, int32_t chunk);
/* tuning of the split launch (threads = 0 above): workgroup size of the
 * replicas with a CU of their own (256 or 512, default 512) and of those
 * sharing CUs (64, 128 or 256, default 128); a replica may get a CU of its
 * own when its cost exceeds heavy_factor times the median (default 2).
 * Results do not depend on these, only the accumulation order of forces
 * does.  No reference counterpart: the autotuner's knobs. */
int mdn_set_batch_tuning(mdn_ctx *c, int32_t own_threads, int32_t shared_threads, double heavy_factor);

/* LDS capacity (in-domain atoms per replica), LDS bytes per workgroup and the
 * number of replicas given a CU of their own, of the last mdn_run_batched
 * launch.  No reference counterpart. */
int mdn_get_batch_info(mdn_ctx *c, int32_t *lds_cap_atoms, int32_t *lds_bytes, int32_t *heavy);

/* launch plan of the last chunk of the last mdn_run_batched call:
 * counts[4] = replicas in the own-CU (512-thread), three-per-CU and
 * four-per-CU (128-thread) classes of a split launch, and in a uniform
 * launch (mdn_set_batch_shape threads != 0, or no more replicas than CUs);
 * order (may be NULL; at least as many entries as replicas) receives the
 * replica ids in launch order: own-CU class first, then the three-per-CU
 * and the four-per-CU classes.  No reference counterpart. */
int mdn_get_batch_plan(mdn_ctx *c, int32_t *counts, int32_t *order, int32_t cap);

/* ---- differentiable vi term (fitting ion-vapor parameters) ----
 * PotentialVaporIon::compute (src/potential/VaporIon.cpp:5-43) over S
 * snapshots with the ion's per-atom sigma, eps, q as inputs (mixing of
 * setCrossPotential.cpp:13-20), FP64 on the device.  Host pointers (copied
 * in and out): x [S][n_ion+n_vap][3] (ion atoms first), w [S][n_ion+n_vap]
 * blend weights, vap_in [S][n_vap] 1 where the vapor atom's molecule is in
 * the domain (region flag; NULL = all), E [S], F [S][n_ion+n_vap][3].
 * mdn_vi_adjoint writes
 * dL/dsigma, dL/deps, dL/dq of the ion atoms given dL/dE (gE, may be NULL)
 * and dL/dF (gF, may be NULL), by a hand-written reverse pass.  No reference
 * counterpart (the reference has no gradients): the fitting path of the
 * force-matching workload. */
int mdn_vi_forward(const double *x, const double *w, int32_t S, int32_t n_ion,
                   int32_t n_vap, const double *ion_sigma,
                   const double *ion_eps, const double *ion_q,
                   const double *vap_sigma, const double *vap_eps,
                   const double *vap_q, const uint8_t *vap_in, double *E,
                   double *F);
int mdn_vi_adjoint(const double *x, const double *w, int32_t S, int32_t n_ion,
                   int32_t n_vap, const double *ion_sigma,
                   const double *ion_eps, const double *ion_q,
                   const double *vap_sigma, const double *vap_eps,
                   const double *vap_q, const uint8_t *vap_in,
                   const double *gE, const double *gF,
                   double *g_sigma, double *g_eps, double *g_q);
/* derivative of mdn_vi_forward's forces along the displacement u
 * ([S][n_ion + n_vap][3]) of the snapshots' atoms, into hu (same shape):
 * the Hessian of -E applied to u, closed form.  No reference counterpart:
 * the second-order terms of the trajectory-segment adjoint. */
int mdn_vi_jvp(const double *x, const double *w, int32_t S, int32_t n_ion, int32_t n_vap,
               const double *ion_sigma, const double *ion_eps, const double *ion_q,
               const double *vap_sigma, const double *vap_eps, const double *vap_q,
               const uint8_t *vap_in, const double *u, double *hu);
/* Trajectory-segment gradient, every state array in device memory: nsteps
 * steps of MD::verlet as shipped (verlet.cpp:20-39: half kick, drift,
 * forces, half kick; no SHAKE, no rebuild, every vapor molecule in the
 * domain with blend weight 1) of the segment atoms (the ion's atoms, then
 * every vapor atom in molecule order; N of them) of replica r, from x0, v0
 * ([N][3] host arrays; NULL = the replica's current ones), with the vi term
 * (VaporIon.cpp:5-43) taking the ion's per-atom sigma, eps, q as inputs and
 * the other selected ion and vapor terms as mdn_compute runs them.  Loss:
 * sum over the observed steps obs[k] (strictly ascending in [0, nsteps]) of
 * mean over atoms of |x_obs[k] - targets[k]|^2 (targets [n_obs][N][3]).
 * Writes *loss and, when g_sigma, g_eps, g_q are given (all three or none),
 * dloss/d(sigma, eps, q) of the ion atoms by a checkpointed reverse pass
 * (checkpoints every `every` steps; exact Hessian-vector products); x_end,
 * v_end (may be NULL) receive the last state.  The replica's own state is
 * left as it was.  No reference counterpart (the reference has no
 * gradients): the segment gradients of the fitting workload. */
int mdn_segment_grad(mdn_ctx *c, int32_t replica, int32_t nsteps, int32_t every,
                     const double *x0, const double *v0,
                     const double *ion_sigma, const double *ion_eps,
                     const double *ion_q, const int32_t *obs, int32_t n_obs,
                     const double *targets, double *loss, double *g_sigma,
                     double *g_eps, double *g_q, double *x_end, double *v_end);
/* mdn_vi_adjoint for callers that cannot allocate (code under HIP graph
 * capture, tight fitting loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Every pointer is a DEVICE pointer on the
 * current device (same layouts as mdn_vi_adjoint); the kernels are queued on
 * `stream` (a hipStream_t; NULL = the null stream) and the call returns
 * without waiting for them.  It allocates nothing: the partial gradients go
 * to a fixed 4 MiB device array of the library, which bounds
 * S * n_ion <= 174762 (larger batches return MDN_E_ARG; use mdn_vi_adjoint).
 * Each GPU holds one such array for the whole process and every call on that
 * GPU uses it, so calls that overlap on one GPU (concurrent streams or host
 * threads) race; calls in one stream are ordered and safe. */
int mdn_vi_adjoint_noalloc(const double *x, const double *w, int32_t S,
                           int32_t n_ion, int32_t n_vap,
                           const double *ion_sigma, const double *ion_eps,
                           const double *ion_q, const double *vap_sigma,
                           const double *vap_eps, const double *vap_q,
                           const uint8_t *vap_in, const double *gE,
                           const double *gF, double *g_sigma, double *g_eps,
                           double *g_q, void *stream);
/* Coulomb constant of the reference (constants.hpp:50), kcal A/(mol e^2) */
double mdn_qqrd2e(void);

/* counters: in-domain atom-steps since commit, rebuilds, max SHAKE
 * iterations, SHAKE calls that failed (cap reached or diverged: that molecule
 * then stays unconstrained for the step).  Returns MDN_E_CAPACITY when a
 * list or the batched kernel's LDS capacity overflowed. */
int mdn_get_counters(mdn_ctx *c, int32_t replica, int64_t *atom_steps,
                     int64_t *rebuilds, int32_t *shake_iters_max,
                     int64_t *shake_failures);
/* device synchronisation and timing of the last mdn_run_batched (sum of its
 * launches) / mdn_verlet call (HIP events on the library's stream), ms */
int mdn_sync(mdn_ctx *c);
int mdn_last_kernel_ms(mdn_ctx *c, double *ms);

#ifdef __cplusplus
}
#endif
#endif
