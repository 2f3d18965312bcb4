/*
 * mdn_oracle.h -- FP64 CPU oracle for the per-step force and integration path
 * of tamadate/MD_nucleation.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is a plain-C restatement of the
 * reference's algorithms (each function cites the reference file:line it
 * follows).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and there only as the checker / CPU baseline.  The product
 * (libmdn, md_nucleation_amd/) never links or calls it.
 *
 * Parity pinning (what checks each function, see tests/test_oracle_*.py):
 *   - forces/energies of every pair and bonded term, kick, drift, COM,
 *     weights, classification, pair lists, loop_update: golden vectors
 *     produced by the reference itself (oracle/_ref/ref_harness built from
 *     the reference sources) on identical states, plus closed forms.
 *   - SHAKE: constraint residual + uniqueness of the fixed point (the
 *     reference never calls shake(); the harness calls it directly).
 *   - Tersoff / Stillinger-Weber: golden vectors from the reference's own
 *     compute() with its neighbour list built by its own make_pair().
 *   - RNG stream (Philox4x32-10): the published Random123 known-answer
 *     vectors; uniforms and Box-Muller normals by moments and a
 *     Kolmogorov-Smirnov test over 10^6 draws; the resampled Maxwell
 *     velocities and N2 spins by their standard deviation sqrt(kB T / m)
 *     (updateGasIn.cpp:18, 43-45, 68-69); the draws themselves have no
 *     reference counterpart (the reference seeds std::random_device), so the
 *     stream is this build's own specification, shared by oracle and GPU.
 *   - gas-gas: the reference's PotentialGasGas::compute run on the intended
 *     in-domain pair list (golden gg.npz), the list by brute force.
 *   - SHAKE mode 1 (textbook update + velocity correction): closed-form
 *     solution of a diatomic constraint.
 */
#ifndef MDN_ORACLE_H
#define MDN_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* interaction bits, same meaning as the reference FLAG fields
 * (flags.hpp:3-66, setCondition.cpp:186-228) */
enum {
  ORC_GI = 1 << 0,          /* gi LJ           (force_lj)      */
  ORC_VI = 1 << 1,          /* vi LJcoul       (inter_vi)      */
  ORC_GG = 1 << 2,          /* gg LJ           (inter_gg)      */
  ORC_GV = 1 << 3,          /* gv LJ           (inter_vg)      */
  ORC_VV = 1 << 4,          /* vv LJcoul       (inter_vv)      */
  ORC_ION_AMBER = 1 << 5,   /* intra_AMBER                     */
  ORC_ION_BMH = 1 << 6,     /* force_born                      */
  ORC_ION_TERSOFF = 1 << 7, /* force_ters                      */
  ORC_ION_SW = 1 << 8,      /* force_sw                        */
  ORC_EFIELD = 1 << 9,      /* efield                          */
  ORC_VAPOR_INTRA = 1 << 10,/* vapor_intra (always 1 in ref)   */
  ORC_GAS_INTRA = 1 << 11   /* gas_intra   (always 1 in ref)   */
};

/* energy buckets in the order of struct Potentials (struct.hpp:21-30) */
enum { U_ION = 0, U_GAS, U_VAP, U_GI, U_GG, U_VG, U_VI, U_VV, U_N };

typedef struct orc_system {
  /* parameter tables */
  int n_types;
  const double *pair_coeff; /* [n_types*n_types*2]: 48 e s^12, 24 e s^6 */
  int n_btypes;
  const double *btype;      /* [n*2]: k, r0 */
  int n_ctypes;
  const double *ctype;      /* [n*2]: k, theta0 (rad) */
  int n_dtypes;
  const int *dmulti;        /* [n] */
  const double *dcoeff;     /* [n*15] (readIonFile.cpp:86-103 layout) */
  /* species 0 = ion, 1 = gas, 2 = vapor */
  int natoms[3];
  const double *mass[3];
  const double *charge[3];
  const int *type[3];
  const double *tmpl_x[3];  /* entry template positions (gas, vapor) */
  int nbond[3];  const int *bond[3];   /* i,j,type   */
  int nangle[3]; const int *angle[3];  /* i,j,k,type */
  int ndih[3];   const int *dih[3];    /* i,j,k,l,type */
  int nrigid[3]; const int *rigid[3];  /* i,j,type (SHAKE) */
  double molmass[3];        /* declared molecule mass (vars->mass) */
  int n_gas, n_vap;
  int ion_type_base;        /* global type of ion-local type 0 (BMH) */
  double dt, T, d_size, margin, ML2;
  double m_gas_kg, m_vap_kg;/* pp->mgas, pp->mvapor */
  int gas_spin;             /* gastype==2: N2 spin at entry */
  uint32_t flags;
  double efield[3];
  int shake;                /* 0: as shipped (verlet.cpp:24); 1: SHAKE with the
                               velocity correction; 2: shake.cpp positions only */
  int shake_max_iter;
  uint64_t seed;
  uint32_t replica;
  int nvt_ion;              /* relaxation thermostat of the ion (NVTion,
                               setCondition.cpp:158-172): 0 off, 1 velocity
                               scaling to 300 K, 2 Nose-Hoover at Tnh_ion */
  double Tnh_ion;
} orc_system;

typedef struct orc_state {
  double *mol_x, *mol_v;    /* [n_mol*3] */
  uint8_t *region;          /* [n_mol] */
  double *atom_x, *atom_v, *atom_f; /* [n_atoms*3] */
  double *atom_w;           /* [n_atoms] */
  int loopPair, loop_update;
  int64_t rebuild_count;
  int cap_gi, cap_gv, cap_gg;
  int n_gi, n_gv, n_gg;
  int *gi, *gv, *gg;        /* pair lists [cap*2] */
  double U[U_N];
  int shake_iters_max;      /* diagnostics: max SHAKE iterations seen */
  int64_t shake_failures;   /* SHAKE calls that hit the cap or diverged */
  int status;               /* 0 ok, <0 error */
  double zeta_ion;          /* Nose-Hoover friction of the ion (vars->zeta_ion) */
} orc_state;

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_create(const orc_system *sys);
void orc_destroy(orc_ctx *c);
int orc_n_atoms(const orc_ctx *c);
int orc_atom_offset(const orc_ctx *c, int mol);
int orc_n_long(const orc_ctx *c);
const int *orc_long_pairs(const orc_ctx *c);

/* setCrossPotential.cpp:4-23 */
void orc_cross_potentials(int n, const double *eps, const double *sigma,
                          double *pair_coeff);
/* Random stream specification shared with the GPU path. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);
void orc_rng_normal2(const orc_system *s, uint32_t mol, int64_t rebuild,
                     uint32_t purpose, double z[2]);
void orc_rng_uniform2(const orc_system *s, uint32_t mol, int64_t rebuild,
                      uint32_t purpose, double u[2]);
void orc_rng_many(const orc_system *s, uint32_t mol0, int n, int64_t rebuild,
                  uint32_t purpose, int normal, double *out);
/* variables.hpp:94-98 */
void orc_rotation(double *X, double *Y, double *Z, double A, double B, double C,
                  double x, double y, double z);
double orc_weight(double r2);

/* pieces of MD::verlet (verlet.cpp) */
void orc_velocity_calculation(orc_ctx *c, orc_state *st);
void orc_update_position(orc_ctx *c, orc_state *st);
void orc_update_position_shake(orc_ctx *c, orc_state *st);
void orc_ion_com_weights(orc_ctx *c, orc_state *st);
/* ion thermostat (nvt.cpp): Tion of Observer::computeProps (observer.cpp:
 * 27-36), MD::velocity_scaling, MD::nosehoover_zeta, MD::nosehoover_ion */
double orc_ion_temperature(const orc_ctx *c, const orc_state *st);
void orc_velocity_scaling(orc_ctx *c, orc_state *st);
void orc_nosehoover_zeta(orc_ctx *c, orc_state *st);
void orc_nosehoover_ion(orc_ctx *c, orc_state *st);
int orc_make_pair(orc_ctx *c, orc_state *st);
int orc_check_pairlist(orc_ctx *c, orc_state *st);
/* every registered Potential::compute in registration order; adds to atom_f */
int orc_forces(orc_ctx *c, orc_state *st, int eflag);
/* one potential by bit (ORC_*) -- adds to atom_f */
int orc_force_term(orc_ctx *c, orc_state *st, uint32_t bit, int eflag);
/* nsteps of MD::verlet; energies on steps with (step+1) % eflag_every == 0 */
int orc_verlet(orc_ctx *c, orc_state *st, int nsteps, int eflag_every,
               int64_t step0, int64_t *atom_steps);
int orc_shake_molecule(orc_ctx *c, double *x, const double *old,
                       int species);
/* one replica per OpenMP thread; seconds (may be NULL) gets each replica's
 * loop time */
int orc_verlet_batch(orc_ctx **ctx, orc_state **st, int n, int nsteps,
                     int eflag_every, int64_t step0, int64_t *atom_steps,
                     int nthreads, double *seconds);

#ifdef __cplusplus
}
#endif
#endif
