/*
 * mdn_oracle.c -- FP64 CPU oracle (TEST INFRASTRUCTURE ONLY, see header).
 *
 * Plain-C restatement of the reference's per-step path.  Loop orders and
 * expression association follow the reference so that results agree with it
 * to rounding; every function cites the reference lines it restates.
 * Build with -ffp-contract=off (the reference is built for x86-64 without FMA).
 */
#include "mdn_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* constants.hpp:28-32, 44-50 */
static const double Nw = 6.02e23;
static const double kb = 1.38e-23;
static const double e0 = 8.85e-12;
static const double echarge = 1.602e-19;
static const double Rinter = 100;
static const double Overlap = 10;
#define Rcenter (Rinter - Overlap * 0.5)
#define Ri2 (Rinter * Rinter)
#define RI2 ((Rinter + Overlap) * (Rinter + Overlap))

/* constants.hpp:30, 49-50; observer.hpp:6 */
static const double kb_real = 1.38e-23 * 6.02e23 / 4184.0;
static const double real_to_kcalmol = 10000 / 4.184;
#define kb_real_inv (1 / kb_real)

static double qqrd2e(void) {
  return echarge * echarge / 4.0 / M_PI / e0 * Nw * 1e10 / 4184.0;
}

struct orc_ctx {
  orc_system s;
  int n_mol, n_atoms;
  int off_gas, off_vap; /* atom offset of first gas / vapor molecule */
  int n_long;
  int *long_pairs;
  double qqrd2e;
};

static int mol_offset(const orc_ctx *c, int i) {
  if (i == 0) return 0;
  if (i <= c->s.n_gas) return c->off_gas + (i - 1) * c->s.natoms[1];
  return c->off_vap + (i - 1 - c->s.n_gas) * c->s.natoms[2];
}
static int mol_species(const orc_ctx *c, int i) {
  return i == 0 ? 0 : (i <= c->s.n_gas ? 1 : 2);
}
int orc_atom_offset(const orc_ctx *c, int mol) { return mol_offset(c, mol); }
int orc_n_atoms(const orc_ctx *c) { return c->n_atoms; }
int orc_n_long(const orc_ctx *c) { return c->n_long; }
const int *orc_long_pairs(const orc_ctx *c) { return c->long_pairs; }

/* PotentialAMBER::initialAMBER, compute_AMBER.cpp:256-298 */
static void initial_amber(orc_ctx *c) {
  const orc_system *s = &c->s;
  int ne = s->nbond[0] + s->nrigid[0] + s->nangle[0] + s->ndih[0];
  int *ex = (int *)malloc(sizeof(int) * 2 * (ne > 0 ? ne : 1));
  int k = 0;
  for (int b = 0; b < s->nbond[0]; b++) { ex[2*k] = s->bond[0][3*b]; ex[2*k+1] = s->bond[0][3*b+1]; k++; }
  for (int b = 0; b < s->nrigid[0]; b++) { ex[2*k] = s->rigid[0][3*b]; ex[2*k+1] = s->rigid[0][3*b+1]; k++; }
  for (int b = 0; b < s->nangle[0]; b++) { ex[2*k] = s->angle[0][4*b]; ex[2*k+1] = s->angle[0][4*b+2]; k++; }
  for (int b = 0; b < s->ndih[0]; b++) { ex[2*k] = s->dih[0][5*b]; ex[2*k+1] = s->dih[0][5*b+3]; k++; }
  int is = s->natoms[0];
  c->long_pairs = (int *)malloc(sizeof(int) * 2 * (is > 1 ? (size_t)is * (is - 1) / 2 : 1));
  c->n_long = 0;
  for (int i = 0; i < is - 1; i++)
    for (int j = i + 1; j < is; j++) {
      int flag = 1;
      for (int a = 0; a < ne; a++)
        if ((ex[2*a] == i && ex[2*a+1] == j) || (ex[2*a] == j && ex[2*a+1] == i)) { flag = 0; break; }
      if (flag) { c->long_pairs[2*c->n_long] = i; c->long_pairs[2*c->n_long+1] = j; c->n_long++; }
    }
  free(ex);
}

orc_ctx *orc_create(const orc_system *sys) {
  orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
  c->s = *sys;
  c->n_mol = 1 + sys->n_gas + sys->n_vap;
  c->off_gas = sys->natoms[0];
  c->off_vap = c->off_gas + sys->n_gas * sys->natoms[1];
  c->n_atoms = c->off_vap + sys->n_vap * sys->natoms[2];
  c->qqrd2e = qqrd2e();
  if (sys->flags & ORC_ION_AMBER) initial_amber(c);
  return c;
}

void orc_destroy(orc_ctx *c) {
  if (!c) return;
  free(c->long_pairs);
  free(c);
}

/* setCrossPotential.cpp:4-23 (Lorentz-Berthelot) */
void orc_cross_potentials(int n, const double *eps, const double *sigma, double *pc) {
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double epu = sqrt(eps[i] * eps[j]);
      double sig = (sigma[i] + sigma[j]) * 0.5;
      pc[(i * n + j) * 2 + 0] = 48 * epu * pow(sig, 12.0);
      pc[(i * n + j) * 2 + 1] = 24 * epu * pow(sig, 6.0);
    }
}

/* ---------------- random stream (this build's specification) ------------- */
/* Philox4x32-10 (Salmon et al., SC'11), counter-based. */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; r++) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void rng_words(const orc_system *s, uint32_t mol, int64_t rebuild, uint32_t purpose, uint32_t w[4]) {
  uint32_t ctr[4] = {mol, (uint32_t)(uint64_t)rebuild, (uint32_t)((uint64_t)rebuild >> 32), purpose};
  uint32_t key[2] = {(uint32_t)s->seed, (uint32_t)(s->seed >> 32) ^ s->replica};
  orc_philox4x32_10(ctr, key, w);
}
/* [0,1) with 53 random bits */
static double u01(uint32_t hi, uint32_t lo) {
  uint64_t v = ((uint64_t)hi << 32) | lo;
  return (double)(v >> 11) * (1.0 / 9007199254740992.0);
}
void orc_rng_uniform2(const orc_system *s, uint32_t mol, int64_t rebuild, uint32_t purpose, double u[2]) {
  uint32_t w[4];
  rng_words(s, mol, rebuild, purpose, w);
  u[0] = u01(w[0], w[1]);
  u[1] = u01(w[2], w[3]);
}
/* Box-Muller on (0,1] x [0,1) */
void orc_rng_normal2(const orc_system *s, uint32_t mol, int64_t rebuild, uint32_t purpose, double z[2]) {
  uint32_t w[4];
  rng_words(s, mol, rebuild, purpose, w);
  uint64_t v = ((uint64_t)w[0] << 32) | w[1];
  double u1 = (double)((v >> 11) + 1) * (1.0 / 9007199254740992.0);
  double u2 = u01(w[2], w[3]);
  double r = sqrt(-2.0 * log(u1));
  double th = 2.0 * M_PI * u2;
  z[0] = r * cos(th);
  z[1] = r * sin(th);
}

/* n draws of one purpose for molecules mol0 .. mol0 + n - 1 (the moment
 * tests): two uniforms or two normals per molecule into out[2n] */
void orc_rng_many(const orc_system *s, uint32_t mol0, int n, int64_t rebuild, uint32_t purpose, int normal,
                  double *out) {
  for (int k = 0; k < n; k++) {
    if (normal) orc_rng_normal2(s, mol0 + (uint32_t)k, rebuild, purpose, out + 2 * k);
    else orc_rng_uniform2(s, mol0 + (uint32_t)k, rebuild, purpose, out + 2 * k);
  }
}

/* variables.hpp:94-98 */
void orc_rotation(double *X, double *Y, double *Z, double A, double B, double C, double x, double y, double z) {
  *X = cos(C) * (x * cos(B) + y * sin(A) * sin(B) - z * cos(A) * sin(B)) + sin(C) * (y * cos(A) + z * sin(A));
  *Y = -sin(C) * (x * cos(B) + y * sin(A) * sin(B) - z * cos(A) * sin(B)) + cos(C) * (y * cos(A) + z * sin(A));
  *Z = x * sin(B) - y * sin(A) * cos(B) + z * cos(A) * cos(B);
}

/* verlet.cpp:117-130 */
double orc_weight(double r2) {
  if (r2 < Ri2) return 1;
  if (r2 > RI2) return 0;
  double dr = (sqrt(r2) - Rcenter);
  double C = cos(M_PI / Overlap * 0.5 * dr);
  return C * C;
}

/* ---------------- integrator -------------------------------------------- */
/* MD::velocity_calculation, verlet.cpp:47-65 */
void orc_velocity_calculation(orc_ctx *c, orc_state *st) {
  const orc_system *s = &c->s;
  double const Coeff = 0.5 * s->dt * 4.184e-4;
  for (int i = 0; i < c->n_mol; i++) {
    if (!st->region[i]) continue;
    int sp = mol_species(c, i), o = mol_offset(c, i);
    for (int ii = 0; ii < s->natoms[sp]; ii++) {
      double Coeff2 = Coeff / s->mass[sp][ii];
      double *v = st->atom_v + 3 * (o + ii), *f = st->atom_f + 3 * (o + ii);
      v[0] += f[0] * Coeff2;
      v[1] += f[1] * Coeff2;
      v[2] += f[2] * Coeff2;
    }
  }
}

/* ion COM + blend weights, verlet.cpp:90-134 */
void orc_ion_com_weights(orc_ctx *c, orc_state *st) {
  const orc_system *s = &c->s;
  double *x = st->mol_x, *v = st->mol_v;
  x[0] = 0; x[1] = 0; x[2] = 0;
  v[0] = 0; v[1] = 0; v[2] = 0;
  for (int ii = 0; ii < s->natoms[0]; ii++) {
    const double *xi = st->atom_x + 3 * ii, *vi = st->atom_v + 3 * ii;
    double m = s->mass[0][ii];
    x[0] += xi[0] * m; x[1] += xi[1] * m; x[2] += xi[2] * m;
    v[0] += vi[0] * m; v[1] += vi[1] * m; v[2] += vi[2] * m;
  }
  double mInv = 1 / s->molmass[0];
  x[0] *= mInv; x[1] *= mInv; x[2] *= mInv;
  v[0] *= mInv; v[1] *= mInv; v[2] *= mInv;
  for (int i = 1; i < c->n_mol; i++) {
    if (!st->region[i]) continue;
    int sp = mol_species(c, i), o = mol_offset(c, i);
    for (int ii = 0; ii < s->natoms[sp]; ii++) {
      const double *xi = st->atom_x + 3 * (o + ii);
      double dx = xi[0] - x[0], dy = xi[1] - x[1], dz = xi[2] - x[2];
      double r2 = dx * dx + dy * dy + dz * dz;
      st->atom_w[o + ii] = orc_weight(r2);
    }
  }
}

/* MD::update_position, verlet.cpp:73-136 */
void orc_update_position(orc_ctx *c, orc_state *st) {
  const orc_system *s = &c->s;
  double dt = s->dt;
  for (int i = 0; i < c->n_mol; i++) {
    if (!st->region[i]) continue;
    int sp = mol_species(c, i), o = mol_offset(c, i);
    for (int ii = 0; ii < s->natoms[sp]; ii++) {
      double *x = st->atom_x + 3 * (o + ii), *v = st->atom_v + 3 * (o + ii), *f = st->atom_f + 3 * (o + ii);
      x[0] += v[0] * dt; x[1] += v[1] * dt; x[2] += v[2] * dt;
      f[0] = f[1] = f[2] = 0.0;
    }
  }
  orc_ion_com_weights(c, st);
}

/* MD::shake, shake.cpp:44-92 (positions only; iteration cap added) */
int orc_shake_molecule(orc_ctx *c, double *x, const double *r, int sp) {
  const orc_system *s = &c->s;
  const double tolerance = 1e-20, rho_zero = 0.01;
  int l = s->nrigid[sp];
  const double *m = s->mass[sp];
  int it = 0, converge = 0;
  while (!converge) {
    if (it >= s->shake_max_iter) return -it;
    it++;
    converge = 1;
    for (int k = 0; k < l; k++) {
      int i = s->rigid[sp][3 * k], j = s->rigid[sp][3 * k + 1], ty = s->rigid[sp][3 * k + 2];
      double dk = s->btype[2 * ty + 1];
      double rkx = r[3*i] - r[3*j], rky = r[3*i+1] - r[3*j+1], rkz = r[3*i+2] - r[3*j+2];
      double rk2 = rkx * rkx + rky * rky + rkz * rkz;
      double _rkx = x[3*i] - x[3*j], _rky = x[3*i+1] - x[3*j+1], _rkz = x[3*i+2] - x[3*j+2];
      double _rk2 = _rkx * _rkx + _rky * _rky + _rkz * _rkz;
      double dot = rkx * _rkx + rky * _rky + rkz * _rkz;
      /* mode 1 tests each constraint before its update (a later update in
       * the sweep can break it again); mode 2 tests after it, as shake.cpp */
      if (s->shake != 2) {
        double d0 = _rk2 - dk * dk;
        if (d0 * d0 > tolerance) converge = 0;
      }
      double mred_inv = 1 / m[i] + 1 / m[j];
      /* mode 2 restates shake.cpp:67 (multiplies by the inverse reduced
       * mass); mode 1 divides by it (textbook SHAKE, as a product with
       * 0.5 / mred_inv), same fixed point */
      double ramk = s->shake == 2 ? (dk * dk - _rk2) / 2 * mred_inv : (dk * dk - _rk2) * (0.5 / mred_inv);
      if (dot > 1e-10) ramk /= dot;
      else ramk = 0.5 * rho_zero + ramk / (rk2 * rho_zero) * mred_inv;
      double coeffi = ramk / m[i], coeffj = ramk / m[j];
      x[3*i] += rkx * coeffi; x[3*i+1] += rky * coeffi; x[3*i+2] += rkz * coeffi;
      x[3*j] -= rkx * coeffj; x[3*j+1] -= rky * coeffj; x[3*j+2] -= rkz * coeffj;
      _rkx = x[3*i] - x[3*j]; _rky = x[3*i+1] - x[3*j+1]; _rkz = x[3*i+2] - x[3*j+2];
      _rk2 = _rkx * _rkx + _rky * _rky + _rkz * _rkz;
      double dr = _rk2 - dk * dk;
      if (!(dr * dr <= 1e30)) return -it; /* diverged (or non-finite) */
      if (s->shake == 2 && dr * dr > tolerance) converge = 0;
    }
  }
  return it;
}

/* MD::updatePositionSHAKE, shake.cpp:6-40, followed by the COM/weights block
 * of update_position (verlet.cpp:90-134) */
void orc_update_position_shake(orc_ctx *c, orc_state *st) {
  const orc_system *s = &c->s;
  double dt = s->dt;
  double old[3 * 256];
  for (int i = 0; i < c->n_mol; i++) {
    if (!st->region[i]) continue;
    int sp = mol_species(c, i), o = mol_offset(c, i), na = s->natoms[sp];
    double *oldp = old;
    double *heap = NULL;
    if (s->nrigid[sp]) {
      if (na > 256) { heap = (double *)malloc(sizeof(double) * 3 * na); oldp = heap; }
      memcpy(oldp, st->atom_x + 3 * o, sizeof(double) * 3 * na);
    }
    for (int ii = 0; ii < na; ii++) {
      double *x = st->atom_x + 3 * (o + ii), *v = st->atom_v + 3 * (o + ii), *f = st->atom_f + 3 * (o + ii);
      x[0] += v[0] * dt; x[1] += v[1] * dt; x[2] += v[2] * dt;
      f[0] = f[1] = f[2] = 0.0;
    }
    if (s->nrigid[sp]) {
      int it = orc_shake_molecule(c, st->atom_x + 3 * o, oldp, sp);
      if (it < 0) {
        /* a failed SHAKE leaves the molecule unconstrained for this step, as
         * the reference (which never calls shake()) always does */
        st->shake_failures++;
        it = -it;
        if (it > st->shake_iters_max) st->shake_iters_max = it;
        for (int ii = 0; ii < 3 * na; ii++) st->atom_x[3 * o + ii] = oldp[ii] + st->atom_v[3 * o + ii] * dt;
        free(heap);
        continue;
      }
      if (it > st->shake_iters_max) st->shake_iters_max = it;
      /* mode 1: the half-step velocity of every constrained atom becomes the
       * constrained displacement over dt (SHAKE in velocity Verlet); mode 2
       * keeps shake.cpp's positions-only update */
      if (s->shake == 1)
        for (int ii = 0; ii < na; ii++) {
          int cons = 0;
          for (int k = 0; k < s->nrigid[sp]; k++)
            if (s->rigid[sp][3 * k] == ii || s->rigid[sp][3 * k + 1] == ii) { cons = 1; break; }
          if (!cons) continue;
          double *x = st->atom_x + 3 * (o + ii), *v = st->atom_v + 3 * (o + ii);
          const double *xo = oldp + 3 * ii;
          v[0] = (x[0] - xo[0]) / dt; v[1] = (x[1] - xo[1]) / dt; v[2] = (x[2] - xo[2]) / dt;
        }
    }
    free(heap);
  }
  orc_ion_com_weights(c, st);
}

/* ---------------- rebuild ------------------------------------------------ */
/* MD::updateInCenters, updateCenters.cpp:6-57 (for the current in-lists) */
static void update_in_centers(orc_ctx *c, orc_state *st) {
  const orc_system *s = &c->s;
  for (int i = 1; i < c->n_mol; i++) {
    if (!st->region[i]) continue;
    int sp = mol_species(c, i), o = mol_offset(c, i);
    double *m = st->mol_x + 3 * i, *mv = st->mol_v + 3 * i;
    m[0] = m[1] = m[2] = mv[0] = mv[1] = mv[2] = 0;
    for (int ii = 0; ii < s->natoms[sp]; ii++) {
      const double *xi = st->atom_x + 3 * (o + ii), *vi = st->atom_v + 3 * (o + ii);
      double mi = s->mass[sp][ii];
      m[0] += xi[0] * mi; m[1] += xi[1] * mi; m[2] += xi[2] * mi;
      mv[0] += vi[0] * mi; mv[1] += vi[1] * mi; mv[2] += vi[2] * mi;
    }
    double massInv = 1 / s->molmass[sp];
    m[0] *= massInv; m[1] *= massInv; m[2] *= massInv;
    mv[0] *= massInv; mv[1] *= massInv; mv[2] *= massInv;
  }
}

/* makeDiatomicProp_in (updateGasIn.cpp:51-98) / makePolyatomicProp_in
 * (updateVaporIn.cpp:55-82): template atoms, random rotation, COM shift */
static void enter_molecule(orc_ctx *c, orc_state *st, int i) {
  const orc_system *s = &c->s;
  int sp = mol_species(c, i), o = mol_offset(c, i), na = s->natoms[sp];
  double sigma = sqrt(kb * s->T / (sp == 1 ? s->m_gas_kg : s->m_vap_kg));
  double *x = st->atom_x + 3 * o, *v = st->atom_v + 3 * o;
  memcpy(x, s->tmpl_x[sp], sizeof(double) * 3 * na);
  for (int ii = 0; ii < 3 * na; ii++) v[ii] = 0;
  if (sp == 2)
    for (int ii = 0; ii < 3 * na; ii++) st->atom_f[3 * o + ii] = 0;
  if (sp == 1 && s->gas_spin && na >= 2) {
    double z[2];
    orc_rng_normal2(s, (uint32_t)i, st->rebuild_count, 4, z);
    double vy_rot = sigma * z[0] * 1e-5;
    v[0] = 0; v[1] = vy_rot; v[2] = vy_rot;
    v[3] = 0; v[4] = -vy_rot; v[5] = -vy_rot;
  }
  double u[2], u3[2];
  orc_rng_uniform2(s, (uint32_t)i, st->rebuild_count, 2, u);
  orc_rng_uniform2(s, (uint32_t)i, st->rebuild_count, 3, u3);
  double a = u[0] * 2 * M_PI, b = u[1] * 2 * M_PI, cc = u3[0] * 2 * M_PI;
  const double *X = st->mol_x + 3 * i, *V = st->mol_v + 3 * i;
  for (int ii = 0; ii < na; ii++) {
    double *p = x + 3 * ii;
    /* gas passes (x, y, y) (updateGasIn.cpp:84) */
    double zarg = sp == 1 ? p[1] : p[2];
    orc_rotation(&p[0], &p[1], &p[2], a, b, cc, p[0], p[1], zarg);
    p[0] += X[0]; p[1] += X[1]; p[2] += X[2];
  }
  for (int ii = 0; ii < na; ii++) {
    double *p = v + 3 * ii;
    double zarg = sp == 1 ? p[1] : p[2];
    orc_rotation(&p[0], &p[1], &p[2], a, b, cc, p[0], p[1], zarg);
    p[0] += V[0]; p[1] += V[1]; p[2] += V[2];
  }
}

/* update_gas_in (updateGasIn.cpp:10-48) and update_vapor_in
 * (updateVaporIn.cpp:10-52) */
static void classify(orc_ctx *c, orc_state *st, int first, int count, double m_kg) {
  const orc_system *s = &c->s;
  double HL = s->d_size * 0.5;
  double sigma = sqrt(kb * s->T / m_kg);
  double *x = st->mol_x, *v = st->mol_v;
  for (int i = first; i < first + count; i++) {
    double dx = x[3*i] - x[0], dy = x[3*i+1] - x[1], dz = x[3*i+2] - x[2];
    double r2 = (dx * dx + dy * dy + dz * dz);
    if (r2 < RI2) {
      if (!st->region[i]) enter_molecule(c, st, i);
      st->region[i] = 1;
    } else {
      st->region[i] = 0;
      if (dx < -HL) x[3*i] += s->d_size;
      if (dy < -HL) x[3*i+1] += s->d_size;
      if (dz < -HL) x[3*i+2] += s->d_size;
      if (dx > HL) x[3*i] -= s->d_size;
      if (dy > HL) x[3*i+1] -= s->d_size;
      if (dz > HL) x[3*i+2] -= s->d_size;
      double z[2], z2[2];
      orc_rng_normal2(s, (uint32_t)i, st->rebuild_count, 0, z);
      orc_rng_normal2(s, (uint32_t)i, st->rebuild_count, 1, z2);
      v[3*i] = sigma * z[0] * 1e-5;
      v[3*i+1] = sigma * z[1] * 1e-5;
      v[3*i+2] = sigma * z2[0] * 1e-5;
    }
  }
}

static int push_pair(int *list, int *n, int cap, int i, int j) {
  if (*n < cap) { list[2 * *n] = i; list[2 * *n + 1] = j; }
  (*n)++;
  return *n <= cap;
}

/* Potential::makePair of the registered InterInter potentials
 * (GasIon.cpp:41-60, VaporGas.cpp:47-64, GasGas.cpp:44-63; VaporVapor's
 * all-pairs list is implicit in the in-domain vapor set) */
static int make_lists(orc_ctx *c, orc_state *st) {
  const orc_system *s = &c->s;
  int ok = 1;
  double *mol = st->mol_x;
  st->n_gi = st->n_gv = st->n_gg = 0;
  int g0 = 1, g1 = 1 + s->n_gas, v1 = c->n_mol;
  if (s->flags & ORC_GI) {
    for (int i = g0; i < g1; i++) {
      if (!st->region[i]) continue;
      for (int j = 0; j < s->natoms[0]; j++) {
        const double *xj = st->atom_x + 3 * j;
        double dx = mol[3*i] - xj[0], dy = mol[3*i+1] - xj[1], dz = mol[3*i+2] - xj[2];
        double r2 = (dx * dx + dy * dy + dz * dz);
        if (r2 < s->ML2) ok &= push_pair(st->gi, &st->n_gi, st->cap_gi, i, j);
      }
    }
  }
  if (s->flags & ORC_GG) {
    /* intended gas-gas list (the reference indexes loop counters, see
     * GasGas.cpp:50; this restates the evident intent) */
    for (int i = g0; i < g1; i++) {
      if (!st->region[i]) continue;
      for (int j = i + 1; j < g1; j++) {
        if (!st->region[j]) continue;
        double dx = mol[3*i] - mol[3*j], dy = mol[3*i+1] - mol[3*j+1], dz = mol[3*i+2] - mol[3*j+2];
        double r2 = (dx * dx + dy * dy + dz * dz);
        if (r2 < s->ML2) ok &= push_pair(st->gg, &st->n_gg, st->cap_gg, i, j);
      }
    }
  }
  if (s->flags & ORC_GV) {
    for (int i = g0; i < g1; i++) {
      if (!st->region[i]) continue;
      for (int j = g1; j < v1; j++) {
        if (!st->region[j]) continue;
        double dx = mol[3*i] - mol[3*j], dy = mol[3*i+1] - mol[3*j+1], dz = mol[3*i+2] - mol[3*j+2];
        double r2 = (dx * dx + dy * dy + dz * dz);
        if (r2 < s->ML2) ok &= push_pair(st->gv, &st->n_gv, st->cap_gv, i, j);
      }
    }
  }
  return ok ? 0 : -2;
}

/* MD::make_pair, pairlist.cpp:16-49 */
int orc_make_pair(orc_ctx *c, orc_state *st) {
  const orc_system *s = &c->s;
  double *mol = st->mol_x, *molv = st->mol_v;
  for (int i = 1; i < c->n_mol; i++) {
    if (st->region[i]) continue;
    mol[3*i] += molv[3*i] * s->dt * st->loopPair;
    mol[3*i+1] += molv[3*i+1] * s->dt * st->loopPair;
    mol[3*i+2] += molv[3*i+2] * s->dt * st->loopPair;
  }
  update_in_centers(c, st);
  classify(c, st, 1, s->n_gas, s->m_gas_kg);
  classify(c, st, 1 + s->n_gas, s->n_vap, s->m_vap_kg);
  int rc = make_lists(c, st);
  double vmax2 = 0.0;
  for (int i = 0; i < c->n_mol; i++) {
    const double *v = molv + 3 * i;
    double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (vmax2 < v2) vmax2 = v2;
  }
  double vmax = sqrt(vmax2);
  double lu = s->margin / (vmax * s->dt);
  st->loop_update = (lu >= (double)INT_MAX) ? INT_MAX : (int)lu;
  st->loopPair = 0;
  st->rebuild_count++;
  if (rc) st->status = rc;
  return rc;
}

/* MD::check_pairlist, pairlist.cpp:58-63 */
int orc_check_pairlist(orc_ctx *c, orc_state *st) {
  st->loopPair++;
  if (st->loopPair > st->loop_update) return orc_make_pair(c, st) == 0 ? 1 : -1;
  return 0;
}

/* ---------------- force terms -------------------------------------------- */
#define PC(t1, t2, k) (s->pair_coeff[((t1) * s->n_types + (t2)) * 2 + (k)])

/* PotentialGasIon::compute, GasIon.cpp:4-38 */
static void f_gas_ion(orc_ctx *c, orc_state *st, int eflag) {
  const orc_system *s = &c->s;
  double *x = st->atom_x, *f = st->atom_f;
  for (int p = 0; p < st->n_gi; p++) {
    int i = st->gi[2*p], j = st->gi[2*p+1];
    int o = mol_offset(c, i);
    for (int ii = 0; ii < s->natoms[1]; ii++) {
      double *xi = x + 3 * (o + ii), *fi = f + 3 * (o + ii);
      double dx = xi[0] - x[3*j], dy = xi[1] - x[3*j+1], dz = xi[2] - x[3*j+2];
      double rsq = (dx * dx + dy * dy + dz * dz);
      double r2inv = 1 / rsq;
      int t1 = s->type[1][ii], t2 = s->type[0][j];
      double r6inv = r2inv * r2inv * r2inv;
      double force_pair = r6inv * (PC(t1, t2, 0) * r6inv - PC(t1, t2, 1)) * r2inv;
      fi[0] += force_pair * dx; fi[1] += force_pair * dy; fi[2] += force_pair * dz;
      f[3*j] -= force_pair * dx; f[3*j+1] -= force_pair * dy; f[3*j+2] -= force_pair * dz;
      if (eflag) st->U[U_GI] += r6inv * (PC(t1, t2, 0) / 12.0 * r6inv - PC(t1, t2, 1) / 6.0);
    }
  }
}

/* weighted LJ + Coulomb between atoms a (species spa, local ia) and b */
static inline void ljcoul_w(orc_ctx *c, orc_state *st, int a, int spa, int ia, int b, int spb, int ib,
                            int eflag, int bucket) {
  const orc_system *s = &c->s;
  double *x = st->atom_x, *f = st->atom_f;
  double dx = x[3*a] - x[3*b], dy = x[3*a+1] - x[3*b+1], dz = x[3*a+2] - x[3*b+2];
  double rsq = (dx * dx + dy * dy + dz * dz);
  double r2inv = 1 / rsq;
  int t1 = s->type[spa][ia], t2 = s->type[spb][ib];
  double r6inv = r2inv * r2inv * r2inv;
  double force_lj = r6inv * (PC(t1, t2, 0) * r6inv - PC(t1, t2, 1));
  double force_coul = c->qqrd2e * s->charge[spa][ia] * s->charge[spb][ib] * sqrt(r2inv);
  double w = st->atom_w[a] * st->atom_w[b];
  double force_pair = (force_lj + force_coul) * r2inv * w;
  f[3*a] += force_pair * dx; f[3*a+1] += force_pair * dy; f[3*a+2] += force_pair * dz;
  f[3*b] -= force_pair * dx; f[3*b+1] -= force_pair * dy; f[3*b+2] -= force_pair * dz;
  if (eflag) {
    st->U[bucket] += r6inv * (PC(t1, t2, 0) / 12.0 * r6inv - PC(t1, t2, 1) / 6.0);
    st->U[bucket] += force_coul;
  }
}

/* PotentialVaporIon::compute, VaporIon.cpp:5-43 */
static void f_vapor_ion(orc_ctx *c, orc_state *st, int eflag) {
  const orc_system *s = &c->s;
  for (int i = 1 + s->n_gas; i < c->n_mol; i++) {
    if (!st->region[i]) continue;
    int o = mol_offset(c, i);
    for (int ii = 0; ii < s->natoms[2]; ii++)
      for (int jj = 0; jj < s->natoms[0]; jj++)
        ljcoul_w(c, st, o + ii, 2, ii, jj, 0, jj, eflag, U_VI);
  }
}

/* PotentialVaporVapor::compute, VaporVapor.cpp:5-45 over the list of
 * VaporVapor::makePair (VaporVapor.cpp:48-61) */
static void f_vapor_vapor(orc_ctx *c, orc_state *st, int eflag) {
  const orc_system *s = &c->s;
  int v0 = 1 + s->n_gas, na = s->natoms[2];
  for (int i = v0; i < c->n_mol; i++) {
    if (!st->region[i]) continue;
    int oi = mol_offset(c, i);
    for (int j = i + 1; j < c->n_mol; j++) {
      if (!st->region[j]) continue;
      int oj = mol_offset(c, j);
      for (int ii = 0; ii < na; ii++)
        for (int jj = 0; jj < na; jj++)
          ljcoul_w(c, st, oi + ii, 2, ii, oj + jj, 2, jj, eflag, U_VV);
    }
  }
}

/* unweighted LJ between atoms (VaporGas.cpp:19-35 / GasGas form) */
static inline void lj_pair(orc_ctx *c, orc_state *st, int a, int spa, int ia, int b, int spb, int ib,
                           double w, int eflag, int bucket) {
  const orc_system *s = &c->s;
  double *x = st->atom_x, *f = st->atom_f;
  double dx = x[3*a] - x[3*b], dy = x[3*a+1] - x[3*b+1], dz = x[3*a+2] - x[3*b+2];
  double rsq = (dx * dx + dy * dy + dz * dz);
  double r2inv = 1 / rsq;
  int t1 = s->type[spa][ia], t2 = s->type[spb][ib];
  double r6inv = r2inv * r2inv * r2inv;
  double force_lj = r6inv * (PC(t1, t2, 0) * r6inv - PC(t1, t2, 1));
  double force_pair = (force_lj) * r2inv;
  if (w != 1.0) force_pair = force_lj * r2inv * w;
  f[3*a] += force_pair * dx; f[3*a+1] += force_pair * dy; f[3*a+2] += force_pair * dz;
  f[3*b] -= force_pair * dx; f[3*b+1] -= force_pair * dy; f[3*b+2] -= force_pair * dz;
  if (eflag) st->U[bucket] += r6inv * (PC(t1, t2, 0) / 12.0 * r6inv - PC(t1, t2, 1) / 6.0);
}

/* PotentialVaporGas::compute, VaporGas.cpp:5-44 (inner bound Nj = gas atom
 * count, VaporGas.cpp:15, clamped to the vapor atom count) */
static void f_vapor_gas(orc_ctx *c, orc_state *st, int eflag) {
  const orc_system *s = &c->s;
  int ni = s->natoms[1];
  int nj = s->natoms[1] < s->natoms[2] ? s->natoms[1] : s->natoms[2];
  for (int p = 0; p < st->n_gv; p++) {
    int i = st->gv[2*p], j = st->gv[2*p+1];
    int oi = mol_offset(c, i), oj = mol_offset(c, j);
    for (int ii = 0; ii < ni; ii++)
      for (int jj = 0; jj < nj; jj++)
        lj_pair(c, st, oi + ii, 1, ii, oj + jj, 2, jj, 1.0, eflag, U_VG);
  }
}

/* gas-gas LJ, intended semantics of GasGas.cpp:4-41 (weighted w_i w_j,
 * energy to Ugg) */
static void f_gas_gas(orc_ctx *c, orc_state *st, int eflag) {
  const orc_system *s = &c->s;
  int na = s->natoms[1];
  for (int p = 0; p < st->n_gg; p++) {
    int i = st->gg[2*p], j = st->gg[2*p+1];
    int oi = mol_offset(c, i), oj = mol_offset(c, j);
    for (int ii = 0; ii < na; ii++)
      for (int jj = 0; jj < na; jj++) {
        double w = st->atom_w[oi + ii] * st->atom_w[oj + jj];
        double *x = st->atom_x, *f = st->atom_f;
        int a = oi + ii, b = oj + jj;
        double dx = x[3*a] - x[3*b], dy = x[3*a+1] - x[3*b+1], dz = x[3*a+2] - x[3*b+2];
        double rsq = (dx * dx + dy * dy + dz * dz);
        double r2inv = 1 / rsq;
        int t1 = s->type[1][ii], t2 = s->type[1][jj];
        double r6inv = r2inv * r2inv * r2inv;
        double force_lj = r6inv * (PC(t1, t2, 0) * r6inv - PC(t1, t2, 1));
        double force_pair = (force_lj) * r2inv * w;
        f[3*a] += force_pair * dx; f[3*a+1] += force_pair * dy; f[3*a+2] += force_pair * dz;
        f[3*b] -= force_pair * dx; f[3*b+1] -= force_pair * dy; f[3*b+2] -= force_pair * dz;
        if (eflag) st->U[U_GG] += r6inv * (PC(t1, t2, 0) / 12.0 * r6inv - PC(t1, t2, 1) / 6.0);
      }
  }
}

/* PotentialEfield::compute, electricField.cpp:54-64 */
static void f_efield(orc_ctx *c, orc_state *st) {
  const orc_system *s = &c->s;
  for (int i = 0; i < s->natoms[0]; i++) {
    double coeff = 5.55e-5 * s->charge[0][i];
    st->atom_f[3*i] += coeff * s->efield[0];
    st->atom_f[3*i+1] += coeff * s->efield[1];
    st->atom_f[3*i+2] += coeff * s->efield[2];
  }
}

/* harmonic bond, compute_AMBER.cpp:52-77 / compute_intra.cpp:14-41 */
static void bond_term(const orc_system *s, double *x, double *f, int i, int j, int type, int eflag, double *U) {
  double dx = x[3*i] - x[3*j], dy = x[3*i+1] - x[3*j+1], dz = x[3*i+2] - x[3*j+2];
  double rsq = (dx * dx + dy * dy + dz * dz);
  double r = sqrt(rsq);
  double dr = (r - s->btype[2 * type + 1]);
  double rk = s->btype[2 * type] * dr;
  double fb = -2.0 * rk / r;
  f[3*i] += fb * dx; f[3*i+1] += fb * dy; f[3*i+2] += fb * dz;
  f[3*j] -= fb * dx; f[3*j+1] -= fb * dy; f[3*j+2] -= fb * dz;
  if (eflag) *U += rk * dr;
}

/* harmonic angle, compute_AMBER.cpp:80-127 */
static void angle_term(const orc_system *s, double *x, double *f, int i, int j, int k, int type, int eflag, double *U) {
  double f1[3], f3[3];
  double dx1 = x[3*i] - x[3*j], dy1 = x[3*i+1] - x[3*j+1], dz1 = x[3*i+2] - x[3*j+2];
  double rsq1 = (dx1 * dx1 + dy1 * dy1 + dz1 * dz1);
  double r1 = sqrt(rsq1);
  double dx2 = x[3*k] - x[3*j], dy2 = x[3*k+1] - x[3*j+1], dz2 = x[3*k+2] - x[3*j+2];
  double rsq2 = (dx2 * dx2 + dy2 * dy2 + dz2 * dz2);
  double r2 = sqrt(rsq2);
  double C = dx1 * dx2 + dy1 * dy2 + dz1 * dz2;
  C /= r1 * r2;
  double Cs = 1 / (sqrt(1.0 - C * C));
  double dtheta = acos(C) - s->ctype[2 * type + 1];
  double tk = s->ctype[2 * type] * dtheta;
  double a = -2.0 * tk * Cs;
  double a11 = a * C / rsq1;
  double a12 = -a / (r1 * r2);
  double a22 = a * C / rsq2;
  f1[0] = a11 * dx1 + a12 * dx2; f1[1] = a11 * dy1 + a12 * dy2; f1[2] = a11 * dz1 + a12 * dz2;
  f3[0] = a22 * dx2 + a12 * dx1; f3[1] = a22 * dy2 + a12 * dy1; f3[2] = a22 * dz2 + a12 * dz1;
  f[3*i] += f1[0]; f[3*i+1] += f1[1]; f[3*i+2] += f1[2];
  f[3*j] -= f1[0] + f3[0]; f[3*j+1] -= f1[1] + f3[1]; f[3*j+2] -= f1[2] + f3[2];
  f[3*k] += f3[0]; f[3*k+1] += f3[1]; f[3*k+2] += f3[2];
  if (eflag) *U += tk * dtheta;
}

/* multi-term dihedral, compute_AMBER.cpp:130-253 */
static void dihedral_term(const orc_system *s, double *x, double *f, int i, int j, int k, int l, int type,
                          int eflag, double *U) {
  double ff1[3], ff2[3], ff3[3], ff4[3];
  double vb1x = x[3*i] - x[3*j], vb1y = x[3*i+1] - x[3*j+1], vb1z = x[3*i+2] - x[3*j+2];
  double vb2x = x[3*k] - x[3*j], vb2y = x[3*k+1] - x[3*j+1], vb2z = x[3*k+2] - x[3*j+2];
  double vb2xm = -vb2x, vb2ym = -vb2y, vb2zm = -vb2z;
  double vb3x = x[3*l] - x[3*k], vb3y = x[3*l+1] - x[3*k+1], vb3z = x[3*l+2] - x[3*k+2];
  double ax = vb1y * vb2zm - vb1z * vb2ym;
  double ay = vb1z * vb2xm - vb1x * vb2zm;
  double az = vb1x * vb2ym - vb1y * vb2xm;
  double bx = vb3y * vb2zm - vb3z * vb2ym;
  double by = vb3z * vb2xm - vb3x * vb2zm;
  double bz = vb3x * vb2ym - vb3y * vb2xm;
  double rasq = ax * ax + ay * ay + az * az;
  double rbsq = bx * bx + by * by + bz * bz;
  double rgsq = vb2xm * vb2xm + vb2ym * vb2ym + vb2zm * vb2zm;
  double rg = sqrt(rgsq);
  double rginv = 0.0, ra2inv = 0.0, rb2inv = 0.0;
  if (rg > 0) rginv = 1.0 / rg;
  if (rasq > 0) ra2inv = 1.0 / rasq;
  if (rbsq > 0) rb2inv = 1.0 / rbsq;
  double rabinv = sqrt(ra2inv * rb2inv);
  double c = (ax * bx + ay * by + az * bz) * rabinv;
  double sn = rg * rabinv * (ax * vb3x + ay * vb3y + az * vb3z);
  double df = 0.0;
  const double *co = s->dcoeff + 15 * type;
  for (int JJ = 0; JJ < s->dmulti[type]; JJ++) {
    int J5 = JJ * 5;
    double p_ = 1.0, df1 = 0.0, ddf1 = 0.0;
    for (int loop = 0; loop < co[J5 + 1]; loop++) {
      ddf1 = p_ * c - df1 * sn;
      df1 = p_ * sn + df1 * c;
      p_ = ddf1;
    }
    p_ = p_ * co[J5 + 3] + df1 * co[J5 + 4];
    df1 = df1 * co[J5 + 3] - ddf1 * co[J5 + 4];
    df1 *= -co[J5 + 1];
    p_ += 1.0;
    if (co[1] == 0) { /* checks the first term's n for every term (:199) */
      p_ = 1.0 + co[J5 + 3];
      df1 = 0.0;
    }
    df += (-co[J5] * df1);
    if (eflag) *U += co[J5] * p_;
  }
  double fg = vb1x * vb2xm + vb1y * vb2ym + vb1z * vb2zm;
  double hg = vb3x * vb2xm + vb3y * vb2ym + vb3z * vb2zm;
  double fga = fg * ra2inv * rginv;
  double hgb = hg * rb2inv * rginv;
  double gaa = -ra2inv * rg;
  double gbb = rb2inv * rg;
  double dtfx = gaa * ax, dtfy = gaa * ay, dtfz = gaa * az;
  double dtgx = fga * ax - hgb * bx, dtgy = fga * ay - hgb * by, dtgz = fga * az - hgb * bz;
  double dthx = gbb * bx, dthy = gbb * by, dthz = gbb * bz;
  double sx2 = df * dtgx, sy2 = df * dtgy, sz2 = df * dtgz;
  ff1[0] = df * dtfx; ff1[1] = df * dtfy; ff1[2] = df * dtfz;
  ff2[0] = sx2 - ff1[0]; ff2[1] = sy2 - ff1[1]; ff2[2] = sz2 - ff1[2];
  ff4[0] = df * dthx; ff4[1] = df * dthy; ff4[2] = df * dthz;
  ff3[0] = -sx2 - ff4[0]; ff3[1] = -sy2 - ff4[1]; ff3[2] = -sz2 - ff4[2];
  f[3*i] += ff1[0]; f[3*i+1] += ff1[1]; f[3*i+2] += ff1[2];
  f[3*j] += ff2[0]; f[3*j+1] += ff2[1]; f[3*j+2] += ff2[2];
  f[3*k] += ff3[0]; f[3*k+1] += ff3[1]; f[3*k+2] += ff3[2];
  f[3*l] += ff4[0]; f[3*l+1] += ff4[1]; f[3*l+2] += ff4[2];
}

/* PotentialAMBER::compute, compute_AMBER.cpp:10-17 */
static void f_amber(orc_ctx *c, orc_state *st, int eflag) {
  const orc_system *s = &c->s;
  double *x = st->atom_x, *f = st->atom_f;
  /* computeLong, :21-49 */
  for (int p = 0; p < c->n_long; p++) {
    int i = c->long_pairs[2*p], j = c->long_pairs[2*p+1];
    double dx = x[3*i] - x[3*j], dy = x[3*i+1] - x[3*j+1], dz = x[3*i+2] - x[3*j+2];
    double rsq = (dx * dx + dy * dy + dz * dz);
    double r2inv = 1 / rsq;
    int t1 = s->type[0][i], t2 = s->type[0][j];
    double r6inv = r2inv * r2inv * r2inv;
    double force_lj = r6inv * (PC(t1, t2, 0) * r6inv - PC(t1, t2, 1));
    double force_coul = c->qqrd2e * s->charge[0][i] * s->charge[0][j] * sqrt(r2inv);
    double force_pair = (force_lj + force_coul) * r2inv;
    f[3*i] += force_pair * dx; f[3*i+1] += force_pair * dy; f[3*i+2] += force_pair * dz;
    f[3*j] -= force_pair * dx; f[3*j+1] -= force_pair * dy; f[3*j+2] -= force_pair * dz;
    if (eflag) {
      st->U[U_ION] += r6inv * (PC(t1, t2, 0) / 12.0 * r6inv - PC(t1, t2, 1) / 6.0);
      st->U[U_ION] += force_coul;
    }
  }
  for (int b = 0; b < s->nbond[0]; b++)
    bond_term(s, x, f, s->bond[0][3*b], s->bond[0][3*b+1], s->bond[0][3*b+2], eflag, &st->U[U_ION]);
  for (int a = 0; a < s->nangle[0]; a++)
    angle_term(s, x, f, s->angle[0][4*a], s->angle[0][4*a+1], s->angle[0][4*a+2], s->angle[0][4*a+3], eflag,
               &st->U[U_ION]);
  for (int d = 0; d < s->ndih[0]; d++)
    dihedral_term(s, x, f, s->dih[0][5*d], s->dih[0][5*d+1], s->dih[0][5*d+2], s->dih[0][5*d+3],
                  s->dih[0][5*d+4], eflag, &st->U[U_ION]);
}

/* PotentialVaporIntra::compute, compute_intra.cpp:45-245 (bonds and angles
 * to Ugas, dihedrals to Uion, as labelled there) */
static void f_vapor_intra(orc_ctx *c, orc_state *st, int eflag) {
  const orc_system *s = &c->s;
  for (int I = 1 + s->n_gas; I < c->n_mol; I++) {
    if (!st->region[I]) continue;
    int o = mol_offset(c, I);
    double *x = st->atom_x + 3 * o, *f = st->atom_f + 3 * o;
    for (int b = 0; b < s->nbond[2]; b++)
      bond_term(s, x, f, s->bond[2][3*b], s->bond[2][3*b+1], s->bond[2][3*b+2], eflag, &st->U[U_GAS]);
    for (int a = 0; a < s->nangle[2]; a++)
      angle_term(s, x, f, s->angle[2][4*a], s->angle[2][4*a+1], s->angle[2][4*a+2], s->angle[2][4*a+3], eflag,
                 &st->U[U_GAS]);
    for (int d = 0; d < s->ndih[2]; d++)
      dihedral_term(s, x, f, s->dih[2][5*d], s->dih[2][5*d+1], s->dih[2][5*d+2], s->dih[2][5*d+3],
                    s->dih[2][5*d+4], eflag, &st->U[U_ION]);
  }
}

/* PotentialGasIntra::compute, compute_intra.cpp:14-41 */
static void f_gas_intra(orc_ctx *c, orc_state *st, int eflag) {
  const orc_system *s = &c->s;
  for (int I = 1; I <= s->n_gas; I++) {
    if (!st->region[I]) continue;
    int o = mol_offset(c, I);
    double *x = st->atom_x + 3 * o, *f = st->atom_f + 3 * o;
    for (int b = 0; b < s->nbond[1]; b++)
      bond_term(s, x, f, s->bond[1][3*b], s->bond[1][3*b+1], s->bond[1][3*b+2], eflag, &st->U[U_GAS]);
  }
}

/* PotentialBorn::compute, BMHFT/compute_born.cpp:10-56, coefficients
 * potentialBMHFT.hpp:13-41 indexed by ion-local type (Na=0, Cl=1) */
static void born_coeff(double bc[2][2][5]) {
  bc[0][0][0] = 25.4435 * 3.1546 / 4.184;
  bc[0][1][0] = bc[1][0][0] = 20.3548 * 3.1546 / 4.184;
  bc[1][1][0] = 15.2661 * 3.1546 / 4.184;
  bc[0][0][1] = 6 * 101.1719 / 4.184;
  bc[0][1][1] = bc[1][0][1] = 6 * 674.4793 / 4.184;
  bc[1][1][1] = 6 * 6985.6786 / 4.184;
  bc[0][0][2] = 8 * 48.1771 / 4.184;
  bc[0][1][2] = bc[1][0][2] = 8 * 837.077 / 4.184;
  bc[1][1][2] = 8 * 14031.5785 / 4.184;
  bc[0][0][3] = 2.340;
  bc[0][1][3] = bc[1][0][3] = 2.755;
  bc[1][1][3] = 3.170;
  bc[0][0][4] = 3.1546;
  bc[0][1][4] = bc[1][0][4] = 3.1546;
  bc[1][1][4] = 3.1546;
}

static void f_born(orc_ctx *c, orc_state *st, int eflag) {
  const orc_system *s = &c->s;
  double bc[2][2][5];
  born_coeff(bc);
  int is = s->natoms[0];
  double *x = st->atom_x, *f = st->atom_f;
  for (int i = 0; i < is - 1; i++)
    for (int j = i + 1; j < is; j++) {
      double dx = x[3*i] - x[3*j], dy = x[3*i+1] - x[3*j+1], dz = x[3*i+2] - x[3*j+2];
      double rsq = (dx * dx + dy * dy + dz * dz);
      double r2inv = 1 / rsq;
      double r6inv = r2inv * r2inv * r2inv;
      double r = sqrt(rsq);
      int t1 = s->type[0][i] - s->ion_type_base, t2 = s->type[0][j] - s->ion_type_base;
      t1 = t1 < 0 ? 0 : (t1 > 1 ? 1 : t1);
      t2 = t2 < 0 ? 0 : (t2 > 1 ? 1 : t2);
      const double *B = bc[t1][t2];
      double rexp = exp(B[4] * (B[3] - r));
      double force1 = B[4] * B[0] * r * rexp;
      double force2 = -B[1] * r6inv;
      double force3 = -B[2] * r6inv * r2inv;
      double force_coul = c->qqrd2e * s->charge[0][i] * s->charge[0][j] * sqrt(r2inv);
      double force_pair = (force1 + force2 + force3 + force_coul) * r2inv;
      f[3*i] += force_pair * dx; f[3*i+1] += force_pair * dy; f[3*i+2] += force_pair * dz;
      f[3*j] -= force_pair * dx; f[3*j+1] -= force_pair * dy; f[3*j+2] -= force_pair * dz;
      if (eflag) {
        st->U[U_ION] += force_coul;
        st->U[U_ION] += rexp * B[0];
        st->U[U_ION] -= B[1] / 6.0 * r6inv;
        st->U[U_ION] -= B[2] / 8.0 * r6inv * r2inv;
      }
    }
}

/* ---- Tersoff (Tersoff/Tersoff.cpp:13-297, constants potentialTersoff.hpp) */
typedef struct { double gamma, lamda3, c2, d2, costheta0, n, beta, lamda2, B, R, D, cut, cut2, cutl, lamda1, A, ML2; } ters_p;
static ters_p ters_params(void) {
  ters_p p;
  p.gamma = 1.0; p.lamda3 = 0; double cc = 100390; p.c2 = cc * cc; double d = 16.218; p.d2 = d * d;
  p.costheta0 = -0.59826; p.n = 0.78734; p.beta = 1.0999e-6; p.lamda2 = 1.7322; p.B = 471.18;
  p.R = 2.85; p.D = 0.15; p.cut = p.R + p.D; p.cut2 = p.cut * p.cut; p.cutl = p.R - p.D;
  p.lamda1 = 2.4799; p.A = 1830.8; p.ML2 = (p.R + p.D + 5) * (p.R + p.D + 5);
  return p;
}
static double t_fA(const ters_p *p, double r) { return -p->B * exp(-p->lamda2 * r); }
static double t_fA_d(const ters_p *p, double r) { return p->B * p->lamda2 * exp(-p->lamda2 * r); }
static double t_fR(const ters_p *p, double r) { return p->A * exp(-p->lamda1 * r); }
static double t_fR_d(const ters_p *p, double r) { return -p->A * p->lamda1 * exp(-p->lamda1 * r); }
static double t_fC(const ters_p *p, double r) {
  double fc = 0.5 * (1 - sin(M_PI * 0.5 * (r - p->R) / p->D));
  if (r < p->cutl) fc = 1;
  if (r > p->cut) fc = 0;
  return fc;
}
static double t_fC_d(const ters_p *p, double r) {
  double fc = -0.25 * M_PI / p->D * cos(M_PI * 0.5 * (r - p->R) / p->D);
  if (r < p->cutl || r > p->cut) fc = 0;
  return fc;
}
static double t_gijk(const ters_p *p, double ct) {
  double h = p->costheta0 - ct;
  return p->gamma * (1.0 + p->c2 / p->d2 - p->c2 / (p->d2 + h * h));
}
static double t_gijk_d(const ters_p *p, double ct) {
  double h = p->costheta0 - ct;
  double num = -2.0 * p->c2 * h;
  double den = 1 / (p->d2 + h * h);
  return p->gamma * num * den * den;
}
static double t_bij(const ters_p *p, double zeta) {
  double tmp = p->beta * zeta;
  if (zeta == 0) return 1.0;
  return pow(1.0 + pow(tmp, p->n), -1.0 / (2.0 * p->n));
}
static double t_bij_d(const ters_p *p, double zeta) {
  double tmp = p->beta * zeta;
  double tmp_n = pow(tmp, p->n);
  return -0.5 * pow(1.0 + tmp_n, -1.0 - (1.0 / (2.0 * p->n))) * tmp_n / zeta;
}
static double t_zeta(const ters_p *p, double rij, double rik, double a, double b, double cz) {
  double costheta = (a + b + cz) / (rij * rik);
  double arg = exp(p->lamda3 * p->lamda3 * p->lamda3 * pow(rij - rik, 3));
  return arg * t_gijk(p, costheta) * t_fC(p, rik);
}
static void t_zeta_d(const ters_p *p, double prefactor, const double *rij_hat, double rij, const double *rik_hat,
                     double rik, double *dri, double *drj, double *drk) {
  double fc = t_fC(p, rik), dfc = t_fC_d(p, rik);
  double ct = rij_hat[0] * rik_hat[0] + rij_hat[1] * rik_hat[1] + rij_hat[2] * rik_hat[2];
  double g = t_gijk(p, ct), g_d = t_gijk_d(p, ct);
  double dcdri[3], dcdrj[3], dcdrk[3];
  for (int i = 0; i < 3; i++) {
    dcdrj[i] = (rik_hat[i] - ct * rij_hat[i]) / rij;
    dcdrk[i] = (rij_hat[i] - ct * rik_hat[i]) / rik;
    dcdri[i] = -dcdrj[i] - dcdrk[i];
  }
  for (int i = 0; i < 3; i++) { dri[i] = -dfc * g * rik_hat[i]; dri[i] += fc * g_d * dcdri[i]; dri[i] *= prefactor; }
  for (int i = 0; i < 3; i++) { drj[i] = fc * g_d * dcdrj[i]; drj[i] *= prefactor; }
  for (int i = 0; i < 3; i++) { drk[i] = dfc * g * rik_hat[i]; drk[i] += fc * g_d * dcdrk[i]; drk[i] *= prefactor; }
}

/* Tersoff with its neighbour list built (the reference never calls
 * make_pair, pairlist.cpp:61-62); its forces and energy converted eV->kcal/mol
 * by 23.06 (Tersoff.cpp:104-106) applied to its own contribution only. */
static void f_tersoff(orc_ctx *c, orc_state *st, int eflag) {
  const orc_system *s = &c->s;
  ters_p P = ters_params();
  const ters_p *p = &P;
  int is = s->natoms[0];
  double *x = st->atom_x;
  double *f = (double *)calloc(3 * (size_t)is, sizeof(double));
  int *nb = (int *)malloc(sizeof(int) * (is > 0 ? is : 1));
  int *ns = (int *)malloc(sizeof(int) * (is > 0 ? is : 1));
  double U = 0;
  for (int i = 0; i < is; i++) {
    int nn = 0, nsh = 0;
    for (int j = 0; j < is; j++) {
      if (i == j) continue;
      double dx = x[3*i] - x[3*j], dy = x[3*i+1] - x[3*j+1], dz = x[3*i+2] - x[3*j+2];
      if (dx * dx + dy * dy + dz * dz < p->ML2) nb[nn++] = j;
    }
    for (int a = 0; a < nn; a++) {
      int j = nb[a];
      double delr[3] = {x[3*i] - x[3*j], x[3*i+1] - x[3*j+1], x[3*i+2] - x[3*j+2]};
      double rsq = delr[0] * delr[0] + delr[1] * delr[1] + delr[2] * delr[2];
      if (rsq > p->cut2) continue;
      double r = sqrt(rsq);
      double fpair = -(t_fC(p, r) * t_fR_d(p, r) + t_fR(p, r) * t_fC_d(p, r)) / r * 0.5;
      for (int d = 0; d < 3; d++) { f[3*i+d] += fpair * delr[d]; f[3*j+d] -= fpair * delr[d]; }
      ns[nsh++] = j;
      U += t_fC(p, r) * t_fR(p, r) * 0.5;
    }
    for (int a = 0; a < nsh; a++) {
      int j = ns[a];
      double delr[3] = {x[3*i] - x[3*j], x[3*i+1] - x[3*j+1], x[3*i+2] - x[3*j+2]};
      double rsq = delr[0] * delr[0] + delr[1] * delr[1] + delr[2] * delr[2];
      double r = sqrt(rsq);
      double zeta_ij = 0;
      for (int b = 0; b < nsh; b++) {
        int k = ns[b];
        if (k == j) continue;
        double dx2 = x[3*i] - x[3*k], dy2 = x[3*i+1] - x[3*k+1], dz2 = x[3*i+2] - x[3*k+2];
        double rsq2 = dx2 * dx2 + dy2 * dy2 + dz2 * dz2;
        if (rsq2 > p->cut2) continue;
        double r2 = sqrt(rsq2);
        zeta_ij += t_zeta(p, r, r2, delr[0] * dx2, delr[1] * dy2, delr[2] * dz2);
      }
      double prefactor = t_fA(p, r) * t_fC(p, r) * t_bij_d(p, zeta_ij);
      if (zeta_ij == 0) prefactor = 0;
      double fz = t_bij(p, zeta_ij) * (t_fA(p, r) * t_fC_d(p, r) + t_fA_d(p, r) * t_fC(p, r)) / r;
      double fpair = -0.5 * fz;
      for (int d = 0; d < 3; d++) { f[3*i+d] += fpair * delr[d]; f[3*j+d] -= fpair * delr[d]; }
      U += t_bij(p, zeta_ij) * t_fC(p, r) * t_fA(p, r) * 0.5;
      if (zeta_ij == 0) continue;
      for (int b = 0; b < nsh; b++) {
        int k = ns[b];
        if (k == j) continue;
        double delr2[3] = {x[3*i] - x[3*k], x[3*i+1] - x[3*k+1], x[3*i+2] - x[3*k+2]};
        double rsq2 = delr2[0] * delr2[0] + delr2[1] * delr2[1] + delr2[2] * delr2[2];
        if (rsq2 > p->cut2) continue;
        double r2 = sqrt(rsq2);
        double rij_hat[3], rik_hat[3], fi[3], fj[3], fk[3];
        for (int d = 0; d < 3; d++) { rij_hat[d] = delr[d] / r; rik_hat[d] = delr2[d] / r2; }
        t_zeta_d(p, prefactor, rij_hat, r, rik_hat, r2, fi, fj, fk);
        for (int d = 0; d < 3; d++) { f[3*i+d] += fi[d] * 0.5; f[3*j+d] += fj[d] * 0.5; f[3*k+d] += fk[d] * 0.5; }
      }
    }
  }
  for (int i = 0; i < 3 * is; i++) st->atom_f[i] += f[i] * 23.06;
  if (eflag) st->U[U_ION] += U * 23.06;
  free(f); free(nb); free(ns);
}

/* ---- Stillinger-Weber (StillingerWeber/StillingerWeber.cpp:11-198,
 * constants potentialSW.hpp:13-37); same list/unit treatment as Tersoff */
typedef struct { double c1, c2, c3, c4, c5, c6, sigma, sigma_gamma, lambda_epsilon, lambda_epsilon2, costheta0, p, q, cut, cut2, ML2; } sw_p;
static sw_p sw_params(void) {
  sw_p s;
  double epsilon = 2.1683, sigma = 2.0951, a = 1.80, lambda = 21.0, gamma = 1.20;
  double A = 7.049556277, B = 0.6022245584, p = 4.0, q = 0.0;
  s.costheta0 = -0.333333333333; s.p = p; s.q = q; s.sigma = sigma;
  s.c1 = A * epsilon * p * B * pow(sigma, p);
  s.c2 = A * epsilon * q * pow(sigma, q);
  s.c3 = A * epsilon * B * pow(sigma, p + 1);
  s.c4 = A * epsilon * pow(sigma, q + 1);
  s.c5 = A * epsilon * B * pow(sigma, p);
  s.c6 = A * epsilon * pow(sigma, q);
  s.sigma_gamma = sigma * gamma; s.lambda_epsilon = lambda * epsilon; s.lambda_epsilon2 = 2 * lambda * epsilon;
  s.cut = a * sigma; s.cut2 = s.cut * s.cut; s.ML2 = (s.cut + 5) * (s.cut + 5);
  return s;
}
static double sw_twobody(const sw_p *p, double rsq, double *U) {
  double r = sqrt(rsq), rinvsq = 1.0 / rsq;
  double rp = pow(r, -p->p), rq = pow(r, -p->q);
  double rainv = 1.0 / (r - p->cut);
  double rainvsq = rainv * rainv * r;
  double expsrainv = exp(p->sigma * rainv);
  *U += (p->c5 * rp - p->c6 * rq) * expsrainv;
  return (p->c1 * rp - p->c2 * rq + (p->c3 * rp - p->c4 * rq) * rainvsq) * expsrainv * rinvsq;
}
static void sw_threebody(const sw_p *p, double rsq1, double rsq2, const double *rij, const double *rik, double *fj,
                         double *fk, double *U) {
  double r1 = sqrt(rsq1), rinvsq1 = 1.0 / rsq1, rainv1 = 1.0 / (r1 - p->cut);
  double gsrainv1 = p->sigma_gamma * rainv1, gsrainvsq1 = gsrainv1 * rainv1 / r1, expgsrainv1 = exp(gsrainv1);
  double r2 = sqrt(rsq2), rinvsq2 = 1.0 / rsq2, rainv2 = 1.0 / (r2 - p->cut);
  double gsrainv2 = p->sigma_gamma * rainv2, gsrainvsq2 = gsrainv2 * rainv2 / r2, expgsrainv2 = exp(gsrainv2);
  double rinv12 = 1.0 / (r1 * r2);
  double cs = (rij[0] * rik[0] + rij[1] * rik[1] + rij[2] * rik[2]) * rinv12;
  double delcs = cs - p->costheta0, delcssq = delcs * delcs;
  double facexp = expgsrainv1 * expgsrainv2;
  double facrad = p->lambda_epsilon * facexp * delcssq;
  double frad1 = facrad * gsrainvsq1, frad2 = facrad * gsrainvsq2;
  double facang = p->lambda_epsilon2 * facexp * delcs;
  double facang12 = rinv12 * facang;
  double csfacang = cs * facang;
  double csfac1 = rinvsq1 * csfacang;
  for (int d = 0; d < 3; d++) fj[d] = rij[d] * (frad1 + csfac1) - rik[d] * facang12;
  double csfac2 = rinvsq2 * csfacang;
  for (int d = 0; d < 3; d++) fk[d] = rik[d] * (frad2 + csfac2) - rij[d] * facang12;
  *U += facrad;
}
static void f_sw(orc_ctx *c, orc_state *st, int eflag) {
  const orc_system *s = &c->s;
  sw_p P = sw_params();
  const sw_p *p = &P;
  int is = s->natoms[0];
  double *x = st->atom_x;
  double *f = (double *)calloc(3 * (size_t)is, sizeof(double));
  int *nb = (int *)malloc(sizeof(int) * (is > 0 ? is : 1));
  int *ns = (int *)malloc(sizeof(int) * (is > 0 ? is : 1));
  double U = 0;
  for (int i = 0; i < is; i++) {
    int nn = 0, nsh = 0;
    for (int j = 0; j < is; j++) {
      if (i == j) continue;
      double dx = x[3*i] - x[3*j], dy = x[3*i+1] - x[3*j+1], dz = x[3*i+2] - x[3*j+2];
      if (dx * dx + dy * dy + dz * dz < p->ML2) nb[nn++] = j;
    }
    for (int a = 0; a < nn; a++) {
      int j = nb[a];
      double delr[3] = {x[3*i] - x[3*j], x[3*i+1] - x[3*j+1], x[3*i+2] - x[3*j+2]};
      double rsq = delr[0] * delr[0] + delr[1] * delr[1] + delr[2] * delr[2];
      if (rsq > p->cut2) continue;
      double fpair = sw_twobody(p, rsq, &U) * 0.5;
      for (int d = 0; d < 3; d++) { f[3*i+d] += fpair * delr[d]; f[3*j+d] -= fpair * delr[d]; }
      ns[nsh++] = j;
    }
    for (int a = 0; a < nsh; a++) {
      int j = ns[a];
      double delr[3] = {x[3*i] - x[3*j], x[3*i+1] - x[3*j+1], x[3*i+2] - x[3*j+2]};
      double rsq = delr[0] * delr[0] + delr[1] * delr[1] + delr[2] * delr[2];
      for (int b = 0; b < nsh; b++) {
        int k = ns[b];
        if (k == j) continue;
        double delr2[3] = {x[3*i] - x[3*k], x[3*i+1] - x[3*k+1], x[3*i+2] - x[3*k+2]};
        double rsq2 = delr2[0] * delr2[0] + delr2[1] * delr2[1] + delr2[2] * delr2[2];
        if (rsq2 > p->cut2) continue;
        double fj[3], fk[3];
        sw_threebody(p, rsq, rsq2, delr, delr2, fj, fk, &U);
        for (int d = 0; d < 3; d++) {
          f[3*i+d] += (fj[d] + fk[d]) * 0.5;
          f[3*j+d] -= fj[d] * 0.5;
          f[3*k+d] -= fk[d] * 0.5;
        }
      }
    }
  }
  for (int i = 0; i < 3 * is; i++) st->atom_f[i] += f[i] * 23.06;
  if (eflag) st->U[U_ION] += U * 0.5 * 23.06;
  free(f); free(nb); free(ns);
}

int orc_force_term(orc_ctx *c, orc_state *st, uint32_t bit, int eflag) {
  switch (bit) {
    case ORC_GI: f_gas_ion(c, st, eflag); break;
    case ORC_VI: f_vapor_ion(c, st, eflag); break;
    case ORC_GG: f_gas_gas(c, st, eflag); break;
    case ORC_GV: f_vapor_gas(c, st, eflag); break;
    case ORC_VV: f_vapor_vapor(c, st, eflag); break;
    case ORC_EFIELD: f_efield(c, st); break;
    case ORC_ION_TERSOFF: f_tersoff(c, st, eflag); break;
    case ORC_ION_SW: f_sw(c, st, eflag); break;
    case ORC_ION_AMBER: f_amber(c, st, eflag); break;
    case ORC_ION_BMH: f_born(c, st, eflag); break;
    case ORC_VAPOR_INTRA: f_vapor_intra(c, st, eflag); break;
    case ORC_GAS_INTRA: f_gas_intra(c, st, eflag); break;
    default: return -1;
  }
  return 0;
}

/* registration order of MD::setPotential, setPotential.cpp:13-108:
 * InterInter = gi, vi, gg, gv, vv, (ion dipole: no-op), efield;
 * IntraInter = tersoff, sw, amber, born, vapor intra, gas intra */
static const uint32_t ORDER[] = {ORC_GI, ORC_VI, ORC_GG, ORC_GV, ORC_VV, ORC_EFIELD, ORC_ION_TERSOFF,
                                 ORC_ION_SW, ORC_ION_AMBER, ORC_ION_BMH, ORC_VAPOR_INTRA, ORC_GAS_INTRA};

/* ---------------- ion thermostat (relaxation phase) --------------------- */
/* Tion of Observer::computeProps, observer.cpp:27-36 */
double orc_ion_temperature(const orc_ctx *c, const orc_state *st) {
  const orc_system *s = &c->s;
  int n = s->natoms[0];
  double K = 0;
  for (int i = 0; i < n; i++) {
    const double *v = st->atom_v + 3 * i;
    double m = s->mass[0][i];
    K += v[0] * v[0] * m;
    K += v[1] * v[1] * m;
    K += v[2] * v[2] * m;
  }
  K *= 0.5 * real_to_kcalmol;
  return K / (double)n * (kb_real_inv / 1.5);
}

static void scale_ion_velocities(const orc_ctx *c, orc_state *st, double r) {
  for (int i = 0; i < 3 * c->s.natoms[0]; i++) st->atom_v[i] *= r;
}

/* MD::velocity_scaling, nvt.cpp:5-16 (target 300 K) */
void orc_velocity_scaling(orc_ctx *c, orc_state *st) {
  double Tp = 300;
  scale_ion_velocities(c, st, sqrt(Tp / orc_ion_temperature(c, st)));
}

/* MD::nosehoover_zeta, nvt.cpp:18-24 */
void orc_nosehoover_zeta(orc_ctx *c, orc_state *st) {
  double Tion = orc_ion_temperature(c, st);
  int g = c->s.natoms[0] * 3;
  double Q_inv = 0.0001;
  st->zeta_ion += (Tion - c->s.Tnh_ion) * g * kb_real * Q_inv * c->s.dt;
}

/* MD::nosehoover_ion, nvt.cpp:27-35 */
void orc_nosehoover_ion(orc_ctx *c, orc_state *st) {
  scale_ion_velocities(c, st, exp(-st->zeta_ion * 0.5 * c->s.dt));
}

int orc_forces(orc_ctx *c, orc_state *st, int eflag) {
  for (unsigned k = 0; k < sizeof(ORDER) / sizeof(ORDER[0]); k++)
    if (c->s.flags & ORDER[k]) orc_force_term(c, st, ORDER[k], eflag);
  return 0;
}

/* MD::verlet, verlet.cpp:20-39 (+ energy cadence of run.cpp:68-71), with
 * the ion thermostat where the reference calls it (verlet.cpp:22, 27, 36) */
int orc_verlet(orc_ctx *c, orc_state *st, int nsteps, int eflag_every, int64_t step0, int64_t *atom_steps) {
  const orc_system *s = &c->s;
  int64_t acc = 0;
  int n_in = -1;
  for (int it = 0; it < nsteps; it++) {
    int eflag = eflag_every > 0 && ((step0 + it + 1) % eflag_every) == 0;
    if (eflag) for (int k = 0; k < U_N; k++) st->U[k] = 0;
    orc_velocity_calculation(c, st);
    if (s->nvt_ion == 1) orc_velocity_scaling(c, st);
    if (s->shake) orc_update_position_shake(c, st);
    else orc_update_position(c, st);
    if (s->nvt_ion == 2) orc_nosehoover_zeta(c, st);
    int rb = orc_check_pairlist(c, st);
    if (rb < 0) return -2;
    if (rb || n_in < 0) {
      n_in = s->natoms[0];
      for (int i = 1; i < c->n_mol; i++)
        if (st->region[i]) n_in += s->natoms[mol_species(c, i)];
    }
    orc_forces(c, st, eflag);
    orc_velocity_calculation(c, st);
    if (s->nvt_ion == 2) orc_nosehoover_ion(c, st);
    acc += n_in;
    if (st->status) return st->status;
  }
  if (atom_steps) *atom_steps = acc;
  return 0;
}
