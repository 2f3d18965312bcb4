// mdn_adjoint.hip -- the vi LJ + Coulomb term (PotentialVaporIon::compute,
// src/potential/VaporIon.cpp:5-43) as a differentiable function of the ion's
// per-atom parameters, over a batch of snapshots, in FP64.
//
//   pair (ion atom i, vapor atom j), d = x_j - x_i (vapor minus ion):
//     sigma_ij = (sigma_i + sigma_j)/2, eps_ij = sqrt(eps_i eps_j)
//       (setCrossPotential.cpp:13-20), c0 = 48 eps s^12, c1 = 24 eps s^6
//     f = [r^-6 (c0 r^-6 - c1) + K q_i q_j / r] r^-2 w_i w_j
//     F_j += f d, F_i -= f d;   E += r^-6 (c0/12 r^-6 - c1/6) + K q_i q_j / r
//
// Forward: E[s] and F[s][atom][3] for every snapshot s.  Adjoint (reverse
// mode, hand-written): given gE[s] = dL/dE[s] and gF[s][a][3] = dL/dF[s][a],
//     dL/dtheta_i = sum_s sum_j [ (gF_j - gF_i) . d  df/dtheta_i
//                                 + gE[s] dE_pair/dtheta_i ]
// for theta in {sigma, eps, q} of ion atom i.  Every output element is owned
// by one team of lanes (i-centric over partners, a fixed reduction tree, then
// a fixed-order sum over snapshots), so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "mdn.h"
#include "mdn_internal.h"

namespace mdn {

extern "C" double mdn_qqrd2e(void);

__device__ inline bool vap_in(const ViBatch &b, int s, int j) {
  return !b.vin || b.vin[(size_t)s * b.n_vap + j];
}

// pair quantities and their parameter derivatives (at the pair level)
struct PairD {
  double f, e;                 // force factor, energy
  double df_ds, df_de, df_dq;  // d f / d sigma_ij, d eps_ij, d q_i
  double de_ds, de_de, de_dq;  // d E_pair / ...
};

__device__ inline PairD vi_pair(double r2, double s, double e, double qi, double qj, double K, double w) {
  PairD p;
  double r2inv = 1.0 / r2, r6inv = r2inv * r2inv * r2inv, rinv = sqrt(r2inv);
  double s2 = s * s, s6 = s2 * s2 * s2, s12 = s6 * s6;
  double c0 = 48 * e * s12, c1 = 24 * e * s6;
  double fc = K * qi * qj * rinv;
  p.f = (r6inv * (c0 * r6inv - c1) + fc) * r2inv * w;
  p.e = r6inv * (c0 / 12.0 * r6inv - c1 / 6.0) + fc;
  // d/d sigma: c0' = 12 c0 / s, c1' = 6 c1 / s; d/d eps: c0 / e, c1 / e
  double dc0_ds = 12 * c0 / s, dc1_ds = 6 * c1 / s;
  p.df_ds = r6inv * (dc0_ds * r6inv - dc1_ds) * r2inv * w;
  p.de_ds = r6inv * (dc0_ds / 12.0 * r6inv - dc1_ds / 6.0);
  p.df_de = r6inv * (48 * s12 * r6inv - 24 * s6) * r2inv * w;
  p.de_de = r6inv * (4 * s12 * r6inv - 4 * s6);
  p.df_dq = K * qj * rinv * r2inv * w;
  p.de_dq = K * qj * rinv;
  return p;
}

// d f / d(r^2) of the force factor f = (c0 r^-14 - c1 r^-8 + K q_i q_j r^-3) w
__device__ inline double vi_df_dr2(double r2, double s, double e, double qi, double qj, double K, double w) {
  double r2inv = 1.0 / r2, r6inv = r2inv * r2inv * r2inv, rinv = sqrt(r2inv), r4inv = r2inv * r2inv;
  double s2 = s * s, s6 = s2 * s2 * s2, s12 = s6 * s6;
  double c0 = 48 * e * s12, c1 = 24 * e * s6;
  return (-7.0 * c0 * r6inv * r6inv * r4inv + 4.0 * c1 * r6inv * r4inv - 1.5 * K * qi * qj * rinv * r4inv) * w;
}

// Work layout of the kernels below: a team of T lanes (a power of two up to
// the wave's 64, aligned within the wave) owns one (snapshot, atom) and
// strides over that atom's partners; the team's partial sums are reduced by
// lane exchanges in a fixed tree, so results do not depend on scheduling.
// The host picks T from the partner count (about eight partners per lane),
// so a one-atom ion's vapor side runs one lane per atom and a 150-atom
// chain's ion side a whole wave per atom.
__device__ inline double team_sum(double v, int T) {
  for (int o = T >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// the team's (snapshot, atom) index and lane in workgroup bx of a side's
// grid; false for a team past n
__device__ inline bool team_of(int n, int T, int bx, int &k, int &lane) {
  const int t = bx * blockDim.x + threadIdx.x;
  k = t / T;
  lane = t - k * T;
  return k < n;
}

// derivative of the pair forces along the displacement u of the atoms (the
// Hessian of -E applied to u): with d = x_j - x_i and du = u_j - u_i,
//   dF_j = f du + 2 (d . du) (df/dr^2) d,   dF_i = -dF_j
// ion side: one team per (snapshot, ion atom)
__device__ void vi_jvp_ion(const ViBatch &b, const double *U, double *HU, int T, int bx) {
  int k, lane;
  const bool on = team_of(b.S * b.n_ion, T, bx, k, lane);
  const int s = on ? k / b.n_ion : 0, i = on ? k - s * b.n_ion : 0, n = b.n_ion + b.n_vap;
  const double *X = b.x + (size_t)s * n * 3, *W = b.w + (size_t)s * n, *Us = U + (size_t)s * n * 3;
  const double xi = X[3 * i], yi = X[3 * i + 1], zi = X[3 * i + 2];
  const double ux = Us[3 * i], uy = Us[3 * i + 1], uz = Us[3 * i + 2];
  double hx = 0, hy = 0, hz = 0;
  for (int j = lane; on && j < b.n_vap; j += T) {
    if (!vap_in(b, s, j)) continue;
    int a = b.n_ion + j;
    double dx = X[3 * a] - xi, dy = X[3 * a + 1] - yi, dz = X[3 * a + 2] - zi;
    double ex = Us[3 * a] - ux, ey = Us[3 * a + 1] - uy, ez = Us[3 * a + 2] - uz;
    double sg = (b.sig_i[i] + b.sig_v[j]) * 0.5, ep = sqrt(b.eps_i[i] * b.eps_v[j]);
    double r2 = dx * dx + dy * dy + dz * dz, w = W[i] * W[a];
    double f = vi_pair(r2, sg, ep, b.q_i[i], b.q_v[j], b.K, w).f;
    double g = 2.0 * (dx * ex + dy * ey + dz * ez) * vi_df_dr2(r2, sg, ep, b.q_i[i], b.q_v[j], b.K, w);
    hx -= f * ex + g * dx; hy -= f * ey + g * dy; hz -= f * ez + g * dz;
  }
  hx = team_sum(hx, T); hy = team_sum(hy, T); hz = team_sum(hz, T);
  if (on && lane == 0) {
    double *H = HU + ((size_t)s * n + i) * 3;
    H[0] = hx; H[1] = hy; H[2] = hz;
  }
}

// vapor side: one team per (snapshot, vapor atom)
__device__ void vi_jvp_vap(const ViBatch &b, const double *U, double *HU, int T, int bx) {
  int k, lane;
  const bool on = team_of(b.S * b.n_vap, T, bx, k, lane);
  const int s = on ? k / b.n_vap : 0, j = on ? k - s * b.n_vap : 0, n = b.n_ion + b.n_vap, a = b.n_ion + j;
  const double *X = b.x + (size_t)s * n * 3, *W = b.w + (size_t)s * n, *Us = U + (size_t)s * n * 3;
  const double xj = X[3 * a], yj = X[3 * a + 1], zj = X[3 * a + 2];
  const double ux = Us[3 * a], uy = Us[3 * a + 1], uz = Us[3 * a + 2];
  const bool in = on && vap_in(b, s, j);
  double hx = 0, hy = 0, hz = 0;
  for (int i = lane; in && i < b.n_ion; i += T) {
    double dx = xj - X[3 * i], dy = yj - X[3 * i + 1], dz = zj - X[3 * i + 2];
    double ex = ux - Us[3 * i], ey = uy - Us[3 * i + 1], ez = uz - Us[3 * i + 2];
    double sg = (b.sig_i[i] + b.sig_v[j]) * 0.5, ep = sqrt(b.eps_i[i] * b.eps_v[j]);
    double r2 = dx * dx + dy * dy + dz * dz, w = W[i] * W[a];
    double f = vi_pair(r2, sg, ep, b.q_i[i], b.q_v[j], b.K, w).f;
    double g = 2.0 * (dx * ex + dy * ey + dz * ez) * vi_df_dr2(r2, sg, ep, b.q_i[i], b.q_v[j], b.K, w);
    hx += f * ex + g * dx; hy += f * ey + g * dy; hz += f * ez + g * dz;
  }
  hx = team_sum(hx, T); hy = team_sum(hy, T); hz = team_sum(hz, T);
  if (on && lane == 0) {
    double *H = HU + ((size_t)s * n + a) * 3;
    H[0] = hx; H[1] = hy; H[2] = hz;
  }
}

// forward, ion side: one team per (snapshot, ion atom): F_i and the
// snapshot's energy partial of atom i
__device__ void vi_fwd_ion(const ViBatch &b, double *F, double *Epart, int T, int bx) {
  int k, lane;
  const bool on = team_of(b.S * b.n_ion, T, bx, k, lane);
  const int s = on ? k / b.n_ion : 0, i = on ? k - s * b.n_ion : 0, n = b.n_ion + b.n_vap;
  const double *X = b.x + (size_t)s * n * 3, *W = b.w + (size_t)s * n;
  const double xi = X[3 * i], yi = X[3 * i + 1], zi = X[3 * i + 2];
  double fx = 0, fy = 0, fz = 0, E = 0;
  for (int j = lane; on && j < b.n_vap; j += T) {
    if (!vap_in(b, s, j)) continue;
    int a = b.n_ion + j;
    double dx = X[3 * a] - xi, dy = X[3 * a + 1] - yi, dz = X[3 * a + 2] - zi;
    double sg = (b.sig_i[i] + b.sig_v[j]) * 0.5, ep = sqrt(b.eps_i[i] * b.eps_v[j]);
    PairD p = vi_pair(dx * dx + dy * dy + dz * dz, sg, ep, b.q_i[i], b.q_v[j], b.K, W[i] * W[a]);
    fx -= p.f * dx; fy -= p.f * dy; fz -= p.f * dz;
    E += p.e;
  }
  fx = team_sum(fx, T); fy = team_sum(fy, T); fz = team_sum(fz, T); E = team_sum(E, T);
  if (on && lane == 0) {
    double *Fi = F + ((size_t)s * n + i) * 3;
    Fi[0] = fx; Fi[1] = fy; Fi[2] = fz;
    Epart[k] = E;
  }
}

// forward, vapor side: one team per (snapshot, vapor atom): F_j
__device__ void vi_fwd_vap(const ViBatch &b, double *F, int T, int bx) {
  int k, lane;
  const bool on = team_of(b.S * b.n_vap, T, bx, k, lane);
  const int s = on ? k / b.n_vap : 0, j = on ? k - s * b.n_vap : 0, n = b.n_ion + b.n_vap, a = b.n_ion + j;
  const double *X = b.x + (size_t)s * n * 3, *W = b.w + (size_t)s * n;
  const double xj = X[3 * a], yj = X[3 * a + 1], zj = X[3 * a + 2];
  const bool in = on && vap_in(b, s, j);
  double fx = 0, fy = 0, fz = 0;
  for (int i = lane; in && i < b.n_ion; i += T) {
    double dx = xj - X[3 * i], dy = yj - X[3 * i + 1], dz = zj - X[3 * i + 2];
    double sg = (b.sig_i[i] + b.sig_v[j]) * 0.5, ep = sqrt(b.eps_i[i] * b.eps_v[j]);
    PairD p = vi_pair(dx * dx + dy * dy + dz * dz, sg, ep, b.q_i[i], b.q_v[j], b.K, W[i] * W[a]);
    fx += p.f * dx; fy += p.f * dy; fz += p.f * dz;
  }
  fx = team_sum(fx, T); fy = team_sum(fy, T); fz = team_sum(fz, T);
  if (on && lane == 0) {
    double *Fj = F + ((size_t)s * n + a) * 3;
    Fj[0] = fx; Fj[1] = fy; Fj[2] = fz;
  }
}

// The ion and vapor sides in one launch (they write the forces of disjoint
// atoms): workgroups [0, gi) take the ion side, the rest the vapor side, so
// that the two run side by side rather than one after the other.
__global__ void k_vi_fwd(ViBatch b, double *F, double *Epart, int Ti, int Tv, int gi) {
  if ((int)blockIdx.x < gi) vi_fwd_ion(b, F, Epart, Ti, blockIdx.x);
  else vi_fwd_vap(b, F, Tv, blockIdx.x - gi);
}
__global__ void k_vi_jvp(ViBatch b, const double *U, double *HU, int Ti, int Tv, int gi) {
  if ((int)blockIdx.x < gi) vi_jvp_ion(b, U, HU, Ti, blockIdx.x);
  else vi_jvp_vap(b, U, HU, Tv, blockIdx.x - gi);
}

// E[s] = sum over ion atoms of the partials (fixed order)
__global__ void k_vi_sum_e(const double *Epart, int S, int n_ion, double *E) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  double e = 0;
  for (int i = 0; i < n_ion; i++) e += Epart[(size_t)s * n_ion + i];
  E[s] = e;
}

// adjoint: one team per (snapshot, ion atom): the snapshot's contribution
// to dL/d(sigma_i, eps_i, q_i)
__global__ void k_vi_adj(ViBatch b, const double *gE, const double *gF, double *part, int T) {
  int k, lane;
  const bool on = team_of(b.S * b.n_ion, T, blockIdx.x, k, lane);
  const int s = on ? k / b.n_ion : 0, i = on ? k - s * b.n_ion : 0, n = b.n_ion + b.n_vap;
  const double *X = b.x + (size_t)s * n * 3, *W = b.w + (size_t)s * n;
  const double *G = gF ? gF + (size_t)s * n * 3 : nullptr;
  const double ge = gE ? gE[s] : 0.0;
  const double xi = X[3 * i], yi = X[3 * i + 1], zi = X[3 * i + 2];
  const double gix = G ? G[3 * i] : 0, giy = G ? G[3 * i + 1] : 0, giz = G ? G[3 * i + 2] : 0;
  double gs = 0, ge_ = 0, gq = 0;
  for (int j = lane; on && j < b.n_vap; j += T) {
    if (!vap_in(b, s, j)) continue;
    int a = b.n_ion + j;
    double dx = X[3 * a] - xi, dy = X[3 * a + 1] - yi, dz = X[3 * a + 2] - zi;
    double sg = (b.sig_i[i] + b.sig_v[j]) * 0.5, ep = sqrt(b.eps_i[i] * b.eps_v[j]);
    PairD p = vi_pair(dx * dx + dy * dy + dz * dz, sg, ep, b.q_i[i], b.q_v[j], b.K, W[i] * W[a]);
    // dL/df_pair = (gF_j - gF_i) . d
    double gf = G ? (G[3 * a] - gix) * dx + (G[3 * a + 1] - giy) * dy + (G[3 * a + 2] - giz) * dz : 0.0;
    double ds = 0.5;                                   // d sigma_ij / d sigma_i
    double de = ep > 0 ? 0.5 * b.eps_v[j] / ep : 0.0;  // d eps_ij / d eps_i
    gs += (gf * p.df_ds + ge * p.de_ds) * ds;
    ge_ += (gf * p.df_de + ge * p.de_de) * de;
    gq += gf * p.df_dq + ge * p.de_dq;
  }
  gs = team_sum(gs, T); ge_ = team_sum(ge_, T); gq = team_sum(gq, T);
  if (on && lane == 0) {
    double *o = part + (size_t)k * 3;
    o[0] = gs; o[1] = ge_; o[2] = gq;
  }
}

__global__ void k_vi_sum_g(const double *part, int S, int n_ion, double *g_sig, double *g_eps, double *g_q) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_ion) return;
  double a = 0, b = 0, c = 0;
  for (int s = 0; s < S; s++) {
    const double *p = part + ((size_t)s * n_ion + i) * 3;
    a += p[0]; b += p[1]; c += p[2];
  }
  g_sig[i] = a; g_eps[i] = b; g_q[i] = c;
}

// team size for atoms with `partners` partners each: about eight partners
// per lane, a power of two from 1 to 64
static int team_size(int partners) {
  int T = 1;
  while (T < 64 && 8 * T < partners) T <<= 1;
  return T;
}
// grid of 256-thread workgroups for `teams` teams of T lanes
static dim3 team_grid(int teams, int T) { return dim3((unsigned)(((size_t)teams * T + 255) / 256)); }

// the kernels of the three entry points on device buffers, queued on a
// stream (the trajectory-segment gradient drives them from device memory)
int vi_forward_dev(const ViBatch &b, double *F, double *Epart, double *E, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int ni = b.S * b.n_ion, nv = b.S * b.n_vap, Ti = team_size(b.n_vap), Tv = team_size(b.n_ion);
  const int gi = (int)team_grid(ni, Ti).x, gv = nv ? (int)team_grid(nv, Tv).x : 0;
  hipLaunchKernelGGL(k_vi_fwd, dim3(gi + gv), dim3(256), 0, st, b, F, Epart, Ti, Tv, gi);
  if (E) hipLaunchKernelGGL(k_vi_sum_e, dim3((b.S + 255) / 256), dim3(256), 0, st, Epart, b.S, b.n_ion, E);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MDN_OK : set_error(MDN_E_HIP, std::string("vi forward: ") + hipGetErrorString(e));
}

int vi_jvp_dev(const ViBatch &b, const double *U, double *HU, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int ni = b.S * b.n_ion, nv = b.S * b.n_vap, Ti = team_size(b.n_vap), Tv = team_size(b.n_ion);
  const int gi = (int)team_grid(ni, Ti).x, gv = nv ? (int)team_grid(nv, Tv).x : 0;
  hipLaunchKernelGGL(k_vi_jvp, dim3(gi + gv), dim3(256), 0, st, b, U, HU, Ti, Tv, gi);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MDN_OK : set_error(MDN_E_HIP, std::string("vi jvp: ") + hipGetErrorString(e));
}

int vi_adjoint_dev(const ViBatch &b, const double *gE, const double *gF, double *part, double *g_sigma,
                   double *g_eps, double *g_q, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int ni = b.S * b.n_ion, Ti = team_size(b.n_vap);
  hipLaunchKernelGGL(k_vi_adj, team_grid(ni, Ti), dim3(256), 0, st, b, gE, gF, part, Ti);
  hipLaunchKernelGGL(k_vi_sum_g, dim3((b.n_ion + 255) / 256), dim3(256), 0, st, part, b.S, b.n_ion, g_sigma, g_eps,
                     g_q);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MDN_OK : set_error(MDN_E_HIP, std::string("vi adjoint: ") + hipGetErrorString(e));
}

static int check_vi(const double *x, const double *w, int S, int n_ion, int n_vap, const double *si, const double *ei,
                    const double *qi, const double *sv, const double *ev, const double *qv) {
  if (!x || !w || !si || !ei || !qi || !sv || !ev || !qv) return set_error(MDN_E_ARG, "mdn_vi: null argument");
  if (S <= 0 || n_ion <= 0 || n_vap < 0) return set_error(MDN_E_ARG, "mdn_vi: bad sizes");
  return MDN_OK;
}

// device copies of a host batch; freed by the destructor
struct DevBatch {
  std::vector<double *> bufs;
  ~DevBatch() {
    for (double *p : bufs) (void)hipFree(p);
  }
  int up(const double *h, size_t n, const double **out) {
    double *d = nullptr;
    if (hipMalloc(&d, sizeof(double) * (n ? n : 1)) != hipSuccess) return set_error(MDN_E_HIP, "hipMalloc failed");
    bufs.push_back(d);
    if (h && n && hipMemcpy(d, h, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
      return set_error(MDN_E_HIP, "hipMemcpy failed");
    *out = d;
    return MDN_OK;
  }
  int alloc(size_t n, double **out) {
    const double *c = nullptr;
    int rc = up(nullptr, n, &c);
    *out = const_cast<double *>(c);
    return rc;
  }
};

static int upload_batch(DevBatch &db, ViBatch &b, const double *x, const double *w, int S, int n_ion, int n_vap,
                        const double *si, const double *ei, const double *qi, const double *sv, const double *ev,
                        const double *qv, const uint8_t *vin) {
  size_t n = (size_t)n_ion + n_vap;
  int rc;
  b.vin = nullptr;
  if (vin) {
    uint8_t *d = nullptr;
    size_t nb = (size_t)S * n_vap;
    if ((rc = db.alloc((nb + 7) / 8, reinterpret_cast<double **>(&d)))) return rc;
    if (nb && hipMemcpy(d, vin, nb, hipMemcpyHostToDevice) != hipSuccess)
      return set_error(MDN_E_HIP, "hipMemcpy failed");
    b.vin = d;
  }
  if ((rc = db.up(x, (size_t)S * n * 3, &b.x)) || (rc = db.up(w, (size_t)S * n, &b.w)) ||
      (rc = db.up(si, n_ion, &b.sig_i)) || (rc = db.up(ei, n_ion, &b.eps_i)) || (rc = db.up(qi, n_ion, &b.q_i)) ||
      (rc = db.up(sv, n_vap, &b.sig_v)) || (rc = db.up(ev, n_vap, &b.eps_v)) || (rc = db.up(qv, n_vap, &b.q_v)))
    return rc;
  b.S = S; b.n_ion = n_ion; b.n_vap = n_vap;
  return MDN_OK;
}

}  // namespace mdn

using namespace mdn;

extern "C" {

int mdn_vi_forward(const double *x, const double *w, int32_t S, int32_t n_ion, int32_t n_vap,
                   const double *ion_sigma, const double *ion_eps, const double *ion_q, const double *vap_sigma,
                   const double *vap_eps, const double *vap_q, const uint8_t *vap_in, double *E,
                   double *F) {
  int rc = check_vi(x, w, S, n_ion, n_vap, ion_sigma, ion_eps, ion_q, vap_sigma, vap_eps, vap_q);
  if (rc) return rc;
  if (!E || !F) return set_error(MDN_E_ARG, "mdn_vi_forward: null output");
  DevBatch db;
  ViBatch b;
  b.K = mdn_qqrd2e();
  if ((rc = upload_batch(db, b, x, w, S, n_ion, n_vap, ion_sigma, ion_eps, ion_q, vap_sigma, vap_eps, vap_q, vap_in)))
    return rc;
  size_t n = (size_t)n_ion + n_vap;
  double *dE, *dF, *work;
  if ((rc = db.alloc(S, &dE)) || (rc = db.alloc((size_t)S * n * 3, &dF)) || (rc = db.alloc((size_t)S * n_ion, &work)))
    return rc;
  int ni = S * n_ion, nv = S * n_vap, Ti = team_size(n_vap), Tv = team_size(n_ion);
  const int gi = (int)team_grid(ni, Ti).x, gv = nv ? (int)team_grid(nv, Tv).x : 0;
  hipLaunchKernelGGL(k_vi_fwd, dim3(gi + gv), dim3(256), 0, 0, b, dF, work, Ti, Tv, gi);
  hipLaunchKernelGGL(k_vi_sum_e, dim3((S + 255) / 256), dim3(256), 0, 0, work, S, n_ion, dE);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    return set_error(MDN_E_HIP, "mdn_vi_forward: kernel failed");
  if (hipMemcpy(E, dE, sizeof(double) * S, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(F, dF, sizeof(double) * S * n * 3, hipMemcpyDeviceToHost) != hipSuccess)
    return set_error(MDN_E_HIP, "mdn_vi_forward: copy back failed");
  return MDN_OK;
}

int mdn_vi_jvp(const double *x, const double *w, int32_t S, int32_t n_ion, int32_t n_vap, const double *ion_sigma,
               const double *ion_eps, const double *ion_q, const double *vap_sigma, const double *vap_eps,
               const double *vap_q, const uint8_t *vap_in, const double *u, double *hu) {
  int rc = check_vi(x, w, S, n_ion, n_vap, ion_sigma, ion_eps, ion_q, vap_sigma, vap_eps, vap_q);
  if (rc) return rc;
  if (!u || !hu) return set_error(MDN_E_ARG, "mdn_vi_jvp: null direction or output");
  DevBatch db;
  ViBatch b;
  b.K = mdn_qqrd2e();
  if ((rc = upload_batch(db, b, x, w, S, n_ion, n_vap, ion_sigma, ion_eps, ion_q, vap_sigma, vap_eps, vap_q, vap_in)))
    return rc;
  size_t n = (size_t)n_ion + n_vap;
  const double *dU = nullptr;
  double *dH = nullptr;
  if ((rc = db.up(u, (size_t)S * n * 3, &dU)) || (rc = db.alloc((size_t)S * n * 3, &dH))) return rc;
  int ni = S * n_ion, nv = S * n_vap, Ti = team_size(n_vap), Tv = team_size(n_ion);
  const int gi = (int)team_grid(ni, Ti).x, gv = nv ? (int)team_grid(nv, Tv).x : 0;
  hipLaunchKernelGGL(k_vi_jvp, dim3(gi + gv), dim3(256), 0, 0, b, dU, dH, Ti, Tv, gi);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    return set_error(MDN_E_HIP, "mdn_vi_jvp: kernel failed");
  if (hipMemcpy(hu, dH, sizeof(double) * S * n * 3, hipMemcpyDeviceToHost) != hipSuccess)
    return set_error(MDN_E_HIP, "mdn_vi_jvp: copy back failed");
  return MDN_OK;
}

int mdn_vi_adjoint(const double *x, const double *w, int32_t S, int32_t n_ion, int32_t n_vap,
                   const double *ion_sigma, const double *ion_eps, const double *ion_q, const double *vap_sigma,
                   const double *vap_eps, const double *vap_q, const uint8_t *vap_in, const double *gE,
                   const double *gF, double *g_sigma, double *g_eps, double *g_q) {
  int rc = check_vi(x, w, S, n_ion, n_vap, ion_sigma, ion_eps, ion_q, vap_sigma, vap_eps, vap_q);
  if (rc) return rc;
  if (!g_sigma || !g_eps || !g_q) return set_error(MDN_E_ARG, "mdn_vi_adjoint: null output");
  DevBatch db;
  ViBatch b;
  b.K = mdn_qqrd2e();
  if ((rc = upload_batch(db, b, x, w, S, n_ion, n_vap, ion_sigma, ion_eps, ion_q, vap_sigma, vap_eps, vap_q, vap_in)))
    return rc;
  size_t n = (size_t)n_ion + n_vap;
  const double *dgE = nullptr, *dgF = nullptr;
  if (gE && (rc = db.up(gE, S, &dgE))) return rc;
  if (gF && (rc = db.up(gF, (size_t)S * n * 3, &dgF))) return rc;
  double *work, *gs, *ge, *gq;
  if ((rc = db.alloc((size_t)S * n_ion * 3, &work)) || (rc = db.alloc(n_ion, &gs)) || (rc = db.alloc(n_ion, &ge)) ||
      (rc = db.alloc(n_ion, &gq)))
    return rc;
  int ni = S * n_ion, Ti = team_size(n_vap);
  hipLaunchKernelGGL(k_vi_adj, team_grid(ni, Ti), dim3(256), 0, 0, b, dgE, dgF, work, Ti);
  hipLaunchKernelGGL(k_vi_sum_g, dim3((n_ion + 255) / 256), dim3(256), 0, 0, work, S, n_ion, gs, ge, gq);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    return set_error(MDN_E_HIP, "mdn_vi_adjoint: kernel failed");
  if (hipMemcpy(g_sigma, gs, sizeof(double) * n_ion, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(g_eps, ge, sizeof(double) * n_ion, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(g_q, gq, sizeof(double) * n_ion, hipMemcpyDeviceToHost) != hipSuccess)
    return set_error(MDN_E_HIP, "mdn_vi_adjoint: copy back failed");
  return MDN_OK;
}

}  // extern "C"

namespace mdn {

// Scratch of mdn_vi_adjoint_noalloc: the per-(snapshot, ion atom) partial
// gradients, 3 doubles each, in a fixed device array (4 MiB per GPU, held
// until the process ends).  In host code `scratch` names a host-side
// stand-in; its device address is taken once per device (scratch_on) and
// only that address is passed to kernels.
constexpr size_t kViScratchDoubles = (size_t)1 << 19;
static __device__ __align__(256) double scratch[kViScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch[kMaxDevices] = {nullptr};

// the current device's copy of `scratch` (its address resolved on the first
// call on that device, with the device current)
static int scratch_on(double **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return set_error(MDN_E_HIP, "mdn_vi_adjoint_noalloc: hipGetDevice failed");
  if (dev < 0 || dev >= kMaxDevices) return set_error(MDN_E_ARG, "mdn_vi_adjoint_noalloc: device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(MDN_E_HIP, std::string("mdn_vi_adjoint_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<double *>(p);
  }
  *out = g_scratch[dev];
  return MDN_OK;
}

}  // namespace mdn

extern "C" {

int mdn_vi_adjoint_noalloc(const double *x, const double *w, int32_t S, int32_t n_ion, int32_t n_vap,
                           const double *ion_sigma, const double *ion_eps, const double *ion_q,
                           const double *vap_sigma, const double *vap_eps, const double *vap_q,
                           const uint8_t *vap_in, const double *gE, const double *gF, double *g_sigma, double *g_eps,
                           double *g_q, void *stream) {
  int rc = check_vi(x, w, S, n_ion, n_vap, ion_sigma, ion_eps, ion_q, vap_sigma, vap_eps, vap_q);
  if (rc) return rc;
  if (!g_sigma || !g_eps || !g_q) return set_error(MDN_E_ARG, "mdn_vi_adjoint_noalloc: null output");
  const size_t need = (size_t)S * (size_t)n_ion * 3;
  if (need > kViScratchDoubles)
    return set_error(MDN_E_ARG, "mdn_vi_adjoint_noalloc: S * n_ion = " + std::to_string((size_t)S * n_ion) +
                                    " exceeds the scratch capacity of " + std::to_string(kViScratchDoubles / 3) +
                                    " (snapshot, ion atom) pairs; use mdn_vi_adjoint");
  double *part = nullptr;
  if ((rc = scratch_on(&part))) return rc;
  ViBatch b;
  b.x = x; b.w = w; b.S = S; b.n_ion = n_ion; b.n_vap = n_vap;
  b.sig_i = ion_sigma; b.eps_i = ion_eps; b.q_i = ion_q;
  b.sig_v = vap_sigma; b.eps_v = vap_eps; b.q_v = vap_q;
  b.vin = vap_in;
  b.K = mdn_qqrd2e();
  hipStream_t st = static_cast<hipStream_t>(stream);
  // every slot of the partials that k_vi_sum_g reads is written by k_vi_adj
  // first, so the scratch needs no clearing between calls
  const int ni = S * n_ion, Ti = team_size(n_vap);
  hipLaunchKernelGGL(k_vi_adj, team_grid(ni, Ti), dim3(256), 0, st, b, gE, gF, part, Ti);
  hipLaunchKernelGGL(k_vi_sum_g, dim3((n_ion + 255) / 256), dim3(256), 0, st, part, S, n_ion, g_sigma, g_eps, g_q);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(MDN_E_HIP, std::string("mdn_vi_adjoint_noalloc: launch: ") + hipGetErrorString(e));
  return MDN_OK;
}

double mdn_qqrd2e(void) { return 1.602e-19 * 1.602e-19 / 4.0 / M_PI / 8.85e-12 * 6.02e23 * 1e10 / 4184.0; }

}  // extern "C"
