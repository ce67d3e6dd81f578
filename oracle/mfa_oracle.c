/*
 * mfa_oracle.c -- TEST INFRASTRUCTURE (the oracle), not product code.
 *
 * Plain-C FP64 restatement of the reference's per-step routines, in the
 * reference's order of operations.  Compiled with -O2 -ffp-contract=off so
 * every FP64 expression rounds as the Fortran source writes it.  Only tests/,
 * smoke() and bench.py's cpu_baseline load it, as the checker.
 *
 * Parity pinning of each function is described next to it; "parity unpinned"
 * marks what no known-answer or mathematical check covers.
 */
#include "mfa_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct orc_sys {
  orc_params p;
  int n;
  int n_neigh;   /* n_neigh_fl (mfa_common.f90:78) */
  int n_bin_fl;  /* mfa_common.f90:76 */
  double *r0, *v, *a, *force, *r0_old;
  double* force_d; /* DPD_VV: the dissipative force of the last new_dpd_fd (mfa_common.f90:197) */
  int32_t *a_type, *mic, *mic_old;
  int32_t* ff_list;    /* [N][0..n_neigh]: row 0 is the count, rows hold 1-based ids */
  int32_t* bin_fluid;  /* [cells][0..n_bin_fl] */
  int32_t* cell_of;    /* [N][3] */
  int f_skin;
  int n_rebuilds;
  /* Coulomb: q(:) (init_config.f90:221-226), self term (:235-239) and the
   * k-space tables of ewald_k(0) (ewald_k.f90:35-146) */
  double* q;
  double v_coul_self;
  int nk;
  double *k_vecs, *k_vecs_2, *inv_k_vecs_2, *re_rho_k, *im_rho_k;
  double inv_V, lz_half, fac_alpha;
  /* observation stage (init_obser.f90, observation.f90, functions.f90 histo) */
  mfa_obs_sums obs;
  double* dens_z;     /* [3][hist_dim] slab counts */
  double* f_on_heads; /* [n_chain][3] (verlet_velocities.f90:18) */
};

static const double PI_F = (double)3.1415926536f; /* mfa_common.f90:332, a default-real literal */

/* ------------------------------------------------------------------------ */
/* Threefry-2x32 (Salmon et al., SC11).  Pinned by the Random123 known-answer
 * vectors in tests/test_oracle.py. */
static uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

void orc_threefry2x32(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1, int rounds,
                      uint32_t* o0, uint32_t* o1) {
  static const int R[8] = {13, 15, 26, 6, 17, 29, 16, 24};
  uint32_t key[3] = {k0, k1, 0x1BD11BDAu ^ k0 ^ k1};
  uint32_t x0 = c0 + key[0], x1 = c1 + key[1];
  int inj = 0;
  for (int r = 0; r < rounds; ++r) {
    x0 = x0 + x1;
    x1 = rotl(x1, R[r % 8]);
    x1 = x1 ^ x0;
    if (r % 4 == 3) {
      ++inj;
      x0 = x0 + key[inj % 3];
      x1 = x1 + key[(inj + 1) % 3] + (uint32_t)inj;
    }
  }
  *o0 = x0;
  *o1 = x1;
}

/* Normal deviate of the pair (i0, j0) (0-based ids) at `step`: Box-Muller
 * on the two 24-bit uniforms of Threefry-2x32-13 with counter (min, max) and
 * key (seed, step).  Pinned by moment tests (mean, variance, kurtosis) and by
 * the FP64 GPU form on identical keys. */
double orc_pair_normal(uint32_t seed, int64_t step, int i0, int j0) {
  uint32_t w0, w1;
  const uint32_t lo = (uint32_t)(i0 < j0 ? i0 : j0), hi = (uint32_t)(i0 < j0 ? j0 : i0);
  orc_threefry2x32(lo, hi, seed, (uint32_t)step, ORC_TF_ROUNDS, &w0, &w1);
  const double u1 = (double)(w0 >> 8) * (1.0 / 16777216.0) + (1.0 / 33554432.0);
  const double u2 = (double)(w1 >> 8) * (1.0 / 16777216.0);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* Normal deviate d (0, 1, 2) of particle i0 (0-based id) at `step`, for the
 * Langevin random force (lgv_forces.f90:25-31, one rnor() per component):
 * the same Box-Muller transform on Threefry-2x32-13 with counter
 * (i0, 2^31 + d), a counter no pair uses (pair counters stay below 2^27).
 * Pinned by moment tests and by the FP64 GPU form on identical keys. */
double orc_particle_normal(uint32_t seed, int64_t step, int i0, int d) {
  uint32_t w0, w1;
  orc_threefry2x32((uint32_t)i0, 0x80000000u | (uint32_t)d, seed, (uint32_t)step, ORC_TF_ROUNDS, &w0,
                   &w1);
  const double u1 = (double)(w0 >> 8) * (1.0 / 16777216.0) + (1.0 / 33554432.0);
  const double u2 = (double)(w1 >> 8) * (1.0 / 16777216.0);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* ------------------------------------------------------------------------ */
/* Derived constants: init_params.f90:395-845 for SOLVENT 0/1, WALL 2,
 * SYSTEM 2/3 (with and without PARTICLE_4), make_binning_boxes.f90:14-64 and
 * init_config.f90:11-22.  Default-real literals are float values.
 * Pinned by tests/test_oracle.py against hand-evaluated values of the
 * reference's expressions (2**(1/6), 2.2, 0.4 as float32, e_shift closed
 * form) and against the library's own derivation. */
int orc_derive(const mfa_inputs* in, orc_params* p) {
  memset(p, 0, sizeof *p);
  const int nt = ORC_NT;
  const float c6 = powf(2.0f, 1.0f / 6.0f);
  p->system = in->system;
  p->symmetry = in->symmetry;
  p->thermostat = in->thermostat;
  p->particle_4 = in->particle_4;
  p->n_mon = in->n_mon;
  p->n_chain = in->n_chain;
  p->n_mon_d = in->n_mon_d;
  p->n_chain_d = in->n_chain_d;
  p->n_mon_e = in->particle_4 ? in->n_mon_e : 0;
  p->n_chain_e = in->particle_4 ? in->n_chain_e : 0;
  p->part_init_d = in->n_mon * in->n_chain;                          /* init_params.f90:871 */
  p->part_init_e = p->part_init_d + in->n_mon_d * in->n_chain_d;     /* :872 */
  p->n_part = p->part_init_e + p->n_mon_e * p->n_chain_e;            /* init_system.f90:199-217 */
  p->boundary[0] = in->x_space * (double)in->n_cell_w_x;             /* init_config.f90:11-19 */
  p->boundary[1] = in->y_space * (double)in->n_cell_w_y;
  p->boundary[2] = in->z_space_wall;
  for (int d = 0; d < 3; ++d) p->inv_boundary[d] = 1. / p->boundary[d];
  p->dt = in->dt;
  p->dt_2 = in->dt * in->dt;                                         /* init_params.f90:395 */
  p->temp = in->temp;
  p->r_chain_2 = in->r_chain * in->r_chain;                          /* :396 */
  p->inv_r_chain_2 = 1. / p->r_chain_2;                              /* :397 */
  p->k_chain = in->k_chain * p->r_chain_2;                           /* :398 */
  for (int t = 0; t < nt; ++t) p->mass_type[t] = in->mass_type[t];
  /* :407-420 */
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nt; ++j) p->sigma[i][j] = (i == j) ? in->sigma_ii[i] : 0.0;
  for (int i = 0; i < nt - 1; ++i)
    for (int j = i + 1; j < nt; ++j) {
      p->sigma[i][j] = (p->sigma[i][i] + p->sigma[j][j]) / 2;
      p->sigma[j][i] = p->sigma[i][j];
    }
  if (!in->particle_4)
    for (int i = 0; i < nt; ++i) p->sigma[i][nt - 1] = p->sigma[nt - 1][i] = 0.;
  /* :467-472 */
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nt; ++j) {
      p->sigma_2[i][j] = p->sigma[i][j] * p->sigma[i][j];
      p->epsil[i][j] = 4. * in->epsil[i][j];
    }
  /* :502 and :516-517 */
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nt; ++j) {
      double x = (double)c6 * p->sigma[i][j];
      p->range_2[i][j] = x * x;
    }
  {
    double x = (double)2.2f * p->sigma[0][nt - 1];
    p->range_2[0][nt - 1] = x * x;
    x = (double)2.2f * p->sigma[nt - 1][0];
    p->range_2[nt - 1][0] = x * x;
  }
  if (in->solvent == 0) { /* :545-558 */
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j) {
        double x = (double)(2.0f * c6) * p->sigma[i][j];
        p->range_2[i][j] = x * x;
      }
    if (!in->particle_4)
      for (int i = 0; i < nt; ++i) p->range_2[i][nt - 1] = p->range_2[nt - 1][i] = 0.;
  }
  /* :581-595 */
  p->a_w = in->a_w;
  for (int t = 0; t < nt; ++t) {
    p->sigma_wall[t] = (in->sigma_w + p->sigma[t][t]) / 2;
    p->a_wall[t] = in->a_w;
  }
  if (in->particle_4) {
    p->sigma_wall[3] = (in->sigma_w4 + p->sigma[3][3]) / 2.;
    p->a_wall[3] = in->a_w4;
  }
  p->z_space_wall = in->z_space_wall;
  if (!in->particle_4) { /* :604-607 */
    for (int i = 0; i < nt; ++i) p->range_2[i][nt - 1] = p->range_2[nt - 1][i] = 0.;
  } else { /* :629-676 */
    const double cc = (double)c6;
    for (int i = 0; i < 3; ++i) {
      double f = (in->inter4 == 4) ? (double)(2.0f * c6) : cc;
      double x = f * p->sigma[i][3];
      p->range_2[i][3] = x * x;
      x = f * p->sigma[3][i];
      p->range_2[3][i] = x * x;
    }
    double x = cc * p->sigma[3][3];
    if (in->inter4 >= 2) x = 2. * x;
    p->range_2[3][3] = x * x;
  }
  if (in->lj_rc > 0.0) /* survey patch P3 (explicit LJ cut-off) */
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j)
        if (p->range_2[i][j] > 0.0) {
          double x = in->lj_rc * p->sigma[i][j];
          p->range_2[i][j] = x * x;
        }
  if (in->system == 0 || in->system == 1 || in->system == 3) p->range_2[0][0] = 0.; /* :703-712 */
  for (int i = 0; i < nt; ++i) /* :713-724 */
    for (int j = 0; j < nt; ++j) {
      if (p->range_2[i][j] > 0.0) {
        double x = (p->sigma[j][i] * p->sigma[j][i]) / p->range_2[j][i];
        x = x * x * x;
        p->e_shift[j][i] = x * (x - 1.);
      } else {
        p->e_shift[j][i] = 0.0;
      }
    }
  double mx = 0.;
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nt; ++j)
      if (p->range_2[i][j] > mx) mx = p->range_2[i][j];
  p->r_cut_max = sqrt(mx); /* :766 */
  if (in->dpd_cut_off > 0.0) { /* :770-781 */
    p->dpd_cut_off_on = 1;
    p->r_cut_dpd = (double)(float)in->dpd_cut_off;
    if (p->r_cut_dpd > p->r_cut_max) p->r_cut_max = p->r_cut_dpd;
    p->r_cut_dpd_2 = p->r_cut_dpd * p->r_cut_dpd;
    p->r_cut_max_2 = p->r_cut_max * p->r_cut_max;
  }
  { /* a real-space Coulomb cut-off beyond the LJ range widens list and cells */
    int charged = 0;
    for (int t = 0; t < nt; ++t) charged |= in->charge[t] != 0.0;
    if (charged && !in->ewald_compat && in->rc_coul > p->r_cut_max) {
      p->r_cut_max = in->rc_coul;
      p->r_cut_max_2 = p->r_cut_max * p->r_cut_max;
    }
  }
  for (int i = 0; i < nt; ++i) /* :786-795 */
    for (int j = 0; j < nt; ++j)
      p->inv_range_2[i][j] = p->range_2[i][j] != 0 ? 1. / p->range_2[i][j] : 0.;
  p->skin = in->skin > 0.0 ? in->skin : (double)0.4f; /* :800 (P4 when set) */
  p->skin_2 = (p->r_cut_max + p->skin) * (p->r_cut_max + p->skin) -
              p->r_cut_max * p->r_cut_max; /* :808 */
  p->friction1 = in->friction[0];
  p->sig = sqrt(2. * in->temp * in->friction[0] / in->dt); /* :845 */
  p->seed = (uint32_t)in->iseed;
  for (int d = 0; d < 3; ++d) p->va_spring_twall[d] = in->va_spring_twall[d];
  /* make_binning_boxes.f90:14-37 on the provisional grid of init_system.f90:273-275 */
  int prov[3] = {(int)(0.4f * (float)in->n_cell_w_x), (int)(0.7f * (float)in->n_cell_w_y), 20};
  for (int d = 0; d < 3; ++d)
    p->delta_bin[d] = prov[d] >= 3 ? 1 : (prov[d] == 2 ? 0 : -1);
  /* make_binning_boxes.f90:42-64 */
  const double binning_box_skin = (double)0.4f;
  for (int d = 0; d < 3; ++d) {
    p->n_bin[d] = (int)(p->boundary[d] / (p->r_cut_max + binning_box_skin));
    if (p->n_bin[d] < 1) return 1;
    p->r_bin[d] = p->boundary[d] / (double)p->n_bin[d];
    p->inv_r_bin[d] = 1. / p->r_bin[d];
  }
  /* Coulomb constants (init_system.f90:83-98, init_config.f90:221-239).  The
   * reference's real-space force uses inv_sqrt_pi = 1./pi (:230); the compat
   * form keeps it, the default uses 1/sqrt(pi).  The reference evaluates the
   * real space only inside the LJ range (fluid_fluid.f90:94,166); without
   * ewald_compat and with rc_coul set, it is evaluated within rc_coul. */
  for (int t = 0; t < nt; ++t) {
    p->charge[t] = in->charge[t];
    if (in->charge[t] != 0.0) p->coulomb = 1;
  }
  p->ewald_compat = in->ewald_compat;
  p->alpha_c = in->alpha_c;
  p->sqrt_alpha = sqrt(in->alpha_c);
  p->inv_sqrt_pi = in->ewald_compat ? 1. / PI_F : 1. / sqrt(3.14159265358979323846);
  p->n_kmax = in->n_kmax;
  p->rc_coul_2 = (!in->ewald_compat && in->rc_coul > 0.0) ? in->rc_coul * in->rc_coul : 0.0;
  p->bending = in->bending;
  p->dpd_vv = in->dpd_vv ? 1 : 0;
  p->k_bend = in->k_bend;
  p->alpha_eq = in->alpha_eq;
  /* init_config.f90:264-281 and init_force_switch_on.f90:15-26 */
  if (in->force_switch_on && in->s_time == 1) {
    double maxsig = 0.;
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j)
        if (p->sigma[i][j] > maxsig) maxsig = p->sigma[i][j];
    p->r_2_min_init = (maxsig / 0.9f) * (maxsig / 0.9f);
    p->r_2_min_time = (double)in->n_relax * in->dt < 40. ? 40. : (double)in->n_relax * in->dt;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* ewald_k(0) (ewald_k.f90:35-146): the k grid of the half space kz >= 0
 * inside the sphere |k| <= k_c, without k = 0, the lower triangle of the
 * plane kz = 0 and half of its diagonal; with walls (SYMMETRY 0) the z
 * period is the slab 3 Lz (z_slab = 2).  Also inv_V (:126) and lz_half
 * (init_config.f90:246).  Pinned by the compiled reference (charged brush),
 * by the count of a brute enumeration and by Madelung sums. */
static void ewald_init(orc_sys* s) {
  const orc_params* p = &s->p;
  double delta_k[3], z_slab;
  if (p->symmetry == 1) {
    for (int d = 0; d < 3; ++d) delta_k[d] = 2. * PI_F * p->inv_boundary[d];
    z_slab = 0.;
  } else {
    z_slab = 2.;
    delta_k[0] = 2. * PI_F * p->inv_boundary[0];
    delta_k[1] = 2. * PI_F * p->inv_boundary[1];
    delta_k[2] = 2. * PI_F / (p->boundary[2] + z_slab * p->boundary[2]);
  }
  const double k_c = (delta_k[0] + delta_k[1]) * (double)p->n_kmax / 2.;
  const int nkx_max = (int)(k_c / delta_k[0]);
  const int nky_max = (int)(k_c / delta_k[1]);
  const int nkz_max = (int)(k_c / delta_k[2]);
  for (int pass = 0; pass < 2; ++pass) {
    int k_index = 0;
    for (int nkz = 0; nkz <= nkz_max; ++nkz)
      for (int nky = -nky_max; nky <= nky_max; ++nky)
        for (int nkx = -nkx_max; nkx <= nkx_max; ++nkx) {
          const double ax = (double)nkx * delta_k[0], ay = (double)nky * delta_k[1],
                       az = (double)nkz * delta_k[2];
          const double rk_2 = ax * ax + ay * ay + az * az;
          if (rk_2 > k_c * k_c) continue;
          if (nkz == 0 && nky == 0 && nkx == 0) continue;
          if (nkz == 0 && nkx + nky < 0) continue;
          if (nkz == 0 && -nkx == nky && nky > 0) continue;
          if (pass == 1) {
            double* kv = s->k_vecs + 3 * (size_t)k_index;
            kv[0] = delta_k[0] * (double)nkx;
            kv[1] = delta_k[1] * (double)nky;
            kv[2] = delta_k[2] * (double)nkz;
            s->k_vecs_2[k_index] = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
            s->inv_k_vecs_2[k_index] = 1. / s->k_vecs_2[k_index];
          }
          ++k_index;
        }
    if (pass == 0) {
      s->nk = k_index;
      s->k_vecs = (double*)calloc(3 * (size_t)k_index + 3, sizeof(double));
      s->k_vecs_2 = (double*)calloc((size_t)k_index + 1, sizeof(double));
      s->inv_k_vecs_2 = (double*)calloc((size_t)k_index + 1, sizeof(double));
      s->re_rho_k = (double*)calloc((size_t)k_index + 1, sizeof(double));
      s->im_rho_k = (double*)calloc((size_t)k_index + 1, sizeof(double));
    }
  }
  s->fac_alpha = 1. / (4. * p->alpha_c);
  s->inv_V = p->inv_boundary[0] * p->inv_boundary[1] / (p->boundary[2] * (1. + z_slab));
  s->lz_half = p->boundary[2] / 2.;
}

int orc_ewald_nk(orc_sys* s, double* kvecs) {
  if (kvecs && s->nk) memcpy(kvecs, s->k_vecs, 3 * (size_t)s->nk * sizeof(double));
  return s->nk;
}

orc_sys* orc_create(const orc_params* p, int n_neigh) {
  orc_sys* s = (orc_sys*)calloc(1, sizeof(orc_sys));
  s->p = *p;
  s->n = p->n_part;
  s->n_neigh = n_neigh > 0 ? n_neigh : 256;
  s->n_bin_fl = 300;
  const size_t n = (size_t)s->n;
  s->r0 = (double*)calloc(3 * n, sizeof(double));
  s->v = (double*)calloc(3 * n, sizeof(double));
  s->a = (double*)calloc(3 * n, sizeof(double));
  s->force = (double*)calloc(3 * n, sizeof(double));
  s->force_d = (double*)calloc(3 * n, sizeof(double)); /* init_params.f90:839 */
  s->r0_old = (double*)calloc(3 * n, sizeof(double));
  s->a_type = (int32_t*)calloc(n, sizeof(int32_t));
  s->mic = (int32_t*)calloc(3 * n, sizeof(int32_t));
  s->mic_old = (int32_t*)calloc(3 * n, sizeof(int32_t));
  s->ff_list = (int32_t*)calloc(n * (size_t)(s->n_neigh + 1), sizeof(int32_t));
  const size_t nc = (size_t)p->n_bin[0] * p->n_bin[1] * p->n_bin[2];
  s->bin_fluid = (int32_t*)calloc(nc * (size_t)(s->n_bin_fl + 1), sizeof(int32_t));
  s->cell_of = (int32_t*)calloc(3 * n, sizeof(int32_t));
  s->q = (double*)calloc(n, sizeof(double));
  s->f_on_heads = (double*)calloc(3 * (size_t)(p->n_chain > 0 ? p->n_chain : 1), sizeof(double));
  if (p->coulomb) ewald_init(s);
  return s;
}

void orc_destroy(orc_sys* s) {
  if (!s) return;
  free(s->r0); free(s->v); free(s->a); free(s->force); free(s->force_d); free(s->r0_old);
  free(s->a_type); free(s->mic); free(s->mic_old); free(s->ff_list); free(s->bin_fluid);
  free(s->cell_of);
  free(s->q);
  free(s->dens_z);
  free(s->f_on_heads);
  free(s->k_vecs); free(s->k_vecs_2); free(s->inv_k_vecs_2); free(s->re_rho_k); free(s->im_rho_k);
  free(s);
}

void orc_set_state(orc_sys* s, const double* r0, const double* v, const double* a,
                   const int32_t* a_type, const int32_t* mic) {
  const size_t n = (size_t)s->n;
  memcpy(s->r0, r0, 3 * n * sizeof(double));
  memcpy(s->v, v, 3 * n * sizeof(double));
  memcpy(s->a, a, 3 * n * sizeof(double));
  memcpy(s->a_type, a_type, n * sizeof(int32_t));
  if (mic) memcpy(s->mic, mic, 3 * n * sizeof(int32_t));
  else memset(s->mic, 0, 3 * n * sizeof(int32_t));
  s->f_skin = 1;
  /* q(i) = charge(a_type(i)) and the self term (init_config.f90:221-239) */
  s->v_coul_self = 0.;
  for (size_t i = 0; i < n; ++i) {
    s->q[i] = s->p.coulomb ? s->p.charge[a_type[i] - 1] : 0.;
    s->v_coul_self = s->v_coul_self - s->q[i] * s->q[i];
  }
  s->v_coul_self = s->v_coul_self * s->p.sqrt_alpha / sqrt(PI_F);
}

void orc_get_state(orc_sys* s, double* r0, double* v, double* a, double* force, int32_t* mic) {
  const size_t n = (size_t)s->n;
  if (r0) memcpy(r0, s->r0, 3 * n * sizeof(double));
  if (v) memcpy(v, s->v, 3 * n * sizeof(double));
  if (a) memcpy(a, s->a, 3 * n * sizeof(double));
  if (force) memcpy(force, s->force, 3 * n * sizeof(double));
  if (mic) memcpy(mic, s->mic, 3 * n * sizeof(int32_t));
}

#define R(i, d) s->r0[3 * (size_t)(i) + (d)]
#define V(i, d) s->v[3 * (size_t)(i) + (d)]
#define A(i, d) s->a[3 * (size_t)(i) + (d)]
#define F(i, d) s->force[3 * (size_t)(i) + (d)]
#define MIC(i, d) s->mic[3 * (size_t)(i) + (d)]

/* verlet_positions.f90:144-170 (update) and :233-250 (wrap).  Pinned by
 * closed forms: free flight and constant acceleration, tests/test_oracle.py. */
void orc_verlet_positions(orc_sys* s) {
  const orc_params* p = &s->p;
  for (int i = 0; i < s->n; ++i) {
    for (int d = 0; d < 3; ++d) R(i, d) = R(i, d) + p->dt * V(i, d) + 0.5 * p->dt_2 * A(i, d);
    for (int d = 0; d < 3; ++d) V(i, d) = V(i, d) + 0.5 * p->dt * A(i, d);
  }
  const int nd = p->symmetry == 0 ? 2 : 3;
  for (int i = 0; i < s->n; ++i)
    for (int d = 0; d < nd; ++d) {
      if (R(i, d) > p->boundary[d]) {
        R(i, d) = R(i, d) - p->boundary[d];
        MIC(i, d) += 1;
      } else if (R(i, d) < 0.) {
        R(i, d) = R(i, d) + p->boundary[d];
        MIC(i, d) -= 1;
      }
    }
}

/* check_skin.f90:8-21.  Pinned by a closed-form displacement test. */
int orc_check_skin(orc_sys* s) {
  const orc_params* p = &s->p;
  for (int i = 0; i < s->n; ++i) {
    double dr[3], r_dummy = 0.;
    for (int d = 0; d < 3; ++d)
      dr[d] = s->r0_old[3 * (size_t)i + d] + (double)s->mic_old[3 * (size_t)i + d] * p->boundary[d] -
              R(i, d) - (double)MIC(i, d) * p->boundary[d];
    for (int d = 0; d < 3; ++d) r_dummy = r_dummy + dr[d] * dr[d];
    if (r_dummy > (p->skin / 2.) * (p->skin / 2.)) s->f_skin = 1;
  }
  return s->f_skin;
}

static double pbc(double d, double L, double invL) { return d - L * (double)(int)(2 * d * invL); }

/* Compensated (Neumaier) sum for the virial, a quantity of this build (the
 * reference has none, so no summation order to follow): the result is the
 * sum of the terms to about one rounding, whatever the number of terms. */
typedef struct { double s, c; } ksum_t;
static void ksum_add(ksum_t* k, double x) {
  const double t = k->s + x;
  if (fabs(k->s) >= fabs(x))
    k->c += (k->s - t) + x;
  else
    k->c += (x - t) + k->s;
  k->s = t;
}
static double ksum_value(const ksum_t* k) { return k->s + k->c; }

static double list_crit(const orc_params* p, int ti, int tj) {
  double l2 = p->dpd_cut_off_on ? p->r_cut_max_2 + p->skin_2 : p->range_2[ti][tj] + p->skin_2;
  if (p->coulomb && p->rc_coul_2 + p->skin_2 > l2) l2 = p->rc_coul_2 + p->skin_2;
  return l2;
}

/* binning.f90:11-127: cell fill in particle order, then the 27-cell scan with
 * i<j and the balanced insertion into the shorter list.  The pair set is
 * pinned by orc_brute_pairs (tests/test_oracle.py). */
int orc_binning(orc_sys* s) {
  const orc_params* p = &s->p;
  const int nx = p->n_bin[0], ny = p->n_bin[1], nz = p->n_bin[2];
  const int stride = s->n_bin_fl + 1;
  const size_t nc = (size_t)nx * ny * nz;
  int err = 0;
  memset(s->bin_fluid, 0, nc * (size_t)stride * sizeof(int32_t));
#define BIN(x, y, z, k) s->bin_fluid[((((size_t)(z) * ny + (y)) * nx + (x)) * stride) + (k)]
  for (int i = 0; i < s->n; ++i) {
    const int ix = ((int)(R(i, 0) * p->inv_r_bin[0])) % nx;
    const int iy = ((int)(R(i, 1) * p->inv_r_bin[1])) % ny;
    const int iz = ((int)(R(i, 2) * p->inv_r_bin[2])) % nz;
    if (ix < 0 || iy < 0 || iz < 0) return MFA_ERR_CELL;
    s->cell_of[3 * (size_t)i + 0] = ix;
    s->cell_of[3 * (size_t)i + 1] = iy;
    s->cell_of[3 * (size_t)i + 2] = iz;
    const int c = ++BIN(ix, iy, iz, 0);
    if (c > s->n_bin_fl) return MFA_ERR_LIST_OVERFLOW;
    BIN(ix, iy, iz, c) = i + 1;
  }
  const int ls = s->n_neigh + 1;
#define FF(i, k) s->ff_list[(size_t)(i) * ls + (k)]
  for (int i = 0; i < s->n; ++i) FF(i, 0) = 0;
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const int n1 = BIN(ix, iy, iz, 0);
        for (int k1 = 1; k1 <= n1; ++k1) {
          const int ip = BIN(ix, iy, iz, k1) - 1;
          const int ti = s->a_type[ip] - 1;
          for (int jx = ix - 1; jx <= ix + p->delta_bin[0]; ++jx) {
            const int kx = (jx + nx) % nx;
            for (int jy = iy - 1; jy <= iy + p->delta_bin[1]; ++jy) {
              const int ky = (jy + ny) % ny;
              for (int jz = iz - 1; jz <= iz + p->delta_bin[2]; ++jz) {
                const int kz = (jz + nz) % nz;
                const int n2 = BIN(kx, ky, kz, 0);
                for (int k2 = 1; k2 <= n2; ++k2) {
                  const int jp = BIN(kx, ky, kz, k2) - 1;
                  if (ip < jp) {
                    const int tj = s->a_type[jp] - 1;
                    double dr[3];
                    for (int d = 0; d < 3; ++d) dr[d] = R(ip, d) - R(jp, d);
                    dr[0] = pbc(dr[0], p->boundary[0], p->inv_boundary[0]);
                    dr[1] = pbc(dr[1], p->boundary[1], p->inv_boundary[1]);
                    if (p->symmetry == 1) dr[2] = pbc(dr[2], p->boundary[2], p->inv_boundary[2]);
                    double r_2 = 0.;
                    r_2 = r_2 + dr[0] * dr[0];
                    r_2 = r_2 + dr[1] * dr[1];
                    r_2 = r_2 + dr[2] * dr[2];
                    if (r_2 < list_crit(p, ti, tj)) {
                      if (FF(ip, 0) <= FF(jp, 0)) {
                        if (FF(ip, 0) >= s->n_neigh) err = MFA_ERR_LIST_OVERFLOW;
                        else { FF(ip, 0) += 1; FF(ip, FF(ip, 0)) = jp + 1; }
                      } else {
                        if (FF(jp, 0) >= s->n_neigh) err = MFA_ERR_LIST_OVERFLOW;
                        else { FF(jp, 0) += 1; FF(jp, FF(jp, 0)) = ip + 1; }
                      }
                    }
                  }
                }
              }
            }
          }
        }
      }
  s->f_skin = 0; /* :125-127 */
  memcpy(s->r0_old, s->r0, 3 * (size_t)s->n * sizeof(double));
  memcpy(s->mic_old, s->mic, 3 * (size_t)s->n * sizeof(int32_t));
  s->n_rebuilds += 1;
  return err;
#undef BIN
}

void orc_get_cells(orc_sys* s, int32_t* cells) {
  memcpy(cells, s->cell_of, 3 * (size_t)s->n * sizeof(int32_t));
}

static int cmp_pair(const void* x, const void* y) {
  const int32_t* a = (const int32_t*)x;
  const int32_t* b = (const int32_t*)y;
  if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
  if (a[1] != b[1]) return a[1] < b[1] ? -1 : 1;
  return 0;
}

/* Unordered pair set of ff_list as sorted (min,max), 1-based. */
int64_t orc_get_pairs(orc_sys* s, int32_t* pairs, int64_t cap) {
  const int ls = s->n_neigh + 1;
  int64_t m = 0;
  for (int i = 0; i < s->n; ++i)
    for (int k = 1; k <= FF(i, 0); ++k) {
      const int j = FF(i, k) - 1;
      if (pairs && m < cap) {
        pairs[2 * m] = (i < j ? i : j) + 1;
        pairs[2 * m + 1] = (i < j ? j : i) + 1;
      }
      ++m;
    }
  if (pairs && m <= cap) qsort(pairs, (size_t)m, 2 * sizeof(int32_t), cmp_pair);
  return m;
}

/* O(N^2) version of the list criterion of binning.f90:84-105 over all i<j,
 * independent of the cell grid. */
int64_t orc_brute_pairs(orc_sys* s, int32_t* pairs, int64_t cap) {
  const orc_params* p = &s->p;
  int64_t m = 0;
  for (int i = 0; i < s->n; ++i)
    for (int j = i + 1; j < s->n; ++j) {
      double dr[3];
      for (int d = 0; d < 3; ++d) dr[d] = R(i, d) - R(j, d);
      dr[0] = pbc(dr[0], p->boundary[0], p->inv_boundary[0]);
      dr[1] = pbc(dr[1], p->boundary[1], p->inv_boundary[1]);
      if (p->symmetry == 1) dr[2] = pbc(dr[2], p->boundary[2], p->inv_boundary[2]);
      const double r_2 = (dr[0] * dr[0] + dr[1] * dr[1]) + dr[2] * dr[2];
      if (r_2 < list_crit(p, s->a_type[i] - 1, s->a_type[j] - 1)) {
        if (pairs && m < cap) {
          pairs[2 * m] = i + 1;
          pairs[2 * m + 1] = j + 1;
        }
        ++m;
      }
    }
  return m;
}

void orc_zero_force(orc_sys* s) { memset(s->force, 0, 3 * (size_t)s->n * sizeof(double)); }

/* ewald_real.f90:15-32 for the pair (ip, jp) at r_2 (after the switch-on
 * clamp) and separation dr, in the reference's operation order.  Pinned by
 * the compiled reference (charged brush, compat form) and by the Madelung
 * constant of a rock-salt crystal (correct form, with ewald_k). */
static void ewald_real(orc_sys* s, int ip, int jp, double r_2, const double* dr, double* v_coul) {
  const orc_params* p = &s->p;
  const double r = sqrt(r_2);
  const double inv_r = 1. / r;
  const double qq = s->q[ip] * s->q[jp];
  *v_coul = *v_coul + qq * erfc(p->sqrt_alpha * r) * inv_r;
  const double f_scal = qq * p->sqrt_alpha * p->inv_sqrt_pi * 2. * exp(-p->alpha_c * r_2) * inv_r +
                        qq * erfc(p->sqrt_alpha * r) * inv_r * inv_r;
  for (int d = 0; d < 3; ++d) {
    const double fc = f_scal * dr[d] * inv_r;
    F(ip, d) = F(ip, d) + fc;
    F(jp, d) = F(jp, d) - fc;
  }
}

/* fluid_fluid.f90:52-223 with dpd_forces.f90:2-98 (DPD_WEIGHT 0,
 * BIN_TYPE 0).  LJ is pinned by closed forms (minimum at 2^(1/6) sigma,
 * shifted energy zero at the cut-off) and energy/force finite differences;
 * the pair loop by antisymmetry (total force zero with DPD off); DPD by its
 * fluctuation-dissipation moments. */
void orc_fluid_fluid(orc_sys* s, int64_t step, double* out) {
  const orc_params* p = &s->p;
  const int ls = s->n_neigh + 1;
  const int head_excl = p->symmetry == 0 && (p->system == 0 || p->system == 1 || p->system == 3);
  double v_ff = 0., v_coul = 0.;
  ksum_t vir = {0., 0.};
  /* force switch-on (fluid_fluid.f90:42-49), r_time = step*dt (md_main.f90:69) */
  double r_2_min = 0.;
  const double r_time = (double)step * p->dt;
  if (p->r_2_min_time > 0. && r_time < p->r_2_min_time)
    r_2_min = (1. - r_time / p->r_2_min_time) * p->r_2_min_init;
  for (int ip = 0; ip < s->n; ++ip) {
    const int ti = s->a_type[ip] - 1;
    double f_ipart[3] = {0., 0., 0.};
    for (int k = 1; k <= FF(ip, 0); ++k) {
      const int jp = FF(ip, k) - 1;
      const int tj = s->a_type[jp] - 1;
      const double r_cut2 = p->range_2[ti][tj];
      double dr[3];
      for (int d = 0; d < 3; ++d) dr[d] = R(ip, d) - R(jp, d);
      dr[0] = pbc(dr[0], p->boundary[0], p->inv_boundary[0]);
      dr[1] = pbc(dr[1], p->boundary[1], p->inv_boundary[1]);
      if (p->symmetry == 1) dr[2] = pbc(dr[2], p->boundary[2], p->inv_boundary[2]);
      double r_2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
      const double r_2_raw = r_2;
      int do_dpd = 0;
      double w_arg = 0.;
      double inv_sqrt_r_2 = sqrt(1. / r_2);
      if (r_2 < r_cut2) {
        if (r_2 < r_2_min) r_2 = r_2_min; /* r_2 = max(r_2, r_2_min) (:96) */
        const double inv_r_2 = 1. / r_2;
        inv_sqrt_r_2 = sqrt(1. / r_2);
        const double l_eps = p->epsil[ti][tj];
        const double r_61 = p->sigma_2[ti][tj] * inv_r_2;
        const double r_6 = r_61 * r_61 * r_61;
        const double r_12 = r_6 * r_6;
        const double pot_loc = (r_12 - r_6) - p->e_shift[ti][tj];
        v_ff = v_ff + l_eps * pot_loc;
        const double r_dummy = l_eps * (-12 * r_12 + 6 * r_6) * inv_r_2;
        ksum_add(&vir, -r_dummy * r_2_raw);
        for (int d = 0; d < 3; ++d) {
          const double fl = r_dummy * dr[d];
          f_ipart[d] = f_ipart[d] - fl;
          F(jp, d) = F(jp, d) + fl;
        }
        if (p->thermostat == 0 && !p->dpd_cut_off_on) {
          do_dpd = 1;
          w_arg = r_2 * p->inv_range_2[ti][tj];
        }
      }
      /* ewald_real (fluid_fluid.f90:166, ewald_real.f90:15-32): inside the LJ
       * range with the clamped r_2 (compat), or within rc_coul */
      const int real_space = p->coulomb && (p->rc_coul_2 > 0. ? r_2_raw < p->rc_coul_2 : r_2 < r_cut2);
      if (real_space && !do_dpd) ewald_real(s, ip, jp, r_2, dr, &v_coul);
      if (p->thermostat == 0 && p->dpd_cut_off_on) {
        if (head_excl && (ti == 0 || tj == 0)) continue;
        if (r_2 < p->r_cut_dpd_2) {
          do_dpd = 1;
          w_arg = r_2 / p->r_cut_dpd_2;
        }
      }
      if (do_dpd) { /* dpd_forces.f90:19-94 */
        const double g_rand = orc_pair_normal(p->seed, step, ip, jp);
        double w_d = 1. - sqrt(w_arg);
        const double w_r = w_d;
        w_d = w_d * w_d;
        double rv[3], vd[3], dv[3];
        for (int d = 0; d < 3; ++d) rv[d] = dr[d] * inv_sqrt_r_2;
        for (int d = 0; d < 3; ++d) vd[d] = p->sig * w_r * g_rand * rv[d];
        for (int d = 0; d < 3; ++d) {
          F(ip, d) = F(ip, d) + vd[d];
          F(jp, d) = F(jp, d) - vd[d];
        }
        for (int d = 0; d < 3; ++d) dv[d] = V(ip, d) - V(jp, d);
        const double rdum = dv[0] * rv[0] + dv[1] * rv[1] + dv[2] * rv[2];
        for (int d = 0; d < 3; ++d) vd[d] = -1. * p->friction1 * w_d * rdum * rv[d];
        for (int d = 0; d < 3; ++d) {
          F(ip, d) = F(ip, d) + vd[d];
          F(jp, d) = F(jp, d) - vd[d];
        }
        /* the reference calls ewald_real after dpd_forces inside the LJ branch */
        if (real_space && !p->dpd_cut_off_on) ewald_real(s, ip, jp, r_2, dr, &v_coul);
      }
    }
    for (int d = 0; d < 3; ++d) F(ip, d) = F(ip, d) + f_ipart[d];
    if (p->thermostat == 1) { /* lgv_forces (fluid_fluid.f90:219-221, lgv_forces.f90:25-43) */
      const double m = p->mass_type[ti];
      double vec[3];
      for (int d = 0; d < 3; ++d) vec[d] = p->sig * orc_particle_normal(p->seed, step, ip, d) * sqrt(m);
      for (int d = 0; d < 3; ++d) vec[d] = vec[d] - p->friction1 * m * V(ip, d);
      for (int d = 0; d < 3; ++d) F(ip, d) = F(ip, d) + vec[d];
    }
  }
  out[0] = v_ff;
  out[1] = ksum_value(&vir);
  out[2] = p->coulomb ? v_coul + s->v_coul_self : 0.; /* fluid_fluid.f90:238-240 */
}

/* ewald_k(1) (ewald_k.f90:150-199): structure factor over the charged
 * particles, u_cou_k, and the k-space forces added to F.  Pinned with
 * ewald_init. */
double orc_ewald_k(orc_sys* s) {
  const orc_params* p = &s->p;
  if (!p->coulomb || s->nk == 0) return 0.;
  const int nk = s->nk;
  for (int k = 0; k < nk; ++k) s->re_rho_k[k] = s->im_rho_k[k] = 0.;
  for (int i = 0; i < s->n; ++i)
    for (int k = 0; k < nk; ++k) {
      const double* kv = s->k_vecs + 3 * (size_t)k;
      const double lphase = kv[0] * R(i, 0) + kv[1] * R(i, 1) + kv[2] * R(i, 2);
      s->re_rho_k[k] = s->re_rho_k[k] + s->q[i] * cos(lphase);
      s->im_rho_k[k] = s->im_rho_k[k] + s->q[i] * sin(lphase);
    }
  double u_cou_k = 0.;
  for (int k = 0; k < nk; ++k)
    u_cou_k = u_cou_k + s->inv_k_vecs_2[k] * exp(-s->k_vecs_2[k] * s->fac_alpha) *
                            (s->re_rho_k[k] * s->re_rho_k[k] + s->im_rho_k[k] * s->im_rho_k[k]);
  u_cou_k = u_cou_k * 4. * PI_F * s->inv_V;
  for (int i = 0; i < s->n; ++i) {
    double f_k[3] = {0., 0., 0.};
    for (int k = 0; k < nk; ++k) {
      const double* kv = s->k_vecs + 3 * (size_t)k;
      const double cphase = kv[0] * R(i, 0) + kv[1] * R(i, 1) + kv[2] * R(i, 2);
      const double t = cos(cphase) * s->im_rho_k[k] - sin(cphase) * s->re_rho_k[k];
      for (int d = 0; d < 3; ++d) f_k[d] = f_k[d] + kv[d] * s->inv_k_vecs_2[k] *
                                                     exp(-s->k_vecs_2[k] * s->fac_alpha) * t;
    }
    for (int d = 0; d < 3; ++d) F(i, d) = F(i, d) + -8. * PI_F * s->q[i] * f_k[d] * s->inv_V;
  }
  return u_cou_k;
}

/* dipolar_correction.f90:20-33 (slab geometry). */
double orc_dipolar(orc_sys* s) {
  const orc_params* p = &s->p;
  if (!p->coulomb || p->symmetry != 0) return 0.;
  double m_z = 0.;
  for (int i = 0; i < s->n; ++i) m_z = m_z + s->q[i] * (R(i, 2) - s->lz_half);
  for (int i = 0; i < s->n; ++i) F(i, 2) = F(i, 2) - 4. * PI_F * s->q[i] * m_z * s->inv_V;
  return 2. * PI_F * m_z * m_z * s->inv_V;
}

/* bending(2) (bending.f90:33-127): harmonic angle between consecutive bonds
 * of the brush chains, V = k_bend/2 (alpha - alpha_eq)^2, with the
 * reference's minimum image in x, y and z and its cut-off of the force
 * direction below |d|^2 < 1e-7.  Pinned by the compiled reference (BENDING
 * build), finite differences of v_bend and zero net force per angle. */
double orc_bending(orc_sys* s) {
  const orc_params* p = &s->p;
  if (!p->bending) return 0.;
  double v_bend = 0.;
  for (int l = 0; l < p->n_chain; ++l) {
    const int b = l * p->n_mon; /* 0-based index of bead 1 */
    double r_next[3], r_prev[3];
    for (int j = 0; j < 3; ++j) r_next[j] = R(b + 1, j) - R(b, j);
    for (int j = 0; j < 3; ++j) r_next[j] = pbc(r_next[j], p->boundary[j], p->inv_boundary[j]);
    for (int i = 2; i <= p->n_mon - 1; ++i) { /* 1-based bead i, centre of the angle */
      double dot_pr = 0, r_prev_2 = 0, r_next_2 = 0, dir_prev_2 = 0, dir_next_2 = 0;
      for (int j = 0; j < 3; ++j) r_prev[j] = r_next[j];
      for (int j = 0; j < 3; ++j) r_next[j] = R(b + i, j) - R(b + i - 1, j);
      for (int j = 0; j < 3; ++j) r_next[j] = pbc(r_next[j], p->boundary[j], p->inv_boundary[j]);
      for (int j = 0; j < 3; ++j) {
        dot_pr = dot_pr + r_next[j] * r_prev[j];
        r_prev_2 = r_prev_2 + r_prev[j] * r_prev[j];
        r_next_2 = r_next_2 + r_next[j] * r_next[j];
      }
      const double cos_alpha = dot_pr / sqrt(r_prev_2 * r_next_2);
      double dir_next[3], dir_prev[3];
      for (int j = 0; j < 3; ++j) {
        dir_next[j] = r_prev[j] * r_next_2 - r_next[j] * dot_pr;
        dir_prev[j] = -r_next[j] * r_prev_2 + r_prev[j] * dot_pr;
      }
      for (int j = 0; j < 3; ++j) {
        dir_prev_2 = dir_prev_2 + dir_prev[j] * dir_prev[j];
        dir_next_2 = dir_next_2 + dir_next[j] * dir_next[j];
      }
      const double eps = (double)0.0000001f;
      if (dir_next_2 < eps) {
        for (int j = 0; j < 3; ++j) dir_next[j] = dir_next[j] * 0;
      } else {
        const double nn = sqrt(dir_next_2 * r_next_2);
        for (int j = 0; j < 3; ++j) dir_next[j] = dir_next[j] / nn;
      }
      if (dir_prev_2 < eps) {
        for (int j = 0; j < 3; ++j) dir_prev[j] = dir_prev[j] * 0;
      } else {
        const double np_ = sqrt(dir_prev_2 * r_prev_2);
        for (int j = 0; j < 3; ++j) dir_prev[j] = dir_prev[j] / np_;
      }
      const double delta_alpha = cos_alpha >= 1.0 ? 0. : acos(cos_alpha) - p->alpha_eq;
      const double f_mod = p->k_bend * delta_alpha;
      v_bend = v_bend + .5 * f_mod * delta_alpha;
      for (int j = 0; j < 3; ++j) {
        const double fp = f_mod * dir_prev[j], fn = f_mod * dir_next[j];
        F(b + i - 2, j) = F(b + i - 2, j) + fp;
        F(b + i - 1, j) = F(b + i - 1, j) - fp - fn;
        F(b + i, j) = F(b + i, j) + fn;
      }
    }
  }
  return v_bend;
}

static double ipow(double x, int n) { /* square-and-multiply, lowest bit first */
  double res = 1.0, sq = x;
  int first = 1;
  while (n) {
    if (n & 1) {
      res = first ? sq : res * sq;
      first = 0;
    }
    n >>= 1;
    if (n) sq = sq * sq;
  }
  return res;
}

/* fluid_wall.f90:81-121 (WALL 2).  Pinned by the closed form of the 9-3
 * potential (force = -dV/dz by finite differences, zero of the force). */
void orc_fluid_wall(orc_sys* s, double* out) {
  const orc_params* p = &s->p;
  double v_fw = 0., fzb = 0., fzt = 0.;
  for (int i = 0; i < s->n; ++i) {
    const int t = s->a_type[i] - 1;
    const double sw = p->sigma_wall[t], aw = p->a_wall[t];
    double inv_z = 1. / R(i, 2);
    double r_dummy = sw * inv_z;
    v_fw = v_fw + fabs(aw) * ipow(r_dummy, 9) - p->a_w * ipow(r_dummy, 3);
    double f1 = 9. * fabs(aw) * ipow(sw, 9) * ipow(inv_z, 10);
    double f2 = 3. * aw * ipow(sw, 3) * ipow(inv_z, 4);
    F(i, 2) = F(i, 2) + f1;
    F(i, 2) = F(i, 2) - f2;
    fzb += f1 - f2;
    inv_z = 1. / (p->z_space_wall - R(i, 2));
    r_dummy = sw * inv_z;
    v_fw = v_fw + fabs(aw) * ipow(r_dummy, 9) - p->a_w * ipow(r_dummy, 3);
    f1 = 9. * fabs(aw) * ipow(sw, 9) * ipow(inv_z, 10);
    f2 = 3 * aw * ipow(sw, 3) * ipow(inv_z, 4);
    F(i, 2) = F(i, 2) - f1;
    F(i, 2) = F(i, 2) + f2;
    fzt += -f1 + f2;
  }
  out[0] = v_fw;
  out[1] = fzb;
  out[2] = fzt;
}

static void fene_bond(orc_sys* s, int ip, int jp, double* pot, ksum_t* vir) {
  const orc_params* p = &s->p;
  double dr[3];
  for (int d = 0; d < 3; ++d) dr[d] = R(ip, d) - R(jp, d);
  dr[0] = pbc(dr[0], p->boundary[0], p->inv_boundary[0]);
  dr[1] = pbc(dr[1], p->boundary[1], p->inv_boundary[1]);
  if (p->symmetry == 1) dr[2] = pbc(dr[2], p->boundary[2], p->inv_boundary[2]);
  double r_2 = 0.;
  for (int d = 0; d < 3; ++d) r_2 = r_2 + dr[d] * dr[d];
  double r_dummy = r_2 * p->inv_r_chain_2;
  if (r_dummy > (double)0.98f) r_dummy = (double)0.98f;
  *pot = *pot + (-0.5 * log(1. - r_dummy));
  r_dummy = p->k_chain / (1. - r_dummy) * p->inv_r_chain_2;
  ksum_add(vir, -r_dummy * r_2);
  for (int d = 0; d < 3; ++d) {
    const double fl = r_dummy * dr[d];
    F(ip, d) = F(ip, d) - fl;
    F(jp, d) = F(jp, d) + fl;
  }
}

/* intra_molec.f90:16-63 (brushes), :101-139 and :270 (melt), :280-446
 * (particle-4 chains).  FENE pinned by its closed form and finite
 * differences of the energy. */
void orc_intra_molec(orc_sys* s, double* out) {
  const orc_params* p = &s->p;
  double vb = 0., vd = 0., ve = 0.;
  ksum_t vir = {0., 0.};
  for (int c = 0; c < p->n_chain; ++c)
    for (int m = 0; m < p->n_mon - 1; ++m) {
      const int ip = c * p->n_mon + m;
      fene_bond(s, ip, ip + 1, &vb, &vir);
    }
  vb = p->k_chain * vb;
  for (int c = 0; c < p->n_chain_d; ++c)
    for (int m = 0; m < p->n_mon_d - 1; ++m) {
      const int ip = p->part_init_d + c * p->n_mon_d + m;
      fene_bond(s, ip, ip + 1, &vd, &vir);
    }
  vd = p->k_chain * vd;
  for (int c = 0; c < p->n_chain_e; ++c)
    for (int m = 0; m < p->n_mon_e - 1; ++m) {
      const int ip = p->part_init_e + c * p->n_mon_e + m;
      fene_bond(s, ip, ip + 1, &ve, &vir);
    }
  ve = p->k_chain * ve;
  out[0] = vb + vd + ve;
  out[1] = ksum_value(&vir);
}

/* verlet_velocities.f90:25-36, head pinning :113-134, kinetic energy
 * :155-177.  Pinned by closed forms (constant force, pinned heads). */
double orc_verlet_velocities(orc_sys* s) {
  const orc_params* p = &s->p;
  /* f_on_heads = force(:, 1:part_init_d:n_mon), verlet_velocities.f90:14-19 */
  if (p->symmetry == 0 && (p->system == 0 || p->system == 3))
    for (int c = 0; c < p->n_chain; ++c)
      for (int d = 0; d < 3; ++d) s->f_on_heads[3 * (size_t)c + d] = F(c * p->n_mon, d);
  for (int i = 0; i < s->n; ++i) {
    const double inv_mass = 1. / p->mass_type[s->a_type[i] - 1];
    for (int d = 0; d < 3; ++d) A(i, d) = F(i, d) * inv_mass;
    for (int d = 0; d < 3; ++d) V(i, d) = V(i, d) + 0.5 * p->dt * A(i, d);
  }
  /* with DPD_VV the head block and the energies are left to new_dpd_fd
   * (verlet_velocities.f90:79-230 is #ifndef DPD_VV) */
  if (p->dpd_vv) return 0.;
  if (p->symmetry == 0 && (p->system == 0 || p->system == 3)) {
    for (int c = 0; c < p->n_chain / 2; ++c) {
      const int h = p->n_mon * c;
      V(h, 0) = p->va_spring_twall[0];
      V(h, 1) = p->va_spring_twall[1];
      V(h, 2) = 0.;
      for (int d = 0; d < 3; ++d) F(h, d) = A(h, d) = 0.;
    }
    for (int c = p->n_chain / 2; c < p->n_chain; ++c) {
      const int h = p->n_mon * c;
      V(h, 0) = -p->va_spring_twall[0];
      V(h, 1) = -p->va_spring_twall[1];
      V(h, 2) = 0.;
      for (int d = 0; d < 3; ++d) F(h, d) = A(h, d) = 0.;
    }
  }
  double t_inst = 0.;
  for (int i = 0; i < s->n; ++i) {
    const double m = p->mass_type[s->a_type[i] - 1];
    t_inst = t_inst + m * (V(i, 0) * V(i, 0) + V(i, 1) * V(i, 1) + V(i, 2) * V(i, 2));
  }
  return 0.5 * t_inst;
}

/* new_dpd_fd.f90:18-177 (DPD_VV, DPD_WEIGHT 0, WALL 2): the dissipative
 * force recomputed with the velocities after verlet_velocities over the
 * half list, heads (type 1) excluded, inside range_2 with w_d =
 * (1 - sqrt(r_2/range_2))^2 and images in x and y only (:38-44); swapped
 * into the force (force - force_d_old + force_d_new), then a further half
 * kick v += 0.5 dt F/m and a = F/m (:106-124).  Returns t_fluid = sum m v^2 / 2,
 * summed dimension by dimension (:148-176).  Pinned against the reference's
 * own DPD_VV build (tests/test_ref_pin.py). */
double orc_new_dpd_fd(orc_sys* s) {
  const orc_params* p = &s->p;
  const int ls = s->n_neigh + 1;
  const size_t n3 = 3 * (size_t)s->n;
  for (size_t q = 0; q < n3; ++q) s->force[q] = s->force[q] - s->force_d[q];
  for (size_t q = 0; q < n3; ++q) s->force_d[q] = 0.;
  for (int ip = 0; ip < s->n; ++ip) {
    const int ti = s->a_type[ip] - 1;
    if (ti == 0) continue;
    for (int k = 1; k <= s->ff_list[(size_t)ip * ls]; ++k) {
      const int jp = s->ff_list[(size_t)ip * ls + k] - 1;
      const int tj = s->a_type[jp] - 1;
      if (tj == 0) continue;
      double dr[3];
      for (int d = 0; d < 3; ++d) dr[d] = R(ip, d) - R(jp, d);
      dr[0] = pbc(dr[0], p->boundary[0], p->inv_boundary[0]);
      dr[1] = pbc(dr[1], p->boundary[1], p->inv_boundary[1]);
      const double r_2 = (dr[0] * dr[0] + dr[1] * dr[1]) + dr[2] * dr[2];
      if (!(r_2 < p->range_2[ti][tj])) continue;
      const double wr = 1. - sqrt(r_2 / p->range_2[ti][tj]);
      const double w_d = wr * wr;
      const double rn = sqrt(r_2);
      double rv[3];
      for (int d = 0; d < 3; ++d) rv[d] = dr[d] / rn;
      const double r_dummy = ((V(ip, 0) - V(jp, 0)) * rv[0] + (V(ip, 1) - V(jp, 1)) * rv[1]) +
                             (V(ip, 2) - V(jp, 2)) * rv[2];
      for (int d = 0; d < 3; ++d) {
        const double vec = ((-1. * p->friction1) * w_d * r_dummy) * rv[d];
        s->force_d[3 * (size_t)ip + d] = s->force_d[3 * (size_t)ip + d] + vec;
        s->force_d[3 * (size_t)jp + d] = s->force_d[3 * (size_t)jp + d] - vec;
      }
    }
  }
  for (size_t q = 0; q < n3; ++q) s->force[q] = s->force[q] + s->force_d[q];
  for (int i = 0; i < s->n; ++i) {
    const double inv_mass = 1. / p->mass_type[s->a_type[i] - 1];
    for (int d = 0; d < 3; ++d) V(i, d) = V(i, d) + ((0.5 * p->dt) * F(i, d)) * inv_mass;
  }
  for (int i = 0; i < s->n; ++i) {
    const double inv_mass = 1. / p->mass_type[s->a_type[i] - 1];
    for (int d = 0; d < 3; ++d) A(i, d) = F(i, d) * inv_mass;
  }
  double t_fluid = 0.;
  for (int d = 0; d < 3; ++d)
    for (int i = 0; i < s->n; ++i) {
      const double m = p->mass_type[s->a_type[i] - 1];
      t_fluid = t_fluid + m * (V(i, d) * V(i, d));
    }
  return t_fluid / 2.;
}

/* md_main.f90:69-155 for SYSTEM 2/3, WALL 2, THERMOSTAT 0 or none, BIN_TYPE 0. */
int orc_run_steps(orc_sys* s, int64_t first_step, int n, double* obs) {
  int rc = 0;
  if (s->f_skin) { /* init_config.f90:251-255: bin before the loop */
    rc = orc_binning(s);
    if (rc) return rc;
  }
  const int reb0 = s->n_rebuilds; /* count the rebuilds inside the loop only */
  double e_ff[3] = {0, 0, 0}, e_fw[3] = {0, 0, 0}, e_in[2] = {0, 0}, tk = 0., v_coul = 0.;
  double v_bend = 0.;
  for (int it = 0; it < n; ++it) {
    const int64_t step = first_step + it;
    orc_verlet_positions(s);
    if (orc_check_skin(s)) {
      rc = orc_binning(s);
      if (rc) return rc;
    }
    orc_zero_force(s);
    orc_fluid_fluid(s, step, e_ff);
    if (s->p.symmetry == 0) orc_fluid_wall(s, e_fw);
    orc_intra_molec(s, e_in);
    v_bend = orc_bending(s); /* md_main.f90:110-111 */
    v_coul = e_ff[2];
    if (s->p.coulomb) { /* md_main.f90:128-132 */
      v_coul = v_coul + orc_ewald_k(s);
      if (s->p.symmetry == 0) v_coul = v_coul + orc_dipolar(s);
    }
    tk = orc_verlet_velocities(s);
    if (s->p.dpd_vv) tk = orc_new_dpd_fd(s); /* md_main.f90:149-155 */
  }
  if (obs) {
    obs[0] = e_ff[0];
    obs[1] = e_fw[0];
    obs[2] = e_in[0];
    obs[3] = tk;
    obs[4] = e_ff[1] + e_in[1];
    obs[5] = (double)(s->n_rebuilds - reb0);
    obs[6] = e_fw[1];
    obs[7] = v_coul;
    obs[8] = v_bend;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Observation stage: init_obser.f90:15-91 (the sums), histo mode 0
 * (functions.f90:106-129) with dim_prof_dens = hist_dim. */
void orc_obs_init(orc_sys* s, int hist_dim) {
  memset(&s->obs, 0, sizeof s->obs);
  s->obs.hist_dim = hist_dim;
  free(s->dens_z);
  s->dens_z = hist_dim > 0 ? (double*)calloc(3 * (size_t)hist_dim, sizeof(double)) : NULL;
}

/* observation.f90:17-307 with PROFILES for SYSTEM 2/3, WALL 2 (no wall
 * atoms), after a step whose verlet_velocities gave e = (v_fluid_fluid,
 * v_fluid_wall, v_intra_molec, t_fluid, v_coul):
 *  - the energy totals of verlet_velocities.f90:183-196 (v_total gains
 *    v_coul only on its tenth calls, :199-222: coul_in_total) and the sums
 *    of observation.f90:129-159;
 *  - brushes (SYSTEM 3): the chains unfolded along their bonds, one image
 *    correction per monomer and axis (:163-199), Rg^2 per chain about its
 *    centre with inv_n_mon = 1/n_mon (init_obser.f90:121) for the brush and
 *    1/n_mon_d for the melt chains, the means r_g2 and their sums
 *    (:205-276);
 *  - the total momentum sum m v (:286-291);
 *  - histo mode 2 (functions.f90:143-206): slab counts of z for the brush,
 *    melt and particle-4 monomers, index int(z / z_step) clamped to the
 *    slabs.
 * Pinned against the reference's own observation stage (tests/test_ref_pin.py,
 * the files fort.40-43 and dens_prof.mide of the PROFILES build). */
void orc_observe(orc_sys* s, const double* e, int coul_in_total) {
  const orc_params* p = &s->p;
  mfa_obs_sums* o = &s->obs;
  const double v_ff = e[0], v_fw = e[1], v_in = e[2], t_fluid = e[3], v_coul = e[4];
  o->count_obs = o->count_obs + 1;
  double v_total = ((0. + v_fw) + v_ff) + v_in;
  const double t_total = 0. + t_fluid;
  const double e_total = v_total + t_total;
  if (coul_in_total) v_total = v_total + v_coul;
  const double vf = (v_ff + v_in) + v_fw;
  o->v_fluid_fluid_1 = o->v_fluid_fluid_1 + vf;
  o->t_fluid_1 = o->t_fluid_1 + t_fluid;
  o->e_fluid_fluid_1 = o->e_fluid_fluid_1 + (v_ff + t_fluid);
  o->v_fluid_fluid_2 = o->v_fluid_fluid_2 + vf * vf;
  o->t_fluid_2 = o->t_fluid_2 + t_fluid * t_fluid;
  o->e_fluid_fluid_2 = o->e_fluid_fluid_2 + (v_ff + t_fluid) * (v_ff + t_fluid);
  o->v_total_1 = o->v_total_1 + v_total;
  o->t_total_1 = o->t_total_1 + t_total;
  o->e_total_1 = o->e_total_1 + e_total;
  o->v_total_2 = o->v_total_2 + v_total * v_total;
  o->t_total_2 = o->t_total_2 + t_total * t_total;
  o->e_total_2 = o->e_total_2 + e_total * e_total;
  o->v_tot_intra_1 = o->v_tot_intra_1 + v_in;
  o->v_tot_intra_2 = o->v_tot_intra_2 + v_in * v_in;
  if (p->system == 3) {
    const size_t n = (size_t)s->n;
    double* sp = (double*)malloc(3 * n * sizeof(double));
    memcpy(sp, s->r0, 3 * n * sizeof(double));
    double half[3];
    for (int d = 0; d < 3; ++d) half[d] = p->boundary[d] / 2.; /* init_config.f90:22 */
#define SP(i, d) sp[3 * (size_t)(i) + (d)]
    if (p->symmetry == 0)
      for (int i = 0; i < p->part_init_d; ++i)
        for (int d = 0; d < 3; ++d) {
          if (i % p->n_mon == 0) continue; /* mod(i_part-1, n_mon) */
          if (SP(i, d) - SP(i - 1, d) > half[d]) SP(i, d) = SP(i, d) - p->boundary[d];
          else if (SP(i - 1, d) - SP(i, d) > half[d]) SP(i, d) = SP(i, d) + p->boundary[d];
        }
    for (int i = p->part_init_d; i < p->part_init_e; ++i)
      for (int d = 0; d < 3; ++d) {
        if (i % p->n_mon_d == 0) continue; /* mod(i_part-1, n_mon_d), the global index */
        if (SP(i, d) - SP(i - 1, d) > half[d]) SP(i, d) = SP(i, d) - p->boundary[d];
        else if (SP(i - 1, d) - SP(i, d) > half[d]) SP(i, d) = SP(i, d) + p->boundary[d];
      }
    double r_g2[2][3] = {{0., 0., 0.}, {0., 0., 0.}};
    const double inv_n_mon = 1. / (double)p->n_mon, inv_n_chain = 1. / (double)p->n_chain;
    for (int c = 0; c < p->n_chain; ++c) {
      double cm[3] = {0., 0., 0.}, vd[3] = {0., 0., 0.};
      for (int m = 0; m < p->n_mon; ++m)
        for (int d = 0; d < 3; ++d) cm[d] = cm[d] + SP(c * p->n_mon + m, d);
      for (int d = 0; d < 3; ++d) cm[d] = cm[d] * inv_n_mon;
      for (int m = 0; m < p->n_mon; ++m)
        for (int d = 0; d < 3; ++d) {
          const double x = SP(c * p->n_mon + m, d) - cm[d];
          vd[d] = vd[d] + x * x;
        }
      for (int d = 0; d < 3; ++d) r_g2[0][d] = r_g2[0][d] + vd[d] * inv_n_mon;
    }
    for (int d = 0; d < 3; ++d) r_g2[0][d] = r_g2[0][d] * inv_n_chain;
    for (int c = 0; c < p->n_chain_d; ++c) {
      const int b = p->part_init_d + c * p->n_mon_d;
      double cm[3] = {0., 0., 0.}, vd[3] = {0., 0., 0.};
      for (int m = 0; m < p->n_mon_d; ++m)
        for (int d = 0; d < 3; ++d) cm[d] = cm[d] + SP(b + m, d);
      for (int d = 0; d < 3; ++d) cm[d] = cm[d] / (double)p->n_mon_d;
      for (int m = 0; m < p->n_mon_d; ++m)
        for (int d = 0; d < 3; ++d) {
          const double x = SP(b + m, d) - cm[d];
          vd[d] = vd[d] + x * x;
        }
      for (int d = 0; d < 3; ++d) r_g2[1][d] = r_g2[1][d] + vd[d] / (double)p->n_mon_d;
    }
    for (int d = 0; d < 3; ++d) r_g2[1][d] = r_g2[1][d] / (double)p->n_chain_d; /* 0/0 without melt chains, as in the reference */
#undef SP
    free(sp);
    for (int d = 0; d < 3; ++d) {
      o->r_g2_mean[0][d] = o->r_g2_mean[0][d] + r_g2[0][d];
      o->r_g2_mean[1][d] = o->r_g2_mean[1][d] + r_g2[1][d];
      o->r_g2_mean[2][d] = o->r_g2_mean[2][d] + (r_g2[0][d] + r_g2[1][d]) / 2.0;
    }
    for (int g = 0; g < 2; ++g) {
      const double sum = (r_g2[g][0] + r_g2[g][1]) + r_g2[g][2];
      o->r_par_per2[g] = o->r_par_per2[g] + (r_g2[g][2] - 0.5 * (r_g2[g][0] + r_g2[g][1])) / sum;
    }
  }
  double pv[3] = {0., 0., 0.};
  const int n_mon_tot = p->part_init_e + p->n_mon_e * p->n_chain_e;
  for (int i = 0; i < n_mon_tot; ++i) {
    const double m = p->mass_type[s->a_type[i] - 1];
    for (int d = 0; d < 3; ++d) pv[d] = pv[d] + m * V(i, d);
  }
  for (int d = 0; d < 3; ++d) o->tot_p[d] = o->tot_p[d] + pv[d];
  if (o->hist_dim > 0) {
    const int hd = o->hist_dim;
    o->count_hist = o->count_hist + 1;
    const double z_step = p->boundary[2] / (double)hd, inv_z_step = 1. / z_step;
    for (int i = 0; i < n_mon_tot; ++i) {
      const int g = i < p->part_init_d ? 0 : (i < p->part_init_e ? 1 : 2);
      int zi = (int)(R(i, 2) * inv_z_step) + 1;
      if (zi < 1) zi = 1;
      if (zi > hd) zi = hd;
      s->dens_z[(size_t)g * hd + (zi - 1)] = s->dens_z[(size_t)g * hd + (zi - 1)] + 1.;
    }
  }
}

void orc_obs_read(orc_sys* s, mfa_obs_sums* out, double* dens_z, double* f_on_heads) {
  if (out) *out = s->obs;
  if (dens_z && s->dens_z) memcpy(dens_z, s->dens_z, 3 * (size_t)s->obs.hist_dim * sizeof(double));
  if (f_on_heads) memcpy(f_on_heads, s->f_on_heads, 3 * (size_t)s->p.n_chain * sizeof(double));
}
