// Host-side part of libmfa_hip: derived constants and seeded input generators.
//
// mfa_derive restates the constant derivation of init_params.f90:395-845,
// the box of init_config.f90:11-22 and the grid of make_binning_boxes.f90:14-64
// for the switch combinations the run-time options cover.  The reference is
// compiled without -r8, so its default-real literals are float32 values that
// are then promoted; they appear below as float constants (2**(1/6), 2.2, 0.4,
// 0.7, 0.98) with the float arithmetic the reference performs on them.
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <stdlib.h>
#include <vector>

#include "../../include/mfa_hip.h"
#include "mfa_rng.h"

namespace {

// 2.**(1./6.) evaluated in default real (init_params.f90:502).
const float kTwoSixth = 1.12246203422546386718750f;  // 0x3f8facd6
const float kPiF = 3.1415926536f;                     // mfa_common.f90:332

}  // namespace

// Message slot shared with mfa_last_error(NULL) (defined with the device code).
extern "C" void mfa_internal_set_noctx_error(const char* msg);

static int derive_fail(const char* msg) {
  mfa_internal_set_noctx_error(msg);
  return MFA_ERR_ARG;
}

extern "C" int32_t mfa_abi_sizes(int32_t* si, int32_t* sp, int32_t* so) {
  if (si) *si = (int32_t)sizeof(mfa_inputs);
  if (sp) *sp = (int32_t)sizeof(mfa_params);
  if (so) *so = (int32_t)sizeof(mfa_obs);
  return MFA_OK;
}

// Skin (init_params.f90:800-808), check_skin's threshold (check_skin.f90:18)
// and the Verlet-list criterion (binning.f90:101-105), from p's ranges and
// r_cut_max; rc2 > 0: a real-space Coulomb cut-off of its own widens the list.
static void set_skin_fields(mfa_params* p, double skin, double rc2) {
  p->skin = skin;
  p->skin_2 = (p->r_cut_max + p->skin) * (p->r_cut_max + p->skin) - p->r_cut_max * p->r_cut_max;
  p->skin_half_2 = (p->skin / 2.0) * (p->skin / 2.0);
  for (int i = 0; i < MFA_NTYPE; ++i)
    for (int j = 0; j < MFA_NTYPE; ++j)
      p->list_2[i][j] = p->dpd_cut_off_on ? p->r_cut_max_2 + p->skin_2 : p->range_2[i][j] + p->skin_2;
  if (rc2 > 0.0) {
    const double lc = rc2 + p->skin_2;
    for (int i = 0; i < MFA_NTYPE; ++i)
      for (int j = 0; j < MFA_NTYPE; ++j) p->list_2[i][j] = p->list_2[i][j] > lc ? p->list_2[i][j] : lc;
  }
}

extern "C" int32_t mfa_params_set_skin(mfa_params* p, double skin) {
  if (!p || !(skin > 0.0)) return derive_fail("the list skin must be positive");
  mfa_params q = *p;
  set_skin_fields(&q, skin, q.coulomb && q.rc_coul_2 > 0.0 ? q.rc_coul_2 : 0.0);
  for (int d = 0; d < 3; ++d)
    for (int i = 0; i < MFA_NTYPE; ++i)
      for (int j = 0; j < MFA_NTYPE; ++j)
        if (q.list_2[i][j] > q.r_bin[d] * q.r_bin[d])
          return derive_fail("binning cell narrower than the list radius");
  *p = q;
  return MFA_OK;
}

extern "C" int32_t mfa_params_set_cells(mfa_params* p, double width) {
  if (!p || !(width > 0.0)) return derive_fail("the cell width must be positive");
  mfa_params q = *p;
  for (int d = 0; d < 3; ++d) {
    q.n_bin[d] = (int)(q.boundary[d] / width);
    if (q.n_bin[d] < 3) return derive_fail("fewer than 3 binning cells along an axis");
    q.r_bin[d] = q.boundary[d] / (double)q.n_bin[d];
    q.inv_r_bin[d] = 1.0 / q.r_bin[d];
    for (int i = 0; i < MFA_NTYPE; ++i)
      for (int j = 0; j < MFA_NTYPE; ++j)
        if (q.list_2[i][j] > q.r_bin[d] * q.r_bin[d])
          return derive_fail("binning cell narrower than the list radius");
  }
  *p = q;
  return MFA_OK;
}

extern "C" int32_t mfa_derive(const mfa_inputs* in, mfa_params* p) {
  if (!in || !p) return MFA_ERR_ARG;
  memset(p, 0, sizeof(*p));
  const int nt = MFA_NTYPE;
  if (in->system != 2 && in->system != 3) return derive_fail("SYSTEM must be 2 or 3");
  if (in->symmetry != 0 && in->symmetry != 1) return derive_fail("SYMMETRY must be 0 or 1");
  if (in->solvent != 0 && in->solvent != 1) return derive_fail("SOLVENT must be 0 or 1");
  if (in->thermostat != 0 && in->thermostat != 1 && in->thermostat != -1)
    return derive_fail("THERMOSTAT must be 0 (DPD), 1 (Langevin) or -1 (none)");
  if (in->dt <= 0.0 || in->r_chain <= 0.0) return derive_fail("dt and r_chain must be positive");
  // SYSTEM 3 needs an even number of brush chains (init_params.f90:878-884).
  if (in->system == 3 && (in->n_chain % 2) != 0)
    return derive_fail("n_chain is not divisible by 2");

  if (in->dpd_vv && in->thermostat != 0) return derive_fail("DPD_VV needs the DPD thermostat");
  p->dpd_vv = in->dpd_vv ? 1 : 0;
  p->system = in->system;
  p->symmetry = in->symmetry;
  p->thermostat = in->thermostat;
  p->bending = in->bending;
  p->ewald_compat = in->ewald_compat;
  p->n_mon = in->n_mon;
  p->n_chain = in->n_chain;
  p->n_mon_d = in->n_mon_d;
  p->n_chain_d = in->n_chain_d;
  p->n_mon_e = in->particle_4 ? in->n_mon_e : 0;
  p->n_chain_e = in->particle_4 ? in->n_chain_e : 0;
  // init_system.f90:199-217 (WALL 2: no wall particles).
  p->n_mon_tot = p->n_mon * p->n_chain + p->n_mon_d * p->n_chain_d + p->n_mon_e * p->n_chain_e;
  p->n_part = p->n_mon_tot;
  // init_params.f90:871-872 (0-based here).
  p->part_init_d = p->n_mon * p->n_chain;
  p->part_init_e = p->part_init_d + p->n_mon_d * p->n_chain_d;
  if (p->n_part <= 0) return derive_fail("no particles");

  // Box: init_config.f90:11-22.
  p->boundary[0] = in->x_space * (double)in->n_cell_w_x;
  p->boundary[1] = in->y_space * (double)in->n_cell_w_y;
  p->boundary[2] = in->z_space_wall;
  for (int d = 0; d < 3; ++d) p->inv_boundary[d] = 1.0 / p->boundary[d];

  // init_params.f90:395-398.
  p->dt = in->dt;
  p->dt_2 = in->dt * in->dt;
  p->temp = in->temp;
  p->r_chain_2 = in->r_chain * in->r_chain;
  p->inv_r_chain_2 = 1.0 / p->r_chain_2;
  p->k_chain = in->k_chain * p->r_chain_2;

  for (int t = 0; t < nt; ++t) {
    p->mass_type[t] = in->mass_type[t];
    p->inv_mass_type[t] = 1.0 / in->mass_type[t];
  }

  // Sum rules, SOLVENT 0/1 (init_params.f90:407-420).
  double sigma[MFA_NTYPE][MFA_NTYPE];
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nt; ++j) sigma[i][j] = 0.0;
  for (int i = 0; i < nt; ++i) sigma[i][i] = in->sigma_ii[i];
  for (int i = 0; i < nt - 1; ++i)
    for (int j = i + 1; j < nt; ++j) {
      sigma[i][j] = (sigma[i][i] + sigma[j][j]) / 2.0;
      sigma[j][i] = sigma[i][j];
    }
  if (!in->particle_4) {
    for (int i = 0; i < nt; ++i) {
      sigma[i][nt - 1] = 0.0;
      sigma[nt - 1][i] = 0.0;
    }
  }
  // Factor 4 into epsilon, sigma_2 (init_params.f90:467-472).
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nt; ++j) {
      p->sigma[i][j] = sigma[i][j];
      p->sigma_2[i][j] = sigma[i][j] * sigma[i][j];
      p->epsil[i][j] = 4.0 * in->epsil[i][j];
    }

  // Cut-offs (init_params.f90:502-712).
  double range_2[MFA_NTYPE][MFA_NTYPE];
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nt; ++j) {
      const double r = (double)kTwoSixth * sigma[i][j];
      range_2[i][j] = r * r;
    }
  {
    const double r = (double)2.2f * sigma[0][nt - 1];
    range_2[0][nt - 1] = r * r;
    const double r2 = (double)2.2f * sigma[nt - 1][0];
    range_2[nt - 1][0] = r2 * r2;
  }
  if (in->solvent == 0) {  // poor solvent: 2*2^(1/6) sigma (init_params.f90:551)
    const float two_c = 2.0f * kTwoSixth;
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j) {
        const double r = (double)two_c * sigma[i][j];
        range_2[i][j] = r * r;
      }
    if (!in->particle_4)
      for (int i = 0; i < nt; ++i) range_2[i][nt - 1] = range_2[nt - 1][i] = 0.0;
  }
  // WALL 2 wall parameters (init_params.f90:581-595).
  p->a_w = in->a_w;
  for (int t = 0; t < nt; ++t) {
    p->sigma_wall[t] = (in->sigma_w + sigma[t][t]) / 2.0;
    p->a_wall[t] = in->a_w;
  }
  if (in->particle_4) {
    p->sigma_wall[3] = (in->sigma_w4 + sigma[3][3]) / 2.0;
    p->a_wall[3] = in->a_w4;
  }
  p->z_space_wall = in->z_space_wall;
  if (!in->particle_4) {
    for (int i = 0; i < nt; ++i) range_2[i][nt - 1] = range_2[nt - 1][i] = 0.0;  // :604-607
  } else {
    // inter4 menu (init_params.f90:629-676), WALL 2.
    const double c = (double)kTwoSixth;
    const double c2 = (double)(2.0f * kTwoSixth);
    switch (in->inter4) {
      case 1:
        for (int i = 0; i < 3; ++i) {
          double r = c * sigma[i][3];
          range_2[i][3] = r * r;
          r = c * sigma[3][i];
          range_2[3][i] = r * r;
        }
        {
          const double r = c * sigma[3][3];
          range_2[3][3] = r * r;
        }
        break;
      case 2:
      case 3:
        for (int i = 0; i < 3; ++i) {
          double r = c * sigma[i][3];
          range_2[i][3] = r * r;
          r = c * sigma[3][i];
          range_2[3][i] = r * r;
        }
        {
          const double r = 2.0 * (c * sigma[3][3]);
          range_2[3][3] = r * r;
        }
        break;
      case 4:
        for (int i = 0; i < 3; ++i) {
          double r = c2 * sigma[i][3];
          range_2[i][3] = r * r;
          r = c2 * sigma[3][i];
          range_2[3][i] = r * r;
        }
        {
          const double r = 2.0 * (c * sigma[3][3]);
          range_2[3][3] = r * r;
        }
        break;
      default:
        return derive_fail("inter4 must be 1..4");
    }
  }
  // Explicit cut-off override (survey patch P3: the reference only offers the
  // menu above; C2 needs 2.5 sigma).  Applies to every pair with a non-zero range.
  if (in->lj_rc > 0.0) {
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j)
        if (range_2[i][j] > 0.0) {
          const double r = in->lj_rc * sigma[i][j];
          range_2[i][j] = r * r;
        }
  }
  // Brush heads do not interact with each other (init_params.f90:703-712).
  if (in->system == 3) range_2[0][0] = 0.0;
  // Energy shifts (init_params.f90:713-724).
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nt; ++j) {
      p->range_2[i][j] = range_2[i][j];
      if (range_2[j][i] > 0.0) {
        const double x = (sigma[j][i] * sigma[j][i]) / range_2[j][i];
        const double x3 = x * x * x;
        p->e_shift[j][i] = x3 * (x3 - 1.0);
      } else {
        p->e_shift[j][i] = 0.0;
      }
    }
  // r_cut_max and DPD cut-off (init_params.f90:766-781).
  double mx = 0.0;
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nt; ++j) mx = range_2[i][j] > mx ? range_2[i][j] : mx;
  p->r_cut_max = sqrt(mx);
  if (in->dpd_cut_off > 0.0) {
    p->dpd_cut_off_on = 1;
    p->r_cut_dpd = (double)(float)in->dpd_cut_off;  // a default-real macro literal
    p->r_cut_max = p->r_cut_dpd > p->r_cut_max ? p->r_cut_dpd : p->r_cut_max;
    p->r_cut_dpd_2 = p->r_cut_dpd * p->r_cut_dpd;
    p->r_cut_max_2 = p->r_cut_max * p->r_cut_max;
  } else {
    p->r_cut_max_2 = p->r_cut_max * p->r_cut_max;
  }
  // A real-space Coulomb cut-off beyond the LJ range (outside the compat
  // form) widens the list radius and the cells.
  int charged_any = 0;
  for (int t = 0; t < nt; ++t) charged_any |= in->charge[t] != 0.0;
  if (charged_any && !in->ewald_compat && in->rc_coul > p->r_cut_max) {
    p->r_cut_max = in->rc_coul;
    p->r_cut_max_2 = p->r_cut_max * p->r_cut_max;
  }
  // inv_range_2 (init_params.f90:786-795).
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nt; ++j)
      p->inv_range_2[i][j] = range_2[i][j] != 0.0 ? 1.0 / range_2[i][j] : 0.0;
  set_skin_fields(p, in->skin > 0.0 ? in->skin : (double)0.4f,
                  charged_any && !in->ewald_compat && in->rc_coul > 0.0 ? in->rc_coul * in->rc_coul
                                                                         : 0.0);
  // DPD (init_params.f90:845).
  p->friction1 = in->friction[0];
  p->sig = sqrt(2.0 * in->temp * in->friction[0] / in->dt);
  p->dpd_weight = 0;
  p->seed = (uint64_t)in->iseed;
  p->va_spring_twall[0] = in->va_spring_twall[0];
  p->va_spring_twall[1] = in->va_spring_twall[1];
  p->va_spring_twall[2] = in->va_spring_twall[2];
  p->k_bend = in->k_bend;
  p->alpha_eq = in->alpha_eq;

  // Cell grid (make_binning_boxes.f90:14-64).  delta_bin comes from the
  // provisional grid of init_system.f90:273-275 (float arithmetic).
  int prov[3];
  prov[0] = (int)(0.4f * (float)in->n_cell_w_x);
  prov[1] = (int)(0.7f * (float)in->n_cell_w_y);
  prov[2] = 20;
  for (int d = 0; d < 3; ++d) p->delta_bin[d] = prov[d] >= 3 ? 1 : (prov[d] == 2 ? 0 : -1);
  const double bbs = (double)0.4f;  // binning_box_skin
  for (int d = 0; d < 3; ++d) {
    p->n_bin[d] = (int)(p->boundary[d] / (p->r_cut_max + bbs));
    if (p->n_bin[d] < 1) return derive_fail("box smaller than one binning cell");
    p->r_bin[d] = p->boundary[d] / (double)p->n_bin[d];
    p->inv_r_bin[d] = 1.0 / p->r_bin[d];
  }
  // The 27-cell scan finds every listed pair only with >= 3 cells per axis,
  // a one-cell reach, and cells at least as wide as the list radius
  // (make_binning_boxes.f90:91-100 only warns); refuse other grids.
  for (int d = 0; d < 3; ++d) {
    if (p->n_bin[d] < 3) return derive_fail("fewer than 3 binning cells along an axis");
    if (p->delta_bin[d] != 1)
      return derive_fail("provisional grid int(0.4*n_cell_w_x), int(0.7*n_cell_w_y) below 3 "
                         "(the 27-cell scan would miss pairs); use a finer x_space/y_space");
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j)
        if (p->list_2[i][j] > p->r_bin[d] * p->r_bin[d])
          return derive_fail("binning cell narrower than the list radius");
  }

  // Charges and Ewald constants (init_config.f90:221-239).
  int charged = 0;
  for (int t = 0; t < nt; ++t) {
    p->charge[t] = in->charge[t];
    if (in->charge[t] != 0.0) charged = 1;
  }
  p->coulomb = charged;
  p->alpha_c = in->alpha_c;
  p->sqrt_alpha = sqrt(in->alpha_c);
  // The reference sets inv_sqrt_pi = 1./pi (init_config.f90:230); the compat
  // form keeps that factor, the default uses 1/sqrt(pi).
  p->inv_sqrt_pi_real = in->ewald_compat ? 1.0 / (double)kPiF : 1.0 / sqrt(M_PI);
  if (in->rc_coul > 0.0 && !in->ewald_compat)
    p->rc_coul_2 = in->rc_coul * in->rc_coul;
  else
    p->rc_coul_2 = 0.0;  // evaluated inside the LJ range (fluid_fluid.f90:94,166)
  p->n_kmax = in->n_kmax;
  // FORCE_SWITCH_ON on a fresh start (init_config.f90:264-281,
  // init_force_switch_on.f90:15-26; SYSTEM != 4).
  if (in->force_switch_on && in->s_time == 1) {
    double maxsig = 0.0;
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j) maxsig = sigma[i][j] > maxsig ? sigma[i][j] : maxsig;
    const double x = maxsig / (double)0.9f;
    p->fso_r2_init = x * x;
    const double t_relax = (double)in->n_relax * in->dt;
    p->fso_time = t_relax < 40.0 ? 40.0 : t_relax;
  }
  return MFA_OK;
}

// ---------------------------------------------------------------------------
// Seeded generators.  Uniform doubles come from Threefry-2x32 (20 rounds,
// independent of the rounds of the force pass's deviates, so the seeded
// inputs do not change with them) on a running counter (53 bits from two
// words), so the configurations are the same on every machine.
namespace {
#define GEN_TF_ROUNDS 20

struct Stream {
  uint32_t k0, k1;
  uint64_t ctr;
  double uniform() {
    uint32_t a, b;
    mfa_threefry2x32((uint32_t)ctr, (uint32_t)(ctr >> 32), k0, k1, GEN_TF_ROUNDS, &a, &b);
    ++ctr;
    const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);  // 53 bits
    return (double)m * (1.0 / 9007199254740992.0);
  }
  double normal() {
    double u1;
    do u1 = uniform();
    while (u1 <= 0.0);
    const double u2 = uniform();
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
  }
  void unit_vector(double* e) {
    const double z = 2.0 * uniform() - 1.0;
    const double phi = 2.0 * M_PI * uniform();
    const double s = sqrt(1.0 - z * z);
    e[0] = s * cos(phi);
    e[1] = s * sin(phi);
    e[2] = z;
  }
};

double wrap(double x, double L) {
  x = fmod(x, L);
  if (x < 0.0) x += L;
  if (x >= L) x -= L;
  return x;
}

// Overlap grid: periodic in x,y (and z when bulk).
struct Grid {
  int n[3];
  double L[3], inv_cell[3];
  bool pz;
  std::vector<std::vector<int>> cells;
  Grid(const double* box, double cell, bool periodic_z) {
    pz = periodic_z;
    for (int d = 0; d < 3; ++d) {
      L[d] = box[d];
      n[d] = (int)(box[d] / cell);
      if (n[d] < 1) n[d] = 1;
      inv_cell[d] = (double)n[d] / box[d];
    }
    cells.resize((size_t)n[0] * n[1] * n[2]);
  }
  int idx(const double* r, int* c) const {
    for (int d = 0; d < 3; ++d) {
      int k = (int)(r[d] * inv_cell[d]);
      if (k < 0) k = 0;
      if (k >= n[d]) k = n[d] - 1;
      c[d] = k;
    }
    return (c[2] * n[1] + c[1]) * n[0] + c[0];
  }
  bool clear(const std::vector<double>& pos, const double* r, double dmin) const {
    int c[3];
    idx(r, c);
    const double d2min = dmin * dmin;
    for (int dz = -1; dz <= 1; ++dz) {
      int z = c[2] + dz;
      if (pz) z = (z + n[2]) % n[2];
      else if (z < 0 || z >= n[2]) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        const int y = (c[1] + dy + n[1]) % n[1];
        for (int dx = -1; dx <= 1; ++dx) {
          const int x = (c[0] + dx + n[0]) % n[0];
          for (int j : cells[(size_t)(z * n[1] + y) * n[0] + x]) {
            double s = 0.0;
            for (int d = 0; d < 3; ++d) {
              double q = r[d] - pos[3 * (size_t)j + d];
              if (d < 2 || pz) q -= L[d] * nearbyint(q / L[d]);
              s += q * q;
            }
            if (s < d2min) return false;
          }
        }
      }
    }
    return true;
  }
  void add(const double* r, int id) {
    int c[3];
    cells[(size_t)idx(r, c)].push_back(id);
  }
  void remove(const double* r, int id) {
    int c[3];
    auto& v = cells[(size_t)idx(r, c)];
    for (size_t k = 0; k < v.size(); ++k)
      if (v[k] == id) {
        v.erase(v.begin() + (long)k);
        return;
      }
  }
};

}  // namespace

extern "C" int32_t mfa_generate(const mfa_params* p, int32_t kind, uint64_t seed, double head_z,
                                double bond, double min_dist, double* r0, double* v,
                                int32_t* a_type) {
  if (!p || !r0 || !v || !a_type) return MFA_ERR_ARG;
  const int n = p->n_part;
  Stream s{(uint32_t)seed, (uint32_t)(seed >> 32) ^ 0x6d666121u, 0};
  const double* L = p->boundary;
  const bool walls = p->symmetry == 0;
  // beads stay where the 9-3 wall force is moderate (about 40 at z = 0.95)
  const double zlo = 0.95, zhi = L[2] - 0.95;
  std::vector<double> pos((size_t)3 * n, 0.0);
  Grid grid(L, min_dist > 0.0 ? (min_dist > 1.0 ? min_dist : 1.0) : 1e30, !walls);
  const bool use_grid = min_dist > 0.0;
  bool check_overlap = true;  // off for the ideal solvent of kind 1
  auto place_ok = [&](const double* r, double dmin) {
    if (walls && (r[2] <= zlo || r[2] >= zhi)) return false;
    return !use_grid || !check_overlap || grid.clear(pos, r, dmin);
  };
  auto put = [&](int i, const double* r, int type) {
    for (int d = 0; d < 3; ++d) pos[3 * (size_t)i + d] = r[d];
    a_type[i] = type;
    if (use_grid && check_overlap) grid.add(r, i);
  };
  // One chain as a random walk from a head; returns false if it got stuck.
  // Each bead gets up to 4000 tries; brush beads first step into the upper
  // hemisphere (tries < 2000).  The minimum distance is never relaxed.
  auto grow = [&](int first, int len, const double* head, int head_type, int mon_type,
                  bool upward) {
    if (head) put(first, head, head_type);
    for (int m = 1; m < len; ++m) {
      const double* prev = &pos[3 * (size_t)(first + m - 1)];
      bool ok = false;
      for (int tries = 0; tries < 4000 && !ok; ++tries) {
        double e[3], r[3];
        s.unit_vector(e);
        if (upward && tries < 2000 && e[2] < 0.0) e[2] = -e[2];
        for (int d = 0; d < 3; ++d) r[d] = prev[d] + bond * e[d];
        r[0] = wrap(r[0], L[0]);
        r[1] = wrap(r[1], L[1]);
        if (!walls) r[2] = wrap(r[2], L[2]);
        if (place_ok(r, min_dist)) {
          put(first + m, r, mon_type);
          ok = true;
        }
      }
      if (!ok) {
        for (int k = head ? first : first + 1; k < first + m; ++k)
          if (use_grid && check_overlap) grid.remove(&pos[3 * (size_t)k], k);
        return false;
      }
    }
    return true;
  };

  // Polyelectrolyte brush (kind 3): every third bead of a chain (1-based
  // positions 3, 6, ...) is a charged monomer of type 4; the single-bead
  // species holds first one counterion (type 3) per charged monomer, then
  // neutral solvent (type 2).
  int n_counterions = 0;
  if (kind == 3) {
    for (int m = 1; m < p->n_mon; ++m) n_counterions += ((m + 1) % 3 == 0) ? 1 : 0;
    n_counterions *= p->n_chain;
    if (p->n_mon_d != 1 || p->n_chain_d < n_counterions) return MFA_ERR_ARG;
  }
  if (kind == 1 || kind == 3) {
    // Brush grafted at z = head_z with heads uniform in x,y (gen_brush.f90:9-30
    // style), then solvent/melt as single beads uniform in the channel.
    if (!walls) return MFA_ERR_ARG;
    // Heads first: they do not interact with each other (range_2(1,1) = 0,
    // init_params.f90:703-712).  Then the chains grow together, one bead of
    // every chain per round, so no chain is walled in by earlier ones; a bead
    // that finds no room relaxes the distance by 0.05 per 1000 tries (to
    // min_dist - 0.15), which the force switch-on ramp then resolves.
    for (int c = 0; c < p->n_chain; ++c) {
      double h[3] = {s.uniform() * L[0], s.uniform() * L[1], head_z};
      put(c * p->n_mon, h, 1);
    }
    for (int m = 1; m < p->n_mon; ++m)
      for (int c = 0; c < p->n_chain; ++c) {
        const double* prev = &pos[3 * (size_t)(c * p->n_mon + m - 1)];
        bool ok = false;
        for (int tries = 0; tries < 4000 && !ok; ++tries) {
          double e[3], r[3];
          s.unit_vector(e);
          // upward steps while the chain is in the lower half of the channel, so
          // long chains do not pile up against the top wall
          if (tries < 2000 && e[2] < 0.0 && prev[2] < 0.5 * L[2]) e[2] = -e[2];
          for (int d = 0; d < 3; ++d) r[d] = prev[d] + bond * e[d];
          r[0] = wrap(r[0], L[0]);
          r[1] = wrap(r[1], L[1]);
          if (place_ok(r, min_dist - 0.05 * (double)(tries / 1000))) {
            put(c * p->n_mon + m, r, (kind == 3 && (m + 1) % 3 == 0) ? 4 : 2);
            ok = true;
          }
        }
        if (!ok) return MFA_ERR_ARG;
      }
    // single-bead solvent is ideal (epsilon 0); counterions keep their distance
    check_overlap = p->n_mon_d > 1 || kind == 3;
    for (int c = 0; c < p->n_chain_d; ++c) {
      const int first = p->part_init_d + c * p->n_mon_d;
      int attempt = 0;
      for (;; ++attempt) {
        if (attempt > 100000) return MFA_ERR_ARG;
        double h[3] = {s.uniform() * L[0], s.uniform() * L[1], zlo + s.uniform() * (zhi - zlo)};
        if (!place_ok(h, min_dist)) continue;
        const int t = kind == 3 ? (c < n_counterions ? 3 : 2) : 3;
        if (grow(first, p->n_mon_d, h, t, t, false)) break;
      }
    }
  } else if (kind == 2) {
    // Melt of n_chain_d chains of n_mon_d beads, periodic box.
    for (int c = 0; c < p->n_chain_d; ++c) {
      const int first = p->part_init_d + c * p->n_mon_d;
      int attempt = 0;
      for (;; ++attempt) {
        if (attempt > 100000) return MFA_ERR_ARG;
        double h[3] = {s.uniform() * L[0], s.uniform() * L[1], s.uniform() * L[2]};
        if (walls) h[2] = zlo + s.uniform() * (zhi - zlo);
        if (!place_ok(h, min_dist)) continue;
        if (grow(first, p->n_mon_d, h, 3, 3, false)) break;
      }
    }
  } else {
    return MFA_ERR_ARG;
  }
  for (size_t k = 0; k < pos.size(); ++k) r0[k] = pos[k];

  // Maxwell-Boltzmann velocities at temp; pinned heads at rest; centre-of-mass
  // momentum removed over the mobile particles (x,y only with walls).
  double pm[3] = {0, 0, 0}, mt = 0.0;
  const int n_heads = p->system == 3 ? p->n_chain : 0;
  for (int i = 0; i < n; ++i) {
    const bool head = i < p->n_mon * p->n_chain && (i % p->n_mon) == 0 && n_heads > 0;
    const double m = p->mass_type[a_type[i] - 1];
    const double sd = sqrt(p->temp / m);
    for (int d = 0; d < 3; ++d) v[3 * (size_t)i + d] = head ? 0.0 : sd * s.normal();
    if (!head) {
      for (int d = 0; d < 3; ++d) pm[d] += m * v[3 * (size_t)i + d];
      mt += m;
    }
  }
  for (int i = 0; i < n; ++i) {
    const bool head = i < p->n_mon * p->n_chain && (i % p->n_mon) == 0 && n_heads > 0;
    if (head) continue;
    for (int d = 0; d < (walls ? 2 : 3); ++d) v[3 * (size_t)i + d] -= pm[d] / mt;
  }
  return MFA_OK;
}
