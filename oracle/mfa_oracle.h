/*
 * mfa_oracle.h -- TEST INFRASTRUCTURE, not product code.
 *
 * A plain-C, single-threaded FP64 restatement of MFA_PROG's per-timestep loop
 * (md_main.f90:68-155), used only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, as the checker.  It follows the reference's
 * data layout (module commons: r0(3,N), ff_list(0:n_neigh,N), bin_fluid) and
 * operation order; every function cites the reference lines it restates.
 *
 * The only departure from the reference is the DPD noise: the reference draws
 * from one sequential ziggurat stream (dpd_forces.f90:19), which no parallel
 * code can reproduce; the oracle uses the counter-based normal that the GPU
 * path uses (Threefry-2x32-13 keyed by seed, step and the pair, Box-Muller),
 * implemented here independently.
 */
#ifndef MFA_ORACLE_H
#define MFA_ORACLE_H
#include <stdint.h>
#include "../include/mfa_hip.h" /* mfa_inputs: the raw input container only */

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NT 4

typedef struct orc_params {
  int n_part, n_mon, n_chain, n_mon_d, n_chain_d, n_mon_e, n_chain_e;
  int part_init_d, part_init_e;
  int system, symmetry, thermostat, dpd_cut_off_on, particle_4;
  double boundary[3], inv_boundary[3];
  double dt, dt_2, temp;
  double mass_type[ORC_NT];
  double sigma[ORC_NT][ORC_NT], sigma_2[ORC_NT][ORC_NT], epsil[ORC_NT][ORC_NT];
  double range_2[ORC_NT][ORC_NT], inv_range_2[ORC_NT][ORC_NT], e_shift[ORC_NT][ORC_NT];
  double r_cut_max, r_cut_max_2, r_cut_dpd, r_cut_dpd_2;
  double skin, skin_2;
  double friction1, sig;
  double k_chain, r_chain_2, inv_r_chain_2;
  double a_w, sigma_wall[ORC_NT], a_wall[ORC_NT], z_space_wall;
  double va_spring_twall[3];
  int n_bin[3], delta_bin[3];
  double r_bin[3], inv_r_bin[3];
  uint32_t seed;
  double r_2_min_init, r_2_min_time; /* FORCE_SWITCH_ON ramp, 0 when off */
  /* Coulomb (SYSTEM 2/3): init_system.f90:83-98, init_config.f90:221-239 */
  int coulomb, ewald_compat, n_kmax;
  double charge[ORC_NT], alpha_c, sqrt_alpha, inv_sqrt_pi, rc_coul_2;
  /* BENDING (init_params.f90:211-213) */
  int bending;
  double k_bend, alpha_eq;
  /* DPD_VV (md_main.f90:149-155) */
  int dpd_vv;
} orc_params;

typedef struct orc_sys orc_sys;

/* counter-based RNG (independent implementation, pinned by known-answer tests) */
/* rounds of the pair and particle deviates (the HIP path's MFA_TF_ROUNDS) */
#define ORC_TF_ROUNDS 13
void orc_threefry2x32(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1, int rounds,
                      uint32_t* o0, uint32_t* o1);
double orc_pair_normal(uint32_t seed, int64_t step, int i0, int j0);
double orc_particle_normal(uint32_t seed, int64_t step, int i0, int d);

int orc_derive(const mfa_inputs* in, orc_params* p);
orc_sys* orc_create(const orc_params* p, int n_neigh);
void orc_destroy(orc_sys* s);
void orc_set_state(orc_sys* s, const double* r0, const double* v, const double* a,
                   const int32_t* a_type, const int32_t* mic);
void orc_get_state(orc_sys* s, double* r0, double* v, double* a, double* force, int32_t* mic);

void orc_verlet_positions(orc_sys* s);
int orc_check_skin(orc_sys* s);
int orc_binning(orc_sys* s); /* 0 or MFA_ERR_LIST_OVERFLOW / MFA_ERR_CELL */
void orc_get_cells(orc_sys* s, int32_t* cells);
int64_t orc_get_pairs(orc_sys* s, int32_t* pairs, int64_t cap);
void orc_zero_force(orc_sys* s);
/* energies: out[0]=v_fluid_fluid, out[1]=virial contribution,
 * out[2]=v_coul (real space + self term) */
void orc_fluid_fluid(orc_sys* s, int64_t step, double* out);
/* ewald_k(1): adds the k-space forces; returns u_cou_k.  orc_ewald_nk gives
 * the number of k vectors and (optionally) copies them (nk x 3). */
double orc_ewald_k(orc_sys* s);
int orc_ewald_nk(orc_sys* s, double* kvecs);
/* dipolar_correction: adds the slab forces; returns its energy. */
double orc_dipolar(orc_sys* s);
/* bending(2): adds the angle forces of the brush chains; returns v_bend. */
double orc_bending(orc_sys* s);
/* out[0]=v_fluid_wall, out[1]=sum Fz bottom, out[2]=sum Fz top */
void orc_fluid_wall(orc_sys* s, double* out);
/* out[0]=v_intra_molec, out[1]=virial contribution */
void orc_intra_molec(orc_sys* s, double* out);
double orc_verlet_velocities(orc_sys* s);
double orc_new_dpd_fd(orc_sys* s);
/* Loop body of md_main.f90:69-155 for n steps.  obs[0..8]: v_ff, v_fw,
 * v_intra, t_kin, virial, n_rebuilds, wall Fz bottom, v_coul, v_bend of the
 * last step. */
int orc_run_steps(orc_sys* s, int64_t first_step, int n, double* obs);
/* Observation stage (observation.f90:17-307 with PROFILES): e = (v_ff,
 * v_fw, v_intra, t_fluid, v_coul) of the step just made. */
void orc_obs_init(orc_sys* s, int hist_dim);
void orc_observe(orc_sys* s, const double* e, int coul_in_total);
void orc_obs_read(orc_sys* s, mfa_obs_sums* out, double* dens_z, double* f_on_heads);
/* Brute-force O(N^2) list criterion (for pinning the cell scan). */
int64_t orc_brute_pairs(orc_sys* s, int32_t* pairs, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
