/*
 * mfa_hip.h -- C ABI of libmfa_hip, the MI355X implementation of MFA_PROG's
 * per-timestep force-and-integrate loop (md_main.f90:68-184).
 *
 * The reference passes every array through `module commons` and its hot-path
 * subroutines take no arguments (or one mode integer).  Here the same state
 * lives in an opaque context that owns device memory; each reference
 * subroutine becomes one entry point below, called in the same order by the
 * Fortran driver (mfa_amd/fortran/md_main_hip.f90) through iso_c_binding.
 *
 * Conventions
 *   - every entry point returns int32 status: 0 = OK, otherwise an MFA_ERR_*
 *     code; mfa_last_error(ctx) gives the message (the reference prints and
 *     `stop`s, e.g. init_system.f90:141-143; the Fortran driver does that).
 *   - arrays cross the boundary in the reference's layout: Fortran (3,N)
 *     column-major == C double[N][3]; a_type and mic are int32; types are
 *     1-based (1..4) like a_type (mfa_common.f90:343).
 *   - a context is not thread-safe; calls are ordered on the context's HIP
 *     stream.  The per-subroutine entry points synchronise before returning,
 *     as the reference's subroutines are synchronous.
 *   - energy out-pointers may be NULL: the energy reduction is then skipped.
 */
#ifndef MFA_HIP_H
#define MFA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFA_NTYPE 4 /* n_type with WALL 2 (mfa_common.f90:47-64) */

enum {
  MFA_OK = 0,
  MFA_ERR_ARG = 1,          /* bad argument / unsupported option */
  MFA_ERR_HIP = 2,          /* HIP runtime error */
  MFA_ERR_LIST_OVERFLOW = 3,/* neighbour list capacity exceeded */
  MFA_ERR_CELL = 4,         /* particle outside the cell grid */
  MFA_ERR_NOMEM = 5,
  MFA_ERR_NAN = 6,          /* non-finite force or position */
  MFA_ERR_IO = 7,
  MFA_ERR_CELL_OVERFLOW = 8,/* more particles in a cell than its capacity */
  MFA_ERR_INTERNAL = 9      /* a device-side consistency check failed (a bug) */
};

/* Precision forms of the force pass. */
enum {
  MFA_FP64 = 0,  /* plain FP64 pair/bond math, FP64 accumulation          */
  MFA_MIXED = 1  /* FP32 pair/bond math, int64 fixed-point accumulation;
                    FP64 positions, velocities, integration and reductions */
};

/*
 * Raw run-time inputs: the numbers system_input, mfa_input and the
 * compile-time switches of control_simulation.h provide to the reference.
 * Filled by the Fortran readers (init_system.f90:12-144,
 * init_params.f90:137-244) or by the Python mirror.
 * Type tables are indexed [i][j] with 0-based types.
 */
typedef struct mfa_inputs {
  /* system_input */
  int32_t n_mon, n_chain, n_cell_w_x, n_cell_w_y;  /* init_system.f90:26 */
  int32_t n_mon_d, n_chain_d;                      /* :32 */
  int32_t n_mon_e, n_chain_e, inter4;              /* :41-47 (PARTICLE_4) */
  int32_t pad0;
  double a_w, sigma_w, a_w4, sigma_w4;             /* :65-81 */
  double charge[MFA_NTYPE];                        /* :83-91 */
  double alpha_c;                                  /* :94 */
  int32_t n_kmax, pad1;                            /* :94 */
  /* mfa_input */
  int32_t s_time, n_relax, n_obser, n_safe;        /* init_params.f90:138-141 */
  double dt, temp;                                 /* :142,145 */
  double x_space, y_space, z_space_wall;           /* :148-150 */
  double mass_type[MFA_NTYPE];                     /* :155-162 */
  double epsil[MFA_NTYPE][MFA_NTYPE];              /* as read (before x4) */
  double sigma_ii[MFA_NTYPE];
  double friction[MFA_NTYPE];
  double k_bend, alpha_eq;                         /* :214-215 (BENDING) */
  int64_t iseed;                                   /* :223 */
  double k_chain, r_chain;                         /* :225-226 */
  double va_spring_twall[3];                       /* :232 */
  /* control_simulation.h switches, made run-time */
  int32_t system;       /* SYSTEM: 2 (no brush) or 3 (brush + walls)       */
  int32_t symmetry;     /* SYMMETRY: 0 channel (PBC x,y), 1 bulk (PBC xyz) */
  int32_t solvent;      /* SOLVENT: 0 poor, 1 good                          */
  int32_t thermostat;   /* THERMOSTAT: 0 DPD, -1 none (NVE)                */
  int32_t particle_4;   /* PARTICLE_4 defined                               */
  int32_t bending;      /* BENDING defined                                  */
  double dpd_cut_off;   /* DPD_CUT_OFF value, 0 = undefined                 */
  /* options the reference hard-codes (documented patches of the survey) */
  double skin;          /* init_params.f90:800 (0 -> reference 0.4)          */
  double lj_rc;         /* >0: every LJ range = lj_rc*sigma (P3); 0 = menu   */
  double rc_coul;       /* >0: real-space Coulomb cutoff; 0 = LJ range (P5)  */
  int32_t ewald_compat; /* 1: reference's 1/pi real-space factor (bug)      */
  int32_t force_switch_on; /* FORCE_SWITCH_ON defined: ramp when s_time == 1 */
  int32_t dpd_vv;       /* DPD_VV defined: new_dpd_fd after verlet_velocities */
  int32_t pad2;
} mfa_inputs;

/*
 * Derived constants (init_params.f90:395-845, init_config.f90:11-22,229-239,
 * make_binning_boxes.f90:14-64), produced by mfa_derive().  Default-real
 * literals of the reference are applied as float32 values, as its Makefile
 * (no -r8) compiles them.
 */
typedef struct mfa_params {
  int32_t n_part, n_mon_tot;
  int32_t n_mon, n_chain, n_mon_d, n_chain_d, n_mon_e, n_chain_e;
  int32_t part_init_d, part_init_e;     /* first melt / particle-4 index (0-based) */
  int32_t system, symmetry, thermostat, dpd_weight;
  int32_t dpd_cut_off_on, coulomb, bending, ewald_compat;
  double boundary[3], inv_boundary[3];
  double dt, dt_2, temp;
  double mass_type[MFA_NTYPE], inv_mass_type[MFA_NTYPE];
  double epsil[MFA_NTYPE][MFA_NTYPE];   /* 4*eps */
  double sigma[MFA_NTYPE][MFA_NTYPE];
  double sigma_2[MFA_NTYPE][MFA_NTYPE];
  double range_2[MFA_NTYPE][MFA_NTYPE];
  double inv_range_2[MFA_NTYPE][MFA_NTYPE];
  double e_shift[MFA_NTYPE][MFA_NTYPE];
  double list_2[MFA_NTYPE][MFA_NTYPE];  /* Verlet-list criterion (binning.f90:101-105) */
  double r_cut_max, r_cut_max_2;
  double r_cut_dpd, r_cut_dpd_2;
  double skin, skin_2, skin_half_2;     /* skin_half_2 = (skin/2)^2, check_skin.f90:18 */
  double friction1, sig;                /* sig = sqrt(2 T gamma / dt), init_params.f90:845 */
  double k_chain, r_chain_2, inv_r_chain_2;
  double a_w, sigma_wall[MFA_NTYPE], a_wall[MFA_NTYPE], z_space_wall;
  double va_spring_twall[3];
  double k_bend, alpha_eq;
  int32_t n_bin[3], delta_bin[3];
  double r_bin[3], inv_r_bin[3];
  double charge[MFA_NTYPE];
  double alpha_c, sqrt_alpha, inv_sqrt_pi_real; /* real-space erfc prefactor */
  double rc_coul_2;                     /* real-space Coulomb cutoff^2 */
  double v_coul_self;                   /* filled by mfa_create from charges */
  int32_t n_kmax, pad0;
  uint64_t seed;
  /* FORCE_SWITCH_ON ramp (init_force_switch_on.f90:15-26): r_2 >= r_2_min with
   * r_2_min = (1 - t/fso_time) * fso_r2_init while t = step*dt < fso_time. */
  double fso_r2_init, fso_time;
  int32_t dpd_vv, pad1; /* DPD_VV (md_main.f90:149-155) */
} mfa_params;

/* Observation sums returned by mfa_run_steps / mfa_energies. */
typedef struct mfa_obs {
  double v_fluid_fluid;   /* LJ energy (fluid_fluid.f90:110) */
  double v_fluid_wall;    /* 9-3 wall energy (fluid_wall.f90:98,116) */
  double v_intra_molec;   /* FENE energy incl. melt (intra_molec.f90:63,537) */
  double v_bend;          /* bending energy (bending.f90) */
  double v_coul;          /* Ewald real + k + self + dipolar */
  double t_kin;           /* 0.5 sum m v^2 (verlet_velocities.f90:155-177) */
  double virial;          /* sum of r_ij . F_ij over LJ pairs and FENE bonds only
                             (no DPD, Coulomb, wall or bending terms) */
  double wall_fz_bottom;  /* sum of wall z-forces from the z=0 wall */
  double wall_fz_top;     /* sum of wall z-forces from the z=Lz wall */
  double z_moment;        /* sum of z over brush monomers (type 2) */
  int64_t n_rebuilds;     /* list rebuilds during the call */
} mfa_obs;

/* Running sums of the observation stage (observation.f90:17-307 as built
 * with PROFILES, SYSTEM 2/3, WALL 2: no wall atoms, so the wall-wall sums
 * of the reference stay zero and are left out). */
typedef struct mfa_obs_sums {
  int64_t count_obs;                        /* observation.f90:17 */
  double v_total_1, v_total_2;              /* :146-152 */
  double t_total_1, t_total_2;
  double e_total_1, e_total_2;
  double v_fluid_fluid_1, v_fluid_fluid_2;  /* :138-144, v_ff + v_intra + v_fw */
  double t_fluid_1, t_fluid_2;
  double e_fluid_fluid_1, e_fluid_fluid_2;
  double v_tot_intra_1, v_tot_intra_2;      /* :158-159 */
  double tot_p[3];                          /* :286-291, sum of m v per observation */
  double r_g2_mean[3][3];                   /* :271-273, [brush, melt, (brush+melt)/2][x,y,z] */
  double r_par_per2[2];                     /* :275-276 */
  int64_t count_hist;                       /* functions.f90:144, histo mode 2 calls */
  int32_t hist_dim, pad;                    /* functions.f90:115, dim_prof_dens */
} mfa_obs_sums;

typedef struct mfa_ctx mfa_ctx;

/* ABI self-check: sizes of the structs as compiled into the library. */
int32_t mfa_abi_sizes(int32_t* sizeof_inputs, int32_t* sizeof_params, int32_t* sizeof_obs);

/* Derived constants.  Replaces the derivation part of init_params
 * (init_params.f90:395-845), the box of init_config (init_config.f90:11-22)
 * and make_binning_boxes (make_binning_boxes.f90:14-64).  Host-only. */
int32_t mfa_derive(const mfa_inputs* in, mfa_params* out);

/* The skin-dependent constants of p (skin, skin_2, skin_half_2, list_2) for
 * another list skin, as mfa_derive computes them (init_params.f90:800-808,
 * binning.f90:101-105); the cells stay (make_binning_boxes.f90:42-47), so
 * MFA_ERR_ARG if the list radius would exceed a cell's width.  Host-only. */
int32_t mfa_params_set_skin(mfa_params* p, double skin);

/* Another cell grid for derived params (no counterpart in the reference,
 * whose make_binning_boxes.f90:42-47 takes the finest grid with cells at
 * least r_cut_max + 0.4 wide): n_bin = int(L / width) per axis (n_bin,
 * r_bin, inv_r_bin); MFA_ERR_ARG if an axis would have fewer than 3 cells
 * or a cell would be narrower than the list radius.  Host-only. */
int32_t mfa_params_set_cells(mfa_params* p, double width);

/* Create/destroy a device context for n_part particles.  Replaces the
 * allocations of init_system.f90:228-299 and make_binning_boxes.f90:70-78.
 * precision: MFA_FP64 or MFA_MIXED.  list_capacity: full neighbour-list slots
 * per particle (0 = sized from the density). */
int32_t mfa_create(const mfa_params* p, int32_t precision, int32_t list_capacity, mfa_ctx** out);
int32_t mfa_destroy(mfa_ctx* ctx);
int32_t mfa_set_precision(mfa_ctx* ctx, int32_t precision);
/* Lanes that share one particle's neighbour list in the force pass
 * (1, 2, 4, 8 or 16; default 8).  A tuning knob with no reference
 * counterpart; results are identical for every value in the mixed form. */
int32_t mfa_set_threads_per_particle(mfa_ctx* ctx, int32_t tpp);
/* Steps per captured hipGraph block in mfa_run_steps (default 16, 0 = launch
 * every kernel directly).  No reference counterpart; results are identical. */
int32_t mfa_set_graph_steps(mfa_ctx* ctx, int32_t steps);
const char* mfa_last_error(mfa_ctx* ctx);

/* State upload/download, (3,N) layout.  Replaces the conf_old read
 * (init_config.f90:110-209) and what store_config reads (store_config.f90:19-61).
 * mic may be NULL (zeros).  After set_state the next mfa_binning/run call
 * rebuilds the list, as init_config does (init_config.f90:251-255). */
int32_t mfa_set_state(mfa_ctx* ctx, const double* r0, const double* v, const double* a,
                      const int32_t* a_type, const int32_t* mic);
int32_t mfa_get_state(mfa_ctx* ctx, double* r0, double* v, double* a, double* force, int32_t* mic);

/* ewald_k(mode) (ewald_k.f90:22-202; called at md_main.f90:129, mode 0 at
 * init_config.f90:244): mode 0 is a no-op here (mfa_create builds the k
 * grid); mode 1 adds the k-space forces of the charged particles to F and
 * returns u_cou_k in *v_coul.  No-op without charges. */
int32_t mfa_ewald_k(mfa_ctx* ctx, int32_t mode, double* v_coul);
/* The launch plan of ewald_k's k-space stage for a box and a number of
 * charged particles, with every bound the kernels rely on checked (LDS per
 * workgroup within 64 KB, rows and segments inside the k grid and the k index
 * range, force partials inside their buffer, the XCD-aware row grid).  The
 * k grid is the half-space enumeration of ewald_k.f90:104-124.  Host-only;
 * mfa_create and mfa_set_state run the same checks (plus the device's LDS
 * limit and a register-spill check of the row kernel) and refuse a system
 * that fails them.  out[MFA_EWPLAN_N]: k vectors, segments, rows, nx_max,
 * ny_max, nz_max, row kernel (0/1), rows per group, force groups, force
 * blocks, force grid, LDS bytes of the structure factor, LDS bytes of the
 * force kernel, force partials used, force partials available,
 * structure-factor chunks.  Returns MFA_ERR_ARG (message in
 * mfa_last_error(NULL)) when a bound fails; out is filled either way. */
#define MFA_EWPLAN_N 16
int32_t mfa_ewald_plan(const mfa_params* p, int32_t n_charged, int64_t* out, int32_t n_out);
/* dipolar_correction() (dipolar_correction.f90:20-33, md_main.f90:131): adds
 * -4 pi q M_z / V to F_z of the charged particles (walls only) and returns
 * 2 pi M_z^2 / V in *v_coul. */
int32_t mfa_dipolar_correction(mfa_ctx* ctx, double* v_coul);

/* verlet_positions (verlet_positions.f90:144-250). */
int32_t mfa_verlet_positions(mfa_ctx* ctx);
/* check_skin (check_skin.f90:4-24): *f_skin = 1 if a rebuild is due. */
int32_t mfa_check_skin(mfa_ctx* ctx, int32_t* f_skin);
/* binning (binning.f90:11-127): cell list + Verlet list, resets the skin
 * reference (binning.f90:125-127). */
int32_t mfa_binning(mfa_ctx* ctx);
/* Parity exports of the list: cell index per particle (x fastest,
 * binning.f90:27-29), and the unordered pair set as (i<j), 1-based. */
int32_t mfa_get_cells(mfa_ctx* ctx, int32_t* cell_xyz /* [N][3] */);
int32_t mfa_get_pairs(mfa_ctx* ctx, int32_t* pairs /* [cap][2] */, int64_t cap, int64_t* n_pairs);
/* force(:,:) = 0 (md_main.f90:88). */
int32_t mfa_zero_force(mfa_ctx* ctx);
/* fluid_fluid (fluid_fluid.f90:52-240) with dpd_forces (dpd_forces.f90:2-98)
 * and ewald_real (ewald_real.f90:15-32).  The DPD noise of pair (i,j) at
 * step `step` is a counter-based normal keyed by (seed, step, min, max). */
int32_t mfa_fluid_fluid(mfa_ctx* ctx, int64_t step, double* v_ff, double* v_coul);
/* fluid_wall(2) (fluid_wall.f90:81-123).  Only inter_type 2 is provided. */
int32_t mfa_fluid_wall(mfa_ctx* ctx, int32_t inter_type, double* v_fw);
/* intra_molec (intra_molec.f90:16-139): FENE, one thread per bond. */
int32_t mfa_intra_molec(mfa_ctx* ctx, double* v_intra);
/* bending(mode_bend) (bending.f90:2-127, md_main.f90:108-111): mode 2 adds
 * the harmonic angle forces of the brush chains, k_bend/2 (alpha -
 * alpha_eq)^2, and returns v_bend; modes 0 and 1 do nothing, as in the
 * reference.  No-op unless the inputs enable bending. */
int32_t mfa_bending(mfa_ctx* ctx, int32_t mode, double* v_bend);
/* verlet_velocities (verlet_velocities.f90:18-199); *t_kin = 0.5 sum m v^2. */
int32_t mfa_verlet_velocities(mfa_ctx* ctx, double* t_kin);
/* verlet_velocities on arrays the caller owns, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same arithmetic as
 * mfa_verlet_velocities (a = F/m, v += dt/2 a, pinned heads) on device
 * arrays force, v, a (double[N][3]) and a_type (int32[N], 1-based), N =
 * p->n_part <= MFA_NOALLOC_MAX_N, on `stream` (a hipStream_t; NULL: the null
 * stream).  t_kin (host pointer, may be NULL) receives 0.5 sum m v^2 and then
 * the call synchronises the stream; without it the call is asynchronous.
 * Scratch: one file-scope device array per GPU (about 400 KB), shared by
 * every call on that GPU: calls that overlap on one GPU race. */
#define MFA_NOALLOC_MAX_N (1 << 20)
int32_t mfa_verlet_velocities_noalloc(const mfa_params* p, double* force, double* v, double* a,
                                      const int32_t* a_type, void* stream, double* t_kin);
/* md_main.f90:149-155 with DPD_VV: new_dpd_fd.f90:18-124, the dissipative
 * force recomputed with the new velocities, swapped into F (F - F_D,old +
 * F_D,new), a further half kick and a = F/m; t_kin as new_dpd_fd computes it. */
int32_t mfa_new_dpd_fd(mfa_ctx* ctx, double* t_kin);

/* Whole loop body (md_main.f90:69-155) for n steps, steps first_step ..
 * first_step+n-1, device resident: fused verlet_velocities + next
 * verlet_positions, skin test as a wavefront max-reduction, counting-sort
 * rebuild when due, one fused force pass.  obs (may be NULL) receives the
 * energies of the last step. */
int32_t mfa_run_steps(mfa_ctx* ctx, int64_t first_step, int32_t n, mfa_obs* obs);

/* Energies and observation sums of the current state (forces recomputed for
 * step `step`, the state is not advanced). */
int32_t mfa_energies(mfa_ctx* ctx, int64_t step, mfa_obs* obs);

/* List statistics: out[0] = largest neighbour count seen since create,
 * out[1] = list capacity, out[2] = device error code, out[3] = rebuilds in the
 * current run_steps call. */
int32_t mfa_stats(mfa_ctx* ctx, int64_t* out /* [4] */);

/* Average duration (ms) of the force pass at the current state: `reps`
 * back-to-back launches captured in one graph, bracketed by two HIP events on
 * the context's stream.  F is recomputed (same values). */
int32_t mfa_time_force(mfa_ctx* ctx, int32_t reps, double* ms_avg);

/* The force pass as mfa_run_steps runs it, timed like mfa_time_force.  In
 * the mixed form without the k-space stage or DPD_VV (and with kernel timing
 * off), the step's force kernel also integrates (verlet_velocities of this
 * step, verlet_positions and check_skin of the next); it is timed with what
 * it writes to the velocities, image counts, flags and cell counts sent to
 * copies, so the state is left as it was.  Otherwise mfa_time_force. */
int32_t mfa_time_force_step(mfa_ctx* ctx, int32_t reps, double* ms_avg);

/* Average duration (ms) of the k-space Ewald stage of a step (ewald_k(1) and,
 * with walls, the slab dipole correction: ewald_k.f90:150-199,
 * dipolar_correction.f90:20-33) at the current state, timed like
 * mfa_time_force; F is left as it was.  MFA_ERR_ARG without charges. */
int32_t mfa_time_ewald(mfa_ctx* ctx, int32_t reps, double* ms_avg);

/* Average durations (ms) of the five kernels of that stage, each timed alone
 * like mfa_time_force (back-to-back launches of that kernel between two HIP
 * events): ms[0] the phase tables, ms[1] the structure factor, ms[2] its sum,
 * ms[3] the forces, ms[4] adding them to F.  F is left as it was.
 * Measurement helper with no reference counterpart. */
int32_t mfa_time_ewald_parts(mfa_ctx* ctx, int32_t reps, double* ms /* [5] */);

/* Average durations (ms) of the kernels of a rebuild: ms[0] the cell fill,
 * ms[1] the cell sort, ms[2] the list build, each bracketed by HIP events on
 * the context's stream, over `reps` rebuilds forced at the current state.
 * Measurement helper with no reference counterpart. */
int32_t mfa_time_rebuild(mfa_ctx* ctx, int32_t reps, double* ms /* [3] */);

/* The same three kernels with no replica due for a rebuild: what a step
 * without a rebuild pays for them (each returns at once).  Measurement
 * helper with no reference counterpart. */
int32_t mfa_time_rebuild_idle(mfa_ctx* ctx, int32_t reps, double* ms /* [3] */);

/* Measurement only, no reference counterpart and no part of the step: a
 * device-to-device copy of `bytes` (read `bytes`, write `bytes`) timed with
 * HIP events, the calibration point for the HBM byte counters
 * (FETCH_SIZE/WRITE_SIZE) and the bandwidth the bench reports.  It allocates
 * and frees its own two buffers. */
int32_t mfa_copy_probe(int64_t bytes, double* ms);

/* Kernel timing of the force pass: average duration (ms) over the launches
 * since the last reset, measured with HIP events on the context's stream. */
int32_t mfa_timing(mfa_ctx* ctx, int32_t enable, double* force_ms_avg, int64_t* force_launches,
                   double* integrate_ms_avg, int64_t* integrate_launches);

/* ------------------------------------------------------------------------
 * Batches of independent replicas (no counterpart in the reference, which
 * runs one system per process).  Replica q owns its own box, cell grid,
 * chains and seed; all replicas share the compile-time switches (system,
 * symmetry, thermostat, DPD cut-off) and carry no charges.  One
 * launch of each kernel advances every replica, so small replicas share the
 * GPU instead of queueing behind each other.  mfa_run_steps then advances all
 * replicas in lockstep; its obs argument receives replica 0's observations
 * and mfa_replica_obs the others'.  A single system is a batch of one. */
int32_t mfa_create_batch(const mfa_params* params, int32_t n_rep, int32_t precision,
                         int32_t list_capacity, mfa_ctx** ctx);
int32_t mfa_set_state_replica(mfa_ctx* ctx, int32_t replica, const double* r0, const double* v,
                              const double* a, const int32_t* a_type, const int32_t* mic);
int32_t mfa_get_state_replica(mfa_ctx* ctx, int32_t replica, double* r0, double* v, double* a,
                              double* force, int32_t* mic);
int32_t mfa_replica_obs(mfa_ctx* ctx, int32_t replica, mfa_obs* obs);
/* Observation stage on the device.  Replaces observation.f90:17-307 (with
 * PROFILES: the density histogram of functions.f90:94-215) and the head-force
 * capture of verlet_velocities.f90:14-19, without copying the state to the
 * host: the sums stay in device memory until mfa_obs_read.
 *   mfa_obs_init: zeroes the sums and sets up a histogram of hist_dim slabs
 *     (init_obser.f90:15-91, histo mode 0; the reference uses 200).
 *   mfa_observe: one observation of the current state, after a step
 *     (md_main.f90:159-161).  e holds that step's energies (from
 *     mfa_run_steps / mfa_verlet_velocities); coul_in_total = 1 on the steps
 *     where the reference's verlet_velocities has just added v_coul to
 *     v_total (every tenth call, verlet_velocities.f90:199-222, SYSTEM 2/3).
 *   mfa_obs_read: the sums; dens_z (may be NULL) receives the raw slab counts
 *     [3][hist_dim] (brush, melt, particle 4; histo mode 2), f_on_heads (may
 *     be NULL) the force on each brush head [n_chain][3] as the last
 *     verlet_velocities found it.
 * Single-system contexts only. */
int32_t mfa_obs_init(mfa_ctx* ctx, int32_t hist_dim);
int32_t mfa_observe(mfa_ctx* ctx, const mfa_obs* e, int32_t coul_in_total);
int32_t mfa_obs_read(mfa_ctx* ctx, mfa_obs_sums* out, double* dens_z, double* f_on_heads);

/* Pair queue of the mixed force pass (no counterpart in the reference;
 * systems without charges, bending or the Langevin force): each wave pushes
 * the list candidates that interact into a queue in LDS and evaluates them
 * 64 at a time, so no lane idles where few candidates interact.  mode 0:
 * no queue; 1: every entry through the queue; 2: near/shell rows (the list
 * build puts the entries inside the interaction range at the build in
 * front, those in the skin behind them), the near part walked directly and
 * the skin part, where few entries interact, through the queue.  The
 * forces are bit-identical in every mode (exact fixed-point sums).  0 by
 * default; the autotuner chooses. */
int32_t mfa_set_pair_queue(mfa_ctx* ctx, int32_t mode);

/* Half-list form of the mixed force pass (no counterpart in the reference,
 * whose binning.f90:81 list is a half list applied with +/- in
 * fluid_fluid.f90:120-126): on, each pair is evaluated once, from the row
 * of its lower sorted position, in a first pass that keeps the pair's scaled
 * force scalar per list slot; a second pass gives the other particle its
 * term from that scalar and its own (negated) separation, through the slot
 * position found once per rebuild, so no atomics are needed and the forces
 * are bit-identical to the full-list pass.  Applies at 4 or more threads per
 * particle; allocates 6 bytes per list slot and 24 per particle (refused,
 * MFA_ERR_ARG, beyond 2^31 slots).  Off by default; the autotuner chooses. */
int32_t mfa_set_half_list(mfa_ctx* ctx, int32_t on);

/* The context's current choices: threads per particle, pair queue, list skin
 * (replica 0), half-list form.  Any out-pointer may be NULL. */
int32_t mfa_get_tuning(mfa_ctx* ctx, int32_t* tpp, int32_t* pair_queue, double* skin,
                       int32_t* half_list);

/* Autotuner: for each number of threads per particle in {1, 2, 4, 8, 16},
 * and in the mixed form each way of walking the list (directly; the skin
 * part of near/shell rows through the pair queue and, for DPD_CUT_OFF, all
 * of it, for systems without charges, bending or the Langevin force; the
 * half-list form at 4 or more threads, also with the queue for
 * DPD_CUT_OFF), runs graph_steps + 1 untimed and then n_steps timed whole
 * steps from first_step (advancing the state, so call it during warm-up),
 * rebuilds included, in two sweeps over the candidates (each keeps its
 * faster trial); keeps the fastest and returns it with its time per step
 * (ms). */
int32_t mfa_autotune(mfa_ctx* ctx, int64_t first_step, int32_t n_steps, int32_t* best_tpp,
                     double* ms_per_step);

/* Another list skin for every replica of a running context (no counterpart
 * in the reference, whose skin is fixed at 0.4): the constants as
 * mfa_params_set_skin gives them, a longer list buffer if the list grows,
 * and a fresh list (with its skin reference) before the next force pass.
 * Pairs within the cut-offs, and so the forces, do not depend on the skin;
 * how often the list is rebuilt and how long it is do. */
int32_t mfa_set_skin(mfa_ctx* ctx, double skin);

/* Skin autotuner: for each candidate skin that fits the cells, runs
 * n_steps whole steps (rebuilds included) from first_step, advancing the
 * state, timed by HIP events; keeps the fastest and returns it with its time
 * per step (ms).  Call it during warm-up. */
int32_t mfa_autotune_skin(mfa_ctx* ctx, int64_t first_step, int32_t n_steps, int32_t n_cand,
                          const double* skins, double* best_skin, double* ms_per_step);

/* Another cell grid for a running single-system context (no counterpart in
 * the reference; replica batches are refused, MFA_ERR_ARG): the grid
 * mfa_params_set_cells gives for `width`, the cell arrays sized for it, and
 * fresh cells and list before the next force pass.  The pair set within the
 * list radius does not depend on the grid; the list build's work (27 cells
 * of candidates per particle), the largest list skin the cells hold and, in
 * the mixed form, the rounding of the FP32 view (positions relative to the
 * cell origin) do. */
int32_t mfa_set_cells(mfa_ctx* ctx, double width);

/* Cell-size and skin autotuner (single-system contexts, as mfa_set_cells):
 * for each candidate cell width (at most 16) that gives at least 3 cells
 * per axis, and each candidate
 * skin (at most 64) that those cells hold, runs n_steps whole steps
 * (rebuilds included) from first_step, advancing the state, timed by HIP
 * events; keeps the fastest pair and returns it with its time per step (ms)
 * and the step after the last one run.  Call it during warm-up. */
int32_t mfa_autotune_cells(mfa_ctx* ctx, int64_t first_step, int32_t n_steps, int32_t n_width,
                           const double* widths, int32_t n_skin, const double* skins,
                           double* best_width, double* best_skin, double* ms_per_step,
                           int64_t* next_step);

/*
 * Adjoint of the force-and-integrate step (no counterpart in the reference,
 * which has no derivatives).  Over a window of n_steps steps from the
 * context's state (as mfa_run_steps(first_step, n_steps) would advance it)
 * the call computes two time-averaged observables of a brush in a channel,
 *   height  = (1/W) sum_n  mean z of the type-2 beads at r_n,
 *   p_wall  = (1/W) sum_n  -(sum_i F_z^top(z_i)) / (Lx Ly), the normal
 *             pressure the particles exert on the z = Lz wall,
 * and, with want_grad, their derivatives with respect to the run-time
 * parameters theta = (epsilon_LJ, sigma, k_FENE, R0, gamma_DPD) of
 * the mfa_input file (init_params.f90:395-845): epsilon is the common value of every
 * non-zero epsil(i,j), sigma the common sigma of every type (it moves the
 * LJ ranges, the mixed sigma(i,j) and the wall sigma_wall(t)), k_FENE and R0
 * the FENE constants, gamma_DPD = friction(1).  The DPD noise is the
 * counter-based deviate of (seed, step, pair), fixed under theta.
 * Reverse mode: the state is checkpointed every checkpoint_every steps, each
 * segment is recomputed forward and the adjoint of the step (pair, bond and
 * wall Jacobians applied to the adjoint vectors of both observables at once,
 * FP64 sums) runs backward through it.  The forward pass computes in the
 * context's precision form; the adjoint is FP64.  On return the context holds
 * the state at the end of the window.  Systems with charges, bending or an
 * active switch-on ramp are refused; a batch must hold one replica.
 */
typedef struct mfa_adjoint_out {
  double height, p_wall;
  double d_height[5];   /* d height / d(epsilon_LJ, sigma, k_FENE, R0, gamma_DPD) */
  double d_p_wall[5];
  int64_t n_rebuilds;   /* list rebuilds of the forward pass */
} mfa_adjoint_out;
int32_t mfa_adjoint_window(mfa_ctx* ctx, int64_t first_step, int32_t n_steps,
                           int32_t checkpoint_every, int32_t want_grad, mfa_adjoint_out* out);

/* Seeded input generators (the reference's conf_default/gen_brush/gen_droplet
 * are replaced by these; conf_default.f90:144-217, gen_brush.f90:101-161).
 * kind 1: brush + ideal solvent in a channel; kind 2: melt in a periodic box;
 * kind 3: polyelectrolyte brush (every third bead type 4), one counterion
 * (type 3) per charged bead and neutral solvent (type 2) in a channel.
 * Writes r0, v (Maxwell-Boltzmann at temp, COM momentum removed), a_type.
 * Arrays are sized n_part = p->n_part.  Host-only. */
int32_t mfa_generate(const mfa_params* p, int32_t kind, uint64_t seed, double head_z,
                     double bond, double min_dist, double* r0, double* v, int32_t* a_type);

#ifdef __cplusplus
}
#endif
#endif /* MFA_HIP_H */
